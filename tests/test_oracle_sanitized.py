"""The oracle's own checks rerun on its AddressSanitizer + UndefinedBehavior-
Sanitizer build (oracle/Makefile target `san`): every heap access of the
restatement stays in bounds and no undefined behaviour (signed overflow,
bad shifts, misaligned loads) is reached on the pinned cases, the golden
fixtures, the loss-code paths and the beam file round trip."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _runtime():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                           text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_checks_pass_under_asan_and_ubsan():
    asan = _runtime()
    if asan is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, HPSIM_ORACLE_SANITIZED="1",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    files = ["test_oracle_pins.py", "test_oracle_pins_sc.py", "test_golden_fixtures.py",
             "test_loss_codes.py", "test_beam_io.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-k", "not openmp",
                        "-p", "no:cacheprovider"] + [os.path.join(HERE, f) for f in files],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "runtime error" not in r.stdout + r.stderr, tail
