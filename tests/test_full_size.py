"""Parity at the headline size.

The OpenMP build of the oracle (the CPU baseline of bench.py) is first
pinned to the serial oracle bit for bit on a beam large enough to spread its
loops over every thread; then the GPU engine is compared with it on the
whole 1,048,576-particle, 800-element workload with SCHEFF 32 x 128
(plain-fp64-1048576x200), which the OpenMP oracle finishes in seconds.
The OpenMP oracle runs in a process of its own (tests/oracle_omp_run.py):
the serial and OpenMP libraries cannot share one."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tests import helpers as H

HERE = os.path.dirname(os.path.abspath(__file__))


def _omp_oracle(tmp_path, idx, num, periods, sig_phi_deg=None, timeout=600):
    out = str(tmp_path / f"omp_{idx}_{num}_{periods}.npz")
    cmd = [sys.executable, os.path.join(HERE, "oracle_omp_run.py"), str(idx), str(num),
           str(periods), out]
    if sig_phi_deg is not None:
        cmd.append(str(sig_phi_deg))
    # at least four threads, so that the parallel loops are what is checked
    threads = max(4, min(16, os.cpu_count() or 4))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run(cmd, check=True, timeout=timeout, env=env)
    res = dict(np.load(out))
    assert int(res["threads"]) == threads
    return res


@pytest.mark.parametrize("sig_phi_deg", [None, 70.0])
def test_openmp_oracle_is_the_serial_oracle(tmp_path, sig_phi_deg):
    """40,000 particles x 5 periods with SCHEFF (and a 70-degree phase spread
    that flags longitudinal losses): every coordinate, loss flag and SCHEFF
    counter of the OpenMP oracle equals the serial oracle's."""
    import math
    omp = _omp_oracle(tmp_path, 0, 40000, 5, sig_phi_deg)
    wl = H.workload(0, num=40000, periods=5, sc=True)
    if sig_phi_deg is not None:
        wl.beam.sig_phi = math.radians(sig_phi_deg)
    ob, osc, _ = H.run_oracle(wl)
    err = H.compare_arrays(omp, ob.arrays(), "openmp vs serial: ")
    assert err is None, err
    st = osc.state()
    assert int(omp["remesh_count"]) == st["remesh_count"]
    assert int(omp["kick_count"]) == st["kick_count"]
    if sig_phi_deg is not None:
        assert (omp["lloss"] != 0).sum() > 100


@pytest.mark.gpu
def test_headline_workload_bit_exact(tmp_path):
    """The benchmarked configuration itself: 1,048,576 particles through the
    800 elements of 200 periods, SCHEFF after every element; the GPU's final
    beam equals the OpenMP oracle's bit for bit, with the same remesh count."""
    omp = _omp_oracle(tmp_path, 1, 1048576, 200)
    wl = H.workload(1, num=1048576, periods=200, sc=True)
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), omp, "headline: ")
    assert err is None, err
    st = gsc.state()
    assert st["remesh_count"] == int(omp["remesh_count"])
    assert st["kick_count"] == int(omp["kick_count"])
    # the run is not trivially empty: most of the beam is transmitted and
    # the mesh was rebuilt at most kicks
    assert (gb.arrays()["loss"] == 0).sum() > 0.5 * 1048576
    assert st["remesh_count"] > 400
