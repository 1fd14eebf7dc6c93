"""The C-ABI library loads and exports every symbol include/hpsim_amd.h
declares (no GPU needed)."""
import os
import re

from hpsim_amd import _native as N


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(os.path.dirname(N.__file__), "..", "include", "hpsim_amd.h")).read()
    declared = set(re.findall(r"\b(hps_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(N.EXPORTS)
    L = N.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_beam_size_limits_are_checked_before_any_allocation():
    """hps_beam_create rejects an empty beam and one beyond the 4,194,304
    particles the cross-chunk reductions fold, with a message, before it
    touches a device (so this runs without a GPU)."""
    import ctypes as C
    L = N.lib()
    h = C.c_void_p()
    for num, msg in ((0, "num > 0"), (4096 * 1024 + 1, "at most 4,194,304")):
        rc = L.hps_beam_create(C.byref(h), num, 939.294, 1.0, 0.0)
        assert rc != 0
        assert msg in L.hps_last_error().decode()
