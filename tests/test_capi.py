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
