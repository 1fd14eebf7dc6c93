"""LANSCE control-system PV names (the helpers of pylib/lcsutil.py that the
reference's scan scripts use, e.g. pytest/sim-lbeg-scan-pv.py:100-120).

A full PV name is AADDNNNCMM: area (2 characters), device type (2), device
number (3 digits), channel type (1), channel number (2).
"""
from __future__ import annotations

_POS = ["TA", "H+", "LA", "IP", "+"]
_NEG = ["TB", "H-", "LB", "-"]


def expand_pv(pv):
    """The PV in full AADDNNNCMM form, zero-padding the numeric parts; a
    shorter input gives the leading parts only (lcsutil.py:5-54).  A leading
    digit area ('5ks1e2' or '05ks1e2') is module 05."""
    s = str(pv)
    pos = 0
    if s[0].isdigit():
        width = 2 if len(s) > 1 and s[1].isdigit() else 1
        area = s[:width].zfill(2)
        pos = width
    else:
        area = s[:2].upper()
        pos = 2
    dev = s[pos:pos + 2].upper()
    pos += 2
    end = pos
    while end < len(s) and s[end].isdigit():
        end += 1
    num = s[pos:end].zfill(3) if pos < len(s) else ""
    pos = end
    ch = s[pos:pos + 1].upper() if pos < len(s) else ""
    pos += 1
    cnum = s[pos:].zfill(2) if pos < len(s) else ""
    return area + dev + num + ch + cnum


def get_pv_area(pv):
    return expand_pv(pv)[0:2]


def get_pv_device(pv):
    return expand_pv(pv)[2:4]


def get_pv_type(pv):
    """Channel type character (e.g. 'E', 'D', 'L', 'V')."""
    return expand_pv(pv)[7:8]


def get_pos_beams():
    return list(_POS)


def get_neg_beams():
    return list(_NEG)


def _area(n):
    return str(n).upper().zfill(2)


def get_pv_psp(n, beam="+"):
    """Phase set point PV of RF module or buncher n (lcsutil.py:80-102)."""
    a = _area(n)
    if a == "TA" or (a == "PB" and beam in _POS):
        return expand_pv("tadb1e2")
    if a == "TB" or (a == "PB" and beam in _NEG):
        return expand_pv("tbdb2e2")
    if a in ("TD", "MB"):
        return expand_pv("tddb1e2")
    if a in ("T1", "T2", "T3", "T4"):
        return expand_pv(a[1].zfill(2) + "js1d2")
    return expand_pv(a + ("js1d2" if int(a) < 5 else "ks1e1"))


def get_pv_asp(n, beam="+"):
    """Amplitude set point PV of RF module or buncher n (lcsutil.py:114-142)."""
    a = _area(n)
    if a == "TA" or (a == "PB" and beam in _POS):
        return expand_pv("tadb1e1")
    if a == "TB" or (a == "PB" and beam in _NEG):
        return expand_pv("tbdb2e4")
    if a in ("TD", "MB"):
        if beam in _POS:
            return expand_pv("tddb1e7")
        if beam in _NEG:
            return expand_pv("tddb1e4")
        raise ValueError(f"beam must be one of {_POS} or {_NEG}")
    if a in ("T1", "T2", "T3", "T4"):
        return expand_pv(a[1].zfill(2) + "js1d1")
    return expand_pv(a + ("js1d1" if int(a) < 5 else "ks1e2"))


def get_pv_asp_n(n):
    """As the reference (lcsutil.py:144-151): the '+' species PV."""
    return get_pv_asp(n, "+")


def get_pv_asp_p(n):
    """As the reference (lcsutil.py:153-160): the '-' species PV."""
    return get_pv_asp(n, "-")


def get_pv_rf(n, beam="+"):
    """RF on (in time) / off (delayed) PV of module or buncher n (lcsutil.py:162-187)."""
    a = _area(n)
    if a == "TA" or (a == "PB" and beam in _POS):
        return expand_pv("tadb1l3")
    if a == "TB" or (a == "PB" and beam in _NEG):
        return expand_pv("tbdb2l3")
    if a in ("TD", "MB"):
        return expand_pv("tddb1l3")
    if a in ("T1", "T2", "T3", "T4"):
        return expand_pv(a[1].zfill(2) + "tm1l1")
    return expand_pv(a + "tm1l1")


def rf_on_val(n):
    """0: RF drive on (bunchers) / RF in time (modules)."""
    return 0


def rf_off_val(n):
    """1: RF drive off (bunchers) / RF delayed (modules)."""
    return 1
