"""The PV-scan API of the reference (hpsim.py:1783-1832, hpsim_module.cpp,
src/server/src/pv_observer.cpp): set_db_epics writes a PV into the lattice
database, whose SQL triggers recompute the model fields, and the PV's
observer copies them into the beamline the engine tracks."""
import os
import shutil

import numpy as np
import pytest

from hpsim_amd import lcsutil as lcs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _facade(tmp_path, names):
    from hpsim_amd import hpsim as hps
    for n in names:
        shutil.copyfile(os.path.join(GOLDEN, n), tmp_path / n)
    conn = hps.DBConnection(str(tmp_path), names)
    bl = hps.BeamLine()
    return hps, conn, bl.beamline


def _snapshot(beamline):
    return [(e.name, e.gradient, dict(e.gap), dict(e.params), e.length, e.aper_xl, e.aper_xr)
            for e in beamline]


def _changed(before, after):
    return {b[0] for b, a in zip(before, after) if b != a}


def test_pv_names_exist_in_the_databases(tmp_path):
    """The LCS name helpers give the PVs the lattice databases hold, with
    the PV types the observers expect."""
    import sqlite3
    types = {}
    for db in ("tbtd.db", "dtl.db", "ccl.db"):
        c = sqlite3.connect(os.path.join(GOLDEN, db))
        types.update(dict(c.execute("select lcs_name, value_type from epics_channel")))
        c.close()
    for m in range(1, 5):
        assert types[lcs.get_pv_asp(m)] == "rf_amp"
        assert types[lcs.get_pv_psp(m)] == "rf_ph"
        assert types[lcs.get_pv_rf(m)] == "delay"
    for m in range(5, 49):
        assert types[lcs.get_pv_asp(m)] == "rf_amp"
        assert types[lcs.get_pv_psp(m)] == "rf_ph"
    assert types[lcs.get_pv_asp("TD", "-")] == "buncher_amp"
    assert types[lcs.get_pv_psp("TD", "-")] == "buncher_ph"
    assert types[lcs.get_pv_asp("TB", "-")] == "buncher_amp"
    assert lcs.get_pv_type("05KS001E02") == "E" and lcs.expand_pv("5ks1e2") == "05KS001E02"


def test_dtl_amplitude_pv_scales_one_module(tmp_path):
    """A DTL module amplitude PV (the scan of sim-lbeg-scan-pv.py) changes the
    amplitude and reference phase of that module's gaps only, all by one
    operating-amplitude fraction; setting the PV back restores the lattice
    bit for bit."""
    hps, conn, bl = _facade(tmp_path, ["dtl.db"])
    pv = lcs.get_pv_asp(2)
    v0 = hps.get_db_epics(pv)
    gaps2 = [n for n in hps.get_element_list(elem_type="dtl-gap") if n.startswith("02")]
    before = _snapshot(bl)
    hps.set_db_epics(pv, v0 * 1.02)
    assert hps.get_db_epics(pv) == pytest.approx(v0 * 1.02, rel=1e-12)
    after = _snapshot(bl)
    changed = _changed(before, after)
    assert changed and changed <= set(gaps2)
    ratios = []
    for name in changed:
        e = bl[name]
        assert e.gap["amplitude"] == hps.get_db_model(name, "amplitude_model")
        assert e.gap["phase_ref"] == hps.get_db_model(name, "ref_phase_model")
        old = dict(before[bl.index(name)][2])
        ratios.append(e.gap["amplitude"] / old["amplitude"])
    assert np.ptp(ratios) < 1e-12 * np.mean(ratios) and abs(np.mean(ratios) - 1) > 1e-4
    hps.set_db_epics(pv, v0)
    assert _snapshot(bl) == before
    conn.dbconnection.close()


def test_quad_pv_moves_its_quad_only(tmp_path):
    hps, conn, bl = _facade(tmp_path, ["tbtd.db"])
    pv = "TBQL002V02"
    v0 = hps.get_db_epics(pv)
    before = _snapshot(bl)
    hps.set_db_epics(pv, v0 * 0.9)
    changed = _changed(before, _snapshot(bl))
    assert len(changed) == 1
    (name,) = changed
    assert bl[name].kind == "Quad"
    assert bl[name].gradient == hps.get_db_model(name, "gradient_model")
    assert bl[name].gradient == pytest.approx(0.9 * before[bl.index(name)][1], rel=0.05)
    conn.dbconnection.close()


def test_master_phase_pv_moves_every_ccl_module(tmp_path):
    """The CCL master reference phase PV drives the phase PVs of all CCL
    modules (MasterPVObserver): every gap's phase shift follows the database."""
    hps, conn, bl = _facade(tmp_path, ["ccl.db"])
    pv = "MRPH001D01"
    v0 = hps.get_db_epics(pv)
    before = _snapshot(bl)
    hps.set_db_epics(pv, v0 + 12.0)
    changed = _changed(before, _snapshot(bl))
    gaps = hps.get_element_list(elem_type="ccl-gap")
    modules = {n[:2] for n in changed}
    assert len(modules) > 10 and changed <= set(gaps)
    for name in list(changed)[:50]:
        assert bl[name].gap["phase_shift"] == hps.get_db_model(name, "beam_phase_shift_model")
    conn.dbconnection.close()


def test_set_db_model_module_phase_offset(tmp_path):
    """set_db_model on an RF module's phase_offset_cal re-phases its gaps by
    the offset (the trigger adds it to the PV phase, in degrees)."""
    hps, conn, bl = _facade(tmp_path, ["dtl.db"])
    mod = conn.dbconnection.conn.execute("select name, phase_offset_cal from rf_module "
                                         "where id = 3").fetchone()
    gaps3 = [r[0] for r in conn.dbconnection.conn.execute(
        "select name from rf_gap where module_id = 3")]
    before = {n: bl[n].gap["phase_shift"] for n in gaps3}
    hps.set_db_model(mod[0], "phase_offset_cal", mod[1] + 4.0)
    for n in gaps3:
        assert bl[n].gap["phase_shift"] == hps.get_db_model(n, "beam_phase_shift_model")
        assert bl[n].gap["phase_shift"] - before[n] == pytest.approx(np.radians(4.0), rel=1e-9)
    conn.dbconnection.close()


def test_unknown_pv_and_record(tmp_path, capsys):
    from hpsim_amd import HPSim
    hps, conn, bl = _facade(tmp_path, ["dtl.db"])
    before = _snapshot(bl)
    HPSim.set_db_epics("XXNO001Z01", "1.0", conn.dbconnection, bl)   # no observer: no change
    assert _snapshot(bl) == before
    assert HPSim.get_db_epics("XXNO001Z01", conn.dbconnection) is None
    assert HPSim.get_db_model("NOPE", "length_model", conn.dbconnection) is None
    conn.dbconnection.close()
