"""Lattice databases (hpsim_amd.db, init.cpp:20-429) on the LANSCE DTL and
CCL sections the reference's own scripts load (fixtures in tests/golden/).

CPU: the loader against direct SQL on the same file, the model_index
write-back, and a physics pin of the whole chain (database fields -> gap
maps): the synchronous particle, started at the phase that reproduces the
first gap's design output energy, must follow the database's design
energies (energy_out_model, computed by the lattice designers' own code)
through every gap.  GPU: the engine against the oracle, bit for bit, on
these real lattices with space charge."""
import collections
import os
import shutil

import numpy as np
import pytest

from hpsim_amd import configs as cfg
from hpsim_amd import db

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASS, FREQ = 939.294, 201.25


def _copy(tmp_path, name):
    dst = tmp_path / name
    shutil.copyfile(os.path.join(GOLDEN, name), dst)
    return str(dst)


@pytest.fixture
def dtl(tmp_path):
    c = db.DBConnection(_copy(tmp_path, "dtl.db"))
    c.clear_model_index()
    yield c, db.beamline_elements(c)
    c.close()


@pytest.fixture
def ccl(tmp_path):
    c = db.DBConnection(_copy(tmp_path, "ccl.db"))
    c.clear_model_index()
    yield c, db.beamline_elements(c)
    c.close()


def test_dtl_loader_matches_sql(dtl):
    c, els = dtl
    kinds = collections.Counter(e.kind for e in els)
    assert kinds["RFGap-DTL"] == 165 and kinds["Quad"] > 0 and kinds["Drift"] > 0
    # every element of the linac view in order, model_index = position
    rows = c.conn.execute("select name, model_type from linac").fetchall()
    tracked = [r for r in rows if r[1] in ("quad", "drift", "dtl_gap", "ccl_gap", "diagnostics",
                                           "rotation", "caperture", "raperture", "buncher",
                                           "dipole", "spch_comp", "steerer")]
    assert [e.name for e in els] == [r[0] for r in tracked]
    for tbl in ("quad", "drift", "rf_gap"):
        for name, mi in c.conn.execute(f"select name, model_index from {tbl}"):
            if mi is not None:
                assert els[mi].name == name
    q = next(e for e in els if e.kind == "Quad")
    g, l, a = c.conn.execute("select gradient_model, length_model, aperture_model from quad "
                             "where name = ?", (q.name,)).fetchone()
    # read as text with 15 significant digits, then converted (std::atof)
    assert (q.gradient, q.length, q.aperture) == tuple(float(f"{v:.15g}") for v in (g, l, a))
    gap = [e for e in els if e.kind == "RFGap-DTL"][7]
    row = c.conn.execute("select amplitude_model, ref_phase_model, beam_phase_shift_model, "
                         "energy_out_model, frequency_model, length_model from rf_gap "
                         "where name = ?", (gap.name,)).fetchone()
    assert (gap.gap["amplitude"], gap.gap["phase_ref"], gap.gap["phase_shift"],
            gap.gap["energy_out"], gap.gap["frequency"], gap.gap["length"]) == tuple(
                float(f"{v:.15g}") for v in row)
    assert len(c.get_epics_channels()) > 0


def _sync_energies(O, els, w0, phi0, freq):
    """Track particle 0 (on axis) gap by gap; returns its energy after each gap."""
    b = O.Beam(2, MASS, -1.0, 0.0)
    z = np.zeros(2)
    b.set_distribution(z, z, z, z, z + phi0, z + w0)
    b.set_frequency(freq)
    b.set_ref_w(w0)
    sim = O.Simulator(b, els, None)
    out, prev = [], 0
    for i, e in enumerate(els):
        if e.kind.startswith("RFGap"):
            sim.simulate(prev, i)
            prev = i + 1
            out.append(b.arrays()["w"][0])
    return np.array(out)


def _entry_phase(O, els, w0, freq):
    """Entrance phase whose first-gap output energy is the design value."""
    first = next(i for i, e in enumerate(els) if e.kind.startswith("RFGap"))
    target = els[first].gap["energy_out"]
    lo, hi = -np.pi, np.pi
    for _ in range(3):
        phis = np.linspace(lo, hi, 241)
        b = O.Beam(len(phis) + 1, MASS, -1.0, 0.0)
        z = np.zeros(len(phis) + 1)
        b.set_distribution(z, z, z, z, np.concatenate([[phis[0]], phis]), z + w0)
        b.set_frequency(freq)
        b.set_ref_w(w0)
        O.Simulator(b, els, None).simulate(0, first)
        k = int(np.argmin(np.abs(b.arrays()["w"][1:] - target)))
        step = phis[1] - phis[0]
        lo, hi = phis[k] - step, phis[k] + step
    return phis[k]


def test_dtl_synchronous_particle_follows_design_energy(oracle_lib, dtl):
    _, els = dtl
    phi0 = _entry_phase(oracle_lib, els, 0.75, FREQ)
    w = _sync_energies(oracle_lib, els, 0.75, phi0, FREQ)
    design = np.array([e.gap["energy_out"] for e in els if e.kind.startswith("RFGap")])
    # the entrance phase is matched on the first gap only, so the particle
    # makes small synchrotron oscillations about the design (a few % early
    # on); a wrong sign, index or field mapping loses synchronism instead
    assert w[-1] == pytest.approx(design[-1], rel=0.01)        # ~100 MeV
    assert np.max(np.abs(w / design - 1)) < 0.06


def test_ccl_synchronous_particle_follows_design_energy(oracle_lib, ccl):
    """4959 CCL gaps at 805 MHz after the 201.25 MHz DTL: exercises the CCL
    gap maps, cell numbering and the beam-frequency change."""
    _, els = ccl
    w_in = 99.9636448918256   # the DTL's last design energy
    phi0 = _entry_phase(oracle_lib, els, w_in, FREQ)
    w = _sync_energies(oracle_lib, els, w_in, phi0, FREQ)
    design = np.array([e.gap["energy_out"] for e in els if e.kind.startswith("RFGap")])
    assert len(w) == 4959
    assert w[-1] == pytest.approx(design[-1], rel=0.01)        # ~800 MeV
    assert np.max(np.abs(w / design - 1)) < 0.06


def test_facade_database_beamline(tmp_path):
    from hpsim_amd import hpsim as hps
    _copy(tmp_path, "dtl.db")
    conn = hps.DBConnection(str(tmp_path), ["dtl.db"])
    hps.BeamLine()
    names = hps.get_element_list()
    assert len(names) == len(db.beamline_elements(conn.dbconnection))
    gaps = hps.get_element_list(elem_type="dtl-gap")
    assert len(gaps) == 165 and all("RG" in n for n in gaps)
    assert hps.get_next_element(names[0]) == names[1]
    conn.dbconnection.close()


def test_unsupported_element_kinds_are_refused():
    from hpsim_amd import HPSim
    els = [cfg.Element("S1", "Solenoid"), cfg.Element("D1", "Drift", length=0.1)]
    with pytest.raises(NotImplementedError, match="Solenoid"):
        HPSim.BeamLine(elements=els)


def test_full_lansce_lattice_loads(tmp_path, oracle_lib):
    """The four sections of the reference's LBEG script (sim-lbeg.py:40-44)
    as one beamline: transport with bunchers and dipoles, DTL, transition
    region, CCL; the engine and the oracle accept every element."""
    from hpsim_amd import HPSim
    c = db.DBConnection(_copy(tmp_path, "tbtd.db"))
    for k, name in enumerate(("dtl.db", "trst.db", "ccl.db")):
        c.attach_db(_copy(tmp_path, name), f"db{k + 1}")
    c.clear_model_index()
    bl = HPSim.BeamLine(c)
    els = list(bl)
    kinds = collections.Counter(e.kind for e in els)
    assert len(els) == 6540 and kinds["RFGap-DTL"] == 165 and kinds["RFGap-CCL"] == 4959
    assert kinds["Dipole"] == 8 and kinds["Buncher"] == 3
    oracle_lib.Simulator(oracle_lib.Beam(4, MASS, -1.0, 0.0), els, None)
    c.close()


def _beam(n, w0, phi0, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.normal(0, 1e-3, (2, n))
    xp, yp = rng.normal(0, 1e-3, (2, n))
    phi = phi0 + rng.normal(0, 0.05, n)
    w = w0 * (1 + rng.normal(0, 2e-4, n))
    return x, xp, y, yp, phi, w


def _gpu_vs_oracle(oracle_lib, els, n, w0, phi0, nr, nz, interval=0.05, adj=0, adj_w=0.0,
                   remesh=None):
    from hpsim_amd import HPSim
    from tests.helpers import compare_arrays
    O = oracle_lib
    dist = _beam(n, w0, phi0, 5)
    ob = O.Beam(n, MASS, -1.0, 0.015)
    ob.set_distribution(*dist)
    ob.set_frequency(FREQ)
    ob.set_ref_w(w0)
    osc = O.SpaceCharge(nr, nz, interval, adj)
    if adj_w:
        osc.set_adj_bunch_cutoff_w(adj_w)
    if remesh is not None:
        osc.set_remesh_threshold(remesh)
    O.Simulator(ob, els, osc).simulate()
    gb = HPSim.Beam(mass=MASS, charge=-1.0, current=0.015, num=n)
    gb.set_distribution(*dist)
    gb.set_frequency(FREQ)
    gb.set_ref_w(w0)
    gsc = HPSim.SpaceCharge(nr=nr, nz=nz, interval=interval, adj_bunch=adj)
    if adj_w:
        gsc.set_adj_bunch_cutoff_w(adj_w)
    if remesh is not None:
        gsc.set_remesh_threshold(remesh)
    HPSim.Simulator(gb, HPSim.BeamLine(elements=els), gsc).simulate()
    err = compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err
    assert gsc.state()["remesh_count"] == osc.state()["remesh_count"]
    return gb


@pytest.mark.gpu
def test_gpu_dtl_lattice_with_space_charge_bit_exact(oracle_lib, dtl):
    _, els = dtl
    gb = _gpu_vs_oracle(oracle_lib, els, 2048, 0.75, -2.758, 16, 64)
    assert gb.get_loss_num("t") < 2048 // 2


def _section(tmp_path, name):
    c = db.DBConnection(_copy(tmp_path, name))
    c.clear_model_index()
    els = db.beamline_elements(c)
    c.close()
    return els


@pytest.mark.gpu
def test_gpu_tbtd_transport_with_bunchers_and_dipoles_bit_exact(oracle_lib, tmp_path):
    """Low-energy transport: bunchers (one off), dipoles with fringe fields
    and pole-face rotations, steerers, apertures, space-charge
    compensation."""
    els = _section(tmp_path, "tbtd.db")
    _gpu_vs_oracle(oracle_lib, els, 2048, 0.7518, 0.0, 16, 64)


@pytest.mark.gpu
def test_gpu_tbtd_with_adjacent_bunches_bit_exact(oracle_lib, tmp_path):
    """The space-charge settings of the reference's LBEG script
    (sim-lbeg.py:72-78): three adjacent bunches each side below 0.8 MeV,
    interval 0.025 m, remesh threshold 0.02."""
    els = _section(tmp_path, "tbtd.db")
    _gpu_vs_oracle(oracle_lib, els, 2048, 0.7518, 0.0, 32, 128, interval=0.025, adj=3,
                   adj_w=0.8, remesh=0.02)


@pytest.mark.gpu
def test_gpu_trst_transition_dipoles_bit_exact(oracle_lib, tmp_path):
    """Transition region at 100 MeV: defocusing field index (hyperbolic
    branch of the dipole map), rotations, rectangular apertures."""
    els = _section(tmp_path, "trst.db")
    _gpu_vs_oracle(oracle_lib, els, 2048, 99.9636448918256, 0.0, 16, 64)


@pytest.mark.gpu
def test_gpu_ccl_lattice_with_space_charge_bit_exact(oracle_lib, ccl):
    _, els = ccl
    _gpu_vs_oracle(oracle_lib, els[:700], 2048, 99.9636448918256, -2.295, 16, 64)


@pytest.mark.gpu
def test_gpu_tracks_the_lattice_after_pv_changes(oracle_lib, tmp_path):
    """PV scan step: set_db_epics on the first DTL module's amplitude and
    phase updates the engine's beamline in place; the GPU then tracks the
    changed lattice bit for bit like the oracle on the same element records."""
    from hpsim_amd import HPSim
    from hpsim_amd import hpsim as hps
    from hpsim_amd import lcsutil as lcs
    from tests.helpers import compare_arrays
    O = oracle_lib
    _copy(tmp_path, "dtl.db")
    conn = hps.DBConnection(str(tmp_path), ["dtl.db"])
    bl = hps.BeamLine().beamline
    n, w0 = 2048, 0.75
    dist = _beam(n, w0, -2.758, 9)
    gb = HPSim.Beam(mass=MASS, charge=-1.0, current=0.015, num=n)
    gb.set_distribution(*dist)
    gb.set_frequency(FREQ)
    gb.set_ref_w(w0)
    gb.save_initial_beam()
    gsc = HPSim.SpaceCharge(nr=16, nz=64, interval=0.05)
    sim = HPSim.Simulator(gb, bl, gsc)
    end = bl.index(hps.get_element_list(elem_type="dtl-gap")[40])
    # the oracle side keeps its beam and SCHEFF state across scan steps too
    ob = O.Beam(n, MASS, -1.0, 0.015)
    ob.set_distribution(*dist)
    ob.set_frequency(FREQ)
    ob.set_ref_w(w0)
    ob.save_initial_beam()
    osc = O.SpaceCharge(16, 64, 0.05)
    for scale, dphi in ((0.97, 0.0), (1.03, 5.0)):
        gb.restore_initial_beam()
        for pv, v in ((lcs.get_pv_asp(1), scale), (lcs.get_pv_psp(1), dphi)):
            v0 = float(HPSim.get_db_epics(pv, conn.dbconnection))
            HPSim.set_db_epics(pv, str(v0 * v if pv.endswith("D01") else v0 + v),
                               conn.dbconnection, bl)
        sim.simulate(0, end)
        ob.restore_initial_beam()
        # the whole lattice, as the engine sees it: the last gap's map reads
        # the quad after it
        O.Simulator(ob, list(bl), osc).simulate(0, end)
        err = compare_arrays(gb.arrays(), ob.arrays())
        assert err is None, (scale, err)
    conn.dbconnection.close()
