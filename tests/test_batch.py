"""Batched launches (hps_batch_*): several simulators, each with its own beam
size and lattice settings, run as one launch sequence; every member must
match its own oracle run bit for bit."""
import numpy as np
import pytest

from hpsim_amd import configs as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _members(nums, periods=3, seed=11, freq=None):
    base = cfg.config(0, periods=periods)
    lattice = cfg.build_lattice(base.lattice)
    out = []
    for k, n in enumerate(nums):
        spec = cfg.BeamSpec(num=n, seed=seed * 1000 + k)
        out.append((spec, cfg.perturb_lattice(lattice, seed + k)))
    return base, out


def _gpu(members, nr=32, nz=128, interval=0.05, adj=0, ranges=((None, None),), sc=True):
    from hpsim_amd import HPSim
    objs = []
    for spec, els in members:
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = HPSim.SpaceCharge(nr=nr, nz=nz, interval=interval, adj_bunch=adj) if sc else None
        m = HPSim.Simulator(b, HPSim.BeamLine(elements=els), s)
        objs.append((b, s, m))
    batch = HPSim.Batch([m for _, _, m in objs])
    for r in ranges:
        if r == (None, None):
            batch.simulate()
        else:
            batch.simulate(*r)
    return objs, batch


def _oracle(members, nr=32, nz=128, interval=0.05, adj=0, ranges=((None, None),), sc=True):
    from oracle import oracle as O
    out = []
    for spec, els in members:
        b = O.Beam(spec.num, spec.mass, spec.charge, spec.current)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = O.SpaceCharge(nr, nz, interval, adj) if sc else None
        m = O.Simulator(b, els, s)
        for r in ranges:
            m.simulate(*r)
        out.append((b, s))
    return out


def _check(objs, ref):
    for k, ((gb, gs, _), (ob, os_)) in enumerate(zip(objs, ref)):
        err = H.compare_arrays(gb.arrays(), ob.arrays(), f"member {k}: ")
        assert err is None, err
        if gs is not None:
            g, o = gs.state(), os_.state()
            for f in ("remesh_count", "kick_count", "num_on_mesh", "dr", "dz"):
                assert g[f] == o[f], (k, f, g[f], o[f])


def test_batch_matches_separate_oracle_runs():
    """Five beams of different sizes, each on its own perturbed lattice."""
    _, members = _members([1000, 3000, 4096, 7000, 2500])
    objs, batch = _gpu(members)
    _check(objs, _oracle(members))
    assert batch.last_time_ms() > 0


def test_batch_partial_ranges_and_repeat():
    """simulate(start, end) twice (the second call continues the first), then
    the whole lattice again (a reset, as for a single simulator)."""
    _, members = _members([2048, 5000, 3333], periods=4)
    ranges = ((0, 6), (7, 15), (None, None))
    objs, _ = _gpu(members, ranges=ranges)
    _check(objs, _oracle(members, ranges=ranges))


def test_batch_without_space_charge_and_small_mesh():
    _, members = _members([1500, 4096])
    objs, _ = _gpu(members, sc=False)
    _check(objs, _oracle(members, sc=False))
    objs, _ = _gpu(members, nr=7, nz=33, adj=2)
    _check(objs, _oracle(members, nr=7, nz=33, adj=2))


def test_batch_with_frequency_change_on_the_ccl(tmp_path):
    """Beams at 201.25 MHz entering the 805 MHz CCL: the frequency change is a
    host step between batched launches."""
    import os
    import shutil
    from hpsim_amd import db
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ccl.db")
    shutil.copyfile(src, tmp_path / "ccl.db")
    c = db.DBConnection(str(tmp_path / "ccl.db"))
    c.clear_model_index()
    els = db.beamline_elements(c)[:120]
    c.close()
    members = []
    for k, n in enumerate((2048, 3000)):
        spec = cfg.BeamSpec(num=n, seed=50 + k, w0=99.9636448918256, phi0=-2.295,
                            sig_phi=0.05, sig_w=0.02)
        members.append((spec, cfg.perturb_lattice(els, 77 + k)))
    objs, _ = _gpu(members, nr=16, nz=64)
    _check(objs, _oracle(members, nr=16, nz=64))


def test_batch_rejects_different_layouts_and_keeps_simulators_usable():
    from hpsim_amd import HPSim
    _, members = _members([2000, 2000])
    members[1] = (members[1][0], members[1][1][:-4])      # one period shorter
    objs = []
    for spec, els in members:
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = HPSim.SpaceCharge(nr=32, nz=128, interval=0.05)
        objs.append((b, s, HPSim.Simulator(b, HPSim.BeamLine(elements=els), s)))
    batch = HPSim.Batch([m for _, _, m in objs])
    with pytest.raises(Exception, match="same elements|different programs"):
        batch.simulate()
    with pytest.raises(Exception, match="already in the batch"):
        batch.add(objs[0][2])
    objs[1][2].simulate()
    ref = _oracle(members[1:])
    assert H.compare_arrays(objs[1][0].arrays(), ref[0][0].arrays()) is None


@pytest.mark.parametrize("nr,nz,nb", [(32, 128, 40), (30, 60, 40), (64, 256, 16)])
def test_batch_of_many_small_beams(nr, nz, nb):
    """Many beams: the batched field sum takes its deepest register window
    (at 30 x 60 with padded partial lanes and a partial last window; at
    64 x 256 reading the tables from global memory, as they exceed LDS)."""
    _, members = _members([600 + 37 * k for k in range(nb)], periods=2, seed=23)
    objs, batch = _gpu(members, nr=nr, nz=nz)
    _check(objs, _oracle(members, nr=nr, nz=nz))
    shape, times = batch.tuning()          # the field shapes were timed and one kept
    assert shape is not None and len(times) == 9 and times[shape] == min(times)
    for k in (0, nb // 2, nb - 1):
        g, o = objs[k][1].tables(), _oracle(members[k:k + 1], nr=nr, nz=nz)[0][1].tables()
        for name, a, b in zip(("greens", "field", "bins"), g, o):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (k, name)


def test_pv_scan_as_one_batch_matches_the_scan_loop(tmp_path):
    """hpsim_amd.scan.pv_scan (the loop of sim-lbeg-scan-pv.py:140-160 as one
    batch) against the loop on the GPU: per value set_db_epics, then simulate
    a fresh copy of the initial beam with a fresh SpaceCharge."""
    import os
    import shutil
    from hpsim_amd import HPSim
    from hpsim_amd import lcsutil as lcs
    from hpsim_amd.scan import _copy_beam, pv_scan
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dtl.db")
    shutil.copyfile(src, tmp_path / "dtl.db")
    dbc = HPSim.DBConnection(str(tmp_path / "dtl.db"))
    dbc.clear_model_index()
    bl = HPSim.BeamLine(dbc)
    spec = cfg.BeamSpec(num=3000, seed=5, w0=0.75, phi0=-2.758, sig_phi=0.05, sig_w=2e-4)
    beam = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=0.015, num=spec.num)
    beam.set_distribution(*cfg.make_distribution(spec))
    beam.set_frequency(201.25)
    beam.set_ref_w(0.75)
    beam.save_initial_beam()
    sc = HPSim.SpaceCharge(nr=16, nz=64, interval=0.05)
    pv = lcs.get_pv_asp(1)
    v0 = float(HPSim.get_db_epics(pv, dbc))
    values = [v0 * f for f in (0.97, 1.0, 1.03)]
    stop = bl.index([n for n in bl.get_element_names() if n.startswith("01RG")][30])
    out = pv_scan(pv, values, beam, dbc, bl, sc, 0, stop)
    assert float(HPSim.get_db_epics(pv, dbc)) == v0
    for v, got in zip(values, out):
        b = _copy_beam(beam)
        HPSim.set_db_epics(pv, str(v), dbc, bl)
        sim = HPSim.Simulator(b, bl, HPSim.SpaceCharge(nr=16, nz=64, interval=0.05))
        sim.simulate(0, stop)
        assert H.compare_arrays(got.arrays(), b.arrays(), f"value {v}: ") is None
    HPSim.set_db_epics(pv, str(v0), dbc, bl)
    dbc.close()


def _run_twice(members, graph):
    """save, simulate, restore, simulate: the first call tunes the field
    launch shape (eager), the second replays the lattice periods as a HIP
    graph when graph is on."""
    from hpsim_amd import HPSim
    objs = []
    for spec, els in members:
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        b.save_initial_beam()
        s = HPSim.SpaceCharge(nr=32, nz=128, interval=0.05)
        objs.append((b, s, HPSim.Simulator(b, HPSim.BeamLine(elements=els), s)))
    batch = HPSim.Batch([m for _, _, m in objs])
    batch.set_graph(graph)
    batch.simulate()
    for b, _, _ in objs:
        b.restore_initial_beam()
    batch.simulate()
    return objs, batch


def test_batch_period_graph_replays_are_bit_exact():
    """Four members on six periods: the second simulate() replays one lattice
    period's launches as a HIP graph (checked by its replay count) and ends
    bit-identical to the eager launches and to the oracle doing the same."""
    from oracle import oracle as O
    _, members = _members([1500, 4096, 6000, 2222], periods=6)
    g_objs, g_batch = _run_twice(members, graph=True)
    e_objs, e_batch = _run_twice(members, graph=False)
    replays, per = g_batch.graph_info()
    assert replays >= 4 and per >= 6, (replays, per)
    assert e_batch.graph_info() == (0, 0)
    for k, ((gb, gs, _), (eb, es, _)) in enumerate(zip(g_objs, e_objs)):
        err = H.compare_arrays(gb.arrays(), eb.arrays(), f"member {k}, graph vs eager: ")
        assert err is None, err
        assert gs.state() == es.state()
    ref = []
    for spec, els in members:
        b = O.Beam(spec.num, spec.mass, spec.charge, spec.current)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        b.save_initial_beam()
        s = O.SpaceCharge(32, 128, 0.05)
        m = O.Simulator(b, els, s)
        m.simulate()
        b.restore_initial_beam()
        m.simulate()
        ref.append((b, s))
    _check(g_objs, ref)
