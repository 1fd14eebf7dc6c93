"""GPU parity: the HIP engine against the serial FP64 oracle on the same
seeded inputs.  Coordinates and loss flags must match bit for bit."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fn,lo,hi", [("exp", -30, 30), ("log", 1e-12, 10), ("sin", -1e4, 1e4),
                                      ("cos", -1e4, 1e4), ("sinh", -3, 3), ("cosh", -3, 3),
                                      ("sqrt", 0, 100), ("fmod2pi", 0, 1e4), ("bessi0", -6, 6),
                                      ("bessi1", -6, 6), ("bessi1p", -6, 6)])
def test_device_math_matches_oracle(oracle_lib, fn, lo, hi):
    from hpsim_amd import HPSim
    import ctypes as C
    rng = np.random.default_rng(3)
    x = rng.uniform(lo, hi, 200000)
    g = HPSim.probe_math(fn, x)
    L = oracle_lib.lib()
    if fn in ("sin", "cos"):
        s, c = C.c_double(), C.c_double()
        ref = []
        for v in x[:20000]:
            L.orc_t_sincos(v, C.byref(s), C.byref(c))
            ref.append(s.value if fn == "sin" else c.value)
        ref = np.array(ref)
        g = g[:20000]
    elif fn == "sqrt":
        ref = np.sqrt(x)
    else:
        f = {"exp": L.orc_t_exp, "log": L.orc_t_log, "sinh": L.orc_t_sinh, "cosh": L.orc_t_cosh,
             "fmod2pi": lambda v: L.orc_t_fmod_pos(v, 2 * np.pi), "bessi0": L.orc_t_bessi0,
             "bessi1": L.orc_t_bessi1, "bessi1p": L.orc_t_bessi1p}[fn]
        ref = np.array([f(v) for v in x[:20000]])
        g = g[:20000]
    assert np.array_equal(g.view(np.uint64), ref.view(np.uint64)), fn


def test_device_division_is_ieee():
    from hpsim_amd import HPSim
    rng = np.random.default_rng(4)
    x = rng.standard_normal(500000) * 10.0 ** rng.uniform(-20, 20, 500000)
    y = rng.standard_normal(500000) * 10.0 ** rng.uniform(-20, 20, 500000)
    g = HPSim.probe_math("div", x, y)
    assert np.array_equal(g.view(np.uint64), (x / y).view(np.uint64))


@pytest.mark.parametrize("num,periods", [(1000, 3), (4096, 4)])
def test_tracking_no_space_charge_bit_exact(num, periods):
    wl = H.workload(0, num=num, periods=periods, sc=False)
    ob, _, _ = H.run_oracle(wl)
    gb, _, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err


# beams up to 524,288 particles push one particle per thread (k_segment_wide),
# larger ones four per thread (k_segment): both paths are covered
@pytest.mark.parametrize("num,periods", [(3000, 2), (4096, 5), (600000, 2)])
def test_tracking_space_charge_bit_exact(num, periods):
    wl = H.workload(0, num=num, periods=periods, sc=True)
    ob, osc, _ = H.run_oracle(wl)
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    gs, os_ = gsc.state(), osc.state()
    assert err is None, err + f"\nspch gpu {gs}\nspch oracle {os_}"
    for k in ("remesh_count", "kick_count", "num_on_mesh", "dr", "dz", "sigr_old", "sigz_old"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])


def test_graph_replays_match_eager_and_oracle():
    """restore + simulate three times on one simulator (the space-charge
    state carries over between calls): eager launches, then the captured
    HIP graph replayed with eager work in between; every call must equal
    the oracle's same sequence."""
    from hpsim_amd import HPSim, configs as cfg
    from oracle import oracle as O
    wl = H.workload(0, num=3000, periods=3, sc=True)
    elems = cfg.build_lattice(wl.lattice)
    dist = cfg.make_distribution(wl.beam)
    b = wl.beam
    ob = O.Beam(b.num, b.mass, b.charge, b.current)
    ob.set_distribution(*dist)
    ob.set_frequency(b.freq)
    ob.set_ref_w(b.w0)
    ob.save_initial_beam()
    osim = O.Simulator(ob, elems, O.SpaceCharge(wl.nr, wl.nz, wl.interval))
    refs = []
    for _ in range(3):
        ob.restore_initial_beam()
        osim.simulate()
        refs.append(ob.arrays())
    for graph in (False, True):
        beam = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
        beam.set_distribution(*dist)
        beam.set_frequency(b.freq)
        beam.set_ref_w(b.w0)
        beam.save_initial_beam()
        sim = HPSim.Simulator(beam, HPSim.BeamLine(elements=elems),
                              HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval))
        sim.set_graph(graph)
        for rep in range(3):
            beam.restore_initial_beam()
            sim.simulate()
            HPSim.probe_math("exp", np.linspace(0, 1, 1000))   # eager work between replays
            err = H.compare_arrays(beam.arrays(), refs[rep], f"graph={graph} rep={rep} ")
            assert err is None, err
        if graph:
            assert sim.graph_replays() == 2


def test_graph_replay_same_simulator():
    """Same simulator, restore + simulate three times: the second and third
    calls replay the captured graph."""
    from hpsim_amd import HPSim, configs as cfg
    wl = H.workload(0, num=2048, periods=2, sc=False)
    elems = cfg.build_lattice(wl.lattice)
    dist = cfg.make_distribution(wl.beam)
    ob, _, _ = H.run_oracle(wl, elems=elems, dist=dist)
    ref = ob.arrays()
    b = wl.beam
    beam = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
    beam.set_distribution(*dist)
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    beam.save_initial_beam()
    sim = HPSim.Simulator(beam, HPSim.BeamLine(elements=elems))
    for k in range(3):
        beam.restore_initial_beam()
        sim.reset()
        sim.simulate()
        err = H.compare_arrays(beam.arrays(), ref, f"call {k} ")
        assert err is None, err
    assert sim.graph_replays() == 2


@pytest.mark.parametrize("num", [5000, 150000, 600000])
def test_longitudinal_loss_rescan_bit_exact(num):
    """A 70-degree phase spread pushes particles out of the bucket, so the
    lloss flags and the chunk rescan path are exercised: 5 chunks (one per
    wavefront of the rescanning workgroup, one wave taking two), 147 chunks
    (more than the 64 one rescan round lists) and 586 chunks (the
    four-particles-per-thread segment kernel)."""
    import math
    wl = H.workload(0, num=num, periods=6, sc=True)
    wl.beam.sig_phi = math.radians(70.0)
    ob, osc, _ = H.run_oracle(wl)
    gb, gsc, _ = H.run_gpu(wl)
    oa = ob.arrays()
    assert (oa["lloss"] != 0).sum() > 10
    err = H.compare_arrays(gb.arrays(), oa)
    assert err is None, err


@pytest.mark.parametrize("num", [400_000, 1_500_000, 2_500_000])
def test_cross_chunk_tree_widths_bit_exact(num):
    """Beams of 391, 1,465 and 2,442 chunks: the cross-chunk trees run with
    2, 8 and 16 entries per thread (chunk partials padded to 512, 2,048 and
    4,096), each a specialised load-then-fold path, with the 70-degree phase
    spread so the chunk maxima of the longitudinal-loss test are live."""
    import math
    wl = H.workload(0, num=num, periods=2, sc=True)
    wl.beam.sig_phi = math.radians(70.0)
    ob, osc, _ = H.run_oracle(wl)
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err
    gs, os_ = gsc.state(), osc.state()
    for k in ("remesh_count", "kick_count", "num_on_mesh"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])


@pytest.mark.parametrize("num", [4096, 600000])
def test_aperture_elements_and_rotation_steerer(num):
    """Aperture elements (their centroid and loss-count statistics), a
    rotation and a steerer inside the lattice: 4 chunks (one particle per
    thread) and 586 chunks (four particles per thread)."""
    from hpsim_amd import configs as cfg
    wl = H.workload(0, num=num, periods=3, sc=True)
    elems = cfg.build_lattice(wl.lattice)
    extra = [cfg.Element("AC", "ApertureC", aperture=0.003, aperture_in=True),
             cfg.Element("ROT", "Rotation", angle=0.3),
             cfg.Element("ST", "Steerer", bl_h=2e-4, bl_v=1e-4),
             cfg.Element("AR", "ApertureR", aper_xl=0.004, aper_xr=0.002, aper_yt=0.003,
                         aper_yb=0.0025, aperture_in=True),
             cfg.Element("DG", "Diagnostics"),
             cfg.Element("SC", "SpchComp", fraction=0.5)]
    elems = elems[:5] + extra + elems[5:]
    dist = cfg.make_distribution(wl.beam)
    ob, _, _ = H.run_oracle(wl, elems=elems, dist=dist)
    gb, _, _ = H.run_gpu(wl, elems=elems, dist=dist)
    oa = ob.arrays()
    assert (oa["loss"] == 5).sum() > 0 and (oa["loss"] == 8).sum() > 0
    err = H.compare_arrays(gb.arrays(), oa)
    assert err is None, err


def test_getters_match_oracle():
    wl = H.workload(0, num=3000, periods=3, sc=True)
    ob, _, _ = H.run_oracle(wl)
    gb, _, _ = H.run_gpu(wl)
    pairs = [(lambda b: b.get_loss_num("t"), lambda o: o.get_loss_num("t")),
             (lambda b: b.get_emittance_x(), lambda o: o.get_emittance("x")),
             (lambda b: b.get_emittance_y(), lambda o: o.get_emittance("y")),
             (lambda b: b.get_emittance_z(), lambda o: o.get_emittance("z")),
             (lambda b: b.get_avg_x(), lambda o: o.get_avg("x")),
             (lambda b: b.get_avg_y("longitudinal_good"), lambda o: o.get_avg("y", True)),
             (lambda b: b.get_avg_phi("relative"), lambda o: o.get_avg("phi_r")),
             (lambda b: b.get_avg_w("longitudinal_good"), lambda o: o.get_avg("w", True)),
             (lambda b: b.get_sig_x(), lambda o: o.get_sig("x")),
             (lambda b: b.get_sig_phi("longitudinal_good"), lambda o: o.get_sig("phi", True)),
             (lambda b: b.get_sig_w(), lambda o: o.get_sig("w")),
             (lambda b: b.get_loss_num("l"), lambda o: o.get_loss_num("l")),
             (lambda b: b.get_loss_num("all"), lambda o: o.get_loss_num("all"))]
    for k, (g, o) in enumerate(pairs):
        gv, ov = g(gb), o(ob)
        assert gv == ov or (gv != gv and ov != ov), (k, gv, ov)
    gs, os_ = gb.stats(), ob.stats()
    for key, v in os_.items():
        assert gs[key] == v or (v != v and gs[key] != gs[key]), (key, gs[key], v)


@pytest.mark.parametrize("num", [1, 2, 3, 5, 257])
def test_tiny_beams_bit_exact(num):
    """Edge cases of the launch geometry: N = 1 skips quad/gap maps, N <= 2
    skips SCHEFF statistics and kicks, N = 3..4 kicks nobody (Q3)."""
    wl = H.workload(0, num=num, periods=2, sc=True)
    ob, osc, _ = H.run_oracle(wl)
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err
    gs, os_ = gsc.state(), osc.state()
    for k in ("remesh_count", "kick_count", "num_on_mesh"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])


def test_waterbag_and_dc_beams_match_oracle_then_track(oracle_lib):
    """set_waterbag / set_dc draw the same particles as the oracle (host
    sampling with the C library's rand), and the sampled beam then tracks
    bit-exactly through a lattice with space charge."""
    from hpsim_amd import HPSim
    from hpsim_amd import configs as cfg
    O = oracle_lib
    n = 4096
    wb = (-0.3, 20.0, 0.003, 0.4, 18.0, 0.003, 0.1, 30.0, 0.01, -0.52, 0.75, 201.25, 11)
    dc = (0.2, 15.0, 0.002, -0.2, 15.0, 0.002, 1.0, -0.5, 0.75, 5)
    for kind, args in (("waterbag", wb), ("dc", dc)):
        g = HPSim.Beam(mass=939.294, charge=-1.0, current=0.01, num=n)
        o = O.Beam(n, 939.294, -1.0, 0.01)
        getattr(g, "set_" + kind)(*args)
        getattr(o, "set_" + kind)(*args)
        if kind == "dc":
            g.set_frequency(201.25)
            o.set_frequency(201.25)
        assert H.compare_arrays(g.arrays(), o.arrays(), kind) is None
        assert g.get_frequency() == o.get_frequency()
    wl = cfg.config(0, num=n, periods=2)
    elems = cfg.build_lattice(wl.lattice)
    HPSim.Simulator(g, HPSim.BeamLine(elements=elems),
                    HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval)).simulate()
    O.Simulator(o, elems, O.SpaceCharge(wl.nr, wl.nz, wl.interval)).simulate()
    assert H.compare_arrays(g.arrays(), o.arrays(), "dc beam tracked") is None


@pytest.mark.parametrize("nr,nz", [(32, 128), (16, 64), (7, 33), (64, 256)])
def test_space_charge_tables_bit_exact(nr, nz):
    """Green's table, charge histogram and field table of the last remesh,
    for several mesh shapes (the LDS-staged and the direct field kernel)."""
    wl = H.workload(0, num=8192, periods=2, sc=True)
    wl.nr, wl.nz = nr, nz
    ob, osc, _ = H.run_oracle(wl)
    gb, gsc, _ = H.run_gpu(wl)
    assert gsc.state()["remesh_count"] == osc.state()["remesh_count"] > 0
    for name, g, o in zip(("greens", "field", "bins"), gsc.tables(), osc.tables()):
        assert g.shape == o.shape, name
        assert np.array_equal(g.view(np.uint64), o.view(np.uint64)), name
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err


@pytest.mark.parametrize("num,adj,adj_w", [(4096, 2, 0.0), (6000, 3, 1.6), (2, 2, 0.0), (5, 1, 0.0)])
def test_adjacent_bunch_space_charge_bit_exact(num, adj, adj_w):
    """adj_bunch > 0: image rings in the Green's table, image point charges
    for off-mesh particles, the mesh sized by the largest radius and one RF
    period; with a cutoff energy the beam crosses mid-lattice (the images
    switch off there).  N = 2 takes the direct r-max scan."""
    wl = H.workload(0, num=num, periods=6, sc=True)
    ob, osc, _ = H.run_oracle(wl, adj=adj, adj_w=adj_w)
    gb, gsc, _ = H.run_gpu(wl, adj=adj, adj_w=adj_w)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err
    gs, os_ = gsc.state(), osc.state()
    for k in ("remesh_count", "kick_count", "num_on_mesh", "dr", "dz"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    if num > 2:
        for name, g, o in zip(("greens", "field", "bins"), gsc.tables(), osc.tables()):
            assert np.array_equal(g.view(np.uint64), o.view(np.uint64)), name


@pytest.mark.parametrize("remesh", [None, 1e9])
def test_mesh_size_cutoff_bit_exact(remesh):
    """mesh_size_cutoff_w: once the beam passes 1.27 MeV the mesh halves and
    the drift kick interval quadruples, kick point by kick point.  With
    remeshing switched off the fresh (zeroed) half-size tables are what the
    later kicks read, as after the reference's SetMeshSize."""
    wl = H.workload(0, num=4096, periods=6, sc=True)
    wl.interval = 0.01
    opts = {"mesh_size_cutoff_w": 1.27}
    if remesh is not None:
        opts["remesh_threshold"] = remesh
    ob, osc, _ = H.run_oracle(wl, sc_opts=opts)
    gb, gsc, _ = H.run_gpu(wl, sc_opts=opts)
    err = H.compare_arrays(gb.arrays(), ob.arrays())
    assert err is None, err
    gs, os_ = gsc.state(), osc.state()
    for k in ("remesh_count", "kick_count", "num_on_mesh", "dr", "dz", "nx", "nz", "interval"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    assert gs["nx"] == wl.nr // 2 and gs["interval"] == 0.04
