"""The user-unit facade (hpsim_amd.hpsim, mirroring pylib/hpsim.py): helper
closed forms on the CPU; unit conversion and the class-level Simulator(beam)
wiring on the GPU against the engine module called directly."""
import math

import numpy as np
import pytest

from hpsim_amd import configs as cfg
from hpsim_amd import hpsim as hps


def test_betalambda_closed_form():
    m, f, w = 939.294, 201.25, 0.75
    g = 1 + w / m
    assert hps.betalambda(m, f, w) == pytest.approx(
        math.sqrt(1 - 1 / g ** 2) * 299.792458 / f, rel=1e-15)


def test_mismatch_factor():
    t = [0.7, 12.0, 0.002]
    assert hps.get_mmf(t, t) == pytest.approx(0.0, abs=1e-7)
    # upright ellipses with beta ratio k: R = k + 1/k, mmf = sqrt(k) - 1
    k = 2.25
    assert hps.get_mmf([0.0, 1.0, 1.0], [0.0, k, 1.0]) == pytest.approx(math.sqrt(k) - 1,
                                                                       rel=1e-12)


def test_modulo_phase_folds_into_window():
    ph = np.array([-725.0, -181.0, 0.0, 179.0, 181.0, 1000.0])
    out = hps.modulo_phase(ph, 30.0)
    assert np.all(out >= 30.0 - 180.0) and np.all(out < 30.0 + 180.0)
    turns = (out - ph) / 360.0
    assert np.allclose(turns, np.round(turns), atol=1e-12)


def test_twiss_of_sampled_ellipse():
    rng = np.random.default_rng(5)
    a, b, e = -0.8, 3.0, 2.0
    u0, v0 = rng.standard_normal((2, 400000))
    # x = sqrt(e b) u0, x' = sqrt(e/b) (v0 - a u0): rms emittance e, alpha a, beta b
    x = math.sqrt(e * b) * u0
    xp = math.sqrt(e / b) * (v0 - a * u0)
    al, bt, er = hps._twiss(x, xp)
    assert er == pytest.approx(e, rel=0.01)
    assert al == pytest.approx(a, abs=0.02) and bt == pytest.approx(b, rel=0.01)
    assert all(math.isnan(v) for v in hps._twiss(x, 2 * x))


def test_most_freq_value_finds_the_mode():
    rng = np.random.default_rng(2)
    vals = np.concatenate([rng.normal(37.0, 2.0, 50000), rng.uniform(-100, 100, 5000)])
    assert abs(hps.most_freq_value(vals) - 37.0) < 1.0


@pytest.mark.gpu
def test_facade_matches_engine_module_in_user_units():
    from hpsim_amd import HPSim
    wl = cfg.config(0, num=4096, periods=3)
    b = wl.beam
    elems = cfg.build_lattice(wl.lattice)
    x, xp, y, yp, phi, w = cfg.make_distribution(b)
    # facade input in cm, mr, deg, MeV
    ux, uxp, uy, uyp, uphi = x * 100.0, xp * 1000.0, y * 100.0, yp * 1000.0, phi * hps._DEG

    fb = hps.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
    fb.set_distribution(ux, uxp, uy, uyp, uphi, w)
    fb.set_frequency(b.freq)
    fb.set_ref_w(b.w0)
    hps.BeamLine(elements=elems)
    hps.SpaceCharge(wl.nr, wl.nz, wl.interval, 0)
    sim = hps.Simulator(fb)
    names = hps.get_element_list()
    sim.simulate(names[0], names[-1])

    eb = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
    eb.set_distribution(ux / 100.0, uxp / 1000.0, uy / 100.0, uyp / 1000.0, uphi / hps._DEG, w)
    eb.set_frequency(b.freq)
    eb.set_ref_w(b.w0)
    HPSim.Simulator(eb, HPSim.BeamLine(elements=elems),
                    HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval)).simulate()

    a = eb.arrays()
    assert np.array_equal(fb.get_losses(), a["loss"])
    assert np.array_equal(fb.get_x("all"), a["x"] * 100.0)
    assert np.array_equal(fb.get_phi("all"), a["phi"] * hps._DEG)
    assert fb.get_emittance_x() == eb.get_emittance_x() * 100.0 * 1000.0
    assert fb.get_size() == b.num - eb.get_loss_num()
    good = fb.get_good_mask()
    assert good.size == fb.get_size()
    d = hps.Distribution(fb, good)
    assert d.get_size() == good.size and d.get_loss_num() == 0
    assert d.get_twiss("x") == pytest.approx(fb.get_twiss("x", good), rel=1e-12)
    assert hps.get_beamline_length(names[0], names[-1]) == pytest.approx(
        sum(e.length for e in elems), rel=1e-12)
