"""Failure paths of the maps: the loss codes the reference writes when a
particle's coordinates break down.

* RF gap second half: NaN in any coordinate or W <= 0 after the kick ->
  33333333 (simulation_kernel.cu:473-483);
* SCHEFF kick: bunch-frame momentum bgzf <= 0 -> 9999
  (space_charge_kernel.cu:695-696).

The inputs are forced from the host through set_distribution (NaN and
negative energies) and through an absurd beam current (10 A, whose space
charge field reverses particles in the bunch frame).  The CPU tests pin the
oracle's behaviour; the GPU tests require the engine to give the same codes
and coordinates bit for bit."""
import numpy as np
import pytest

from hpsim_amd import configs as cfg
from tests import helpers as H

GAP_LOSS = 33333333
KICK_LOSS = 9999
SLOW = slice(9, 41)


def _broken_beam():
    """64 particles of the 65,536-particle configuration's law; particles
    1..4 carry a NaN coordinate, 5..8 a negative kinetic energy, and 9..40
    sit on axis (no aperture loss) with energies from 1 keV to 0.3 MeV, of
    which those reaching the gap at a decelerating phase leave it with
    W <= 0."""
    spec = cfg.BeamSpec(num=64, seed=3)
    x, xp, y, yp, phi, w = (np.array(a) for a in cfg.make_distribution(spec))
    x[1] = np.nan
    xp[2] = np.nan
    y[3] = np.nan
    phi[4] = np.nan
    w[5:9] = -0.25
    x[SLOW] = xp[SLOW] = y[SLOW] = yp[SLOW] = 0.0
    w[SLOW] = np.geomspace(1e-3, 0.3, SLOW.stop - SLOW.start)
    return spec, (x, xp, y, yp, phi, w)


def _gap_lattice():
    wl = cfg.config(0, periods=1)
    return wl, cfg.build_lattice(wl.lattice)


def _oracle(spec, dist, elems, wl, sc):
    from oracle import oracle as O
    b = O.Beam(spec.num, spec.mass, spec.charge, spec.current)
    b.set_distribution(*dist)
    b.set_frequency(spec.freq)
    b.set_ref_w(spec.w0)
    s = O.SpaceCharge(wl.nr, wl.nz, wl.interval) if sc else None
    O.Simulator(b, elems, s).simulate()
    return b.arrays()


def _gpu(spec, dist, elems, wl, sc):
    from hpsim_amd import HPSim
    b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
    b.set_distribution(*dist)
    b.set_frequency(spec.freq)
    b.set_ref_w(spec.w0)
    s = HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval) if sc else None
    HPSim.Simulator(b, HPSim.BeamLine(elements=elems), s).simulate()
    return b.arrays()


def test_oracle_gap_flags_nan_and_nonpositive_energy():
    """One D-Q-D-G period, no space charge: every particle with a NaN
    coordinate or a negative energy reaches the gap (the drift and quad
    aperture tests are false for NaN) and is flagged there; the slow
    particles at the decelerating phase leave the gap with W <= 0."""
    spec, dist = _broken_beam()
    wl, elems = _gap_lattice()
    a = _oracle(spec, dist, elems, wl, sc=False)
    gap = len(elems) - 1
    assert (a["loss"][1:9] == GAP_LOSS).all(), a["loss"][1:9]
    # a flagged particle keeps its coordinates from before the gap's second half
    assert np.isnan(a["x"][1]) and np.isnan(a["xp"][2])
    assert (a["w"][5:9] == -0.25).all()
    slow = a["loss"][SLOW]
    assert (slow == GAP_LOSS).sum() >= 8, slow
    # the slow particles that pass keep a positive energy
    assert ((slow == GAP_LOSS) | ((slow == 0) & (a["w"][SLOW] > 0.0))).all()
    # particles with sound coordinates at the beam's energy are not flagged
    healthy = np.ones(64, bool)
    healthy[1:SLOW.stop] = False
    assert not (a["loss"][healthy] == GAP_LOSS).any()
    assert gap == 3


def test_oracle_kick_flags_reversed_particles():
    """At 10 A the space-charge kick reverses particles in the bunch frame:
    they are flagged 9999 and keep their momenta (no update after the test)."""
    wl = cfg.config(0, num=4096, periods=2)
    wl.beam.current = 10.0
    elems = cfg.build_lattice(wl.lattice)
    a = _oracle(wl.beam, cfg.make_distribution(wl.beam), elems, wl, sc=True)
    n_kick = int((a["loss"] == KICK_LOSS).sum())
    assert n_kick > 100, n_kick
    wl.beam.current = 0.010
    b = _oracle(wl.beam, cfg.make_distribution(wl.beam), elems, wl, sc=True)
    assert not (b["loss"] == KICK_LOSS).any()


@pytest.mark.gpu
def test_gpu_gap_loss_codes_match_oracle():
    spec, dist = _broken_beam()
    wl, elems = _gap_lattice()
    o = _oracle(spec, dist, elems, wl, sc=False)
    g = _gpu(spec, dist, elems, wl, sc=False)
    err = H.compare_arrays(g, o, "gap loss codes: ")
    assert err is None, err
    assert (g["loss"] == GAP_LOSS).sum() >= 8


@pytest.mark.gpu
def test_gpu_kick_loss_codes_match_oracle():
    wl = cfg.config(0, num=4096, periods=2)
    wl.beam.current = 10.0
    elems = cfg.build_lattice(wl.lattice)
    dist = cfg.make_distribution(wl.beam)
    o = _oracle(wl.beam, dist, elems, wl, sc=True)
    g = _gpu(wl.beam, dist, elems, wl, sc=True)
    err = H.compare_arrays(g, o, "kick loss codes: ")
    assert err is None, err
    assert (g["loss"] == KICK_LOSS).sum() > 100
