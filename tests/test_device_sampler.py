"""The device sampler of synthetic beams (hps_beam_sample_gaussian) draws the
distribution configs.make_distribution gives on the host: the same Philox
counters, rejections and Box-Muller, up to the last bit of log and cos."""
import numpy as np
import pytest

from hpsim_amd import configs as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("num,seed", [(3, 5), (250_000, 1234567), (70_001, 2**32 + 99)])
def test_device_sampler_matches_host_sampler(num, seed):
    from hpsim_amd import HPSim
    spec = cfg.BeamSpec(num=num, seed=seed, phi0=0.3, w0=0.81)
    b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=num)
    b.sample_gaussian(spec)
    got = b.arrays()
    want = cfg.make_distribution(spec)
    sig = (spec.sig_x, spec.sig_xp, spec.sig_y, spec.sig_yp, spec.sig_phi, spec.sig_w)
    for k, w, s in zip(H.COORDS, want, sig):
        d = np.abs(got[k] - w).max()
        assert d <= 1e-14 * s * 4 + 1e-15 * np.abs(w).max(), (k, d)
    assert not got["loss"].any() and not got["lloss"].any()
    assert got["x"][0] == 0.0 and got["phi"][0] == 0.3 and got["w"][0] == 0.81
