"""Committed regression fixtures (tests/golden/oracle_*.npz, written by
tests/golden/make_oracle_fixtures.py): input beam, the serial oracle's final
beam and SCHEFF counters for three small cases.  The oracle must still
reproduce them bit for bit (CPU), and so must the GPU engine (GPU), so a
change that moved the oracle and the engine together would still be seen."""
import os

import numpy as np
import pytest

from tests import helpers as H
from tests.golden import make_oracle_fixtures as F

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    return dict(np.load(os.path.join(GOLDEN, f"oracle_{name}.npz")))


def _outputs(fx):
    return {k[4:]: v for k, v in fx.items() if k.startswith("out_")}


@pytest.mark.parametrize("name", F.CASES)
def test_oracle_reproduces_fixture(name):
    fx = _fixture(name)
    got = F.run_oracle(name)
    for k in ("x", "xp", "y", "yp", "phi", "w"):
        assert np.array_equal(got["in_" + k].view(np.uint64), fx["in_" + k].view(np.uint64)), k
    err = H.compare_arrays(_outputs(got), _outputs(fx), f"{name}: ")
    assert err is None, err
    assert int(got["remesh_count"]) == int(fx["remesh_count"])
    assert int(got["kick_count"]) == int(fx["kick_count"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", F.CASES)
def test_gpu_reproduces_fixture(name):
    from hpsim_amd import HPSim
    fx = _fixture(name)
    els, spec, _, scs = F.case_setup(name)
    dist = [fx["in_" + k] for k in ("x", "xp", "y", "yp", "phi", "w")]
    b = HPSim.Beam(mass=spec["mass"], charge=spec["charge"], current=spec["current"],
                   num=len(dist[0]))
    b.set_distribution(*dist)
    b.set_frequency(spec["freq"])
    b.set_ref_w(spec["w0"])
    sc = HPSim.SpaceCharge(nr=scs["nr"], nz=scs["nz"], interval=scs["interval"],
                           adj_bunch=scs["adj"])
    if scs["adj_w"]:
        sc.set_adj_bunch_cutoff_w(scs["adj_w"])
    if scs["remesh"] is not None:
        sc.set_remesh_threshold(scs["remesh"])
    HPSim.Simulator(b, HPSim.BeamLine(elements=els), sc).simulate()
    err = H.compare_arrays(b.arrays(), _outputs(fx), f"{name}: ")
    assert err is None, err
    assert sc.state()["remesh_count"] == int(fx["remesh_count"])
    assert sc.state()["kick_count"] == int(fx["kick_count"])
