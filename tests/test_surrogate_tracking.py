"""The learned SCHEFF surrogate in tracking: the 65,536-particle
configuration on its 40-period lattice with every remesh's field served by
the trained network (committed weights, hpsim_amd/weights), audited against
the exact grouped product every 20 kick points with fallback above 2 %
relative L2 field error, compared with the FP64 oracle; and the ablation --
the same beam with the untrained network and with no space charge."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from hpsim_amd import configs as cfg
from tests import helpers as H

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def _oracle_c1(tmp_path):
    sp = tmp_path / "c1.json"
    out = tmp_path / "c1.npz"
    sp.write_text(json.dumps({"config": 0, "num": 65536, "periods": 40}))
    env = dict(os.environ, OMP_NUM_THREADS=str(max(4, min(16, os.cpu_count() or 4))))
    subprocess.run([sys.executable, os.path.join(HERE, "oracle_omp_run.py"), str(sp), str(out)],
                   check=True, timeout=600, env=env)
    return dict(np.load(out))


def _track(net=None, audit_every=20, sc=True):
    """The configuration's beam alone in a batch: grouped field solve with
    the network attached (or the exact product when net is None, or no
    space charge); returns (arrays, surrogate counters)."""
    from hpsim_amd import surrogate as S
    wl = cfg.config(0)
    bw = cfg.BatchWorkload(wl.name, [wl.beam], [cfg.build_lattice(wl.lattice)], nr=wl.nr,
                           nz=wl.nz, interval=wl.interval)
    batch, objs = S.make_batch(bw, sc=sc)
    if net is not None:
        S.attach(batch, net, audit_every=audit_every, max_err=0.02)
    batch.simulate()
    info = S.info(batch) if net is not None else None
    return objs[0][0].arrays(), info


def test_surrogate_tracks_the_configuration_within_the_error_bounds(tmp_path):
    from hpsim_amd import surrogate as S
    mass = cfg.config(0).beam.mass
    ref = H.beam_metrics(_oracle_c1(tmp_path), mass)
    net = S.load(S.weights_path(32, 128))
    arrays, info = _track(net)
    e = H.metric_errors(H.beam_metrics(arrays, mass), ref)
    print("surrogate, audited:", info, e, flush=True)
    assert info["audits"] >= 1 and info["network_remeshes"] > 0
    assert max(e["ex"], e["ey"]) <= 0.01, e
    assert e["transmission"] <= 0.001, e
    assert e["w_avg"] <= 0.001, e
    # the same network without audits: every field from the network
    arrays_na, info_na = _track(net, audit_every=0)
    e_na = H.metric_errors(H.beam_metrics(arrays_na, mass), ref)
    print("surrogate, one audit only:", info_na, e_na, flush=True)
    # ablation: an untrained network of the same shape, and no space charge
    raw = S.Network(nr=32, nz=128, widths=tuple(int(v) for v in
                                                np.load(S.weights_path(32, 128))["dims"][3:7]),
                    nbot=int(np.load(S.weights_path(32, 128))["dims"][2]), seed=9)
    raw.set_norm(*net.norm())
    arrays_raw, info_raw = _track(raw, audit_every=0)
    e_raw = H.metric_errors(H.beam_metrics(arrays_raw, mass), ref)
    arrays_off, _ = _track(sc=False)
    e_off = H.metric_errors(H.beam_metrics(arrays_off, mass), ref)
    print("untrained network:", info_raw, e_raw, flush=True)
    print("no space charge:", e_off, flush=True)
    trained = max(e_na["ex"], e_na["ey"])
    assert trained < 0.1 * max(e_raw["ex"], e_raw["ey"])
    assert trained < 0.1 * max(e_off["ex"], e_off["ey"])
