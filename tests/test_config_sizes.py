"""GPU parity at the sizes the configurations are specified at.

* the 65,536-particle single beam on its whole 40-period lattice;
* members of the 256-beam batched scan (log-uniform 16,384 .. 1,048,576
  particles, perturbed lattices) at their own sizes, several above 536,576
  particles -- the size from which a batch member's rebuild grid has its
  deposit workgroups fill the Green's table -- run as one batch;
* one 4,194,304-particle beam on the 64 x 256 mesh, the member shape of the
  combined-stress configuration;
* three members of the combined-stress configuration as one batch (12,288
  chunks: the batch form of the phase pass that runs one chunk per
  wavefront).

Each is compared bit for bit with the OpenMP oracle (itself pinned to the
serial oracle in test_full_size.py), run in a process of its own by
tests/oracle_omp_run.py."""
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


def _omp(tmp_path, spec, name, timeout=600):
    sp = tmp_path / f"{name}.json"
    out = tmp_path / f"{name}.npz"
    sp.write_text(json.dumps(spec))
    threads = max(4, min(16, os.cpu_count() or 4))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([sys.executable, os.path.join(HERE, "oracle_omp_run.py"), str(sp), str(out)],
                   check=True, timeout=timeout, env=env)
    return dict(np.load(out))


def _check_state(label, st, ref):
    for f in ("remesh_count", "kick_count", "num_on_mesh", "dr", "dz"):
        assert st[f] == ref[f].item(), (label, f, st[f], ref[f])


def test_single_beam_65536x40_bit_exact(tmp_path):
    """The 65,536-particle configuration on its whole 160-element lattice."""
    ref = _omp(tmp_path, {"config": 0, "num": 65536, "periods": 40}, "c1")
    wl = H.workload(0, num=65536, periods=40, sc=True)
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ref, "65536x40: ")
    assert err is None, err
    _check_state("65536x40", gsc.state(), ref)
    assert gsc.state()["remesh_count"] > 40


def test_scan_members_at_their_sizes_as_one_batch(tmp_path):
    """Six members of the 256-beam scan on two periods, run as one batch:
    the three largest (all above 536,576 particles), the smallest and two
    from the middle of the size range."""
    from hpsim_amd import HPSim
    nbeams, periods = 256, 2
    bw = cfg.batched_scan(nbeams=nbeams, periods=periods)
    nums = np.array([b.num for b in bw.beams])
    order = np.argsort(nums)
    pick = [int(order[-1]), int(order[-2]), int(order[-3]), int(order[0]),
            int(order[nbeams // 2]), int(order[nbeams // 4])]
    assert (nums[pick[:3]] > 536576).all() and nums[pick[3]] < 17000
    print("members", [(k, int(nums[k])) for k in pick], flush=True)
    objs = []
    for k in pick:
        spec = bw.beams[k]
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = HPSim.SpaceCharge(nr=bw.nr, nz=bw.nz, interval=bw.interval)
        objs.append((b, s, HPSim.Simulator(b, HPSim.BeamLine(elements=bw.lattices[k]), s)))
    HPSim.Batch([m for _, _, m in objs]).simulate()
    for k, (b, s, _) in zip(pick, objs):
        ref = _omp(tmp_path, {"scan_member": k, "nbeams": nbeams, "periods": periods},
                   f"member{k}")
        err = H.compare_arrays(b.arrays(), ref, f"member {k} ({nums[k]}): ")
        assert err is None, err
        _check_state(f"member {k}", s.state(), ref)


def test_stress_member_4194304_on_64x256_bit_exact(tmp_path):
    """One beam of the combined-stress shape: 4,194,304 particles (the most
    a Beam holds), SCHEFF on the 64 x 256 mesh, two periods."""
    ref = _omp(tmp_path, {"config": 0, "num": 4194304, "periods": 2, "nr": 64, "nz": 256},
               "c5member")
    wl = H.workload(0, num=4194304, periods=2, sc=True)
    wl.nr, wl.nz = 64, 256
    gb, gsc, _ = H.run_gpu(wl)
    err = H.compare_arrays(gb.arrays(), ref, "4194304 on 64x256: ")
    assert err is None, err
    _check_state("4194304 on 64x256", gsc.state(), ref)


def test_stress_members_as_one_batch_bit_exact(tmp_path):
    """The first three beams of the combined-stress configuration (4,194,304
    particles each, 64 x 256 mesh) on one period as one batch: 12,288
    chunks, so the phase pass runs one chunk per wavefront."""
    from hpsim_amd import HPSim
    nb, periods = 3, 1
    bw = cfg.combined_stress(nbeams=nb, num=4194304, periods=periods)
    assert nb * (bw.beams[0].num // 1024) >= 8192
    objs = []
    for spec, lat in zip(bw.beams, bw.lattices):
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = HPSim.SpaceCharge(nr=bw.nr, nz=bw.nz, interval=bw.interval)
        objs.append((b, s, HPSim.Simulator(b, HPSim.BeamLine(elements=lat), s)))
    HPSim.Batch([m for _, _, m in objs]).simulate()
    for k in (0, nb - 1):
        b, s, _ = objs[k]
        ref = _omp(tmp_path, {"stress_member": k, "nbeams": nb, "num": 4194304,
                              "periods": periods}, f"stress{k}")
        err = H.compare_arrays(b.arrays(), ref, f"stress member {k}: ")
        assert err is None, err
        _check_state(f"stress member {k}", s.state(), ref)
