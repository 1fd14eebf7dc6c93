"""Runs the OpenMP build of the oracle on one single-beam workload in a
process of its own (the serial and OpenMP oracle libraries cannot share a
process) and saves the final particle arrays and SCHEFF counters:

    python tests/oracle_omp_run.py IDX NUM PERIODS OUT.npz [SIG_PHI_DEG]

Used by the tests as the checker; it touches no GPU.
"""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from hpsim_amd import configs as cfg  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    idx, num, periods, out = int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]), sys.argv[4]
    O.use_openmp()
    wl = cfg.config(idx, num=num, periods=periods)
    if len(sys.argv) > 5:
        wl.beam.sig_phi = math.radians(float(sys.argv[5]))
    b = wl.beam
    beam = O.Beam(b.num, b.mass, b.charge, b.current)
    beam.set_distribution(*cfg.make_distribution(b))
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    sc = O.SpaceCharge(wl.nr, wl.nz, wl.interval)
    O.Simulator(beam, cfg.build_lattice(wl.lattice), sc).simulate()
    st = sc.state()
    np.savez(out, threads=O.lib().orc_threads(), remesh_count=st["remesh_count"],
             kick_count=st["kick_count"], **beam.arrays())


if __name__ == "__main__":
    main()
