"""Shared helpers for the parity tests: run the same workload through the
oracle and through the HIP engine and compare."""
from __future__ import annotations

import numpy as np

from hpsim_amd import configs as cfg

COORDS = ("x", "xp", "y", "yp", "phi", "w")


def workload(idx=0, num=4096, periods=4, current=None, sc=True, seed=None):
    wl = cfg.config(idx, num=num, periods=periods)
    if current is not None:
        wl.beam.current = current
    if seed is not None:
        wl.beam.seed = seed
    wl.space_charge = sc
    return wl


def _configure(sc, opts):
    """SpaceCharge setters by name, e.g. {"mesh_size_cutoff_w": 1.2}."""
    for k, v in (opts or {}).items():
        getattr(sc, "set_" + k)(v)


def run_oracle(wl, elems=None, dist=None, adj=0, adj_w=0.0, sc_opts=None):
    from oracle import oracle as O
    b = wl.beam
    elems = elems if elems is not None else cfg.build_lattice(wl.lattice)
    beam = O.Beam(b.num, b.mass, b.charge, b.current)
    beam.set_distribution(*(dist if dist is not None else cfg.make_distribution(b)))
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    sc = O.SpaceCharge(wl.nr, wl.nz, wl.interval, adj) if wl.space_charge else None
    if sc is not None and adj_w:
        sc.set_adj_bunch_cutoff_w(adj_w)
    if sc is not None:
        _configure(sc, sc_opts)
    sim = O.Simulator(beam, elems, sc)
    sim.simulate()
    return beam, sc, sim


def run_gpu(wl, elems=None, dist=None, adj=0, adj_w=0.0, sc_opts=None):
    from hpsim_amd import HPSim
    b = wl.beam
    elems = elems if elems is not None else cfg.build_lattice(wl.lattice)
    beam = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
    beam.set_distribution(*(dist if dist is not None else cfg.make_distribution(b)))
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    bl = HPSim.BeamLine(elements=elems)
    sc = (HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval, adj_bunch=adj)
          if wl.space_charge else None)
    if sc is not None and adj_w:
        sc.set_adj_bunch_cutoff_w(adj_w)
    if sc is not None:
        _configure(sc, sc_opts)
    sim = HPSim.Simulator(beam, bl, sc)
    sim.simulate()
    return beam, sc, sim


def compare_arrays(ga, oa, label=""):
    """Bit-exact comparison; returns a description of the first mismatch."""
    for k in COORDS:
        g, o = ga[k], oa[k]
        same = (g.view(np.uint64) == o.view(np.uint64)) | (np.isnan(g) & np.isnan(o))
        if not same.all():
            i = int(np.argmin(same))
            rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-300)
            return (f"{label}{k}: {int((~same).sum())} differ, first at {i}: gpu {g[i]!r} "
                    f"oracle {o[i]!r}, max rel {np.nanmax(rel):.3e}")
    for k in ("loss", "lloss"):
        if not np.array_equal(ga[k], oa[k]):
            i = int(np.argmax(ga[k] != oa[k]))
            return f"{label}{k}: differ at {i}: gpu {ga[k][i]} oracle {oa[k][i]}"
    return None


from hpsim_amd.metrics import beam_metrics, metric_errors  # noqa: E402,F401
