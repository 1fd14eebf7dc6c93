#!/usr/bin/env python3
"""Benchmark: particle-element pushes/s including SCHEFF space charge.

One step = one simulate() of the whole lattice for the 1,048,576-particle
DTL-like workload (200 Drift/Quad/Drift/DTL-gap periods, gap voltage ramped
0.3 -> 0.6 MV, quad gradients scaled with beta*gamma, SCHEFF 32x128 after
every element), started from the saved initial beam.  value = particles x
elements x steps / wall time of the timed steps.

The JSON line also carries
  roofline:      HBM roofline of the fused element-map kernel (k_segment):
                 algorithmic bytes per launch / average launch time, the
                 average measured with HIP events (on the engine's stream)
                 around every launch of one extra eager step run after the
                 timed steps, which replay a captured HIP graph;
  cpu_baseline:  the FP64 oracle built with OpenMP (oracle/liboracle_omp.so,
                 bit-identical to the serial checker) on all host cores
                 (OMP_NUM_THREADS) over the whole workload, timed on this
                 host; "serial" inside it is the one-core oracle over the
                 first 10 periods.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
SEG_BYTES_PER_PARTICLE = 104   # k_segment: reads 6 FP64 + 2 u32, writes 6 FP64
# HBM bytes per k_segment launch at 1,048,576 particles, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE counter runs (calibrated on a plain device copy)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_segment_traffic_plain-fp64-1048576x200.md")


def measured_traffic(num):
    """Traffic per k_segment launch from the committed counter summary, for
    the particle count it was measured at; None otherwise."""
    if num != 1048576 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--num", type=int, default=None, help="override particle count")
    p.add_argument("--periods", type=int, default=None, help="override lattice periods")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--eager", action="store_true",
                   help="launch every kernel from the host instead of replaying a HIP graph "
                        "(for profilers that do not follow graph launches)")
    p.add_argument("--cpu-periods", type=int, default=200,
                   help="lattice periods of the all-core CPU-baseline sample")
    p.add_argument("--cpu-serial-periods", type=int, default=10,
                   help="lattice periods of the one-core CPU-baseline sample")
    p.add_argument("--workload", default="plain",
                   choices=("plain", "small", "batched-scan", "stress"),
                   help="plain: the 1,048,576-particle single beam (default); small: the "
                        "65,536-particle single beam on 40 periods; batched-scan: 256 beams "
                        "of 16K-1M particles with per-beam lattice settings in one batch; "
                        "stress: 64 beams of 4,194,304 particles, 400 periods, 64x256 mesh, "
                        "in one batch")
    p.add_argument("--beams", type=int, default=256, help="beams of the batched scan")
    return p.parse_args()


def _oracle_run(wl, periods, openmp):
    """The oracle on the first `periods` lattice periods of the workload;
    returns (pushes/s, threads, seconds, elements)."""
    import copy

    from hpsim_amd import configs as cfg
    from oracle import oracle as O

    if openmp:
        O.use_openmp()
    w2 = copy.deepcopy(wl)
    elems = cfg.build_lattice(w2.lattice)[: 4 * periods]
    b = w2.beam
    dist = cfg.make_distribution(b)
    beam = O.Beam(b.num, b.mass, b.charge, b.current)
    beam.set_distribution(*dist)
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    sc = O.SpaceCharge(w2.nr, w2.nz, w2.interval)
    sim = O.Simulator(beam, elems, sc)
    t0 = time.perf_counter()
    sim.simulate()
    dt = time.perf_counter() - t0
    return b.num * len(elems) / dt, O.lib().orc_threads(), dt, len(elems)


def cpu_baseline(wl, periods, serial_periods, idx=1):
    """OpenMP oracle on all host cores (OMP_NUM_THREADS) over `periods`
    periods, plus the serial oracle on `serial_periods` periods in a child
    process (one process holds one oracle build)."""
    import subprocess

    periods = min(periods, wl.lattice.periods)
    rate, threads, dt, nel = _oracle_run(wl, periods, True)
    out = {"value": rate, "unit": "particle-element pushes/s", "cores": threads,
           "kind": "port",
           "sample": f"{wl.beam.num} particles x first {nel} elements ({periods} periods) of "
                     f"the same lattice, SCHEFF on, OpenMP oracle, {threads} threads, {dt:.1f} s"}
    code = ("import json, sys; sys.path.insert(0, %r); import bench; "
            "from hpsim_amd import configs as cfg; "
            "print(json.dumps(bench._oracle_run(cfg.config(%d, num=%d, periods=%d), %d, False)))"
            % (ROOT, idx, wl.beam.num, wl.lattice.periods, serial_periods))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True)
    srate, _, sdt, snel = json.loads(r.stdout.strip().splitlines()[-1])
    out["serial"] = {"value": srate, "cores": 1,
                     "sample": f"first {snel} elements, serial oracle, {sdt:.1f} s"}
    return out


def _member_dist(spec):
    from hpsim_amd import configs as cfg
    return cfg.make_distribution(spec)


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batched_cpu_baseline(w, budget_s=20.0, max_elements=None):
    """OpenMP oracle over whole members of the batch, in order, until about
    budget_s seconds of tracking; pushes/s of that sample (simulate() time
    only).  max_elements: track only the first elements of each member."""
    from oracle import oracle as O
    O.use_openmp()
    done = pushes = 0
    busy = 0.0
    for spec, els in zip(w.beams, w.lattices):
        if max_elements:
            els = els[:max_elements]
        beam = O.Beam(spec.num, spec.mass, spec.charge, spec.current)
        beam.set_distribution(*_member_dist(spec))
        beam.set_frequency(spec.freq)
        beam.set_ref_w(spec.w0)
        sim = O.Simulator(beam, els, O.SpaceCharge(w.nr, w.nz, w.interval))
        t0 = time.perf_counter()
        sim.simulate()
        busy += time.perf_counter() - t0
        pushes += spec.num * len(els)
        done += 1
        if busy > budget_s:
            break
    nel = min(len(w.lattices[0]), max_elements or len(w.lattices[0]))
    return {"value": pushes / busy, "unit": "particle-element pushes/s",
            "cores": O.lib().orc_threads(), "kind": "port",
            "sample": f"first {done} of {len(w.beams)} beams of the batch "
                      f"({pushes // nel} particles x first {nel} elements), "
                      f"OpenMP oracle, one beam after the other, {busy:.1f} s"}


def main_batched(args):
    import multiprocessing as mp

    from hpsim_amd import HPSim
    from hpsim_amd import configs as cfg

    if args.workload == "stress":
        w = cfg.combined_stress(nbeams=args.beams if args.beams != 256 else 64,
                                num=args.num or 4194304, periods=args.periods or 400)
    else:
        w = cfg.batched_scan(nbeams=args.beams, periods=args.periods or 40)
    total = sum(b.num for b in w.beams)
    nel = len(w.lattices[0])
    _log(f"{len(w.beams)} beams, {total} particles, {nel} elements: sampling distributions")
    with mp.get_context("spawn").Pool(min(16, os.cpu_count() or 1)) as pool:
        dists = pool.map(_member_dist, w.beams, chunksize=1)
    _log("uploading beams")
    sims = []
    for spec, els, d in zip(w.beams, w.lattices, dists):
        beam = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        beam.set_distribution(*d)
        beam.set_frequency(spec.freq)
        beam.set_ref_w(spec.w0)
        beam.save_initial_beam()
        sc = HPSim.SpaceCharge(nr=w.nr, nz=w.nz, interval=w.interval)
        sims.append(HPSim.Simulator(beam, HPSim.BeamLine(elements=els), sc))
    del dists
    batch = HPSim.Batch(sims)

    def step():
        for m in sims:
            m.beam.restore_initial_beam()
        batch.simulate()

    for k in range(args.warmup):
        step()
        _log(f"warm-up step {k + 1} done")
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        _log(f"timed step {k + 1} done")
    dt = time.perf_counter() - t0
    batch.set_profiling(True)
    step()
    prof = batch.profile()
    batch.set_profiling(False)
    seg_ms, seg_n = prof["segment"]
    avg_s = seg_ms / 1e3 / max(seg_n, 1)
    achieved = SEG_BYTES_PER_PARTICLE * total / avg_s / 1e9
    transmitted = sum(m.beam.get_size() - m.beam.get_loss_num("t") for m in sims)
    out = {
        "metric": "particle-element pushes/s incl. space charge",
        "value": total * nel * args.steps / dt,
        "unit": "particle-element pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": ("synthetic (Philox-seeded truncated Gaussian beams, generated lattice"
                 + (")" if args.workload == "stress" else " with seeded per-beam perturbations)")),
        "config": {"workload": w.name, "beams": len(w.beams), "particles": total,
                   "particles_min": min(b.num for b in w.beams),
                   "particles_max": max(b.num for b in w.beams), "elements": nel,
                   "mesh": f"{w.nr}x{w.nz}", "transmitted": transmitted},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_segment_batch", "avg_launch_us": avg_s * 1e6,
                     "launches": seg_n, "bytes_per_launch": SEG_BYTES_PER_PARTICLE * total,
                     "timing": "HIP events around every launch of one extra step"},
        "kernel_ms_per_step": {k: v[0] for k, v in prof.items()},
        "launch": "eager batched",
        "field_shape_autotuned": dict(zip(("shape", "candidate_ms"), batch.tuning())),
    }
    if not args.no_cpu_baseline:
        _log("cpu baseline")
        out["cpu_baseline"] = batched_cpu_baseline(
            w, max_elements=40 if args.workload == "stress" else None)
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent beams "
              "shard across GPUs without communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("batched-scan", "stress"):
        return main_batched(args)
    from hpsim_amd import HPSim
    from hpsim_amd import configs as cfg

    wl = cfg.config(0 if args.workload == "small" else 1, num=args.num, periods=args.periods)
    b = wl.beam
    elems = cfg.build_lattice(wl.lattice)
    dist = cfg.make_distribution(b)
    beam = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
    beam.set_distribution(*dist)
    beam.set_frequency(b.freq)
    beam.set_ref_w(b.w0)
    beam.save_initial_beam()
    bl = HPSim.BeamLine(elements=elems)
    sc = HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval)
    sim = HPSim.Simulator(beam, bl, sc)
    sim.set_graph(not args.eager)
    nel = len(elems)

    def step():
        beam.restore_initial_beam()
        sim.simulate()

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = time.perf_counter() - t0
    replays = sim.graph_replays()
    # one more step launched eagerly with HIP events around every kernel
    # launch (on the engine's stream) for the per-kernel times
    sim.set_profiling(True)
    step()
    prof = sim.profile()
    sim.set_profiling(False)

    pushes = b.num * nel * args.steps
    seg_ms, seg_n = prof["segment"]
    avg_s = seg_ms / 1e3 / max(seg_n, 1)
    kernel_ms = {k: v[0] for k, v in prof.items()}
    achieved = SEG_BYTES_PER_PARTICLE * b.num / avg_s / 1e9
    lost = beam.get_loss_num("t")
    out = {
        "metric": "particle-element pushes/s incl. space charge",
        "value": pushes / dt,
        "unit": "particle-element pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-seeded truncated Gaussian beam, generated lattice)",
        "config": {"workload": wl.name, "particles": b.num, "elements": nel,
                   "mesh": f"{wl.nr}x{wl.nz}", "sc_kicks_per_step": prof["phase_final"][1],
                   "transmitted": b.num - lost},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(b.num),
                     # beams of at most 512 chunks of 1024 particles run the
                     # one-particle-per-thread form of the segment kernel
                     "kernel": "k_segment_wide" if b.num <= 512 * 1024 else "k_segment",
                     "avg_launch_us": avg_s * 1e6,
                     "launches": seg_n, "bytes_per_launch": SEG_BYTES_PER_PARTICLE * b.num,
                     "timing": "HIP events around every launch of one eager step after "
                               "the timed steps"},
        "kernel_ms_per_step": kernel_ms,
        "graph_replays": replays,
        "launch": "eager" if args.eager else "hip-graph",
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_periods, args.cpu_serial_periods,
                                           0 if args.workload == "small" else 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
