#!/usr/bin/env python3
"""Benchmark: particle-element pushes/s including SCHEFF space charge.

Default workload: the combined stress configuration -- 64 beams of
4,194,304 particles (drawn on the device), the 400-period Drift/Quad/Drift/
DTL-gap lattice, SCHEFF on a 64x256 mesh after every element, all beams in
one batch whose fields come from the grouped SCHEFF solve on the FP32
matrix cores (Green's tables shared over aspect classes, the product of
every remeshing beam's charges in one launch sequence; its launch shape,
direct or in the frequency domain along z, chosen by timing).  One step =
one simulate() over the next 40 lattice periods (160 elements) of every
beam; the beams go on from where the previous step ended and start again
at the lattice entrance after the last period.  value = particles x
elements x steps / wall time of the timed steps.  The same run times the
learned-surrogate form (network fields audited against the grouped solve
every 20 kick points, fallback above 2 % relative field error) and reports
it under field_surrogate with its own accuracy.

The JSON line also carries
  roofline:      HBM roofline of the batched element-map kernel
                 (k_segment_batch), algorithmic bytes per launch / average
                 launch time from HIP events around every launch of one extra
                 step; roofline_gemm: the grouped field product's FP32 matrix-
                 core rate (its algorithmic flops per product column: 2 M K for
                 the direct product, the per-frequency product and transforms
                 for the spectral one);
  accuracy:      rms emittance, transmission and mean output energy of a few
                 beams of the workload tracked through the whole lattice with
                 the benchmarked field mode, against the same beams with the
                 FP64 per-beam SCHEFF solve (bit-identical to the FP64 oracle);
  plain_fp64:    the plain FP64 single-beam paths of the same box (1,048,576
                 particles x 800 elements and 65,536 x 160) and the ratio;
  cpu_baseline:  the FP64 oracle built with OpenMP (bit-identical to the
                 serial checker) on whole beams of the workload over their
                 first elements, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
GEMM_ASPECT_STEP = 0.05        # aspect grid of the grouped field solve
SURROGATE_AUDIT_EVERY = 20     # kick points between audits of the surrogate
SURROGATE_MAX_ERR = 0.02       # relative L2 field error that triggers the fallback
SEG_BYTES_PER_PARTICLE = 104   # k_segment: reads 6 FP64 + 2 u32, writes 6 FP64
SPECTRAL_SHAPE = 8             # product shape that runs in the frequency domain along z
DEFAULT_FIELD = "gemm"         # field form of the batched workloads' headline figure
MI355X_CUS = 256
MI355X_CLOCK_HZ = 2.4e9
# HBM bytes per segment-kernel launch, measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE counter runs (calibrated on a plain device copy), and the segment
# kernel's VALU instructions per launch, for the launch shapes they were
# measured at
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILES = {
    ("plain-fp64-1048576x200", 1048576):
        os.path.join(PROFILES, "k_segment_traffic_plain-fp64-1048576x200.md"),
    ("stress-64x4194304x400", 268435456):
        os.path.join(PROFILES, "k_segment_batch_traffic_stress-64x4194304x400.md"),
}
VALU_FILES = {
    ("stress-64x4194304x400", 268435456):
        os.path.join(PROFILES, "counters_stress-64x4194304x400.md"),
}


def _profile_value(files, key, field):
    path = files.get(key)
    if path is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(field + ":"):
            return int(line.split(":", 1)[1])
    return None


def measured_traffic(workload, num):
    """HBM bytes per segment-kernel launch from the committed counter summary
    of this workload and particle count; None when not measured."""
    return _profile_value(TRAFFIC_FILES, (workload, num), "traffic_bytes_per_launch")


def valu_issue(workload, num, avg_launch_s):
    """Fraction of the segment kernel's time its VALU instructions (committed
    counter summary) need at one wave64 instruction per CU per clock."""
    n = _profile_value(VALU_FILES, (workload, num), "valu_insts_per_launch")
    if n is None:
        return None
    return n / (MI355X_CUS * MI355X_CLOCK_HZ) / avg_launch_s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--num", type=int, default=None, help="override particle count")
    p.add_argument("--periods", type=int, default=None, help="override lattice periods")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--eager", action="store_true",
                   help="launch every kernel from the host instead of replaying HIP graphs "
                        "(for profilers that do not follow graph launches)")
    p.add_argument("--cpu-periods", type=int, default=200,
                   help="lattice periods of the all-core CPU-baseline sample (single beam)")
    p.add_argument("--cpu-serial-periods", type=int, default=10,
                   help="lattice periods of the one-core CPU-baseline sample (single beam)")
    p.add_argument("--workload", default="stress",
                   choices=("stress", "batched-scan", "plain", "small"),
                   help="stress: 64 beams of 4,194,304 particles, 400 periods, 64x256 mesh, in "
                        "one batch (default); batched-scan: 256 beams of 16K-1M particles with "
                        "per-beam lattice settings in one batch; plain: the 1,048,576-particle "
                        "single beam; small: the 65,536-particle single beam on 40 periods")
    p.add_argument("--compare-field", default=None, choices=("fp64", "gemm", "surrogate", "none"),
                   help="second field form timed in the same run (default: the other of gemm "
                        "and surrogate)")
    p.add_argument("--field", default=None, choices=("fp64", "gemm", "surrogate"),
                   help="batched field solve: fp64 (per-beam FP64 tables, bit-exact), gemm "
                        "(grouped FP32 matrix-core product, Green's tables shared over aspect "
                        "classes) or surrogate (the trained network, audited by gemm); default "
                        "surrogate")
    p.add_argument("--beams", type=int, default=None, help="beams of a batched workload")
    p.add_argument("--slice-periods", type=int, default=None,
                   help="lattice periods per step of a batched workload (stress: 40, scan: all)")
    p.add_argument("--validate-beams", type=int, default=4,
                   help="beams tracked through the whole lattice for the accuracy figures")
    p.add_argument("--no-plain", action="store_true",
                   help="skip the plain FP64 single-beam reference runs")
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


def _batch_workload(args):
    from hpsim_amd import configs as cfg
    if args.workload == "stress":
        return cfg.combined_stress(nbeams=args.beams or 64, num=args.num or 4194304,
                                   periods=args.periods or 400)
    return cfg.batched_scan(nbeams=args.beams or 256, periods=args.periods or 40)


def _make_batch(w, members, mode, net=None):
    """Batch of the workload's members with beams drawn on the device; the
    field mode and, for 'surrogate', the network attached."""
    from hpsim_amd import HPSim
    from hpsim_amd import surrogate as S
    sims = []
    for k in members:
        spec = w.beams[k]
        beam = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        beam.sample_gaussian(spec)
        beam.set_frequency(spec.freq)
        beam.set_ref_w(spec.w0)
        beam.save_initial_beam()
        sc = HPSim.SpaceCharge(nr=w.nr, nz=w.nz, interval=w.interval)
        sims.append(HPSim.Simulator(beam, HPSim.BeamLine(elements=w.lattices[k]), sc))
    batch = HPSim.Batch(sims)
    if mode in ("gemm", "surrogate"):
        batch.set_field_mode("gemm", GEMM_ASPECT_STEP)
    if mode == "surrogate":
        S.attach(batch, net, audit_every=SURROGATE_AUDIT_EVERY, max_err=SURROGATE_MAX_ERR)
    return batch, sims


def _accuracy(w, mode, net, nval):
    """Beams 0 .. nval-1 of the workload through the whole lattice with the
    field mode and with the FP64 per-beam solve; relative errors of the rms
    emittances, transmission and mean energy (max over the beams)."""
    from hpsim_amd.metrics import beam_metrics, metric_errors
    members = list(range(min(nval, len(w.beams))))
    out = {}
    runs = {}
    for m in (mode, "fp64"):
        batch, sims = _make_batch(w, members, m, net)
        batch.simulate()
        runs[m] = [beam_metrics(s.beam.arrays(), s.beam.get_mass()) for s in sims]
        if m == "surrogate":
            from hpsim_amd import surrogate as S
            out["surrogate_counters"] = S.info(batch)
        del batch, sims
    errs = [metric_errors(g, r) for g, r in zip(runs[mode], runs["fp64"])]
    out.update({k: max(e[k] for e in errs) for k in errs[0]})
    out["beams"] = len(members)
    out["reference"] = ("the same beams with the FP64 per-beam SCHEFF solve, whole lattice "
                        "(bit-identical to the FP64 oracle)")
    return out


def _plain_rates():
    """pushes/s of the plain FP64 single-beam path on this box: the
    1,048,576-particle 800-element and the 65,536-particle 160-element
    workloads, graph-launched, from the saved initial beam."""
    from hpsim_amd import HPSim
    from hpsim_amd import configs as cfg
    out = {}
    for idx, steps in ((1, 4), (0, 20)):
        wl = cfg.config(idx)
        b = wl.beam
        elems = cfg.build_lattice(wl.lattice)
        beam = HPSim.Beam(mass=b.mass, charge=b.charge, current=b.current, num=b.num)
        beam.sample_gaussian(b)
        beam.set_frequency(b.freq)
        beam.set_ref_w(b.w0)
        beam.save_initial_beam()
        sim = HPSim.Simulator(beam, HPSim.BeamLine(elements=elems),
                              HPSim.SpaceCharge(nr=wl.nr, nz=wl.nz, interval=wl.interval))
        for k in range(steps + 1):
            if k == 1:
                t0 = time.perf_counter()
            beam.restore_initial_beam()
            sim.simulate()
        dt = time.perf_counter() - t0
        out[wl.name] = b.num * len(elems) * steps / dt
    return out


def _load_net(w):
    from hpsim_amd import surrogate as S
    path = S.weights_path(w.nr, w.nz)
    if not os.path.exists(path):
        sys.exit(f"bench.py: no trained surrogate for the {w.nr}x{w.nz} mesh ({path}); "
                 f"train it with python -m hpsim_amd.surrogate --nr {w.nr} --nz {w.nz}, or "
                 "pass --field gemm")
    return S.load(path, max_batch=64)


def _timed_batch(w, mode, net, sel, steps, warmup, eager, profile):
    """All beams of the workload in one batch with the given field mode:
    warmup + timed steps (each the next `sel` elements of the lattice, back
    to the entrance after the last), then, with profile, one more step with
    HIP events around every launch.  Returns the figures of the run."""
    total = sum(b.num for b in w.beams)
    nel = len(w.lattices[0])
    _log(f"{len(w.beams)} beams, {total} particles, {nel} elements, field {mode}: "
         "drawing beams on the device")
    batch, sims = _make_batch(w, range(len(w.beams)), mode, net)
    batch.set_graph(not eager)
    nslice = nel // sel
    pos = [0]

    def step():
        if pos[0] == 0:
            for m in sims:
                m.beam.restore_initial_beam()
        batch.simulate(pos[0] * sel, pos[0] * sel + sel - 1)
        pos[0] = (pos[0] + 1) % nslice

    for k in range(warmup):
        step()
        _log(f"{mode}: warm-up step {k + 1} done")
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        _log(f"{mode}: timed step {k + 1} done")
    dt = time.perf_counter() - t0
    r = {"value": total * sel * steps / dt, "ms_per_step": dt / steps * 1e3, "steps": steps,
         "warmup": warmup}
    r["graph"] = batch.graph_info()
    if profile:
        gi0 = batch.gemm_info() if mode != "fp64" else None
        batch.set_profiling(True)
        step()
        r["prof"] = batch.profile()
        batch.set_profiling(False)
        r["gemm"] = (gi0, batch.gemm_info()) if mode != "fp64" else None
    r["transmitted"] = sum(m.beam.get_size() - m.beam.get_loss_num("t") for m in sims)
    if profile:
        # untimed census: particles not lost (loss == 0) at every slice boundary
        # of one pass through the lattice from the initial beams; the timed
        # steps ran slices (warmup + j) mod nslice
        def alive():
            return sum(m.beam.get_size() - m.beam.get_loss_num("t") for m in sims)
        ends = []
        for k in range(nslice):
            if k == 0:
                for m in sims:
                    m.beam.restore_initial_beam()
            ends.append(alive())
            batch.simulate(k * sel, k * sel + sel - 1)
        ends.append(alive())
        mean_alive = sum(0.5 * (ends[q] + ends[q + 1])
                         for q in ((warmup + j) % nslice for j in range(steps))) / steps
        r["alive"] = {"value": mean_alive * sel * steps / dt,
                      "mean_alive_fraction": mean_alive / total,
                      "alive_at_slice_ends": ends}
        _log(f"{mode}: alive census done")
    if mode == "surrogate":
        from hpsim_amd import surrogate as S
        r["surrogate_counters"] = S.info(batch)
    del batch, sims
    return r


def main_batched(args):
    w = _batch_workload(args)
    mode = args.field or DEFAULT_FIELD
    other = args.compare_field
    if other is None:
        other = "surrogate" if mode == "gemm" else ("gemm" if mode == "surrogate" else "none")
    net = _load_net(w) if "surrogate" in (mode, other) else None
    total = sum(b.num for b in w.beams)
    nel = len(w.lattices[0])
    slice_periods = args.slice_periods or (40 if args.workload == "stress" else nel // 4)
    sel = 4 * slice_periods
    if nel % sel:
        sys.exit(f"bench.py: {nel} elements are not a whole number of {slice_periods}-period "
                 "slices")
    run = _timed_batch(w, mode, net, sel, args.steps, args.warmup, args.eager, True)
    dt = run["ms_per_step"] * args.steps / 1e3
    replays, per_period = run["graph"]
    prof = run["prof"]
    gi0, gi1 = run["gemm"] or (None, None)
    seg_ms, seg_n = prof["segment"]
    avg_s = seg_ms / 1e3 / max(seg_n, 1)
    achieved = SEG_BYTES_PER_PARTICLE * total / avg_s / 1e9
    transmitted = run["transmitted"]
    out = {
        "metric": "particle-element pushes/s incl. space charge",
        "value": total * sel * args.steps / dt,
        "unit": "particle-element pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "dtype_note": ("particles, maps, moments and kicks in FP64 (divisions by refined "
                       "reciprocals, square roots from the reciprocal square root, shorter "
                       "sin/cos/sinh/cosh polynomials); the grouped field product in FP32 "
                       "(FP32 operands and accumulation), its field tables stored FP64"
                       if mode in ("gemm", "surrogate") else "FP64 throughout"),
        "data": ("synthetic (Philox-seeded truncated Gaussian beams drawn on the device, "
                 "generated lattice" +
                 (")" if args.workload == "stress" else " with seeded per-beam perturbations)")),
        "config": {"workload": w.name, "beams": len(w.beams), "particles": total,
                   "particles_min": min(b.num for b in w.beams),
                   "particles_max": max(b.num for b in w.beams), "elements": nel,
                   "elements_per_step": sel, "mesh": f"{w.nr}x{w.nz}", "field": mode,
                   "transmitted_after_profiled_step": transmitted},
        "alive_weighted": {
            "value": run["alive"]["value"],
            "unit": "particle-element pushes/s of particles not lost",
            "mean_alive_fraction": run["alive"]["mean_alive_fraction"],
            "definition": ("the timed steps' mean count of particles with loss == 0 (per "
                           "step the mean of its slice's two ends, counted in an untimed "
                           "pass through the lattice from the same initial beams) x "
                           "elements per step / wall time")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(w.name, total),
                     "valu_issue_frac": valu_issue(w.name, total, avg_s),
                     "kernel": "k_segment_batch", "avg_launch_us": avg_s * 1e6,
                     "launches": seg_n, "bytes_per_launch": SEG_BYTES_PER_PARTICLE * total,
                     "timing": "HIP events around every launch of one extra (eager) step"},
        "kernel_ms_per_step": {k: v[0] for k, v in prof.items()},
        "kernel_ms_timing": ("HIP events around every launch of one extra eager step; each "
                             "launch is bracketed on its own, so the families sum to more "
                             "than a graph-launched step (field_product is part of field)"),
        "launch": "eager" if args.eager else "lattice-period HIP graphs",
        "period_graph_replays": replays,
        "launches_per_replayed_period": per_period,
    }
    if mode != "fp64":
        prod_ms, prod_n = prof["field_product"]
        cols = gi1["product_columns"] - gi0["product_columns"]
        direct = 2.0 * (2 * (w.nr + 1) * (w.nz + 1)) * (w.nr * w.nz) * cols
        spectral = gi1["shape"] == SPECTRAL_SHAPE
        if spectral:
            # per column: the per-frequency complex product as a real product of
            # 2 R x 2 nr (R = 2 (nr + 1) field rows) at nz + 1 frequencies, plus the
            # radix-2 transforms (nr / 2 forward, nr + 1 inverse, 5 L log2 L each)
            ln = 2 * w.nz
            fft = 5.0 * ln * math.log2(ln) * (w.nr // 2 + w.nr + 1)
            flops = (2.0 * (4 * (w.nr + 1)) * (2 * w.nr) * (w.nz + 1) + fft) * cols
        else:
            flops = direct
        rate = flops / (prod_ms / 1e3) / 1e12 if prod_ms > 0 else None
        out["roofline_gemm"] = {
            "bound": "mfma",
            "kernel": ("k_fft_product (per-frequency product on FP32 MFMA 16x16x4) with the "
                       "k_fft_* transforms" if spectral else "k_gemm_field (FP32 MFMA 16x16x4)"),
            "achieved": rate, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": rate / FP32_MFMA_PEAK_TFLOPS if rate else None,
            "flops_per_column": flops / max(cols, 1),
            "direct_product_flops_per_column": direct / max(cols, 1),
            "direct_equivalent_tflops": direct / (prod_ms / 1e3) / 1e12 if prod_ms > 0 else None,
            "launches": prod_n, "product_columns": cols,
            "shape": gi1["shape"], "shape_candidate_ms": gi1["times_ms"],
            "class_tables_built": gi1["tables_built"]}
    if "surrogate_counters" in run:
        out["surrogate_counters"] = run["surrogate_counters"]
    if args.validate_beams > 0:
        _log("accuracy run")
        out["accuracy"] = _accuracy(w, mode, net, args.validate_beams)
    if other != "none":
        # the other field form on the same box: a shorter timed run and its
        # accuracy against the same FP64 reference
        o = _timed_batch(w, other, net, sel, min(args.steps, 3), 1, args.eager, False)
        cmp = {"value": o["value"], "ms_per_step": o["ms_per_step"], "steps": o["steps"],
               "warmup": o["warmup"], "ratio_to_value": o["value"] / out["value"]}
        if "surrogate_counters" in o:
            cmp["surrogate_counters"] = o["surrogate_counters"]
        if args.validate_beams > 0:
            _log(f"accuracy run ({other})")
            cmp["accuracy"] = _accuracy(w, other, net, args.validate_beams)
        out["field_" + other] = cmp
    if not args.no_plain:
        _log("plain FP64 single-beam runs")
        plain = _plain_rates()
        out["plain_fp64"] = {k: v for k, v in plain.items()}
        out["plain_fp64"]["ratio_to_" + "plain-fp64-1048576x200"] = (
            out["value"] / plain["plain-fp64-1048576x200"])
        out["plain_fp64"]["ratio_to_" + "plain-fp64-65536x40"] = (
            out["value"] / plain["plain-fp64-65536x40"])
    if not args.no_cpu_baseline:
        _log("cpu baseline")
        out["cpu_baseline"] = batched_cpu_baseline(
            w, max_elements=40 if args.workload == "stress" else None)
        out["cpu_baseline"]["build"] = "gcc -O2 -march=x86-64-v3 -ffp-contract=off -fopenmp"
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent beams "
              "would shard across GPUs by beam (one batch per GPU, no communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in ("batched-scan", "stress"):
        return main_batched(args)
    if args.field:
        sys.exit("bench.py: --field applies to the batched workloads")
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(wl.name, b.num),
                     # beams of at most 512 chunks of 1024 particles run the
                     # one-particle-per-thread form of the segment kernel
                     "kernel": "k_segment_wide" if b.num <= 512 * 1024 else "k_segment",
                     "avg_launch_us": avg_s * 1e6,
                     "launches": seg_n, "bytes_per_launch": SEG_BYTES_PER_PARTICLE * b.num,
                     "timing": "HIP events around every launch of one eager step after "
                               "the timed steps"},
        "kernel_ms_per_step": kernel_ms,
        "kernel_ms_timing": ("HIP events around every launch of one extra eager step; each "
                             "launch is bracketed on its own, so the families sum to more "
                             "than a graph-launched step"),
        "graph_replays": replays,
        "launch": "eager" if args.eager else "hip-graph",
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_periods, args.cpu_serial_periods,
                                           0 if args.workload == "small" else 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
