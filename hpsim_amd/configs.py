"""Synthetic beams and lattices for the benchmark workloads.

The reference drives HPSim from LANSCE SQLite lattices and user-made beams
(pytest/sim-lbeg.py:35-75).  The benchmark workloads are synthetic instead:

* beam: H- macro-particles, 6-D Gaussian truncated at 3 sigma per
  coordinate, drawn with a counter-based Philox4x32-10 generator
  (Salmon et al., SC'11) so every configuration is reproducible from its seed;
* lattice: periods of Drift / Quad / Drift / DTL RF gap whose gap design
  tables (phase shift, output energy, centre velocity) come from tracking the
  synchronous particle through the same maps the engine applies.

Everything here is plain host-side Python producing element descriptions and
coordinate arrays; it does not depend on the GPU library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

CLIGHT = 299.792458          # m * MHz, reference constant.h:7
H_MINUS_MASS = 939.294       # MeV
FREQ = 201.25                # MHz

# ---------------------------------------------------------------------------
# Philox4x32-10
# ---------------------------------------------------------------------------
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Philox4x32 block function on uint32 arrays (vectorised)."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32).copy() for c in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(rounds):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & _MASK32).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & _MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + _W0)
            k1 = np.uint32(k1 + _W1)
    return c0, c1, c2, c3


def _u53(a, b):
    """Two uint32 words -> uniform double in [0, 1) with 53 random bits."""
    return ((a >> np.uint32(5)).astype(np.float64) * 67108864.0
            + (b >> np.uint32(6)).astype(np.float64)) / 9007199254740992.0


def truncated_normal(n, coord, seed, cut=3.0):
    """n standard normals truncated to |z| <= cut; rejection per value with the
    attempt number in the Philox counter (particle i, coordinate, attempt)."""
    idx = np.arange(n, dtype=np.uint64)
    out = np.empty(n, dtype=np.float64)
    todo = np.arange(n)
    attempt = 0
    while todo.size:
        i = idx[todo]
        r0, r1, r2, r3 = philox4x32((i & _MASK32).astype(np.uint32),
                                    (i >> np.uint64(32)).astype(np.uint32),
                                    np.full(todo.size, coord, np.uint32),
                                    np.full(todo.size, attempt, np.uint32),
                                    seed & 0xFFFFFFFF, 0x48505331)
        u1 = 1.0 - _u53(r0, r1)          # (0, 1]
        u2 = _u53(r2, r3)
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
        ok = np.abs(z) <= cut
        out[todo[ok]] = z[ok]
        todo = todo[~ok]
        attempt += 1
        if attempt > 64:
            raise RuntimeError("truncated normal rejection did not converge")
    return out


def uniform(n, stream, seed):
    """n uniforms in [0, 1) from Philox counter (i, stream)."""
    i = np.arange(n, dtype=np.uint64)
    r0, r1, _, _ = philox4x32((i & _MASK32).astype(np.uint32), (i >> np.uint64(32)).astype(np.uint32),
                              np.full(n, stream, np.uint32), np.zeros(n, np.uint32),
                              seed & 0xFFFFFFFF, 0x554E4946)
    return _u53(r0, r1)


@dataclass
class BeamSpec:
    num: int
    mass: float = H_MINUS_MASS
    charge: float = -1.0
    current: float = 0.010          # A
    freq: float = FREQ
    w0: float = 0.75                # MeV
    phi0: float = 0.0               # rad, synchronous (reference) phase at entry
    sig_x: float = 1e-3             # m
    sig_xp: float = 1e-3            # rad
    sig_y: float = 1e-3
    sig_yp: float = 1e-3
    sig_phi: float = math.radians(10.0)
    sig_w: float = 2e-3             # MeV
    seed: int = 1


def make_distribution(spec: BeamSpec):
    """Seeded truncated-Gaussian 6-D distribution in SI/engine units
    (m, rad, rad, MeV).  Particle 0 is the reference particle on axis at the
    synchronous phase and energy (reference beam.cpp:342-345)."""
    n = spec.num
    x = spec.sig_x * truncated_normal(n, 0, spec.seed)
    xp = spec.sig_xp * truncated_normal(n, 1, spec.seed)
    y = spec.sig_y * truncated_normal(n, 2, spec.seed)
    yp = spec.sig_yp * truncated_normal(n, 3, spec.seed)
    phi = spec.phi0 + spec.sig_phi * truncated_normal(n, 4, spec.seed)
    w = spec.w0 + spec.sig_w * truncated_normal(n, 5, spec.seed)
    x[0] = xp[0] = y[0] = yp[0] = 0.0
    phi[0] = spec.phi0
    w[0] = spec.w0
    return x, xp, y, yp, phi, w


# ---------------------------------------------------------------------------
# Lattice description
# ---------------------------------------------------------------------------
GAP_FIELDS = ("length", "frequency", "cell_length_over_beta_lambda", "amplitude",
              "phase_ref", "phase_shift", "energy_out", "beta_center", "dg", "t", "tp",
              "sp", "fit_beta_center", "fit_beta_min", "fit_t0", "fit_t1", "fit_t2",
              "fit_t3", "fit_t4", "fit_t5", "fit_s1", "fit_s2", "fit_s3", "fit_s4", "fit_s5")


@dataclass
class Element:
    """One beamline element (reference beamline_element.h).  ``kind`` is the
    reference type string: Drift, Quad, RFGap-DTL, RFGap-CCL, ApertureC,
    ApertureR, SpchComp, Diagnostics, Rotation, Steerer, Buncher, Dipole.
    Buncher and Dipole model fields are in ``params``."""
    name: str
    kind: str
    length: float = 0.0
    aperture: float = 0.0
    gradient: float = 0.0
    angle: float = 0.0
    bl_h: float = 0.0
    bl_v: float = 0.0
    aper_xl: float = 0.0
    aper_xr: float = 0.0
    aper_yt: float = 0.0
    aper_yb: float = 0.0
    aperture_in: bool = False
    fraction: float = 1.0
    gap: dict = field(default_factory=dict)
    params: dict = field(default_factory=dict)   # Buncher / Dipole model fields


def _beta(w, mass):
    g = w / mass
    return math.sqrt((g + 2.0) * g) / (g + 1.0)


def _ttf(beta, gp):
    """Transit-time factor polynomial, reference simulation_kernel.cu:290-308."""
    b = beta
    if beta > 1.1 * gp["fit_beta_center"]:
        b = gp["fit_beta_center"]
    elif beta < gp["fit_beta_min"]:
        b = gp["fit_beta_min"]
    return gp["fit_t0"] + b * (gp["fit_t1"] + b * (gp["fit_t2"] + b * (gp["fit_t3"] + b * (
        gp["fit_t4"] + b * gp["fit_t5"]))))


@dataclass
class LatticeSpec:
    periods: int = 40
    drift_len: float = 0.05
    quad_len: float = 0.04
    gradient: float = 30.0             # T/m, alternating sign by period
    v_start: float = 0.3               # MV gap voltage (amplitude * gap length)
    v_end: float = 0.3
    phi_s_deg: float = -30.0
    aperture: float = 0.01
    scale_gradient_with_bg: bool = False
    ttf_seed: int = 7
    freq: float = FREQ
    mass: float = H_MINUS_MASS
    charge: float = -1.0
    w0: float = 0.75
    phi0: float = 0.0


def build_lattice(spec: LatticeSpec):
    """Drift / Quad / Drift / DTL-gap periods with design tables from
    synchronous-particle tracking (x = x' = 0, the reference particle).

    Builder conventions (the reference lattice always comes from a database):
    the gap is one 2*pi-mode DTL cell, length beta_c * lambda with
    cell_length_over_beta_lambda = 1 and beta_c the design velocity entering
    the gap; amplitude = V / length; the TTF is a seeded cubic in beta around
    0.8 (t4 = t5 = s* = 0); phase_shift puts the synchronous particle at
    phi_s in every gap; energy_out is the synchronous output energy."""
    lam = CLIGHT / spec.freq
    q = abs(spec.charge)
    phi_s = math.radians(spec.phi_s_deg)
    rng = np.random.default_rng(spec.ttf_seed)
    w = spec.w0
    phi = spec.phi0
    bg0 = None
    elems = []
    twopi = 2.0 * math.pi
    for p in range(spec.periods):
        frac = p / (spec.periods - 1) if spec.periods > 1 else 0.0
        volt = spec.v_start + (spec.v_end - spec.v_start) * frac
        beta = _beta(w, spec.mass)
        # drift 1 (two half steps, as the engine does)
        phi += 2 * (twopi * (0.5 * spec.drift_len) / (lam * beta))
        elems.append(Element(f"D{p}a", "Drift", length=spec.drift_len, aperture=spec.aperture))
        # quad
        g = spec.gradient * (1 if p % 2 == 0 else -1)
        if spec.scale_gradient_with_bg:
            # thin-lens FODO: sin(mu/2) = L Lq q G c / (2 beta gamma m c^2), so
            # the phase advance holds for G proportional to beta*gamma / L
            bg = math.sqrt((w / spec.mass) * (w / spec.mass + 2.0))
            period = 2 * spec.drift_len + spec.quad_len + beta * lam
            if bg0 is None:
                bg0 = (bg, period)
            g *= (bg / bg0[0]) * (bg0[1] / period)
        phi += 2 * (twopi * spec.quad_len * 0.5 / (lam * beta))
        elems.append(Element(f"Q{p}", "Quad", length=spec.quad_len, aperture=spec.aperture,
                             gradient=g))
        phi += 2 * (twopi * (0.5 * spec.drift_len) / (lam * beta))
        elems.append(Element(f"D{p}b", "Drift", length=spec.drift_len, aperture=spec.aperture))
        # gap
        beta_c = beta
        glen = beta_c * lam
        t0 = 0.8 + 0.02 * (rng.random() - 0.5)
        gp = {k: 0.0 for k in GAP_FIELDS}
        gp.update(length=glen, frequency=spec.freq, cell_length_over_beta_lambda=1.0,
                  amplitude=volt / glen, phase_ref=phi_s, beta_center=beta_c,
                  fit_beta_center=beta_c, fit_beta_min=0.5 * beta_c,
                  fit_t0=t0, fit_t1=0.05 * (rng.random() - 0.5),
                  fit_t2=0.2 * (rng.random() - 0.5), fit_t3=0.5 * (rng.random() - 0.5))
        # first half: phase advance of the synchronous particle (dps = 0)
        phi += (1.0 - (beta - beta) / beta) * (math.pi * 1.0 + 0.0)
        shift = phi_s - phi
        shift = (shift + math.pi) % twopi - math.pi
        gp["phase_shift"] = shift
        t = _ttf(beta, gp)
        w_out = w + q * (gp["amplitude"] * glen) * (t * 1.0 * math.cos(phi_s))
        gp["energy_out"] = w_out
        beta_new = _beta(w_out, spec.mass)
        phi += (1.0 - (beta_new - beta_new) / beta_new) * (math.pi * 1.0 + 0.0)
        w = w_out
        elems.append(Element(f"G{p}", "RFGap-DTL", length=glen, gap=gp))
    return elems


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    beam: BeamSpec
    lattice: LatticeSpec
    nr: int = 32
    nz: int = 128
    interval: float = 0.05          # m: one SCHEFF kick per element
    space_charge: bool = True


def config(idx: int, num: int | None = None, periods: int | None = None) -> Workload:
    """Single-beam configurations 0 (65,536 particles, 40 periods) and 1
    (1,048,576 particles, 200 periods, 0.3 -> 0.6 MV ramp, gradients scaled
    with beta*gamma).  ``num`` / ``periods`` override sizes for tests."""
    if idx == 0:
        b = BeamSpec(num=num or 65536)
        lat = LatticeSpec(periods=periods or 40)
        return Workload("plain-fp64-65536x40", b, lat)
    if idx == 1:
        b = BeamSpec(num=num or 1048576)
        lat = LatticeSpec(periods=periods or 200, v_start=0.3, v_end=0.6,
                          scale_gradient_with_bg=True)
        return Workload("plain-fp64-1048576x200", b, lat)
    raise ValueError(f"unknown single-beam configuration {idx}")


@dataclass
class BatchWorkload:
    """Batched scan: one lattice layout, per-beam settings and sizes."""
    name: str
    beams: list            # BeamSpec per member
    lattices: list         # element list per member
    nr: int = 32
    nz: int = 128
    interval: float = 0.05
    space_charge: bool = True


def perturb_lattice(elems, seed, grad=0.05, volt=0.02, phase_deg=3.0):
    """A copy of elems with every quad gradient scaled by U(1 - grad, 1 + grad),
    every gap amplitude by U(1 - volt, 1 + volt) and every gap RF phase
    (phase_shift) moved by U(-phase_deg, phase_deg) degrees, from Philox
    uniforms of stream 16 + k for the k-th setting of each element."""
    import copy
    n = len(elems)
    u = uniform(3 * n, 16, seed).reshape(n, 3)
    out = []
    for i, e in enumerate(elems):
        e = copy.deepcopy(e)
        if e.kind == "Quad":
            e.gradient *= 1.0 + grad * (2.0 * u[i, 0] - 1.0)
        elif e.kind.startswith("RFGap"):
            e.gap = dict(e.gap)
            e.gap["amplitude"] *= 1.0 + volt * (2.0 * u[i, 1] - 1.0)
            e.gap["phase_shift"] += math.radians(phase_deg) * (2.0 * u[i, 2] - 1.0)
        out.append(e)
    return out


def batched_scan(nbeams=256, nmin=16384, nmax=1048576, periods=40, seed=7):
    """The batched-scan configuration: nbeams beams of the 65,536-particle
    configuration's distribution law with log-uniform seeded particle counts
    in [nmin, nmax], each on the 40-period lattice with its own seeded
    perturbations (quad gradients +-5 %, gap voltages +-2 %, RF phases
    +-3 degrees)."""
    base = config(0, periods=periods)
    lattice = build_lattice(base.lattice)
    u = uniform(nbeams, 15, seed)
    nums = np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * u).astype(np.int64)
    beams, lats = [], []
    for k in range(nbeams):
        b = BeamSpec(num=int(nums[k]), seed=seed * 100003 + k + 1)
        beams.append(b)
        lats.append(perturb_lattice(lattice, seed * 7919 + k + 1))
    return BatchWorkload(f"batched-scan-{nbeams}x{periods}", beams, lats)


def combined_stress(nbeams=64, num=4194304, periods=400, seed=11):
    """The combined-stress configuration: nbeams beams of num particles on the
    65,536-particle configuration's lattice extended to `periods` periods,
    SCHEFF on a 64 x 256 mesh; beams differ by their seeds."""
    base = config(0, periods=periods)
    lattice = build_lattice(base.lattice)
    beams = [BeamSpec(num=num, seed=seed * 100003 + k + 1) for k in range(nbeams)]
    return BatchWorkload(f"stress-{nbeams}x{num}x{periods}", beams, [lattice] * nbeams,
                         nr=64, nz=256)
