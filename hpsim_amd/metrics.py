"""Output figures of a tracked beam for the accuracy statements of the
approximate field modes (grouped FP32 product, surrogate): rms emittances as
Beam::UpdateEmittance defines them (beam.cpp:985-994, beam_kernel.cu:
1333-1375), transmission and mean output energy, from downloaded arrays."""
from __future__ import annotations

import numpy as np


def beam_metrics(a, mass):
    """Output figures of a final beam (arrays as Beam.arrays()): rms
    emittances x, y, z as Beam::UpdateEmittance defines them (particles with
    loss == 0; beta*gamma of the reference particle's w[0]; the phase
    relative to phi[0], wrapped), transmission (loss == 0 fraction) and the
    mean energy of the transmitted particles.  Two diagnostics beside the
    reference's figures: ez_core, the same longitudinal emittance over the
    particles not flagged as longitudinally lost (lloss == 0, which drops the
    far-out wrapped phases that dominate ez), and the longitudinally lost
    fraction of the transmitted particles."""
    import math
    sel = a["loss"] == 0
    core = sel & (a["lloss"] == 0)
    g = a["w"][0] / mass
    bg = math.sqrt(g * (g + 2.0))

    def emit(u, v, s=sel):
        u, v = u[s], v[s]
        du, dv = u - u.mean(), v - v.mean()
        return math.sqrt(max((du * du).mean() * (dv * dv).mean() - (du * dv).mean() ** 2, 0.0))

    phr = np.mod(a["phi"] - a["phi"][0] + math.pi, 2 * math.pi) - math.pi
    return {"ex": emit(a["x"], a["xp"]) * bg, "ey": emit(a["y"], a["yp"]) * bg,
            "ez": emit(phr, a["w"]), "transmission": float(sel.mean()),
            "w_avg": float(a["w"][sel].mean()), "ez_core": emit(phr, a["w"], core),
            "lloss": float((sel & (a["lloss"] != 0)).sum() / max(int(sel.sum()), 1))}


def metric_errors(got, ref):
    """Relative errors of beam_metrics(got) against beam_metrics(ref); the
    transmission error is the difference of the transmitted fractions."""
    out = {}
    for k in ("ex", "ey", "ez", "w_avg", "ez_core"):
        if k in got and k in ref:
            out[k] = abs(got[k] / ref[k] - 1.0)
    out["transmission"] = abs(got["transmission"] - ref["transmission"])
    if "lloss" in got and "lloss" in ref:
        out["lloss"] = abs(got["lloss"] - ref["lloss"])
    return out
