"""ctypes binding of the parity oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (built by oracle/Makefile).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, as
the checker or the timed CPU baseline; the hpsim_amd product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_LIB_OMP = os.path.join(_HERE, "liboracle_omp.so")
# HPSIM_ORACLE_SANITIZED=1 loads the ASan/UBSan build instead of the serial
# library (the process must have the ASan runtime preloaded)
_LIB_SAN = os.path.join(_HERE, "liboracle_san.so")
_use_omp = False

EL_TYPES = {"Drift": 1, "Quad": 2, "RFGap-DTL": 3, "RFGap-CCL": 4, "ApertureC": 5,
            "ApertureR": 6, "SpchComp": 7, "Diagnostics": 8, "Rotation": 9, "Steerer": 10,
            "Buncher": 11, "Dipole": 12}

GAP_FIELDS = ("length", "frequency", "cell_length_over_beta_lambda", "amplitude",
              "phase_ref", "phase_shift", "energy_out", "beta_center", "dg", "t", "tp",
              "sp", "fit_beta_center", "fit_beta_min", "fit_t0", "fit_t1", "fit_t2",
              "fit_t3", "fit_t4", "fit_t5", "fit_s1", "fit_s2", "fit_s3", "fit_s4", "fit_s5")


class GapParam(C.Structure):
    _fields_ = [(f, C.c_double) for f in GAP_FIELDS]


# Buncher (MV, MHz, rad) and Dipole model fields after the gap block; the
# Dipole bend angle is ``angle``
ELEMENT_EXTRA = ("voltage", "frequency", "phase", "rho", "edge_angle_in", "edge_angle_out",
                 "half_gap", "k1", "k2", "field_index", "kinetic_energy")


class OrcElement(C.Structure):
    _fields_ = [("type", C.c_int32), ("aperture_in", C.c_int32), ("length", C.c_double),
                ("aperture", C.c_double), ("gradient", C.c_double), ("angle", C.c_double),
                ("bl_h", C.c_double), ("bl_v", C.c_double), ("aper_xl", C.c_double),
                ("aper_xr", C.c_double), ("aper_yt", C.c_double), ("aper_yb", C.c_double),
                ("fraction", C.c_double), ("gap", GapParam)] + [
                    (f, C.c_double) for f in ELEMENT_EXTRA] + [
                    ("on", C.c_int32), ("pad_", C.c_int32)]


STAT_FIELDS = ("x_avg", "x_avg_good", "xp_avg", "y_avg", "y_avg_good", "yp_avg", "phi_avg",
               "phi_avg_r", "phi_avg_good", "w_avg", "w_avg_good", "x_sig", "x_sig_good",
               "xp_sig", "y_sig", "y_sig_good", "yp_sig", "phi_sig", "phi_sig_r",
               "phi_sig_good", "w_sig", "w_sig_good", "x_emit", "y_emit", "z_emit")


class BeamStats(C.Structure):
    _fields_ = [(f, C.c_double) for f in STAT_FIELDS] + [
        ("num_loss", C.c_uint32), ("num_lloss", C.c_uint32), ("num_good", C.c_uint32),
        ("pad_", C.c_uint32)]


class SpchState(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("nz", C.c_uint32), ("nx_init", C.c_uint32),
                ("nz_init", C.c_uint32), ("adj_bunch", C.c_uint32), ("good_cnt_old", C.c_uint32),
                ("fraction", C.c_double), ("interval", C.c_double), ("interval_init", C.c_double),
                ("adj_bunch_w", C.c_double), ("mesh_w", C.c_double), ("sigr_old", C.c_double),
                ("sigz_old", C.c_double), ("gm_old", C.c_double), ("remesh_val", C.c_double),
                ("dr", C.c_double), ("dz", C.c_double), ("num_on_mesh", C.c_uint32),
                ("remesh_count", C.c_uint32), ("kick_count", C.c_uint32), ("pad_", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def use_openmp():
    """Route every later call to liboracle_omp.so (OpenMP loops on all
    OMP_NUM_THREADS cores, bit-identical results).  Must come before the
    first oracle object is created in this process."""
    global _use_omp
    if _lib is not None and not _use_omp:
        raise RuntimeError("the serial oracle library is already in use in this process")
    _use_omp = True


def lib():
    global _lib
    if _lib is None:
        path = _LIB_OMP if _use_omp else _LIB
        if os.environ.get("HPSIM_ORACLE_SANITIZED") == "1":
            if _use_omp:
                raise RuntimeError("the sanitized oracle build is serial only")
            path = _LIB_SAN
            if not os.path.exists(path):
                subprocess.run(["make", "-s", "-C", _HERE, "san"], check=True)
        elif not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        L.orc_last_error.restype = C.c_char_p
        L.orc_threads.restype = C.c_int
        L.orc_beam_create.restype = P
        L.orc_beam_create.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double]
        L.orc_beam_destroy.argtypes = [P]
        L.orc_beam_set_distribution.argtypes = [P, dp, dp, dp, dp, dp, dp, up, up]
        for f in ("orc_beam_set_ref_w", "orc_beam_set_ref_phi", "orc_beam_set_frequency"):
            getattr(L, f).argtypes = [P, C.c_double]
        L.orc_beam_get_frequency.restype = C.c_double
        L.orc_beam_get_frequency.argtypes = [P]
        L.orc_beam_get_design_w.restype = C.c_double
        L.orc_beam_get_design_w.argtypes = [P]
        L.orc_beam_get_arrays.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, up, up]
        L.orc_beam_get_stats.argtypes = [P, C.POINTER(BeamStats)]
        L.orc_beam_set_num_loss.argtypes = [P, C.c_uint32]
        L.orc_beam_loss_num.restype = C.c_uint32
        L.orc_beam_loss_num.argtypes = [P, C.c_int]
        L.orc_beam_emittance.restype = C.c_double
        L.orc_beam_emittance.argtypes = [P, C.c_int]
        L.orc_beam_avg.restype = C.c_double
        L.orc_beam_avg.argtypes = [P, C.c_int, C.c_int]
        L.orc_beam_sig.restype = C.c_double
        L.orc_beam_sig.argtypes = [P, C.c_int, C.c_int]
        for f in ("orc_beam_update_loss", "orc_beam_update_stat_for_space_charge",
                  "orc_beam_save_initial", "orc_beam_restore_initial",
                  "orc_beam_save_intermediate", "orc_beam_restore_intermediate"):
            getattr(L, f).argtypes = [P]
        L.orc_beam_translate.argtypes = [P, C.c_int, C.c_double]
        L.orc_beam_set_waterbag.restype = C.c_int
        L.orc_beam_set_waterbag.argtypes = [P] + [C.c_double] * 12 + [C.c_uint32]
        L.orc_beam_set_dc.restype = C.c_int
        L.orc_beam_set_dc.argtypes = [P] + [C.c_double] * 9 + [C.c_uint32]
        L.orc_beam_apply_cut.argtypes = [P, C.c_char, C.c_double, C.c_double]
        L.orc_spch_create.restype = P
        L.orc_spch_create.argtypes = [C.c_uint32, C.c_uint32, C.c_int]
        L.orc_spch_destroy.argtypes = [P]
        for f in ("orc_spch_set_interval", "orc_spch_set_remesh_threshold",
                  "orc_spch_set_adj_bunch_cutoff_w", "orc_spch_set_mesh_size_cutoff_w"):
            getattr(L, f).argtypes = [P, C.c_double]
        L.orc_spch_set_adj_bunch.argtypes = [P, C.c_uint32]
        L.orc_spch_set_fp32_table.argtypes = [P, C.c_int]
        L.orc_spch_set_aspect_interp.argtypes = [P, C.c_double, C.c_int]
        L.orc_spch_get_state.argtypes = [P, C.POINTER(SpchState)]
        L.orc_spch_get_tables.argtypes = [P, dp, dp, C.POINTER(C.c_int64)]
        L.orc_spch_start.restype = C.c_int
        L.orc_spch_start.argtypes = [P, P, C.c_double]
        L.orc_sim_create.restype = P
        L.orc_sim_create.argtypes = [P, C.POINTER(OrcElement), C.c_uint32, P]
        L.orc_sim_destroy.argtypes = [P]
        L.orc_sim_simulate.restype = C.c_int
        L.orc_sim_simulate.argtypes = [P, C.c_int32, C.c_int32]
        L.orc_sim_reset.argtypes = [P]
        L.orc_sim_set_space_charge.argtypes = [P, C.c_int]
        for f in ("orc_t_exp", "orc_t_log", "orc_t_sinh", "orc_t_cosh", "orc_t_bessi0",
                  "orc_t_bessi1", "orc_t_bessi1p"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        L.orc_t_sincos.argtypes = [C.c_double, dp, dp]
        L.orc_t_tan.restype = C.c_double
        L.orc_t_tan.argtypes = [C.c_double]
        L.orc_t_fmod_pos.restype = C.c_double
        L.orc_t_fmod_pos.argtypes = [C.c_double, C.c_double]
        L.orc_t_ttf.argtypes = [C.c_double, C.POINTER(GapParam), dp, dp, dp]
        L.orc_t_canon_sum.restype = C.c_double
        L.orc_t_canon_sum.argtypes = [dp, C.c_uint32]
        L.orc_t_green_table.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_uint32, C.c_double,
                                        C.c_uint32, dp]
        L.orc_t_offmesh_field.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint32, dp, dp]
        L.orc_t_field.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_int64), dp, dp]
        L.orc_t_kick.argtypes = [P, P, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_uint32]
        L.orc_t_nextpow2.restype = C.c_uint32
        L.orc_t_nextpow2.argtypes = [C.c_uint32]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def _err():
    return lib().orc_last_error().decode()


def to_orc_elements(elements):
    arr = (OrcElement * max(1, len(elements)))()
    for i, e in enumerate(elements):
        o = arr[i]
        o.type = EL_TYPES[e.kind]
        o.aperture_in = 1 if e.aperture_in else 0
        for f in ("length", "aperture", "gradient", "angle", "bl_h", "bl_v", "aper_xl",
                  "aper_xr", "aper_yt", "aper_yb", "fraction"):
            setattr(o, f, float(getattr(e, f)))
        if e.gap:
            for f in GAP_FIELDS:
                setattr(o.gap, f, float(e.gap.get(f, 0.0)))
        params = getattr(e, "params", None) or {}
        for f in ELEMENT_EXTRA:
            setattr(o, f, float(params.get(f, 0.0)))
        o.on = 1 if params.get("on", False) else 0
    return arr


class Beam:
    """Oracle beam with the getter semantics of the reference Python type."""

    def __init__(self, num, mass, charge, current):
        self.n = int(num)
        self.h = lib().orc_beam_create(self.n, mass, charge, current)
        if not self.h:
            raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_beam_destroy(self.h)
            self.h = None

    def set_distribution(self, x, xp, y, yp, phi, w, loss=None, lloss=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, xp, y, yp, phi, w)]
        lo = None if loss is None else np.ascontiguousarray(loss, dtype=np.uint32)
        ll = None if lloss is None else np.ascontiguousarray(lloss, dtype=np.uint32)
        lib().orc_beam_set_distribution(self.h, *[_dp(a) for a in arrs], _up(lo), _up(ll))

    def set_waterbag(self, ax, bx, ex, ay, by, ey, az, bz, ez, phi, w, freq, seed=0):
        if lib().orc_beam_set_waterbag(self.h, ax, bx, ex, ay, by, ey, az, bz, ez, phi, w,
                                       freq, seed):
            raise MemoryError("set_waterbag")

    def set_dc(self, ax, bx, ex, ay, by, ey, dphi, sync_phi, sync_w, seed=0):
        if lib().orc_beam_set_dc(self.h, ax, bx, ex, ay, by, ey, dphi, sync_phi, sync_w, seed):
            raise MemoryError("set_dc")

    def set_ref_w(self, v): lib().orc_beam_set_ref_w(self.h, v)
    def set_ref_phi(self, v): lib().orc_beam_set_ref_phi(self.h, v)
    def set_frequency(self, v): lib().orc_beam_set_frequency(self.h, v)
    def get_frequency(self): return lib().orc_beam_get_frequency(self.h)
    def get_design_w(self): return lib().orc_beam_get_design_w(self.h)

    def arrays(self):
        n = self.n
        out = {k: np.empty(n) for k in ("x", "xp", "y", "yp", "phi", "phi_r", "w")}
        loss = np.empty(n, np.uint32)
        lloss = np.empty(n, np.uint32)
        lib().orc_beam_get_arrays(self.h, *[_dp(out[k]) for k in ("x", "xp", "y", "yp", "phi",
                                                                    "phi_r", "w")],
                                  _up(loss), _up(lloss))
        out["loss"] = loss
        out["lloss"] = lloss
        return out

    def stats(self):
        s = BeamStats()
        lib().orc_beam_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in BeamStats._fields_ if f != "pad_"}

    def set_num_loss(self, n): lib().orc_beam_set_num_loss(self.h, n)
    def get_loss_num(self, opt="t"): return lib().orc_beam_loss_num(self.h, {"t": 0, "l": 1, "all": 2}[opt])
    def get_emittance(self, axis): return lib().orc_beam_emittance(self.h, {"x": 0, "y": 1, "z": 2}[axis])
    def get_avg(self, var, good=False): return lib().orc_beam_avg(self.h, {"x": 0, "y": 1, "phi": 2, "w": 3, "phi_r": 4}[var], int(good))
    def get_sig(self, var, good=False): return lib().orc_beam_sig(self.h, {"x": 0, "y": 1, "phi": 2, "w": 3, "phi_r": 4}[var], int(good))
    def update_loss(self): lib().orc_beam_update_loss(self.h)
    def update_stat_for_space_charge(self): lib().orc_beam_update_stat_for_space_charge(self.h)
    def save_initial_beam(self): lib().orc_beam_save_initial(self.h)
    def restore_initial_beam(self): lib().orc_beam_restore_initial(self.h)
    def save_intermediate_beam(self): lib().orc_beam_save_intermediate(self.h)
    def restore_intermediate_beam(self): lib().orc_beam_restore_intermediate(self.h)

    def translate(self, axis, val):
        lib().orc_beam_translate(self.h, ["x", "xp", "y", "yp", "phi", "w"].index(axis), val)

    def apply_cut(self, coord, lo, hi): lib().orc_beam_apply_cut(self.h, coord.encode(), lo, hi)


class SpaceCharge:
    def __init__(self, nr, nz, interval=0.01, adj_bunch=0):
        self.h = lib().orc_spch_create(nr, nz, adj_bunch)
        if not self.h:
            raise RuntimeError(_err())
        lib().orc_spch_set_interval(self.h, interval)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_spch_destroy(self.h)
            self.h = None

    def set_remesh_threshold(self, v): lib().orc_spch_set_remesh_threshold(self.h, v)
    def set_adj_bunch(self, v): lib().orc_spch_set_adj_bunch(self.h, int(v))
    def set_adj_bunch_cutoff_w(self, v): lib().orc_spch_set_adj_bunch_cutoff_w(self.h, v)
    def set_mesh_size_cutoff_w(self, v): lib().orc_spch_set_mesh_size_cutoff_w(self.h, v)

    def set_fp32_table(self, on=True): lib().orc_spch_set_fp32_table(self.h, 1 if on else 0)

    def set_aspect_interp(self, step, order=1):
        lib().orc_spch_set_aspect_interp(self.h, step, order)

    def state(self):
        s = SpchState()
        lib().orc_spch_get_state(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in SpchState._fields_ if f != "pad_"}

    def tables(self):
        st = self.state()
        nr, nz = st["nx"], st["nz"]
        f1 = np.empty((nr + 1) * nr * nz * 2)
        f2 = np.empty((nz + 1) * (nr + 1) * 2)
        bins = np.empty(nr * nz, np.int64)
        lib().orc_spch_get_tables(self.h, _dp(f1), _dp(f2), bins.ctypes.data_as(C.POINTER(C.c_int64)))
        return f1, f2, bins

    def start(self, beam, length):
        if lib().orc_spch_start(self.h, beam.h, length):
            raise RuntimeError(_err())


class Simulator:
    def __init__(self, beam, elements, spch=None):
        self.beam = beam
        self.spch = spch
        self.elements = list(elements)
        self._names = [e.name.lower() for e in self.elements]
        self._arr = to_orc_elements(self.elements)
        self.h = lib().orc_sim_create(beam.h, self._arr, len(self.elements),
                                      spch.h if spch is not None else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def index(self, name):
        return self._names.index(name.lower())

    def simulate(self, start=None, end=None):
        s = -1 if start is None else (start if isinstance(start, int) else self.index(start))
        e = -1 if end is None else (end if isinstance(end, int) else self.index(end))
        if lib().orc_sim_simulate(self.h, s, e):
            raise RuntimeError(_err())

    def reset(self): lib().orc_sim_reset(self.h)
    def set_space_charge(self, on): lib().orc_sim_set_space_charge(self.h, 1 if on in (True, "on") else 0)


def run_workload(wl, num=None, periods=None, space_charge=None):
    """Build the oracle objects for a configs.Workload and run simulate()."""
    from hpsim_amd import configs as cfg  # host-side workload description only
    bspec = wl.beam
    elems = cfg.build_lattice(wl.lattice)
    beam = Beam(bspec.num, bspec.mass, bspec.charge, bspec.current)
    beam.set_distribution(*cfg.make_distribution(bspec))
    beam.set_frequency(bspec.freq)
    beam.set_ref_w(bspec.w0)
    sc = SpaceCharge(wl.nr, wl.nz, wl.interval) if (wl.space_charge if space_charge is None else space_charge) else None
    sim = Simulator(beam, elems, sc)
    sim.simulate()
    return beam, sc, sim
