"""Loads libhpsim_amd.so (the gfx950 engine) and declares its C ABI.

There is no fallback: if the library is missing or cannot be loaded the
import fails, because every tracking call must run on the GPU kernels.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhpsim_amd.so")

GAP_FIELDS = ("length", "frequency", "cell_length_over_beta_lambda", "amplitude",
              "phase_ref", "phase_shift", "energy_out", "beta_center", "dg", "t", "tp",
              "sp", "fit_beta_center", "fit_beta_min", "fit_t0", "fit_t1", "fit_t2",
              "fit_t3", "fit_t4", "fit_t5", "fit_s1", "fit_s2", "fit_s3", "fit_s4", "fit_s5")

KINDS = {"Drift": 1, "Quad": 2, "RFGap-DTL": 3, "RFGap-CCL": 4, "ApertureC": 5,
         "ApertureR": 6, "SpchComp": 7, "Diagnostics": 8, "Rotation": 9, "Steerer": 10,
         "Buncher": 11, "Dipole": 12}
# Buncher and Dipole model fields of hps_element (after the gap block)
ELEMENT_EXTRA = ("voltage", "frequency", "phase", "rho", "edge_angle_in", "edge_angle_out",
                 "half_gap", "k1", "k2", "field_index", "kinetic_energy")
KIND_NAMES = {v: k for k, v in KINDS.items()}


class GapParam(C.Structure):
    _fields_ = [(f, C.c_double) for f in GAP_FIELDS]


class Element(C.Structure):
    _fields_ = [("kind", C.c_int32), ("aperture_in", C.c_int32), ("length", C.c_double),
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


EXPORTS = (
    "hps_last_error", "hps_set_device",
    "hps_beam_create", "hps_beam_destroy", "hps_beam_set_distribution", "hps_beam_download",
    "hps_beam_set_waterbag", "hps_beam_set_dc", "hps_beam_sample_gaussian",
    "hps_beam_relative_phi", "hps_beam_size", "hps_beam_set_ref_w", "hps_beam_set_ref_phi",
    "hps_beam_set_frequency", "hps_beam_get_scalar", "hps_beam_get_loss_num",
    "hps_beam_get_avg", "hps_beam_get_sig", "hps_beam_get_emittance", "hps_beam_get_stats",
    "hps_beam_save", "hps_beam_restore", "hps_beam_translate", "hps_beam_apply_cut",
    "hps_beamline_create", "hps_beamline_destroy", "hps_beamline_add", "hps_beamline_set",
    "hps_beamline_size", "hps_beamline_index",
    "hps_spch_create", "hps_spch_destroy", "hps_spch_set", "hps_spch_set_mesh_size",
    "hps_spch_get_state", "hps_spch_get_tables",
    "hps_sim_create", "hps_sim_destroy", "hps_sim_simulate", "hps_sim_reset",
    "hps_sim_set_space_charge", "hps_sim_last_time_ms", "hps_sim_set_profiling",
    "hps_sim_get_profile", "hps_sim_set_graph", "hps_sim_graph_replays", "hps_probe_math",
    "hps_batch_create", "hps_batch_destroy", "hps_batch_add", "hps_batch_size",
    "hps_batch_simulate", "hps_batch_last_time_ms", "hps_batch_set_profiling",
    "hps_batch_get_profile", "hps_batch_set_autotune", "hps_batch_get_tuning",
    "hps_nn_create", "hps_nn_destroy", "hps_nn_num_params", "hps_nn_get_params",
    "hps_nn_set_params", "hps_nn_layers", "hps_nn_dims", "hps_nn_set_norm", "hps_nn_get_norm",
    "hps_nn_data_reserve", "hps_nn_data_append", "hps_nn_data_size", "hps_nn_data_get",
    "hps_nn_data_shuffle",
    "hps_nn_train", "hps_nn_eval", "hps_nn_predict", "hps_nn_loss_grad",
    "hps_batch_set_capture", "hps_batch_set_surrogate", "hps_batch_get_surrogate_info",
    "hps_batch_set_field_mode", "hps_batch_set_gemm_shape", "hps_batch_get_gemm_info",
    "hps_batch_set_graph", "hps_batch_graph_info",
    "hps_emittance_noalloc",
)

_lib = None


def _declare(L):
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint32)
    fp = C.POINTER(C.c_float)
    i32 = C.c_int32
    u32 = C.c_uint32
    d = C.c_double
    sig = {
        "hps_last_error": (C.c_char_p, []),
        "hps_set_device": (C.c_int, [C.c_int]),
        "hps_beam_create": (C.c_int, [PP, u32, d, d, d]),
        "hps_beam_destroy": (C.c_int, [P]),
        "hps_beam_set_distribution": (C.c_int, [P, dp, dp, dp, dp, dp, dp, up, up]),
        "hps_beam_download": (C.c_int, [P, dp, dp, dp, dp, dp, dp, up, up]),
        "hps_beam_set_waterbag": (C.c_int, [P] + [d] * 12 + [u32]),
        "hps_beam_set_dc": (C.c_int, [P] + [d] * 9 + [u32]),
        "hps_beam_sample_gaussian": (C.c_int, [P, u32, dp, dp]),
        "hps_beam_relative_phi": (C.c_int, [P, C.c_int, dp]),
        "hps_beam_size": (u32, [P]),
        "hps_beam_set_ref_w": (C.c_int, [P, d]),
        "hps_beam_set_ref_phi": (C.c_int, [P, d]),
        "hps_beam_set_frequency": (C.c_int, [P, d]),
        "hps_beam_get_scalar": (C.c_int, [P, C.c_int, dp]),
        "hps_beam_get_loss_num": (C.c_int, [P, C.c_int, up]),
        "hps_beam_get_avg": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "hps_beam_get_sig": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "hps_beam_get_emittance": (C.c_int, [P, C.c_int, dp]),
        "hps_beam_get_stats": (C.c_int, [P, C.POINTER(BeamStats)]),
        "hps_beam_save": (C.c_int, [P, C.c_int]),
        "hps_beam_restore": (C.c_int, [P, C.c_int]),
        "hps_beam_translate": (C.c_int, [P, C.c_int, d]),
        "hps_beam_apply_cut": (C.c_int, [P, C.c_char, d, d]),
        "hps_beamline_create": (C.c_int, [PP]),
        "hps_beamline_destroy": (C.c_int, [P]),
        "hps_beamline_add": (C.c_int, [P, C.c_char_p, C.POINTER(Element)]),
        "hps_beamline_set": (C.c_int, [P, u32, C.POINTER(Element)]),
        "hps_beamline_size": (u32, [P]),
        "hps_beamline_index": (C.c_int, [P, C.c_char_p, up]),
        "hps_spch_create": (C.c_int, [PP, u32, u32, d, u32]),
        "hps_spch_destroy": (C.c_int, [P]),
        "hps_spch_set": (C.c_int, [P, C.c_int, d]),
        "hps_spch_set_mesh_size": (C.c_int, [P, u32, u32]),
        "hps_spch_get_state": (C.c_int, [P, C.POINTER(SpchState)]),
        "hps_spch_get_tables": (C.c_int, [P, dp, dp, C.POINTER(C.c_int64)]),
        "hps_sim_create": (C.c_int, [PP, P, P, P]),
        "hps_sim_destroy": (C.c_int, [P]),
        "hps_sim_simulate": (C.c_int, [P, i32, i32]),
        "hps_sim_reset": (C.c_int, [P]),
        "hps_sim_set_space_charge": (C.c_int, [P, C.c_int]),
        "hps_sim_last_time_ms": (C.c_int, [P, dp]),
        "hps_sim_set_profiling": (C.c_int, [P, C.c_int]),
        "hps_sim_get_profile": (C.c_int, [P, dp, up]),
        "hps_sim_set_graph": (C.c_int, [P, C.c_int]),
        "hps_sim_graph_replays": (C.c_int, [P, up]),
        "hps_probe_math": (C.c_int, [C.c_int, dp, dp, dp, u32]),
        "hps_batch_create": (C.c_int, [PP]),
        "hps_batch_destroy": (C.c_int, [P]),
        "hps_batch_add": (C.c_int, [P, P]),
        "hps_batch_size": (u32, [P]),
        "hps_batch_simulate": (C.c_int, [P, i32, i32]),
        "hps_batch_last_time_ms": (C.c_int, [P, dp]),
        "hps_batch_set_profiling": (C.c_int, [P, C.c_int]),
        "hps_batch_get_profile": (C.c_int, [P, dp, up]),
        "hps_batch_set_autotune": (C.c_int, [P, C.c_int]),
        "hps_nn_create": (C.c_int, [PP, u32, u32, up, u32, u32, u32]),
        "hps_nn_destroy": (C.c_int, [P]),
        "hps_nn_num_params": (u32, [P]),
        "hps_nn_get_params": (C.c_int, [P, fp]),
        "hps_nn_set_params": (C.c_int, [P, fp]),
        "hps_nn_layers": (C.c_int, [P, C.POINTER(C.c_int32), u32, up]),
        "hps_nn_dims": (C.c_int, [P, up]),
        "hps_nn_set_norm": (C.c_int, [P, C.c_float, C.c_float, C.c_float]),
        "hps_nn_get_norm": (C.c_int, [P, fp]),
        "hps_nn_data_reserve": (C.c_int, [P, u32]),
        "hps_nn_data_append": (C.c_int, [P, fp, fp, fp, u32]),
        "hps_nn_data_size": (C.c_int, [P, up]),
        "hps_nn_data_get": (C.c_int, [P, u32, u32, fp, fp, fp]),
        "hps_nn_data_shuffle": (C.c_int, [P, u32]),
        "hps_nn_train": (C.c_int, [P, u32, u32, u32, u32, C.c_float, C.c_float, u32, fp]),
        "hps_nn_eval": (C.c_int, [P, u32, u32, fp]),
        "hps_nn_predict": (C.c_int, [P, fp, fp, u32, fp]),
        "hps_nn_loss_grad": (C.c_int, [P, fp, fp, fp, u32, fp, fp]),
        "hps_batch_set_capture": (C.c_int, [P, P]),
        "hps_batch_set_surrogate": (C.c_int, [P, P, u32, d]),
        "hps_batch_get_surrogate_info": (C.c_int, [P, dp]),
        "hps_batch_get_tuning": (C.c_int, [P, up, C.POINTER(C.c_float), u32]),
        "hps_batch_set_field_mode": (C.c_int, [P, C.c_int, d]),
        "hps_batch_set_graph": (C.c_int, [P, C.c_int]),
        "hps_batch_graph_info": (C.c_int, [P, C.POINTER(C.c_uint64), up]),
        "hps_batch_set_gemm_shape": (C.c_int, [P, u32]),
        "hps_batch_get_gemm_info": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_float),
                                             u32]),
        "hps_emittance_noalloc": (C.c_int, [dp] * 6 + [up, up, u32, d, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded engine library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m hpsim_amd.build` "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
        atexit.register(_at_exit)
    return _lib


def _at_exit():
    """Objects still alive when the interpreter shuts down (e.g. held by a
    traceback) are not destroyed through the engine: the process exit
    releases their device memory, and the HIP runtime may already be
    finishing."""
    global _lib
    _lib = None


class HPSimError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise HPSimError(lib().hps_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def uptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))
