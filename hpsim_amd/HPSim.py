"""Drop-in for the reference CPython extension module ``HPSim``.

Same types (Beam, BeamLine, SpaceCharge, Simulator), method names, keyword
arguments, option strings and engine units (m, rad, MeV, MHz, A) as
src/pywrapper/src/wrap_*.cpp; every call goes through the C ABI of
include/hpsim_amd.h into the gfx950 engine.  Differences, all deliberate:
errors raise Python exceptions (the reference prints to stderr and returns
NULL/0), coordinate arrays may be numpy arrays as well as lists, and a
BeamLine can be filled programmatically (``BeamLine(elements=...)`` or
``add_element``) besides from a database connection.
"""
from __future__ import annotations

import ctypes as C
import math
import re

import numpy as np

from . import _native as N
from ._native import check, dptr, uptr
from .db import DBConnection, beamline_elements  # noqa: F401  (the reference module type)

_COORD = ("x", "xp", "y", "yp", "phi", "w")


def _lib():
    return N.lib()


_NUM_RE = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")
_ONLY_NUM_RE = re.compile(r"^[\s\d.eE+-]+$")


def write_beam_file(path, msg, charge, mass, current, freq, arrays):
    """The text form of Beam::PrintToFile (beam.cpp:670-707): an info line,
    the charge, mass, current and frequency lines, then one row per particle
    of x, xp, y, yp, phi, w, loss, lloss (15 significant digits)."""
    with open(path, "w") as f:
        f.write(f"Info {msg}\nCharge  {charge:g}\nMass {mass:g}\n"
                f"Current {current:g}\nFrequency {freq:g}\n")
        n = len(arrays["x"])
        for i in range(n):
            f.write("  ".join(f"{arrays[k][i]:.15g}" for k in _COORD)
                    + f"  {arrays['loss'][i]}  {arrays['lloss'][i]}\n")


def read_beam_file(path):
    """Parse a beam file as Beam::InitBeamFromFile / UpdateBeamFromFile do
    (beam.cpp:179-298): a line naming charge, mass, current or frequency
    (case-insensitive) and holding a number sets that value (its first
    number); every line of numbers only is a particle (x, xp, y, yp, phi, w
    [, loss [, lloss]]).  Particle 0 is the reference particle and its
    transverse coordinates are set to 0 (beam.cpp:289).  Returns (values,
    six coordinate arrays, loss or None, lloss or None)."""
    vals = {}
    rows = []
    with open(path) as f:
        for line in f:
            low = line.lower()
            m = _NUM_RE.search(line)
            for key in ("charge", "mass", "current", "frequency"):
                if key in low and m:
                    vals[key] = float(m.group(0))
            if line.strip() and _ONLY_NUM_RE.match(line):
                rows.append(line.split())
    for key in ("charge", "mass", "current", "frequency"):
        if key not in vals:
            raise ValueError(f"Beam input file error: must define {key} value!")
    if not rows:
        raise ValueError("Beam input file error: no particle rows")
    ncol = len(rows[-1])
    cols = [np.array([float(r[k]) for r in rows]) for k in range(6)]
    for k in range(4):
        cols[k][0] = 0.0
    loss = np.array([int(float(r[6])) for r in rows], np.uint32) if ncol > 6 else None
    lloss = np.array([int(float(r[7])) for r in rows], np.uint32) if ncol > 7 else None
    return vals, cols, loss, lloss


class Beam:
    """Beam(mass=, charge=, current=, num=) or Beam(file=), wrap_beam.cpp:22-50."""

    def __init__(self, mass=None, charge=None, current=0.0, num=0, file=None):
        self._h = None
        if file:
            self._init_from_file(file)
            return
        if not num:
            raise ValueError("Beam constructor needs either a file name or num of particles")
        self._create(int(num), float(mass), float(charge), float(current))

    def _create(self, num, mass, charge, current):
        h = C.c_void_p()
        check(_lib().hps_beam_create(C.byref(h), num, mass, charge, current))
        self._h = h
        self._n = num

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_beam_destroy(h)
            self._h = None

    # -- initialisation -------------------------------------------------
    def _init_from_file(self, path):
        """Beam::InitBeamFromFile, beam.cpp:179-298."""
        vals, cols, loss, lloss = read_beam_file(path)
        self._create(len(cols[0]), vals["mass"], vals["charge"], vals["current"])
        self.set_distribution(*cols, loss, lloss)
        check(_lib().hps_beam_set_ref_w(self._h, float(cols[5][0])))
        self.set_frequency(vals["frequency"])

    def set_distribution(self, x, xp, y, yp, phi, w, loss=None, lloss=None):
        """set_distribution(x, xp, y, yp, phi, w[, loss[, lloss]]), wrap_beam.cpp:61-118."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, xp, y, yp, phi, w)]
        if any(a.shape != (self._n,) for a in arrs):
            raise ValueError("Beam.set_distribution: provided distribution size doesn't match "
                             "number of particles allocated")
        lo = None if loss is None else np.ascontiguousarray(loss, dtype=np.uint32)
        # the reference reads lloss from the loss list (wrap_beam.cpp:111); the
        # lloss argument is honoured here
        ll = None if lloss is None else np.ascontiguousarray(lloss, dtype=np.uint32)
        check(_lib().hps_beam_set_distribution(self._h, *[dptr(a) for a in arrs], uptr(lo), uptr(ll)))

    def set_waterbag(self, ax, bx, ex, ay, by, ey, az, bz, ez, phi, w, freq, seed=0):
        """set_waterbag(alpha_x, beta_x, emittance_x, ..., phi, w, frequency[, seed]),
        wrap_beam.cpp:125-144: PARMILA Twiss input (beta in cm, emittance in
        cm*rad), sampled with the C library's srand/rand (seed 0 = time)."""
        check(_lib().hps_beam_set_waterbag(self._h, *[float(v) for v in (
            ax, bx, ex, ay, by, ey, az, bz, ez, phi, w, freq)], int(seed)))

    def set_dc(self, ax, bx, ex, ay, by, ey, dphi, sync_phi, sync_w, seed=0):
        """set_dc(alpha_x, beta_x, emittance_x, alpha_y, beta_y, emittance_y,
        width_phi, synchronous_phi, synchronous_w[, seed]), wrap_beam.cpp:146-165."""
        check(_lib().hps_beam_set_dc(self._h, *[float(v) for v in (
            ax, bx, ex, ay, by, ey, dphi, sync_phi, sync_w)], int(seed)))

    # -- save / restore / output ---------------------------------------------
    def sample_gaussian(self, spec):
        """Draw the truncated-Gaussian beam of a configs.BeamSpec on the device
        (the distribution configs.make_distribution gives, without a host
        upload)."""
        import numpy as np
        center = np.array([0.0, 0.0, 0.0, 0.0, spec.phi0, spec.w0])
        sigma = np.array([spec.sig_x, spec.sig_xp, spec.sig_y, spec.sig_yp, spec.sig_phi,
                          spec.sig_w])
        check(_lib().hps_beam_sample_gaussian(self._h, spec.seed & 0xFFFFFFFF, dptr(center),
                                              dptr(sigma)))

    def save_initial_beam(self): check(_lib().hps_beam_save(self._h, 0))
    def save_intermediate_beam(self): check(_lib().hps_beam_save(self._h, 1))
    def restore_initial_beam(self): check(_lib().hps_beam_restore(self._h, 0))
    def restore_intermediate_beam(self): check(_lib().hps_beam_restore(self._h, 1))

    def print_to(self, output_file, msg=""):
        """Beam::PrintToFile, beam.cpp:670-707."""
        self.get_loss_num()
        write_beam_file(output_file, msg, self.get_charge(), self.get_mass(), self._scalar(5),
                        self.get_frequency(), self._download())

    def print_simple(self):
        a = self._download()
        for i in (0, self._n // 2 - 1, self._n - 2):
            if 0 <= i < self._n:
                print("\t".join(f"{a[k][i]:.15f}" for k in _COORD) + f"\t{a['loss'][i]}\t{a['lloss'][i]}")

    # -- reference values ------------------------------------------------------
    def set_ref_w(self, w): check(_lib().hps_beam_set_ref_w(self._h, float(w)))
    def set_ref_phi(self, phi): check(_lib().hps_beam_set_ref_phi(self._h, float(phi)))

    def set_frequency(self, freq):
        """First call sets the frequency, later calls rescale phases (wrap_beam.cpp:315-341)."""
        check(_lib().hps_beam_set_frequency(self._h, float(freq)))

    def _scalar(self, which):
        v = C.c_double()
        check(_lib().hps_beam_get_scalar(self._h, which, C.byref(v)))
        return v.value

    def get_mass(self): return self._scalar(0)
    def get_charge(self): return self._scalar(1)
    def get_frequency(self): return self._scalar(2)
    def get_ref_w(self): return self._scalar(3)
    def get_ref_phi(self): return self._scalar(4)
    def get_current(self): return self._scalar(7)
    def get_design_w(self): return self._scalar(6)
    def get_size(self): return self._n

    # -- coordinates ------------------------------------------------------------
    def _download(self):
        out = {k: np.empty(self._n) for k in _COORD}
        out["loss"] = np.empty(self._n, np.uint32)
        out["lloss"] = np.empty(self._n, np.uint32)
        check(_lib().hps_beam_download(self._h, *[dptr(out[k]) for k in _COORD],
                                       uptr(out["loss"]), uptr(out["lloss"])))
        return out

    def arrays(self):
        """All coordinates and loss flags as numpy arrays (no filtering)."""
        return self._download()

    def _coord(self, name, option):
        if option not in ("good", "lost", "all"):
            raise ValueError(f"invalid option: {option}, pick from (\"good\", \"lost\", \"all\")")
        self.get_loss_num("t")    # UpdateLoss, as the reference getters do
        a = self._download()
        v = a[name]
        if option == "all":
            return v
        return v[a["loss"] == 0] if option == "good" else v[a["loss"] != 0]

    def get_x(self, option="good"): return self._coord("x", option)
    def get_xp(self, option="good"): return self._coord("xp", option)
    def get_y(self, option="good"): return self._coord("y", option)
    def get_yp(self, option="good"): return self._coord("yp", option)
    def get_w(self, option="good"): return self._coord("w", option)

    def get_phi(self, option="good", option2="absolute"):
        """get_phi(opt, 'absolute'|'relative'), wrap_beam.cpp:712-798."""
        if option2 not in ("absolute", "relative"):
            raise ValueError(f"invalid 2nd option: {option2}")
        if option2 == "absolute":
            return self._coord("phi", option)
        if option not in ("good", "lost", "all"):
            raise ValueError(f"invalid option: {option}")
        self.get_loss_num("t")
        rel = np.empty(self._n)
        check(_lib().hps_beam_relative_phi(self._h, 0, dptr(rel)))
        if option == "all":
            return rel
        loss = self._download()["loss"]
        return rel[loss == 0] if option == "good" else rel[loss != 0]

    def get_losses(self, option="t"):
        """get_losses('t'|'l'), wrap_beam.cpp:868-912."""
        if option not in ("t", "l"):
            raise ValueError(f"invalid option: {option}, pick from (\"t\", \"l\")")
        self.get_loss_num(option)
        a = self._download()
        return a["loss"] if option == "t" else a["lloss"]

    def get_loss_num(self, option="t"):
        """get_loss_num('t'|'l'|'all'), wrap_beam.cpp:1245-1284."""
        which = {"t": 0, "l": 1, "all": 2}.get(option)
        if which is None:
            raise ValueError(f"invalid option: {option}, pick from (\"t\", \"l\", \"all\")")
        v = C.c_uint32()
        check(_lib().hps_beam_get_loss_num(self._h, which, C.byref(v)))
        return v.value

    # -- statistics ---------------------------------------------------------------
    def _avg(self, var, good):
        v = C.c_double()
        check(_lib().hps_beam_get_avg(self._h, var, int(good), C.byref(v)))
        return v.value

    def _sig(self, var, good):
        v = C.c_double()
        check(_lib().hps_beam_get_sig(self._h, var, int(good), C.byref(v)))
        return v.value

    @staticmethod
    def _good(option):
        if option not in ("transverse_good", "longitudinal_good"):
            raise ValueError("option must be \"transverse_good\" or \"longitudinal_good\"")
        return option == "longitudinal_good"

    def get_avg_x(self, option="transverse_good"): return self._avg(0, self._good(option))
    def get_avg_y(self, option="transverse_good"): return self._avg(1, self._good(option))
    def get_avg_w(self, option="transverse_good"): return self._avg(3, self._good(option))
    def get_sig_x(self, option="transverse_good"): return self._sig(0, self._good(option))
    def get_sig_y(self, option="transverse_good"): return self._sig(1, self._good(option))
    def get_sig_w(self, option="transverse_good"): return self._sig(3, self._good(option))

    def get_avg_phi(self, option="absolute"):
        """wrap_beam.cpp:978-1017: 'absolute' | 'relative' | 'longitudinal_good'."""
        if option == "absolute":
            return self._avg(2, False)
        if option == "relative":
            return self._avg(4, False)
        return self._avg(2, True)

    def get_sig_phi(self, option="absolute"):
        """wrap_beam.cpp:1116-1157: 'absolute' | 'relative' | 'longitudinal_good'."""
        if option == "absolute":
            return self._sig(2, False)
        if option == "relative":
            return self._sig(4, False)
        return self._sig(2, True)

    def _emit(self, axis):
        v = C.c_double()
        check(_lib().hps_beam_get_emittance(self._h, axis, C.byref(v)))
        return v.value

    def get_emittance_x(self): return self._emit(0)
    def get_emittance_y(self): return self._emit(1)
    def get_emittance_z(self): return self._emit(2)

    def stats(self):
        s = N.BeamStats()
        check(_lib().hps_beam_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.BeamStats._fields_ if f != "pad_"}

    # -- manipulation ---------------------------------------------------------------
    def apply_cut(self, coord, lo, hi):
        """apply_cut(coord, min, max), wrap_beam.cpp:1353-1371."""
        check(_lib().hps_beam_apply_cut(self._h, str(coord)[0].encode(), float(lo), float(hi)))

    def translate(self, axis, val):
        """translate(axis, val), wrap_beam.cpp:1317-1351 (axis x, xp, y, yp, phi, w)."""
        a = str(axis).lower()
        if a not in _COORD:
            return
        check(_lib().hps_beam_translate(self._h, _COORD.index(a), float(val)))


def _to_element(e) -> N.Element:
    o = N.Element()
    o.kind = N.KINDS[e.kind]
    o.aperture_in = 1 if getattr(e, "aperture_in", False) else 0
    for f in ("length", "aperture", "gradient", "angle", "bl_h", "bl_v", "aper_xl", "aper_xr",
              "aper_yt", "aper_yb", "fraction"):
        setattr(o, f, float(getattr(e, f, 0.0)))
    gap = getattr(e, "gap", None) or {}
    for f in N.GAP_FIELDS:
        setattr(o.gap, f, float(gap.get(f, 0.0)))
    params = getattr(e, "params", None) or {}
    for f in N.ELEMENT_EXTRA:
        setattr(o, f, float(params.get(f, 0.0)))
    o.on = 1 if params.get("on", False) else 0
    return o


class BeamLine:
    """BeamLine(DBConnection), wrap_beamline.cpp:24-42: the lattice of the
    connected databases (hpsim_amd.db, init.cpp:20-429).  A BeamLine can
    also be filled from element descriptions (hpsim_amd.configs.Element or
    any object with the same attributes) or with add_element()."""

    def __init__(self, dbconnection=None, elements=None):
        if dbconnection is not None:
            if elements is not None:
                raise TypeError("BeamLine takes a DBConnection or elements, not both")
            elements = beamline_elements(dbconnection)
        elements = list(elements or ())
        bad = sorted({e.kind for e in elements if e.kind not in N.KINDS})
        if bad:
            raise NotImplementedError(f"element kinds without an engine map: {', '.join(bad)}")
        h = C.c_void_p()
        check(_lib().hps_beamline_create(C.byref(h)))
        self._h = h
        self._elements = []
        for e in elements:
            self.add_element(e)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_beamline_destroy(h)
            self._h = None

    def add_element(self, e):
        el = _to_element(e)
        check(_lib().hps_beamline_add(self._h, e.name.encode(), C.byref(el)))
        self._elements.append(e)

    def update_element(self, index_or_name, e):
        """Replace one element's model parameters (the role of the reference's
        PV -> model path, pv_observer.cpp)."""
        i = index_or_name if isinstance(index_or_name, int) else self.index(index_or_name)
        el = _to_element(e)
        check(_lib().hps_beamline_set(self._h, i, C.byref(el)))
        self._elements[i] = e

    def __len__(self):
        return len(self._elements)

    def __getitem__(self, i):
        return self._elements[i if isinstance(i, int) else self.index(i)]

    def index(self, name):
        v = C.c_uint32()
        check(_lib().hps_beamline_index(self._h, str(name).encode(), C.byref(v)))
        return v.value

    def get_element_names(self, start="", end="", elem_type=""):
        i0 = self.index(start) if start else 0
        i1 = self.index(end) if end else len(self._elements) - 1
        return [e.name for e in self._elements[i0:i1 + 1]
                if not elem_type or elem_type.lower() in e.kind.lower()]

    def print_out(self):
        for e in self._elements:
            print(f"{e.name}: {e.kind}, length = {e.length}, aper = {e.aperture}")

    def print_range(self, start, end):
        for e in self._elements[self.index(start):self.index(end) + 1]:
            print(f"{e.name}: {e.kind}, length = {e.length}, aper = {e.aperture}")


class SpaceCharge:
    """SpaceCharge(nr, nz, interval=0.01, adj_bunch=0, type='scheff'),
    wrap_spacecharge.cpp:31-56."""

    def __init__(self, nr=32, nz=64, interval=0.01, adj_bunch=0, type="scheff"):
        if type != "scheff":
            raise ValueError(f"SpaceCharge type: {type} is not supported.")
        h = C.c_void_p()
        check(_lib().hps_spch_create(C.byref(h), int(nr), int(nz), float(interval), int(adj_bunch)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_spch_destroy(h)
            self._h = None

    def state(self):
        s = N.SpchState()
        check(_lib().hps_spch_get_state(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.SpchState._fields_ if f != "pad_"}

    def tables(self):
        st = self.state()
        nr, nz = st["nx"], st["nz"]
        f1 = np.empty((nr + 1) * nr * nz * 2)
        f2 = np.empty((nz + 1) * (nr + 1) * 2)
        bins = np.empty(nr * nz, np.int64)
        check(_lib().hps_spch_get_tables(self._h, dptr(f1), dptr(f2),
                                         bins.ctypes.data_as(C.POINTER(C.c_int64))))
        return f1, f2, bins

    def _set(self, what, v): check(_lib().hps_spch_set(self._h, what, float(v)))
    def set_interval(self, v): self._set(0, v)
    def get_interval(self): return self.state()["interval"]
    def set_remesh_threshold(self, v): self._set(1, v)
    def get_remesh_threshold(self): return self.state()["remesh_val"]
    def set_adj_bunch(self, v): self._set(2, int(v))
    def get_adj_bunch(self): return float(self.state()["adj_bunch"])
    def set_adj_bunch_cutoff_w(self, v): self._set(3, v)
    def get_adj_bunch_cutoff_w(self): return self.state()["adj_bunch_w"]
    def set_mesh_size_cutoff_w(self, v): self._set(4, v)
    def get_mesh_size_cutoff_w(self): return self.state()["mesh_w"]

    def set_mesh_size(self, nr, nz):
        check(_lib().hps_spch_set_mesh_size(self._h, int(nr), int(nz)))

    def get_mesh_size(self):
        s = self.state()
        return [float(s["nx"]), float(s["nz"])]


class Simulator:
    """Simulator(beam, beamline, spch=None), wrap_simulator.cpp:27-54.  Unlike
    the reference it keeps references to its beam, beamline and space charge."""

    def __init__(self, beam, beamline, spch=None):
        if not isinstance(beam, Beam) or not isinstance(beamline, BeamLine):
            raise TypeError("Simulator needs a Beam and a BeamLine")
        h = C.c_void_p()
        check(_lib().hps_sim_create(C.byref(h), beam._h, beamline._h,
                                    spch._h if isinstance(spch, SpaceCharge) else None))
        self._h = h
        self.beam, self.beamline, self.spch = beam, beamline, spch

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_sim_destroy(h)
            self._h = None

    def simulate(self, *args):
        """simulate() or simulate(start_element, end_element), inclusive."""
        if len(args) == 0:
            i0, i1 = -1, -1
        elif len(args) == 2:
            i0, i1 = (a if isinstance(a, int) else self.beamline.index(a) for a in args)
        else:
            raise TypeError("simulate takes either 0 or 2 arguments")
        check(_lib().hps_sim_simulate(self._h, i0, i1))

    def reset(self):
        check(_lib().hps_sim_reset(self._h))

    def set_space_charge(self, on_off):
        check(_lib().hps_sim_set_space_charge(self._h, 1 if str(on_off).lower() == "on" else 0))

    def set_graph(self, on=True):
        """Launch simulate() as a cached HIP graph (default) or eagerly."""
        check(_lib().hps_sim_set_graph(self._h, 1 if on else 0))

    def graph_replays(self):
        v = C.c_uint32()
        check(_lib().hps_sim_graph_replays(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, on=True):
        """Time every kernel launch of later simulate() calls with HIP events."""
        check(_lib().hps_sim_set_profiling(self._h, 1 if on else 0))

    def profile(self):
        """{family: (total ms, launches)} since profiling was enabled."""
        names = ("segment", "phase", "sc_build", "field", "segment_final", "phase_final",
                 "field_product")
        ms = (C.c_double * len(names))()
        n = (C.c_uint32 * len(names))()
        check(_lib().hps_sim_get_profile(self._h, ms, n))
        return {names[k]: (ms[k], n[k]) for k in range(len(names))}

    def last_time_ms(self):
        v = C.c_double()
        check(_lib().hps_sim_last_time_ms(self._h, C.byref(v)))
        return v.value


class Batch:
    """Several Simulators run as one launch sequence (hps_batch_*): the PV
    scan of pytest/sim-lbeg-scan-pv.py:140-160, one simulate per machine
    setting, with every kernel launch covering all settings.  Results are
    those of separate Simulator.simulate calls.  The simulators' lattices
    must share one layout (element values may differ) and one SCHEFF mesh
    size."""

    _FAMILIES = ("segment", "phase", "sc_build", "field", "segment_final", "phase_final",
                 "field_product")

    def __init__(self, simulators=()):
        h = C.c_void_p()
        check(_lib().hps_batch_create(C.byref(h)))
        self._h = h
        self.simulators = []
        for m in simulators:
            self.add(m)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_batch_destroy(h)
            self._h = None

    def add(self, simulator):
        if not isinstance(simulator, Simulator):
            raise TypeError("a Batch holds Simulators")
        check(_lib().hps_batch_add(self._h, simulator._h))
        self.simulators.append(simulator)

    def __len__(self):
        return len(self.simulators)

    def simulate(self, *args):
        """simulate() or simulate(start_element, end_element) of every member."""
        if len(args) == 0:
            i0, i1 = -1, -1
        elif len(args) == 2:
            bl = self.simulators[0].beamline
            i0, i1 = (a if isinstance(a, int) else bl.index(a) for a in args)
        else:
            raise TypeError("simulate takes either 0 or 2 arguments")
        check(_lib().hps_batch_simulate(self._h, i0, i1))

    def set_profiling(self, on=True):
        check(_lib().hps_batch_set_profiling(self._h, 1 if on else 0))

    def set_graph(self, on=True):
        """Replay repeated lattice periods as one HIP graph (default on)."""
        check(_lib().hps_batch_set_graph(self._h, 1 if on else 0))

    def graph_info(self):
        """(period graph replays so far, launches per replayed period)."""
        r = C.c_uint64()
        p = C.c_uint32()
        check(_lib().hps_batch_graph_info(self._h, C.byref(r), C.byref(p)))
        return r.value, p.value

    def set_field_mode(self, mode, aspect_step=0.05):
        """'fp64' (default: each beam's own Green's table and FP64 field sum,
        results identical to separate simulate calls) or 'gemm' (the grouped
        FP32 matrix-core field solve: shared Green's tables on the aspect grid
        (1 + aspect_step)^c, each beam's field interpolated quadratically
        between its three nearest classes; approximate)."""
        modes = {"fp64": 0, "gemm": 1}
        if mode not in modes:
            raise ValueError("field mode is 'fp64' or 'gemm'")
        check(_lib().hps_batch_set_field_mode(self._h, modes[mode], float(aspect_step)))

    def set_gemm_shape(self, shape=None):
        """Fix the grouped product's launch shape (None: autotune)."""
        check(_lib().hps_batch_set_gemm_shape(self._h, 0xFFFFFFFF if shape is None else shape))

    def gemm_info(self):
        """{shape, tables_built, conflicts, columns_last, product_columns, times_ms}:
        product columns are three per remeshing simulation."""
        info = (C.c_uint64 * 5)()
        ms = (C.c_float * 16)()
        check(_lib().hps_batch_get_gemm_info(self._h, info, ms, 16))
        return {"shape": None if info[0] == 0xFFFFFFFF else info[0], "tables_built": info[1],
                "conflicts": info[2], "columns_last": info[3], "product_columns": info[4],
                "times_ms": [t for t in ms if t > 0]}

    def set_autotune(self, on=True):
        """Time the field-sum launch shapes at the first rebuild (default on)."""
        check(_lib().hps_batch_set_autotune(self._h, 1 if on else 0))

    def tuning(self):
        """(chosen field-sum shape or None, candidate times in ms)."""
        shape = C.c_uint32()
        ms = (C.c_float * 16)()
        check(_lib().hps_batch_get_tuning(self._h, C.byref(shape), ms, 16))
        times = [t for t in ms if t > 0]
        return (None if shape.value == 0xFFFFFFFF else shape.value), times

    def profile(self):
        ms = (C.c_double * len(self._FAMILIES))()
        n = (C.c_uint32 * len(self._FAMILIES))()
        check(_lib().hps_batch_get_profile(self._h, ms, n))
        return {self._FAMILIES[k]: (ms[k], n[k]) for k in range(len(self._FAMILIES))}

    def last_time_ms(self):
        v = C.c_double()
        check(_lib().hps_batch_last_time_ms(self._h, C.byref(v)))
        return v.value


def probe_math(fn, x, y=None):
    """Evaluate one of the engine's device math routines over x (for tests)."""
    codes = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "sinh": 4, "cosh": 5, "sqrt": 6,
             "fmod2pi": 7, "div": 8, "bessi0": 9, "bessi1": 10, "bessi1p": 11, "psqrt": 12,
             "psin": 13, "pcos": 14, "psinh": 15, "pcosh": 16}
    # the same routines as the reciprocal-division build of the grouped-solve
    # push computes them
    codes.update({k + "_fast": v + 32 for k, v in list(codes.items())})
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(y, np.float64)
    out = np.empty_like(x)
    check(_lib().hps_probe_math(codes[fn], dptr(x), None if yy is None else dptr(yy), dptr(out),
                                x.size))
    return out


def set_db_epics(pv_name, value, dbconnection, beamline):
    """set_db_epics(pv, value, DBConnection, BeamLine), hpsim_module.cpp:52-82:
    writes the PV into the database (whose triggers recompute the model
    fields) and updates the elements that follow the PV."""
    from . import pv
    pv.set_db_epics(pv_name, value, dbconnection, beamline)


def get_db_epics(pv_name, dbconnection):
    """get_db_epics(pv, DBConnection) -> value text or None, hpsim_module.cpp:267-307."""
    from . import pv
    return pv.get_db_epics(pv_name, dbconnection)


def set_db_model(record_name, field_name, value, dbconnection, beamline):
    """set_db_model(record, field, value, DBConnection, BeamLine), hpsim_module.cpp:89-265."""
    from . import pv
    pv.set_db_model(record_name, field_name, value, dbconnection, beamline)


def get_db_model(record_name, field_name, dbconnection):
    """get_db_model(record, field, DBConnection) -> value text or None, hpsim_module.cpp:309-366."""
    from . import pv
    return pv.get_db_model(record_name, field_name, dbconnection)


def set_gpu(n):
    """set_gpu(n), hpsim_module.cpp / init.cpp:12-18."""
    check(_lib().hps_set_device(int(n)))


def betalambda(mass, freq, w):
    """hpsim.py:1986-2003: beta * lambda (m) for kinetic energy w (MeV)."""
    gamma = 1.0 + w / mass
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    return beta * 299.792458 / freq
