"""User-unit facade over the engine module, mirroring pylib/hpsim.py.

The reference's scripts drive HPSim through this layer: coordinates in cm,
mrad, deg and MeV (pylib/hpsim.py:44-53), one class-level BeamLine and
SpaceCharge that Simulator(beam) picks up (pylib/hpsim.py:1112-1123), and
numpy analysis helpers (masks, Twiss parameters, rms emittances).  Every
call lands in hpsim_amd.HPSim and from there in the gfx950 engine.

Left out: the matplotlib plotting classes (BeamPlot, DistPlot) and the
EPICS helpers (DBState, set_db_epics, ...).  The lattice comes from the
reference's SQLite databases (DBConnection + BeamLine()) or from element
records (BeamLine(elements=[...])).
"""
from __future__ import annotations

import math

import numpy as np

from . import HPSim
from . import lcsutil as lcs

_COORDINATES = ["x", "xp", "y", "yp", "phi", "w"]
_EMITTANCE = ["x", "y", "z"]
_LOSSES = ["losses"]
# engine units are m, rad, MeV; user units cm, mrad, deg, MeV (hpsim.py:46-53)
_CM = 100.0
_MR = 1000.0
_DEG = 180.0 / math.pi
_RAD = 1 / _DEG
_MA = 1000.0
_USER_UNITS = {"x": _CM, "xp": _MR, "y": _CM, "yp": _MR, "phi": _DEG, "w": 1.0}
_USER_LABELS = {"x": "cm", "xp": "mr", "y": "cm", "yp": "mr", "phi": "deg", "w": "MeV",
                "i": "mA"}


def _twiss(u, v):
    """[alpha, beta, rms emittance] of two coordinate vectors; NaNs when the
    emittance is undefined (hpsim.py:474-501)."""
    su2 = float(np.std(u)) ** 2
    sv2 = float(np.std(v)) ** 2
    cuv = float(np.mean((u - np.mean(u)) * (v - np.mean(v))))
    e2 = su2 * sv2 - cuv * cuv
    if e2 > 0:
        e = math.sqrt(e2)
        return [-cuv / e, su2 / e, e]
    return 3 * [float("nan")]


class Beam:
    """Beam in user units (pylib/hpsim.py:56-697): Beam(file=...) or
    Beam(mass=, charge=, current=, num=)."""

    def __init__(self, **args):
        if "file" in args:
            self.beam = HPSim.Beam(file=args["file"])
            self.initial_current = self.beam.get_current()
            self.initial_frequency = self.beam.get_frequency()
            self.initial_gd_size = self.beam.get_size() - self.beam.get_loss_num()
        else:
            self.beam = HPSim.Beam(mass=float(args["mass"]), charge=float(args["charge"]),
                                   current=float(args["current"]), num=int(args["num"]))
            self.initial_current = float(args["current"])
            self.initial_frequency = 0
            self.initial_gd_size = self.get_initial_size()

    # -- engine calls with unit conversion ---------------------------------
    def get_distribution(self, option="all"):
        loss_all = self.get_losses()
        if option == "all":
            loss = loss_all
        elif option == "bad":
            loss = np.where(loss_all > 0)[0]
        else:
            loss = np.where(loss_all == 0)[0]
        return (self.get_x(option), self.get_xp(option), self.get_y(option),
                self.get_yp(option), self.get_phi(option), self.get_w(option), loss)

    def set_distribution(self, x, xp, y, yp, phi, w, loss=None):
        """Coordinates in cm, mr, deg, MeV; a shorter set leaves zeros behind
        (hpsim.py:137-160)."""
        n = self.beam.get_size()
        arrs = [np.asarray(a, np.float64) for a in (x, xp, y, yp, phi, w)]
        if len(arrs[0]) < n:
            z = np.zeros(n)
            self.beam.set_distribution(z, z, z, z, z, z)
            full = []
            for a in arrs:
                f = np.zeros(n)
                f[: len(a)] = a
                full.append(f)
            arrs = full
            if loss is not None:
                lo = np.zeros(n, np.uint32)
                lo[: len(loss)] = loss
                loss = lo
        self.beam.set_distribution(arrs[0] / _CM, arrs[1] / _MR, arrs[2] / _CM, arrs[3] / _MR,
                                   arrs[4] / _DEG, arrs[5], loss)

    def set_waterbag(self, alpha_x, beta_x, emittance_x, alpha_y, beta_y, emittance_y,
                     alpha_z, beta_z, emittance_z, synch_phi, synch_w, frequency, random_seed=0):
        """PARMILA units; z-plane beta in deg/MeV and emittance in deg*MeV
        (hpsim.py:162-193)."""
        self.beam.set_waterbag(alpha_x, beta_x, emittance_x, alpha_y, beta_y, emittance_y,
                               alpha_z, beta_z * _DEG, emittance_z / _DEG, synch_phi / _DEG,
                               synch_w, frequency, int(random_seed))

    def set_dc(self, alpha_x, beta_x, emittance_x, alpha_y, beta_y, emittance_y,
               delta_phi, synch_phi, synch_w, random_seed=0):
        """hpsim.py:195-223; phases in deg."""
        self.beam.set_dc(alpha_x, beta_x, emittance_x, alpha_y, beta_y, emittance_y,
                         delta_phi / _DEG, synch_phi / _DEG, synch_w, int(random_seed))

    def save_initial_beam(self): self.beam.save_initial_beam()
    def save_intermediate_beam(self): self.beam.save_intermediate_beam()
    def restore_initial_beam(self): self.beam.restore_initial_beam()
    def restore_intermediate_beam(self): self.beam.restore_intermediate_beam()
    def print_simple(self): self.beam.print_simple()
    def print_to(self, output_file_name): self.beam.print_to(output_file_name)

    def set_ref_w(self, w): self.beam.set_ref_w(float(w))
    def get_ref_w(self): return self.beam.get_ref_w()
    def set_ref_phi(self, phi): self.beam.set_ref_phi(float(phi / _DEG))
    def get_ref_phi(self): return self.beam.get_ref_phi() * _DEG

    def set_frequency(self, frequency):
        self.beam.set_frequency(float(frequency))
        if self.initial_frequency == 0:
            self.initial_frequency = self.get_frequency()

    def get_frequency(self): return self.beam.get_frequency()
    def get_mass(self): return self.beam.get_mass()
    def get_charge(self): return self.beam.get_charge()
    def get_initial_size(self): return self.beam.get_size()

    def get_x(self, option="all"): return self.beam.get_x(option) * _CM
    def get_xp(self, option="all"): return self.beam.get_xp(option) * _MR
    def get_y(self, option="all"): return self.beam.get_y(option) * _CM
    def get_yp(self, option="all"): return self.beam.get_yp(option) * _MR
    def get_phi(self, option="all"): return self.beam.get_phi(option) * _DEG
    def get_w(self, option="all"): return self.beam.get_w(option)
    def get_losses(self): return self.beam.get_losses()
    def get_loss_num(self): return self.beam.get_loss_num()

    def get_avg_x(self): return self.beam.get_avg_x() * _CM
    def get_avg_y(self): return self.beam.get_avg_y() * _CM
    def get_avg_phi(self, option="absolute"): return self.beam.get_avg_phi(option) * _DEG
    def get_avg_w(self): return self.beam.get_avg_w()
    def get_sig_x(self): return self.beam.get_sig_x() * _CM
    def get_sig_y(self): return self.beam.get_sig_y() * _CM
    def get_sig_phi(self): return self.beam.get_sig_phi() * _DEG
    def get_sig_w(self): return self.beam.get_sig_w()
    def get_emittance_x(self): return self.beam.get_emittance_x() * _CM * _MR
    def get_emittance_y(self): return self.beam.get_emittance_y() * _CM * _MR
    def get_emittance_z(self): return self.beam.get_emittance_z() * _DEG

    def apply_cut(self, axis, minval, maxval): self.beam.apply_cut(axis, minval, maxval)

    def translate(self, axis, value):
        self.beam.translate(axis, value / _USER_UNITS[axis.lower()])

    # -- mask-based analysis (hpsim.py:398-697) ----------------------------
    def get_coor(self, var, mask=None):
        v = var.lower()
        if v in _COORDINATES:
            c = getattr(self.beam, "get_" + v)("all") * _USER_UNITS[v]
        elif v in _LOSSES:
            c = self.beam.get_losses()
        else:
            raise ValueError(f"variable {var!r} not recognized")
        return c if mask is None else c[mask]

    def get_avg(self, var, mask=None):
        if var.lower() not in _COORDINATES:
            raise ValueError(f"variable {var!r} not recognized")
        if mask is not None and len(mask) == 0:
            return 0.0
        return float(np.mean(self.get_coor(var, mask)))

    def get_sig(self, var, mask=None):
        if var.lower() not in _COORDINATES:
            raise ValueError(f"variable {var!r} not recognized")
        return float(np.std(self.get_coor(var, mask)))

    def get_twiss(self, var, mask=None):
        """[alpha, beta, unnormalised rms emittance] for 'x', 'y' or 'z'."""
        v = var.lower()
        if v not in _EMITTANCE:
            raise ValueError(f"coordinate must be one of {_EMITTANCE}")
        u, up = ("phi", "w") if v == "z" else (v, v + "p")
        return _twiss(self.get_coor(u, mask), self.get_coor(up, mask))

    def get_urms_emit(self, var, mask=None): return self.get_twiss(var, mask)[2]

    def get_nrms_emit(self, var, mask=None):
        e = self.get_urms_emit(var, mask)
        return e * self.get_betagamma(mask) if var in ("x", "y") else e

    def get_mask_with_limits(self, var, lolim, uplim=None):
        a = self.get_coor(var)
        m = a > lolim
        if uplim is not None:
            m &= a < uplim
        return np.where(m)[0]

    def get_good_mask(self, mask=None):
        good = np.where(self.beam.get_losses() == 0)[0]
        return good if mask is None else np.intersect1d(mask, good)

    def get_lost_mask(self, mask=None):
        lost = np.where(self.beam.get_losses() != 0)[0]
        return lost if mask is None else np.intersect1d(mask, lost)

    def get_intersection_mask(self, mask1, mask2): return np.intersect1d(mask1, mask2)

    def get_betagamma(self, mask=None):
        if mask is not None and len(mask) == 0:
            return 0.0
        g = 1.0 + self.get_avg("w", mask) / self.get_mass()
        return math.sqrt(g * g - 1.0)

    def get_betalambda(self, mask=None):
        wavelength = 2.99792458e8 / (self.get_frequency() * 1.0e6)
        if mask is not None and len(mask) == 0:
            return 0.0
        g = 1.0 + self.get_avg("w", mask) / self.get_mass()
        return math.sqrt(1.0 - 1 / (g * g)) * wavelength

    def get_current(self, mask=None):
        """mA; without a mask the engine's transmission-scaled current
        (hpsim.py:622-632)."""
        if mask is None:
            return self.beam.get_current() * _MA
        return (self.initial_current * _MA * self.get_frequency() / self.initial_frequency *
                self.get_size(mask) / self.initial_gd_size)

    def get_size(self, mask=None):
        if mask is None:
            return self.get_initial_size() - self.beam.get_loss_num()
        return np.asarray(mask).size

    def print_results(self, mask=None):
        print(_results_text(self, mask))


class Distribution:
    """Static numpy copy of a (masked) beam in user units (hpsim.py:699-873)."""

    coor_index = {c: i for i, c in enumerate(_COORDINATES + _LOSSES)}

    def __init__(self, beam, mask=None):
        n = beam.get_initial_size() if mask is None else beam.get_size(mask)
        self.coor = np.zeros([len(_COORDINATES + _LOSSES), n])
        for c in _COORDINATES + _LOSSES:
            self.coor[Distribution.coor_index[c]] = beam.get_coor(c, mask=mask)
        self.current = beam.get_current(mask)
        self.frequency = beam.get_frequency()
        self.mass = beam.get_mass()
        self.charge = beam.get_charge()
        self.size = beam.get_size(mask)
        self.betagamma = beam.get_betagamma(mask)
        self.betalambda = beam.get_betalambda(mask)
        self.ref_phi = beam.get_ref_phi()
        self.ref_w = beam.get_ref_w()

    def get_ref_phi(self): return self.ref_phi
    def get_ref_w(self): return self.ref_w
    def get_current(self): return self.current
    def get_frequency(self): return self.frequency
    def get_mass(self): return self.mass
    def get_charge(self): return self.charge
    def get_betagamma(self): return self.betagamma
    def get_betalambda(self): return self.betalambda
    def get_size(self): return self.size

    def get_coor(self, var):
        if var not in Distribution.coor_index:
            raise ValueError(f"variable {var!r} not recognized")
        return self.coor[Distribution.coor_index[var]]

    def get_loss_num(self):
        t = self.get_coor("losses")
        return int((t > 0).sum())

    def get_avg(self, var): return float(np.mean(self.get_coor(var)))
    def get_sig(self, var): return float(np.std(self.get_coor(var)))

    def get_twiss(self, var):
        v = var.lower()
        if v not in _EMITTANCE:
            raise ValueError(f"coordinate must be one of {_EMITTANCE}")
        u, up = ("phi", "w") if v == "z" else (v, v + "p")
        return _twiss(self.get_coor(u), self.get_coor(up))

    def get_urms_emit(self, var): return self.get_twiss(var)[2]

    def get_nrms_emit(self, var):
        e = self.get_urms_emit(var)
        return e * self.get_betagamma() if var in ("x", "y") else e

    def print_results(self):
        print(_results_text(self, None))


def _results_text(src, mask):
    args = () if isinstance(src, Distribution) else (mask,)
    lines = ["Distribution Analysis Results (w/user units)",
             f"Mass = {src.get_mass():.4f}", f"Charge/|e| = {src.get_charge():.0f}",
             f"Ib = {src.get_current(*args):.2f} {_USER_LABELS['i']}",
             f"Frequency = {src.get_frequency():.3f} MHz",
             f"Number of macroparticles = {src.get_size(*args):.0f}",
             f"Ref part. \n phi = {src.get_ref_phi():10.4f} {_USER_LABELS['phi']}\n"
             f"   w = {src.get_ref_w():10.4f} {_USER_LABELS['w']}"]
    if src.get_size(*args) > 0:
        lines += ["", "Centroids and RMS sizes", "            Avg         Sigma"]
        for c in _COORDINATES:
            lines.append(f"{c:3}: {src.get_avg(c, *args):12.4f}    {src.get_sig(c, *args):10.4f} "
                         f"{_USER_LABELS[c]:3}")
        lines += ["", "Twiss parameters", "          Alpha       Beta       Eurms       Enrms"]
        for c in _EMITTANCE:
            a, b, e = src.get_twiss(c, *args)
            lines.append(f"{c:2}: {a:11.4f} {b:11.4f} {e:11.4f} {src.get_nrms_emit(c, *args):11.5f}")
    else:
        lines.append("\n*** No particles remaining ****")
    return "\n".join(lines) + "\n"


class DBConnection:
    """Lattice databases, kept class-level for BeamLine() (hpsim.py:875-932):
    the first database opens as ``main``, the others attach as db1, db2, ...
    in beamline order."""

    dbconnection = None

    def __init__(self, db_dir, databases, libsql_dir=None, libsql_file=None):
        import os
        self.dbconnection = None
        for k, name in enumerate(databases):
            path = os.path.join(os.path.abspath(db_dir), name)
            if k == 0:
                self.dbconnection = HPSim.DBConnection(path)
            else:
                self.dbconnection.attach_db(path, f"db{k}")
        DBConnection.dbconnection = self.dbconnection
        if libsql_dir is not None and libsql_file is not None:
            self.dbconnection.load_lib(os.path.join(os.path.abspath(libsql_dir), libsql_file))
        self.clear_model_index()

    def print_dbs(self): self.dbconnection.print_dbs()
    def print_libs(self): self.dbconnection.print_libs()
    def clear_model_index(self): self.dbconnection.clear_model_index()
    def get_epics_channels(self): return self.dbconnection.get_epics_channels()


class BeamLine:
    """The beamline, kept class-level for Simulator(beam) (hpsim.py:935-980):
    from the class-level DBConnection, or from element records
    (hpsim_amd.configs.Element)."""

    beamline = None

    def __init__(self, elements=None):
        if elements is None:
            if DBConnection.dbconnection is None:
                raise ValueError("create a DBConnection (or pass element records) first")
            self.beamline = HPSim.BeamLine(DBConnection.dbconnection)
        else:
            self.beamline = HPSim.BeamLine(elements=elements)
        BeamLine.beamline = self.beamline

    def print_out(self): self.beamline.print_out()
    def print_range(self, start_element, end_element):
        self.beamline.print_range(start_element, end_element)

    def get_element_names(self, start_element=None, end_element=None, elem_type=None):
        return get_element_list(start_element, end_element, elem_type)


class SpaceCharge:
    """SCHEFF settings, kept class-level for Simulator(beam) (hpsim.py:983-1104)."""

    spacecharge = None

    def __init__(self, nr, nz, interval, adj_bunch, type="scheff"):
        self.spacecharge = HPSim.SpaceCharge(nr=nr, nz=nz, interval=interval,
                                             adj_bunch=adj_bunch)
        SpaceCharge.spacecharge = self.spacecharge

    def get_interval(self): return self.spacecharge.get_interval()
    def set_interval(self, interval): self.spacecharge.set_interval(interval)
    def get_adj_bunch(self): return int(self.spacecharge.get_adj_bunch())
    def set_adj_bunch(self, adj_bunch): self.spacecharge.set_adj_bunch(int(adj_bunch))
    def get_adj_bunch_cutoff_w(self): return self.spacecharge.get_adj_bunch_cutoff_w()
    def set_adj_bunch_cutoff_w(self, w): self.spacecharge.set_adj_bunch_cutoff_w(w)
    def get_mesh_size(self): return self.spacecharge.get_mesh_size()
    def set_mesh_size(self, nr, nz): self.spacecharge.set_mesh_size(nr, nz)
    def get_mesh_size_cutoff_w(self): return self.spacecharge.get_mesh_size_cutoff_w()
    def set_mesh_size_cutoff_w(self, w): self.spacecharge.set_mesh_size_cutoff_w(w)
    def get_remesh_threshold(self): return self.spacecharge.get_remesh_threshold()
    def set_remesh_threshold(self, v): self.spacecharge.set_remesh_threshold(v)


class Simulator:
    """Simulator(beam) over the class-level BeamLine and SpaceCharge
    (hpsim.py:1107-1134)."""

    def __init__(self, beam):
        if BeamLine.beamline is None:
            raise ValueError("create a BeamLine before the Simulator")
        self.sim = HPSim.Simulator(beam.beam, BeamLine.beamline, SpaceCharge.spacecharge)

    def simulate(self, start_elem_name, end_elem_name):
        self.sim.simulate(start_elem_name, end_elem_name)

    def set_space_charge(self, state="off"):
        self.sim.set_space_charge(state)


def set_gpu(n):
    HPSim.set_gpu(n)


def most_freq_value(vals):
    """Peak of a unit-bin histogram refined by a weighted average with its
    neighbours, plus half a bin (hpsim.py:1737-1768)."""
    vals = np.asarray(vals)
    lo, hi = float(np.min(vals)), float(np.max(vals))
    hist, edge = np.histogram(vals, bins=max(int(hi - lo) + 1, 3), density=False)
    width = edge[1] - edge[0]
    k = int(np.argmax(hist))
    if 0 < k < len(hist):
        pk = np.average(edge[k - 1:k + 2], weights=hist[k - 1:k + 2])
    else:
        pk = edge[k]
    return pk + 0.5 * width


def modulo_phase(phase_dist, ref_phs):
    """Phases (deg) folded into ref_phs +- 180 deg (hpsim.py:1770-1781)."""
    return ((phase_dist - ref_phs + math.pi * _DEG) % (2 * math.pi * _DEG)) + ref_phs \
        - (math.pi * _DEG)


_ELEM_TYPES = {"caperture": "ApertureC", "raperture": "ApertureR", "buncher": "Buncher",
               "diagnostics": "Diagnostics", "dipole": "Dipole", "drift": "Drift",
               "quad": "Quad", "dtl-gap": "RFGap-DTL", "ccl-gap": "RFGap-CCL",
               "rotation": "Rotation", "spch_comp": "SpchComp"}


def get_element_list(start_elem_name=None, end_elem_name=None, elem_type=None):
    """Names of the elements from start to end, optionally of one type
    (hpsim.py:1849-1883)."""
    t = _ELEM_TYPES.get(elem_type, elem_type) if elem_type else ""
    return BeamLine.beamline.get_element_names(start_elem_name or "", end_elem_name or "", t)


def set_db_epics(pv_name, value):
    """Change an EPICS PV value in the database and the beamline model
    (hpsim.py:1783-1793); DBConnection and BeamLine must exist."""
    HPSim.set_db_epics(pv_name, str(value), DBConnection.dbconnection, BeamLine.beamline)


def get_db_epics(pv_name):
    """An EPICS PV value from the database (hpsim.py:1795-1805): a float,
    except for PVs of LCS channel type 'L'."""
    value = HPSim.get_db_epics(pv_name, DBConnection.dbconnection)
    if lcs.get_pv_type(pv_name) != "L":
        value = float(value)
    return value


def set_db_model(table_name, field_name, value):
    """Change a model field of a database record (hpsim.py:1807-1817)."""
    HPSim.set_db_model(table_name, field_name, str(value), DBConnection.dbconnection,
                       BeamLine.beamline)


def get_db_model(elem_name, field_name):
    """A model field of a database record (hpsim.py:1819-1832): a float, except
    for the text fields id, name and model_type."""
    value = HPSim.get_db_model(elem_name, field_name, DBConnection.dbconnection)
    if field_name not in ("id", "name", "model_type"):
        value = float(value)
    return value


def get_beamline_direction(start_elem, stop_elem):
    names = get_element_list()
    return -1 if names.index(start_elem) > names.index(stop_elem) else 1


def get_element_length(elem_name):
    """Length (m) of an element (hpsim.py:1886-1902)."""
    bl = BeamLine.beamline
    return float(bl[bl.index(elem_name)].length)


def get_beamline_length(start, end):
    """Summed element length from start to end, signed by direction
    (hpsim.py:1904-1923)."""
    direction = get_beamline_direction(start, end)
    a, b = (start, end) if direction > 0 else (end, start)
    return sum(get_element_length(n) for n in get_element_list(a, b)) * direction


def get_first_element(): return get_element_list()[0]


def get_last_element(): return get_element_list()[-1]


def get_next_element(elem_name):
    names = get_element_list()
    i = names.index(elem_name) + 1
    return names[i] if i < len(names) else None


def get_mmf(twiss1, twiss2):
    """Mismatch factor of two [alpha, beta, emittance] ellipses
    (hpsim.py:1970-1984)."""
    a1, b1, _ = twiss1
    a2, b2, _ = twiss2
    g1 = (1.0 + a1 * a1) / b1
    g2 = (1.0 + a2 * a2) / b2
    r = b1 * g2 + g1 * b2 - 2.0 * a1 * a2
    return math.sqrt(0.5 * (r + math.sqrt(r * r - 4))) - 1.0


def betalambda(mass, freq, w):
    """beta * lambda (m) at kinetic energy w (hpsim.py:1986-2000)."""
    wavelength = 2.99792458e8 / (freq * 1.0e6)
    g = 1.0 + w / mass
    return math.sqrt(1.0 - 1 / (g * g)) * wavelength
