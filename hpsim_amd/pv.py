"""EPICS process variables and model fields of the lattice databases.

The reference changes a machine setting in two steps: the new PV value is
written into the database's ``epics_channel`` table, whose SQL triggers
recompute the model fields (gradients, RF amplitudes and phases, buncher
settings ...), and a *PV observer* then copies the recomputed fields of the
elements it watches into the live beamline (src/server/src/pv_observer.cpp,
observers built by src/server/src/init_pv_observer_list.cpp).  ``set_db_model``
writes a model-side field directly (src/pywrapper/src/hpsim_module.cpp:89-265).

Here the beamline is hpsim_amd.HPSim.BeamLine: an element's new fields go to
the engine with BeamLine.update_element, and the next simulate() compiles
them.  Differences: the reference keeps one global observer list built on
the first call and bound to that call's beamline; here each (database
connection, beamline) pair gets its own list.  Where the reference calls
exit(-1) (an element of the wrong type attached to an observer) a TypeError
is raised; where it prints an error and carries on, so does this module.
"""
from __future__ import annotations

import copy
import sys

from .db import _ident

# observer kinds (init_pv_observer_list.cpp:38-64)
QUAD, RF_PHASE, RF_AMP, MASTER, BUNCHER_PH, BUNCHER_AMP, BUNCHER_ON_OFF, DIPOLE = range(8)

_ELEMENT_KINDS = {
    QUAD: ("Quad",),
    RF_PHASE: ("RFGap-DTL", "RFGap-CCL"),
    RF_AMP: ("RFGap-DTL", "RFGap-CCL"),
    BUNCHER_PH: ("Buncher",),
    BUNCHER_AMP: ("Buncher",),
    BUNCHER_ON_OFF: ("Buncher",),
    DIPOLE: ("Dipole", "ApertureR", "Drift"),
}


def _err(msg):
    print(msg, file=sys.stderr)


def _num(v):
    """std::atof of a field read as text."""
    try:
        return float(v)
    except (TypeError, ValueError):
        return 0.0


class PVObserver:
    """One PV and the beamline elements that follow it (pv_observer.cpp)."""

    def __init__(self, pv, db, kind):
        self.pv, self.db, self.kind = pv, db, kind
        self.elements = []        # element names
        self.children = []        # MasterPVObserver: the module phase observers

    def attach(self, elem):
        """AttachBeamLineElement: only the element types the observer updates."""
        if self.kind == MASTER:
            _err("Cannot directly attach element to MasterPVObserver! ")
            return
        if elem.kind not in _ELEMENT_KINDS[self.kind]:
            raise TypeError(f"cannot attach {elem.name} ({elem.kind}) to the observer of {self.pv}")
        if self.kind in (BUNCHER_PH, BUNCHER_AMP, BUNCHER_ON_OFF):
            self.elements = [elem.name]           # one buncher per observer
        else:
            self.elements.append(elem.name)

    def update(self, conn, beamline, value):
        """Update: the database, then the model (pv_observer.cpp:17-40)."""
        conn.execute(f"update {_ident(self.db)}.epics_channel set value = ? where lcs_name = ?",
                     (_num(value) if _is_number(value) else value, self.pv))
        conn.commit()
        self.update_model(conn, beamline)

    def _row(self, conn, table, cols, name):
        # fields are read as SQLite's text form and converted like atof, as
        # the reference's GetDataFromDB / GetQueryResults + std::atof do
        sel = ", ".join(f"cast({c} as text)" for c in cols)
        return conn.execute(f"select {sel} from {_ident(self.db)}.{table} "
                            "where name = ?", (name,)).fetchone()

    def update_model(self, conn, beamline):
        if self.kind == MASTER:
            for ch in self.children:
                ch.update_model(conn, beamline)
            return
        for name in self.elements:
            e = copy.deepcopy(beamline[name])
            if self.kind == QUAD:
                r = self._row(conn, "quad", ["gradient_model"], name)
                if r is None or r[0] is None:
                    _err(f"QuadPVObserver::UpdateModel() failed, no gradient were found for quad : {name}")
                    continue
                e.gradient = _num(r[0])
            elif self.kind == RF_PHASE:
                r = self._row(conn, "rf_gap", ["beam_phase_shift_model"], name)
                if r is None or r[0] is None:
                    _err(f"RFPhasePVObserver::UpdateModel() failed for RFGap: {name}")
                    continue
                e.gap = dict(e.gap, phase_shift=_num(r[0]))
            elif self.kind == RF_AMP:
                r = self._row(conn, "rf_gap", ["amplitude_model", "ref_phase_model"], name)
                if r is None:
                    _err(f"RFAmplitudePVObserver::UpdateModel() failed for RFGap: {name}")
                    continue
                e.gap = dict(e.gap, amplitude=_num(r[0]), phase_ref=_num(r[1]))
            elif self.kind == BUNCHER_PH:
                r = self._row(conn, "buncher", ["phase_model"], name)
                if r is None or r[0] is None:
                    _err(f"BuncherPhasePVObserver::UpdateModel() failed for buncher : {name}")
                    continue
                e.params = dict(e.params, phase=_num(r[0]))
            elif self.kind == BUNCHER_AMP:
                r = self._row(conn, "buncher", ["voltage_model"], name)
                if r is None or r[0] is None:
                    _err(f"BuncherAmplitudePVObserver::UpdateModel() failed for buncher : {name}")
                    continue
                e.params = dict(e.params, voltage=_num(r[0]))
            elif self.kind == BUNCHER_ON_OFF:
                r = self._row(conn, "buncher", ["on_off"], name)
                if r is None or r[0] is None:
                    _err(f"BuncherOneOffPVObserver::UpdateModel() failed for buncher : {name}")
                    continue
                e.params = dict(e.params, on=_num(r[0]) > 0)
            elif self.kind == DIPOLE:
                if e.kind == "Dipole":
                    r = self._row(conn, "dipole", ["rho_model", "angle_model", "edge_angle1_model",
                                                   "edge_angle2_model", "kenergy_model"], name)
                    if r is None:
                        _err(f"DipolePVObserver::UpdateModel() failed, for dipole : {name}")
                        continue
                    e.angle = _num(r[1])
                    e.params = dict(e.params, rho=_num(r[0]), edge_angle_in=_num(r[2]),
                                    edge_angle_out=_num(r[3]), kinetic_energy=_num(r[4]))
                elif e.kind == "ApertureR":
                    r = self._row(conn, "raperture", ["aperture_xl_model", "aperture_xr_model"], name)
                    if r is None:
                        _err(f"DipolePVObserver::UpdateModel() failed, for rectangular aperture : {name}")
                        continue
                    e.aper_xl, e.aper_xr = _num(r[0]), _num(r[1])
                else:
                    r = self._row(conn, "drift", ["length_model"], name)
                    if r is None:
                        _err(f"DipolePVObserver::UpdateModel() failed, for drift : {name}")
                        continue
                    e.length = _num(r[0])
            beamline.update_element(name, e)


def _is_number(v):
    try:
        float(v)
        return True
    except (TypeError, ValueError):
        return False


def build_observers(dbc, beamline):
    """InitPVObserverList (init_pv_observer_list.cpp:8-121): one observer per
    PV that drives a quad, an RF module, a buncher or a dipole, with the
    elements it updates; master phase PVs drive the module phase PVs."""
    conn = dbc.conn
    names = {e.name for e in beamline}
    obs = {}
    for db in dbc.dbs:
        d = _ident(db)
        masters = {}
        for pv, pv_type in conn.execute(f"select lcs_name, value_type from {d}.epics_channel").fetchall():
            rows = conn.execute(f"select model_type, name from {d}.channel_list where channel1 = ? "
                                "or channel2 = ? or channel3 = ? or channel4 = ?",
                                (pv, pv, pv, pv)).fetchall()
            for elem_type, elem_name in rows:
                if pv not in obs:
                    kind = None
                    if elem_type == "quad":
                        kind = QUAD
                    elif elem_type == "rf_module":
                        kind = {"rf_ph": RF_PHASE, "rf_amp": RF_AMP, "delay": RF_AMP,
                                "rf_ph_master": MASTER}.get(pv_type)
                    elif elem_type == "buncher":
                        kind = {"buncher_ph": BUNCHER_PH, "buncher_amp": BUNCHER_AMP,
                                "buncher_on_off": BUNCHER_ON_OFF}.get(pv_type)
                    elif elem_type == "dipole":
                        kind = DIPOLE
                    if kind is not None:
                        obs[pv] = PVObserver(pv, db, kind)
                        if kind == MASTER:
                            masters[pv] = []
                if elem_type != "rf_module" and elem_name in names:
                    if pv not in obs:
                        raise RuntimeError(f"PVObserver for {pv} does not exist, cannot attach element!")
                    obs[pv].attach(beamline[elem_name])
                elif elem_type == "rf_module":
                    if pv_type != "rf_ph_master":
                        gaps = conn.execute(f"select g.name from {d}.rf_gap g join {d}.rf_module m "
                                            "on m.id = g.module_id where m.name = ?",
                                            (elem_name,)).fetchall()
                        for (g,) in gaps:
                            if pv in obs and g in names:
                                obs[pv].attach(beamline[g])
                    elif pv in masters:
                        masters[pv].append(elem_name)
                else:
                    _err(f"InitPVObserverList error: Cannot find {elem_name} in beamline!")
        for mpv, modules in masters.items():
            for m in modules:
                r = conn.execute(f"select lcs_name from {d}.epics_channel e join {d}.rf_module m "
                                 "where m.name = ? and m.phase_channel = e.id", (m,)).fetchone()
                if r is not None and r[0] in obs:
                    obs[mpv].children.append(obs[r[0]])
    return obs


_OBSERVERS = {}


def _observers(dbc, beamline):
    key = (id(dbc), id(beamline))
    ent = _OBSERVERS.get(key)
    if ent is None or ent[0] is not dbc or ent[1] is not beamline:
        ent = (dbc, beamline, build_observers(dbc, beamline))
        _OBSERVERS[key] = ent
    return ent[2]


def set_db_epics(pv, value, dbc, beamline):
    """SetDbEPICS (hpsim_module.cpp:56-82): a PV without an observer changes
    nothing."""
    ob = _observers(dbc, beamline).get(str(pv))
    if ob is not None:
        ob.update(dbc.conn, beamline, value)


def get_db_epics(pv, dbc):
    """GetDbEPICS (hpsim_module.cpp:271-307): the value as text (SQLite's
    own conversion), or None."""
    for db in dbc.dbs:
        r = dbc.conn.execute(f"select cast(value as text) from {_ident(db)}.epics_channel "
                             "where lcs_name = ?", (str(pv),)).fetchone()
        if r is not None:
            return r[0]
    _err(f"Error in get_db_epics : can't find pv: {pv}, or this record has a NULL value ")
    return None


def get_db_model(record, field, dbc):
    """GetDbModel (hpsim_module.cpp:313-366): the first table (other than
    epics_channel) with the column and a non-empty value for the record."""
    for db in dbc.dbs:
        for t in dbc.tables(db):
            if t == "epics_channel" or field not in dbc.columns(db, t):
                continue
            r = dbc.conn.execute(f"select cast({_ident(field)} as text) from {_ident(db)}.{_ident(t)} "
                                 "where name = ?", (str(record),)).fetchone()
            if r is not None and r[0] not in (None, ""):
                return r[0]
    _err(f"Error in get_db_model: can't find record name: {record} or its field name: {field}")
    return None


_MODEL_TABLES = ("buncher", "rf_module", "spch_comp", "raperture", "caperture")


def set_db_model(record, field, value, dbc, beamline):
    """SetDbModel (hpsim_module.cpp:89-265): writes one field of a buncher,
    RF module, space-charge compensation or aperture record, then copies the
    fields the triggers recompute into the beamline."""
    conn = dbc.conn
    record, field = str(record), str(field)
    for db in dbc.dbs:
        d = _ident(db)
        for t in dbc.tables(db):
            if t not in _MODEL_TABLES or field not in dbc.columns(db, t):
                continue
            if conn.execute(f"select id from {d}.{t} where name = ?", (record,)).fetchone() is None:
                continue
            conn.execute(f"update {d}.{t} set {_ident(field)} = ? where name = ?",
                         (_num(value) if _is_number(value) else value, record))
            conn.commit()

            def one(sql, args=()):
                r = conn.execute(sql, args).fetchone()
                return None if r is None else r[0]

            def text(col):
                return f"cast({col} as text)"

            def update(name, fn):
                e = copy.deepcopy(beamline[name])
                fn(e)
                beamline.update_element(name, e)

            if t in ("spch_comp", "buncher", "raperture", "caperture"):
                mi = one(f"select model_index from {d}.{t} where name = ?", (record,))
                if mi is None:
                    return
                name = beamline[int(mi)].name
                if t == "spch_comp" and field == "fraction_model":
                    v = _num(one(f"select {text('fraction_model')} from {d}.spch_comp where name = ?",
                                 (record,)))
                    update(name, lambda e: setattr(e, "fraction", v))
                elif t == "buncher" and field == "phase_offset_cal":
                    v = _num(one(f"select {text('phase_model')} from {d}.buncher where name = ?", (record,)))
                    update(name, lambda e: setattr(e, "params", dict(e.params, phase=v)))
                elif t == "raperture" and field in ("aperture_xl_model", "aperture_xr_model",
                                                    "aperture_yt_model", "aperture_yb_model",
                                                    "in_out_model"):
                    v = _num(one(f"select {text(field)} from {d}.raperture where name = ?", (record,)))
                    attr = {"aperture_xl_model": "aper_xl", "aperture_xr_model": "aper_xr",
                            "aperture_yt_model": "aper_yt", "aperture_yb_model": "aper_yb"}.get(field)
                    if attr:
                        update(name, lambda e: setattr(e, attr, v))
                    else:
                        update(name, lambda e: setattr(e, "aperture_in", v > 0))
                elif t == "caperture" and field in ("aperture_model", "in_out_model"):
                    v = _num(one(f"select {text(field)} from {d}.caperture where name = ?", (record,)))
                    if field == "aperture_model":
                        update(name, lambda e: setattr(e, "aperture", v))
                    else:
                        update(name, lambda e: setattr(e, "aperture_in", v > 0))
            elif t == "rf_module":
                if field == "phase_offset_cal":
                    rows = conn.execute(
                        f"select model_index, {text('beam_phase_shift_model')} from {d}.rf_gap where module_id = "
                        f"(select id from {d}.rf_module where name = ?)", (record,)).fetchall()
                    for mi, ps in rows:
                        if mi is not None:
                            update(beamline[int(mi)].name,
                                   lambda e, ps=ps: setattr(e, "gap", dict(e.gap, phase_shift=_num(ps))))
                elif field in ("amplitude_scale_cal", "amplitude_tilt_cal", "voltage_cal"):
                    rows = conn.execute(
                        f"select model_index, {text('amplitude_model')}, {text('ref_phase_model')} from {d}.rf_gap where "
                        f"module_id = (select id from {d}.rf_module where name = ?)", (record,)).fetchall()
                    for mi, amp, ph in rows:
                        if mi is not None:
                            update(beamline[int(mi)].name,
                                   lambda e, amp=amp, ph=ph: setattr(
                                       e, "gap", dict(e.gap, amplitude=_num(amp), phase_ref=_num(ph))))
            return
    _err(f"Error in set_db_model: can't make the change for {record}:{field}")
