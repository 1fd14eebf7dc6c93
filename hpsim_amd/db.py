"""Lattice databases: the reference's DBConnection and BeamLine-from-database.

The reference keeps its lattices in SQLite files (one per beamline section,
e.g. tbtd.db, dtl.db, trst.db, ccl.db) and builds the beamline by walking
each database's ``linac`` view (src/beamline/src/init.cpp:20-429).  This
module does the same with Python's sqlite3 and returns element records
(hpsim_amd.configs.Element) for hpsim_amd.HPSim.BeamLine.  Like the
reference (GetDataFromDB + std::atof), every field is read as SQLite's text
form of the value (15 significant digits for a REAL) and converted to double
from that text.

The databases' triggers use SQL math functions that the reference loads
from a compiled SQLite extension (DBConnection::LoadLib,
src/util/src/sql_utility.cpp:251-257).  No compiled library is loaded here:
the functions are registered on the connection from Python's math module.
"""
from __future__ import annotations

import math
import os
import sqlite3

from .configs import GAP_FIELDS, Element


def _sql_functions():
    """Scalar SQL functions the lattice databases' triggers and views call."""
    def power(x, y):
        return None if x is None or y is None else math.pow(x, y)

    def unary(f):
        return lambda x: None if x is None else f(x)

    return {
        ("pi", 0): lambda: math.pi,
        ("power", 2): power,
        ("sqrt", 1): unary(math.sqrt),
        ("sin", 1): unary(math.sin),
        ("cos", 1): unary(math.cos),
        ("tan", 1): unary(math.tan),
        ("asin", 1): unary(math.asin),
        ("acos", 1): unary(math.acos),
        ("atan", 1): unary(math.atan),
        ("exp", 1): unary(math.exp),
        ("log", 1): unary(math.log),
        ("log10", 1): unary(math.log10),
        ("degrees", 1): unary(math.degrees),
        ("radians", 1): unary(math.radians),
        ("square", 1): unary(lambda v: v * v),
        ("sign", 1): unary(lambda v: (v > 0) - (v < 0)),
    }


class DBConnection:
    """DBConnection(db_addr), src/util/src/sql_utility.cpp:211-217: opens the
    first database as ``main``; attach_db adds the others under their
    aliases (pylib/hpsim.py:886-910 uses db1, db2, ...)."""

    def __init__(self, db_addr):
        if not os.path.isfile(db_addr):
            raise FileNotFoundError(f"database {db_addr!r} not found")
        self.conn = sqlite3.connect(db_addr)
        for (name, nargs), f in _sql_functions().items():
            self.conn.create_function(name, nargs, f, deterministic=True)
        self.dbs = ["main"]
        self.db_addrs = [db_addr]
        self.libs = []

    def attach_db(self, db_addr, db_name):
        """AttachDB, sql_utility.cpp:264-271."""
        if not os.path.isfile(db_addr):
            raise FileNotFoundError(f"database {db_addr!r} not found")
        self.conn.execute("attach database ? as " + _ident(db_name), (db_addr,))
        self.dbs.append(db_name)
        self.db_addrs.append(db_addr)

    def load_lib(self, lib_addr):
        """LoadLib, sql_utility.cpp:251-257.  The SQL functions are already
        registered natively; the library is recorded, never loaded."""
        self.libs.append(lib_addr)

    def print_dbs(self):
        for n, a in zip(self.dbs, self.db_addrs):
            print(f"{n} -> {a}")

    def print_libs(self):
        if not self.libs:
            print("No lib is loaded for this DBConnection.")
        for lib in self.libs:
            print(f"{lib} (SQL functions provided natively)")

    def tables(self, db):
        return [r[0] for r in self.conn.execute(
            f"select name from {_ident(db)}.sqlite_master where type = 'table'")]

    def columns(self, db, table):
        return [r[1] for r in self.conn.execute(f"pragma {_ident(db)}.table_info({_ident(table)})")]

    def clear_model_index(self):
        """ClearModelIndex, sql_utility.cpp:332-350."""
        for db in self.dbs:
            for t in self.tables(db):
                if "model_index" in self.columns(db, t):
                    self.conn.execute(f"update {_ident(db)}.{_ident(t)} set model_index = NULL")
        self.conn.commit()

    def get_epics_channels(self):
        """GetEPICSChannels, sql_utility.cpp:355-366."""
        out = []
        for db in self.dbs:
            out += [r[0] for r in self.conn.execute(f"select lcs_name from {_ident(db)}.epics_channel")]
        return out

    def close(self):
        self.conn.close()


def _ident(name):
    """A database, table or column name, checked to be a plain identifier."""
    if not name.replace("_", "").isalnum():
        raise ValueError(f"bad SQL identifier {name!r}")
    return name


def _f(v):
    """std::atof of a field: NULL or text that is not a number reads 0."""
    if v is None:
        return 0.0
    try:
        return float(v)
    except (TypeError, ValueError):
        return 0.0


def _one(conn, sql, args=()):
    r = conn.execute(sql, args).fetchone()
    return None if r is None else r[0]


def beamline_elements(dbc: DBConnection):
    """GenerateBeamLine, src/beamline/src/init.cpp:20-429: the elements of
    every attached database's ``linac`` view in order, with the same model
    fields; each element's model_index is written back as in the reference.
    Buncher and Dipole records are returned too (kind "Buncher" / "Dipole",
    model fields in ``params``)."""
    conn = dbc.conn
    out = []
    model_index = 0
    for db in dbc.dbs:
        d = _ident(db)
        elems = conn.execute(f"select name, view_index, model_type from {d}.linac").fetchall()
        for name, view_index, mtype in elems:
            e = None
            table = None

            def col(tbl, c):
                return _f(_one(conn, f"select cast({_ident(c)} as text) from {d}.{_ident(tbl)} "
                                     "where view_index = ?", (view_index,)))

            if mtype == "quad":
                table = "quad"
                e = Element(name, "Quad", length=col(table, "length_model"),
                            aperture=col(table, "aperture_model"),
                            gradient=col(table, "gradient_model"))
            elif mtype == "drift":
                table = "drift"
                e = Element(name, "Drift", length=col(table, "length_model"),
                            aperture=col(table, "aperture_model"))
            elif mtype in ("dtl_gap", "ccl_gap"):
                table = "rf_gap"
                key = "cell_id" if mtype == "dtl_gap" else "tank_id"
                cols = ", ".join(f"cast(t1.{c} as text)" for c in (
                    "beta_g", "beta_min", "ta0", "ta1", "ta2", "ta3", "ta4", "ta5", "sa1", "sa2",
                    "sa3", "sa4", "sa5"))
                ttf = conn.execute(
                    f"select {cols} from {d}.transit_time_factor t1 "
                    f"join {d}.rf_gap t2 on t1.module_id = t2.module_id and t1.tank_id = t2.{key} "
                    "where t2.view_index = ?", (view_index,)).fetchone()
                if ttf is None:
                    raise ValueError(f"{db}: no transit-time factors for gap {name}")
                g = {k: 0.0 for k in GAP_FIELDS}
                for k, v in zip(("fit_beta_center", "fit_beta_min", "fit_t0", "fit_t1",
                                 "fit_t2", "fit_t3", "fit_t4", "fit_t5", "fit_s1", "fit_s2",
                                 "fit_s3", "fit_s4", "fit_s5"), ttf):
                    g[k] = _f(v)
                g["length"] = col(table, "length_model")
                g["frequency"] = col(table, "frequency_model")
                g["cell_length_over_beta_lambda"] = _f(_one(
                    conn, f"select cast(cell_length_betalambda_design as text) from {d}.rf_module "
                          "where id = "
                          f"(select module_id from {d}.rf_gap where view_index = ?)",
                    (view_index,)))
                g["amplitude"] = col(table, "amplitude_model")
                g["phase_ref"] = col(table, "ref_phase_model")
                g["phase_shift"] = col(table, "beam_phase_shift_model")
                g["energy_out"] = col(table, "energy_out_model")
                g["beta_center"] = col(table, "beta_center_model")
                g["dg"] = col(table, "dg_model")
                g["t"] = col(table, "t_model")
                g["tp"] = col(table, "tp_model")
                g["sp"] = col(table, "sp_model")
                e = Element(name, "RFGap-DTL" if mtype == "dtl_gap" else "RFGap-CCL",
                            length=g["length"], gap=g)
            elif mtype == "rotation":
                table = "rotation"
                e = Element(name, "Rotation", angle=col(table, "angle_model"))
            elif mtype == "caperture":
                table = "caperture"
                e = Element(name, "ApertureC", aperture=col(table, "aperture_model"),
                            aperture_in=int(col(table, "in_out_model")) > 0)
            elif mtype == "raperture":
                table = "raperture"
                e = Element(name, "ApertureR", aper_xl=col(table, "aperture_xl_model"),
                            aper_xr=col(table, "aperture_xr_model"),
                            aper_yt=col(table, "aperture_yt_model"),
                            aper_yb=col(table, "aperture_yb_model"),
                            aperture_in=int(col(table, "in_out_model")) > 0)
            elif mtype == "diagnostics":
                table = "diagnostics"
                e = Element(name, "Diagnostics")
            elif mtype == "buncher":
                table = "buncher"
                e = Element(name, "Buncher", aperture=col(table, "aperture_model"),
                            params={"voltage": col(table, "voltage_model"),
                                    "frequency": col(table, "frequency_model"),
                                    "phase": col(table, "phase_model"),
                                    "on": int(col(table, "on_off")) > 0})
            elif mtype == "dipole":
                table = "dipole"
                e = Element(name, "Dipole", angle=col(table, "angle_model"),
                            params={k: col(table, c) for k, c in (
                                ("rho", "rho_model"), ("edge_angle_in", "edge_angle1_model"),
                                ("edge_angle_out", "edge_angle2_model"),
                                ("half_gap", "half_gap_model"), ("k1", "k1_model"),
                                ("k2", "k2_model"), ("field_index", "field_index_model"),
                                ("kinetic_energy", "kenergy_model"))})
            elif mtype == "spch_comp":
                table = "spch_comp"
                e = Element(name, "SpchComp", fraction=col(table, "fraction_model"))
            elif mtype == "steerer":
                table = "steerer"
                e = Element(name, "Steerer", bl_h=col(table, "bl_h_model"),
                            bl_v=col(table, "bl_v_model"))
            if e is None:
                continue   # other model types are not beamline elements (init.cpp)
            out.append(e)
            conn.execute(f"update {d}.{_ident(table)} set model_index = ? where view_index = ?",
                         (model_index, view_index))
            model_index += 1
    conn.commit()
    return out
