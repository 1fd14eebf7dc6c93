"""PV scans as one batched launch.

The reference scans a machine setting with a host loop (pytest/sim-lbeg-scan-pv.py:140-160):
for every value, restore_initial_beam(), set_db_epics(pv, value), simulate(start, stop), then
analyse the beam.  ``pv_scan`` runs the same scan as one hpsim_amd.HPSim.Batch: the lattice
of every value is taken from the database (set_db_epics, observers), each value tracks its
own copy of the beam's saved initial distribution, and one launch sequence advances all of
them.  Each value starts from a fresh copy of the initial beam and a fresh SpaceCharge
with the given settings; the reference loop instead carries the beam's statistics (such as
the transverse-loss count, which only getters and apertures update) and the SpaceCharge
remesh state from one value to the next, so its later values can differ slightly.  The PV
is restored to its value before the scan.
"""
from __future__ import annotations

import copy

from . import HPSim


def _copy_beam(beam):
    """A new HPSim.Beam holding beam's saved initial distribution and settings
    (beam itself is left restored to it)."""
    beam.restore_initial_beam()
    a = beam.arrays()
    nb = HPSim.Beam(mass=beam.get_mass(), charge=beam.get_charge(), current=beam.get_current(),
                    num=beam.get_size())
    nb.set_distribution(a["x"], a["xp"], a["y"], a["yp"], a["phi"], a["w"], a["loss"], a["lloss"])
    nb.set_frequency(beam.get_frequency())
    nb.set_ref_w(beam.get_ref_w())
    nb.set_ref_phi(beam.get_ref_phi())
    nb.save_initial_beam()
    return nb


def _copy_spch(sc):
    st = sc.state()
    n = HPSim.SpaceCharge(nr=st["nx_init"], nz=st["nz_init"], interval=st["interval_init"],
                          adj_bunch=int(st["adj_bunch"]))
    n.set_remesh_threshold(st["remesh_val"])
    if st["adj_bunch_w"]:
        n.set_adj_bunch_cutoff_w(st["adj_bunch_w"])
    if st["mesh_w"]:
        n.set_mesh_size_cutoff_w(st["mesh_w"])
    return n


def pv_scan(pv_name, values, beam, dbconnection, beamline, spacecharge=None, start=None,
            stop=None):
    """Track beam's saved initial distribution through the lattice at every
    value of pv_name, all values in one batch; returns one HPSim.Beam per
    value (in order).  dbconnection / beamline: the HPSim.DBConnection and
    HPSim.BeamLine the PV belongs to (beamline is left at the PV's original
    value)."""
    values = list(values)
    if not values:
        return []
    v0 = HPSim.get_db_epics(pv_name, dbconnection)
    lattices = []
    try:
        for v in values:
            HPSim.set_db_epics(pv_name, str(v), dbconnection, beamline)
            lattices.append([copy.deepcopy(e) for e in beamline])
    finally:
        if v0 is not None:
            HPSim.set_db_epics(pv_name, v0, dbconnection, beamline)
    sims = []
    for els in lattices:
        b = _copy_beam(beam)
        sc = _copy_spch(spacecharge) if spacecharge is not None else None
        sims.append(HPSim.Simulator(b, HPSim.BeamLine(elements=els), sc))
    batch = HPSim.Batch(sims)
    if start is None and stop is None:
        batch.simulate()
    else:
        s0 = beamline.index(start) if isinstance(start, str) else (0 if start is None else start)
        s1 = beamline.index(stop) if isinstance(stop, str) else (len(beamline) - 1 if stop is None
                                                                 else stop)
        batch.simulate(s0, s1)
    return [m.beam for m in sims]
