"""Air process lists and densities.

The process lists follow the table builders of scripts/swarm.jl:104-139 (coefficient =
density of scatterers: atoms for atomic processes, molecules for the RBEB molecular
orbitals).  Argon extends them: the reference has no argon orbitals (src/rbeb.jl:51-52),
so argon ionisation is Moller(18, tcut) with tcut equal to the energy cut; argon takes
part in every atomic process (elastic, bremsstrahlung, photoelectric, pair production,
Compton, Bhabha, annihilation).
"""
import numpy as np

from . import constants as co
from .physics import (BetheHeitler, Bhaba, Compton, Moller, ORBITALS, PhotoElectric,
                      PositronAnihilation, RelativisticCoulomb, SeltzerBerger)
from .tables import collision_table_from_processes
from ._lib import PT_ELECTRON, PT_PHOTON, PT_POSITRON


def composition(fN2=0.781, fO2=0.210, fAr=0.009, T=273.0, P=co.atm):
    """Number densities (1/m^3) of dry air at pressure P and temperature T."""
    n = P / (co.k * T)
    return {"N2": fN2 * n, "O2": fO2 * n, "Ar": fAr * n}


_SB = {}
_PE = {}


def _sb(Z):
    if Z not in _SB:
        _SB[Z] = SeltzerBerger(Z)
    return _SB[Z]


def _pe(Z):
    if Z not in _PE:
        _PE[Z] = PhotoElectric(Z)
    return _PE[Z]


def _atoms(comp):
    a = [(7, 2 * comp["N2"]), (8, 2 * comp["O2"])]
    if comp.get("Ar", 0.0) > 0:
        a.append((18, comp["Ar"]))
    return a


def electron_processes(comp, tcut):
    atoms = _atoms(comp)
    procs = [(c, RelativisticCoulomb(Z)) for Z, c in atoms]
    procs += [(c, _sb(Z)) for Z, c in atoms]
    procs += [(comp["N2"], o) for o in ORBITALS["N2"]]
    procs += [(comp["O2"], o) for o in ORBITALS["O2"]]
    if comp.get("Ar", 0.0) > 0:
        procs.append((comp["Ar"], Moller(18, tcut)))
    return procs


def positron_processes(comp, tcut):
    atoms = _atoms(comp)
    procs = [(c, RelativisticCoulomb(Z)) for Z, c in atoms]
    procs += [(c, Bhaba(Z, tcut)) for Z, c in atoms]
    procs += [(c, PositronAnihilation(Z)) for Z, c in atoms]
    return procs


def photon_processes(comp):
    atoms = _atoms(comp)
    procs = [(c, _pe(Z)) for Z, c in atoms]
    procs += [(c, BetheHeitler(Z)) for Z, c in atoms]
    procs += [(c, Compton(Z)) for Z, c in atoms]
    return procs


def swarm_tables(comp, efields, dt, cut, safety=1.1):
    """Tables for independent swarms in fields efields[i] (V/m): one set of rate fits, and
    per-swarm rate bounds of electrons and positrons (Fdt_i = e E_i dt, scripts/swarm.jl:47);
    the photon bound does not depend on the field."""
    te, tg, tp = build_tables(comp, efields[0], dt, cut, safety)
    Fdts = [co.elementary_charge * e * dt for e in efields]
    te = te.with_ratebounds(np.stack([te.ratebound_for(F, safety) for F in Fdts]))
    tp = tp.with_ratebounds(np.stack([tp.ratebound_for(F, safety) for F in Fdts]))
    return te, tg, tp


def build_tables(comp, efield, dt, cut, safety=1.1):
    """Electron, photon and positron tables (Fdt = e E dt, scripts/swarm.jl:47)."""
    Fdt = co.elementary_charge * efield * dt
    te = collision_table_from_processes(electron_processes(comp, cut), PT_ELECTRON, Fdt,
                                        safety=safety, energy_cut=cut)
    tg = collision_table_from_processes(photon_processes(comp), PT_PHOTON, 0.0, safety=safety,
                                        energy_cut=cut)
    tp = collision_table_from_processes(positron_processes(comp, cut), PT_POSITRON, Fdt,
                                        safety=safety, energy_cut=cut)
    return te, tg, tp
