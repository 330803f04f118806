"""Shared set-up of the parity tests: product tables, the oracle restatement fed the same
tables, and paired worlds (device MultiPopulation vs oracle OracleWorld)."""
import numpy as np

import oracle as O
import particulator_hip as P
from particulator_hip import air, co
from particulator_hip.physics import (RBEB, BetheHeitler, Bhaba, Compton, KleinNishinaCompton,
                                      Moller, PhotoElectric, PositronAnihilation,
                                      RelativisticCoulomb, SeltzerBerger)

CUT = 1e4 * co.eV
_TABLES = {}


def tables(efield=1e6, dt=1e-12, cut=CUT, comp=None):
    key = (efield, dt, cut, None if comp is None else tuple(sorted(comp.items())))
    if key not in _TABLES:
        _TABLES[key] = air.build_tables(comp or air.composition(), efield, dt, cut)
    return _TABLES[key]


def oracle_proc(p, coef=1.0):
    if isinstance(p, RelativisticCoulomb):
        return O.make_proc(O.ELASTIC, p.Z, coef)
    if isinstance(p, SeltzerBerger):
        sbt = O.OracleSB.from_arrays(p.Z, p.log_energy, p.tcs, p.data)
        return O.make_proc(O.SB, p.Z, coef, sbt=sbt)
    if isinstance(p, RBEB):
        return O.make_proc(O.RBEB, 0, coef, B=p.B, U=p.U, N=p.N)
    if isinstance(p, Moller):
        return O.make_proc(O.MOLLER, p.Z, coef, tcut=p.tcut)
    if isinstance(p, KleinNishinaCompton):
        return O.make_proc(O.KN_COMPTON, p.Z, coef)
    if isinstance(p, Compton):
        return O.make_proc(O.COMPTON, p.Z, coef)
    if isinstance(p, PhotoElectric):
        return O.make_proc(O.PHOTO, p.Z, coef, pet=O.pe(p.Z))
    if isinstance(p, BetheHeitler):
        return O.make_proc(O.BH, p.Z, coef)
    if isinstance(p, Bhaba):
        return O.make_proc(O.BHABHA, p.Z, coef, tcut=p.tcut)
    if isinstance(p, PositronAnihilation):
        return O.make_proc(O.ANNIH, p.Z, coef)
    raise TypeError(p)


def oracle_table(tb):
    """The oracle's view of a product table: same arrays, same process order."""
    procs = [oracle_proc(p) for p in tb.proc]
    return O.OracleTable(procs, tb.species, 0.0, rate=tb.rate, ratebound=tb.ratebound,
                         perm=np.arange(len(procs)))


def soa_of(states):
    """oracle STATE_DTYPE array -> (10, n) SoA, uid, swarm"""
    n = states.size
    soa = np.zeros((10, n))
    soa[0:3] = states["x"].T
    soa[3:6] = states["p"].T
    soa[6], soa[7], soa[8], soa[9] = states["w"], states["t"], states["s"], states["r"]
    return soa, states["uid"].astype(np.uint64), states["swarm"].astype(np.int32)


def make_device(te, tg, tp, seed, caps=(1 << 16, 1 << 16, 1 << 16), tracked=(True, True, True),
                efield=(0.0, 0.0, -1e6), cut=CUT, nswarm=1):
    pops = {}
    for name, tb, cap, tr in (("electron", te, caps[0], tracked[0]), ("photon", tg, caps[1], tracked[1]),
                              ("positron", tp, caps[2], tracked[2])):
        pops[name] = P.Population(cap, [], tb, cut, tracked=tr)
    mp = P.MultiPopulation(seed=seed, nswarm=nswarm, **pops)
    mp.set_pusher(P.RK2Pusher(P.ElectromagneticField(P.HomogeneousField(efield))))
    return mp


def make_oracle(te, tg, tp, seed, caps=(1 << 16, 1 << 16, 1 << 16), tracked=(True, True, True),
                efield=(0.0, 0.0, -1e6), cut=CUT, nthreads=1, tally=None):
    ots = [oracle_table(te), oracle_table(tg), oracle_table(tp)]
    w = O.OracleWorld(ots, seed, caps, (cut, cut, cut), tracked=tracked, nthreads=nthreads,
                      field=O.field_homogeneous(efield), tally=tally)
    w._tables_keep = ots
    return w


def by_uid(arr):
    return arr[np.argsort(arr["uid"], kind="stable")]
