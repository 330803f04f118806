"""Tables built by the library's host code (pt_sb_build, pt_table_build) drive the device
step exactly as the oracle's step with the same tables: a few coupled steps compared
particle by particle."""
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
from particulator_hip import air, co, tables
from particulator_hip.physics import SeltzerBerger

pytestmark = pytest.mark.gpu


def native_air_tables(efield=1e6, dt=1e-12):
    comp = air.composition()
    F = co.e * efield * dt

    def native_sb(procs):
        return [(c, SeltzerBerger(p.Z, native=True) if isinstance(p, SeltzerBerger) else p) for c, p in procs]

    te = tables.native_table(native_sb(air.electron_processes(comp, H.CUT)), 0, F, energy_cut=H.CUT)
    tg = tables.native_table(air.photon_processes(comp), 1, 0.0, energy_cut=H.CUT)
    tp = tables.native_table(air.positron_processes(comp, H.CUT), 2, F, energy_cut=H.CUT)
    return te, tg, tp


def test_native_tables_step_matches_oracle():
    te, tg, tp = native_air_tables()
    seed, n, steps = 77, 4000, 4
    init = O.init_swarm(n, 0, seed, e1=1e5, e2=1e7, loguniform=True)
    ow = H.make_oracle(te, tg, tp, seed)
    ow.add(0, init)
    ow.init()
    mp = H.make_device(te, tg, tp, seed)
    soa, uid, sw = H.soa_of(init)
    mp.add_soa(0, soa, uid=uid, swarm=sw)
    mp.init()
    print("native tables: %d electrons, %d steps" % (n, steps), flush=True)
    sys.stdout.flush()
    t = 0.0
    for _ in range(steps):
        ow.advance(t + 1e-12)
        mp.advance(t + 1e-12)
        t += 1e-12
    for sp in (0, 1):
        a = H.by_uid(ow.particles(sp))
        b = H.by_uid(mp.download(sp))
        common = np.intersect1d(a["uid"], b["uid"])
        assert common.size >= 0.99 * a.size - 2
        ia, ib = np.searchsorted(a["uid"], common), np.searchsorted(b["uid"], common)
        if common.size:
            rel = np.abs(a["p"][ia] - b["p"][ib]).max(axis=1) / np.abs(a["p"][ia]).max(axis=1)
            assert np.mean(rel <= 1e-9) >= 0.99
    mp.close()
