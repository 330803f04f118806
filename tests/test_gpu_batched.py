"""Batched swarms on the device: independent swarms at different fields in one population
(per-swarm field and rate bound), the captured (hipGraph) run, launch shapes and the
autotuner, and per-swarm statistics.

Parity bars: per-swarm step parity against the oracle as in the single-swarm step test
(>= 99% of particles equal to 1e-9 after several steps); captured runs, launch shapes and
the autotuner's trial steps must give bit-identical particles (matched by uid) to the
plain eager run, since every random stream is keyed by (seed, uid, step).
"""
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
import particulator_hip as P
from particulator_hip import air, co
from particulator_hip.population import kinenergy

pytestmark = pytest.mark.gpu

FIELDS = [0.4e6, 1e6, 2e6, 4e6]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def swarm_device(tables, seed, nswarm, fields, caps=(1 << 18, 1 << 16, 1 << 16), tracked=(True, True, True)):
    te, tg, tp = tables
    pops = {}
    for name, tb, cap, tr in (("electron", te, caps[0], tracked[0]), ("photon", tg, caps[1], tracked[1]),
                              ("positron", tp, caps[2], tracked[2])):
        pops[name] = P.Population(cap, [], tb, H.CUT, tracked=tr)
    mp = P.MultiPopulation(seed=seed, nswarm=nswarm, **pops)
    for i, E in enumerate(fields):
        mp.set_pusher(P.RK2Pusher(P.ElectromagneticField(P.HomogeneousField((0.0, 0.0, -E)))), swarm=i)
    return mp


def fill(mp, n_each, seed, loguniform=True):
    for i in range(mp.cfg.nswarm):
        init = O.init_swarm(n_each, 0, seed, uid0=i * n_each, swarm=i, e1=1e5, e2=1e7, loguniform=loguniform)
        soa, uid, sw = H.soa_of(init)
        mp.add_soa(0, soa, uid=uid, swarm=sw)
    mp.init()


def identical(a, b):
    """bit-identical particle sets (matched by uid)"""
    a, b = H.by_uid(a), H.by_uid(b)
    if a.size != b.size or not np.array_equal(a["uid"], b["uid"]):
        return False
    for f in ("x", "p", "w", "t", "s", "r", "swarm"):
        if not np.array_equal(a[f], b[f]):
            return False
    return True


def test_swarms_at_different_fields_match_oracle():
    tables = air.swarm_tables(air.composition(), FIELDS, 1e-12, H.CUT)
    te, tg, tp = tables
    seed, n_each, steps, dt = 11, 1000, 6, 1e-12
    ots = [H.oracle_table(te), H.oracle_table(tg), H.oracle_table(tp)]
    caps = (1 << 16, 1 << 16, 1 << 16)
    ow = O.OracleWorld(ots, seed, caps, (H.CUT,) * 3, fields=[O.field_homogeneous((0.0, 0.0, -E)) for E in FIELDS])
    for i in range(len(FIELDS)):
        ow.add(0, O.init_swarm(n_each, 0, seed, uid0=i * n_each, swarm=i, e1=1e5, e2=1e7, loguniform=True))
    ow.init()
    say("device: %d swarms x %d e-, %d steps" % (len(FIELDS), n_each, steps))
    mp = swarm_device(tables, seed, len(FIELDS), FIELDS, caps=caps)
    fill(mp, n_each, seed)
    t = 0.0
    for _ in range(steps):
        ow.advance(t + dt)
        mp.advance(t + dt)
        t += dt
    a = H.by_uid(ow.particles(0))
    b = H.by_uid(mp.download(0))
    for i in range(len(FIELDS)):
        ai, bi = a[a["swarm"] == i], b[b["swarm"] == i]
        common = np.intersect1d(ai["uid"], bi["uid"])
        x = ai[np.isin(ai["uid"], common)]
        y = bi[np.isin(bi["uid"], common)]
        rel = np.abs(x["p"] - y["p"]).max(axis=1) / np.abs(x["p"]).max(axis=1)
        ok = int((rel <= 1e-9).sum())
        say("swarm %d (E = %.2g V/m): oracle %d device %d common %d within 1e-9 %d"
            % (i, FIELDS[i], ai.size, bi.size, common.size, ok))
        assert abs(ai.size - bi.size) <= 3
        assert common.size >= 0.99 * max(ai.size, bi.size) - 3
        assert ok >= 0.99 * common.size - 3
    # per-swarm statistics against the downloaded particles
    st = mp.swarm_stats(0)
    for i in range(len(FIELDS)):
        bi = b[b["swarm"] == i]
        E = kinenergy(0, bi["p"])
        assert st[i, 0] == bi.size
        assert st[i, 1] == pytest.approx(bi["w"].sum(), rel=1e-12)
        assert st[i, 2] == pytest.approx((bi["w"] * E).sum(), rel=1e-12)
        assert st[i, 3] == E.max()
    mp.close()


def _pair(seed=5, n_each=20000, nswarm=4):
    tables = air.swarm_tables(air.composition(), FIELDS[:nswarm], 1e-12, H.CUT)
    out = []
    for _ in range(2):
        mp = swarm_device(tables, seed, nswarm, FIELDS[:nswarm])
        fill(mp, n_each, seed)
        out.append(mp)
    return out


def test_graph_replays_equal_eager_steps():
    dt = 1e-12
    eager, graph = _pair()
    eager.set_launch(block=256, grid_mult=1, compact_every=8, rounds=0)
    graph.set_launch(block=256, grid_mult=1, compact_every=8, rounds=3, defer_mask=0x7FE)
    t = 0.0
    for nsteps in (8, 16, 11):
        say("eager %d steps from step %d" % (nsteps, eager.step))
        eager.run_steps(t, dt, nsteps, compact_every=8, check_every=nsteps)
        say("graph %d steps from step %d" % (nsteps, graph.step))
        graph.run_graph(t, dt, nsteps, check_every=8)
        t += nsteps * dt
        assert eager.step == graph.step
        for sp in (0, 1, 2):
            a, b = eager.download(sp), graph.download(sp)
            say("  species %d: %d / %d particles" % (sp, a.size, b.size))
            assert identical(a, b)
        # the device clock reached the same time as the host loop
        tg = graph.download(0)["t"]
        assert tg.size and np.all(np.abs(tg - t) <= 2.5e-16)
    assert eager.flags()[0] == 0 and graph.flags()[0] == 0
    eager.close()
    graph.close()


def test_launch_shape_does_not_change_particles():
    dt = 1e-12
    a, b = _pair(seed=6, n_each=10000)
    a.set_launch(block=128, grid_mult=1, compact_every=8, rounds=0)
    b.set_launch(block=512, grid_mult=4, compact_every=8, rounds=4, defer_mask=1 << 2)
    say("block 128 x1 history mode vs block 512 x4 with 4 collision rounds, 10 steps")
    a.run_steps(0.0, dt, 10, compact_every=3)
    b.run_graph(0.0, dt, 10)
    for sp in (0, 1, 2):
        assert identical(a.download(sp), b.download(sp))
    a.close()
    b.close()


def test_autotune_keeps_the_run_exact():
    dt = 1e-12
    a, b = _pair(seed=7, n_each=50000)
    say("autotune, 2 steps per shape")
    best, trials, tend = a.autotune(0.0, dt, trial_steps=2)
    say("best %s" % best)
    for tr in trials:
        say("  %s" % tr)
    nsteps = 4 * 2 + 9 * 2 + 5 * 16
    assert len(trials) == 18
    assert best["rounds"] in (0, 2, 3, 4)
    assert all(tr["rate"] > 0 for tr in trials)
    assert best["block"] in (128, 256, 512) and best["grid_mult"] in (1, 2, 4)
    assert best["compact_every"] in (1, 2, 4, 8, 16)
    assert a.get_launch() == best
    assert a.step == nsteps
    assert tend == pytest.approx(nsteps * dt, rel=1e-12)
    b.run_steps(0.0, dt, nsteps, compact_every=8)
    for sp in (0, 1, 2):
        assert identical(a.download(sp), b.download(sp))
    a.close()
    b.close()
