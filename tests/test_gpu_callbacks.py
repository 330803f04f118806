"""Device callbacks: WallCallback (src/callback.jl:146-184) against the oracle on the same
particles and streams, and the device roulette! / split! (src/population.jl:291-330) by
their expectations."""
import math
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
from particulator_hip import co
from particulator_hip.population import roulette_, split_

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_wall_records_match_oracle():
    te, tg, tp = H.tables()
    n, steps, seed, zw = 4000, 8, 13, 3e-4
    init = O.init_swarm(n, 0, seed, e1=1e6)
    caps = (1 << 16, 1 << 16, 1 << 16)
    ow = O.OracleWorld([H.oracle_table(te), H.oracle_table(tg), H.oracle_table(tp)], seed, caps, (H.CUT,) * 3,
                       field=O.field_homogeneous((0.0, 0.0, -1e6)), walls={0: (2, zw, True)})
    ow.add(0, init)
    ow.init()
    ow.run(0.0, 1e-12, steps)
    mp = H.make_device(te, tg, tp, seed, caps=caps)
    mp.set_launch(rounds=0)
    mp.set_wall(0, 2, zw, drop=True, capacity=1 << 16)
    soa, uid, sw = H.soa_of(init)
    mp.add_soa(0, soa, uid=uid, swarm=sw)
    mp.init()
    say("device with a wall at z = %g m, %d steps" % (zw, steps))
    mp.run_steps(0.0, 1e-12, steps, compact_every=4)
    a = H.by_uid(ow.wall_records())
    b, nb = mp.wall_records()
    b = H.by_uid(b)
    say("wall records: oracle %d device %d" % (a.size, b.size))
    assert a.size > 100
    assert abs(a.size - b.size) <= 3
    common = np.intersect1d(a["uid"], b["uid"])
    x = a[np.isin(a["uid"], common)]
    y = b[np.isin(b["uid"], common)]
    assert common.size >= 0.99 * a.size - 3
    rel = np.abs(x["p"] - y["p"]).max(axis=1) / np.abs(x["p"]).max(axis=1)
    assert np.mean(rel <= 1e-9) >= 0.99
    assert np.allclose(y["x"][:, 2], zw, rtol=1e-12, atol=0)
    assert np.all(y["species"] == 0)
    # dropped: no recorded electron is still active
    alive = set(mp.download(0)["uid"].tolist())
    assert not (alive & set(b["uid"].tolist()))
    mp.close()


def fresh(n, seed):
    te, tg, tp = H.tables()
    mp = H.make_device(te, tg, tp, seed, caps=(1 << 22, 1 << 16, 1 << 16), tracked=(True, False, False))
    mp.init_particles(0, n, 1e6 * co.eV)
    mp.init()
    return mp


def test_device_roulette():
    n, q = 1_000_000, 0.3
    mp = fresh(n, 5)
    roulette_(q, mp.get("electron"))
    st = mp.stats(0)
    sd = math.sqrt(n * q * (1 - q))
    say("roulette: %d of %d survive (expected %.0f +- %.0f); weight %.1f" % (st.nactive, n, n * q, sd, st.weight))
    assert abs(st.nactive - n * q) <= 5 * sd
    assert abs(st.weight - st.nactive / q) <= 1e-6 * st.weight
    assert abs(st.weight - n) <= 5 * sd / q
    # the survivors depend only on (seed, uid, step)
    mp2 = fresh(n, 5)
    roulette_(q, mp2.get("electron"))
    assert np.array_equal(np.sort(mp.download(0)["uid"]), np.sort(mp2.download(0)["uid"]))
    mp.close()
    mp2.close()


def test_device_split():
    n, f = 100_000, 2.5
    mp = fresh(n, 6)
    split_(f, mp.get("electron"))
    st = mp.stats(0)
    sd = math.sqrt(n * f)
    say("split: %d particles (expected %.0f +- %.0f); weight %.2f of %d" % (st.nactive, n * (1 + f), sd, st.weight, n))
    assert abs(st.nactive - n * (1 + f)) <= 5 * sd
    assert abs(st.weight - n) <= 5 * sd / (1 + f)
    arr = mp.download(0)
    assert np.unique(arr["uid"]).size == arr.size
    assert np.allclose(arr["w"], 1.0 / (1 + f))
    mp.close()
