"""The oracle's depth-dose bookkeeping (build-defined deposition: particles dropped below the
cut, secondaries born below it, binding energies) closes the energy balance where the
reference's kinematics conserve energy: electrons with elastic scattering, RBEB and Moller,
photons with Compton and photoabsorption, no field.  (Bremsstrahlung builds the electron
from p - p_gamma, src/seltzer.jl:88, and pair production creates rest mass, so both are left
out of the balance.)"""
import numpy as np

import oracle as O
from particulator_hip import co
from particulator_hip.population import kinenergy


def test_dose_tally_closes_the_energy_balance(O):
    comp = O.air_composition()
    cut = 1e4 * co.eV
    keep_e = (O.ELASTIC, O.RBEB, O.MOLLER)
    keep_g = (O.COMPTON, O.PHOTO)
    pe = [p for p in O.air_processes(O.ELECTRON, comp, cut) if p.kind in keep_e]
    pg = [p for p in O.air_processes(O.PHOTON, comp, cut) if p.kind in keep_g]
    te = O.OracleTable(pe, O.ELECTRON, 0.0)
    tg = O.OracleTable(pg, O.PHOTON, 0.0)
    tp = O.OracleTable(O.air_processes(O.POSITRON, comp, cut), O.POSITRON, 0.0)
    w = O.OracleWorld([te, tg, tp], 7, (1 << 16, 1 << 16, 1 << 16), (cut, cut, cut), nthreads=4,
                      field=O.field_homogeneous((0.0, 0.0, 0.0)), dose=(64, -0.02, 0.02))
    e0 = O.init_swarm(3000, O.ELECTRON, 7, uid0=0, e1=1.5e4, e2=1e5, loguniform=True)
    g0 = O.init_swarm(3000, O.PHOTON, 7, uid0=10_000, e1=1.5e4, e2=1e5, loguniform=True)
    w.add(O.ELECTRON, e0)
    w.add(O.PHOTON, g0)
    w.init()
    ein = kinenergy(0, e0["p"]).sum() + kinenergy(1, g0["p"]).sum()
    w.run(0.0, 1e-12, 200)
    eout = w.dose[-1] + sum(kinenergy(s, w.particles(s)["p"]).sum() for s in (0, 1))
    assert w.dose[-1] > 0.1 * ein  # a good part of the energy was deposited
    assert abs(eout - ein) <= 1e-9 * ein
    # the histogram (with under- and overflow) holds the total
    assert abs(w.dose[:-1].sum() - w.dose[-1]) <= 1e-9 * w.dose[-1]


def test_wall_records_lie_on_the_wall(O):
    """WallCallback (src/callback.jl:146-184) in the oracle: photons of a pencil beam cross
    z = 0.1 m once; the recorded lincomb state lies on the plane at t = z / (c cos theta),
    with the momentum unchanged (no field), and the particle is dropped."""
    comp = O.air_composition()
    cut = 1e4 * co.eV
    tabs = [O.OracleTable(O.air_processes(s, comp, cut), s, 0.0) for s in (0, 1, 2)]
    n = 2000
    st = np.zeros(n, dtype=O.STATE_DTYPE)
    d = np.array([1e-6, 0.0, 1.0]) / np.sqrt(1.0 + 1e-12)
    pn = 1e7 * co.eV / co.c
    st["p"] = pn * d
    st["w"] = 1.0
    st["s"] = -np.log(np.random.default_rng(2).random(n))
    st["active"] = True
    st["uid"] = np.arange(n, dtype=np.uint64)
    w = O.OracleWorld(tabs, 3, (1 << 14, 1 << 14, 1 << 14), (cut, cut, cut),
                      field=O.field_homogeneous((0.0, 0.0, 0.0)), walls={1: (2, 0.1, True)})
    w.add(1, st)
    w.init()
    w.run(0.0, 1e-9, 2)
    rec = w.wall_records()
    assert rec.size >= 0.99 * n and rec.size <= n
    assert np.unique(rec["uid"]).size == rec.size
    assert np.allclose(rec["x"][:, 2], 0.1, rtol=1e-12, atol=0)
    # photons that did not scatter before the wall (nearly all: the mean free path is
    # ~380 m) keep their momentum and reach it at z / (c cos theta)
    straight = np.all(np.abs(rec["p"] - pn * d) <= 1e-12 * pn, axis=1)
    assert straight.sum() >= 0.99 * n
    assert np.allclose(rec["t"][straight], 0.1 / (co.c * d[2]), rtol=1e-12, atol=0)
    alive = set(w.particles(1)["uid"].tolist())
    assert not (alive & set(rec["uid"].tolist()))
