"""Photon pencil beam (scripts/beam.jl-style): monoenergetic 10 MeV photons along +z (with
the 1e-6 tilt of scripts/beam.jl:56) into homogeneous dry air, no field, every secondary
tracked down to 10 keV; depth-dose tally in z and photon birth-energy tally in 512 log bins
over 10 keV - 10 MeV.  The device against the oracle on the same particles and streams."""
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
import particulator_hip as P
from particulator_hip import co

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def beam(n, seed, E=1e7):
    """n photons at E eV along +z from the origin, uid 0..n-1, s = -log(U)"""
    rng = np.random.default_rng(seed)
    st = np.zeros(n, dtype=O.STATE_DTYPE)
    pn = E * co.eV / co.c
    d = np.array([1e-6, 0.0, 1.0])
    d /= np.linalg.norm(d)
    st["p"] = pn * d
    st["w"] = 1.0
    st["s"] = -np.log(rng.random(n))
    st["active"] = True
    st["uid"] = np.arange(n, dtype=np.uint64)
    return st


DOSE = (64, 0.0, 9.0)
TALLY = (512, 1e4 * co.eV, 1e7 * co.eV)


def test_beam_dose_and_photon_tally_match_oracle():
    te, tg, tp = H.tables()
    n, steps, dt, seed = 20000, 30, 1e-9, 31
    init = beam(n, seed)
    caps = (1 << 18, 1 << 16, 1 << 16)
    ow = O.OracleWorld([H.oracle_table(te), H.oracle_table(tg), H.oracle_table(tp)], seed, caps, (H.CUT,) * 3,
                       nthreads=8, field=O.field_homogeneous((0.0, 0.0, 0.0)), dose=DOSE,
                       tally={1: (TALLY[0], TALLY[1], TALLY[2])}, tally_tracked=True)
    ow.add(1, init)
    ow.init()
    say("oracle beam: %d photons, %d steps of %g s" % (n, steps, dt))
    ow.run(0.0, dt, steps)
    pops = {nm: P.Population(cp, [], tb, H.CUT) for nm, tb, cp in
            (("electron", te, caps[0]), ("photon", tg, caps[1]), ("positron", tp, caps[2]))}
    mp = P.MultiPopulation(seed=seed, tally_bins=TALLY[0], tally_range=TALLY[1:], dose=DOSE, tally_tracked=True,
                           **pops)
    mp.set_pusher(P.RK2Pusher(P.ElectromagneticField(P.HomogeneousField((0.0, 0.0, 0.0)))))
    soa, uid, sw = H.soa_of(init)
    mp.add_soa(1, soa, uid=uid, swarm=sw)
    mp.init()
    say("device beam")
    mp.run_steps(0.0, dt, steps, compact_every=8)
    hd, ud, od, td = mp.dose()
    ho = ow.dose
    say("dose total device %.6g oracle %.6g J; under %.3g / %.3g; over %.3g / %.3g"
        % (td, ho[-1], ud, ho[DOSE[0]], od, ho[DOSE[0] + 1]))
    assert ho[-1] > 0
    assert abs(td - ho[-1]) <= 2e-3 * ho[-1]
    assert np.abs(hd - ho[:DOSE[0]]).sum() <= 1e-2 * ho[-1]
    # photon births (tracked photons tallied at birth)
    h1, c1, w1 = mp.tally(1)
    say("photon births device %d oracle %d" % (c1, ow.w.pop[1].tally_n))
    assert abs(c1 - ow.w.pop[1].tally_n) <= max(3, 0.01 * c1)
    assert np.abs(h1 - ow.hists[1]).sum() <= max(3, 0.01 * c1)
    # the particles themselves
    for sp in (0, 1, 2):
        a, b = ow.particles(sp), mp.download(sp)
        say("species %d: oracle %d device %d" % (sp, a.size, b.size))
        assert abs(a.size - b.size) <= max(3, 0.005 * a.size)
    mp.close()
