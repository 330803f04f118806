"""Statistical parity of the device step with the oracle on independent random streams
(different seeds): energy and angle spectra by two-sample KS distance, and avalanche
growth rate and mean energy over 32 seeds.

Workload: the runaway swarm in dry air (N2:O2:Ar 0.781:0.210:0.009, 1 atm, 273 K) at
E = 1 MV/m along -z, 1 MeV isotropic electrons, 10 keV cut, dt = 1 ps, photons tallied at
birth (not tracked).
"""
import math
import os
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
from particulator_hip import co
from particulator_hip.population import kinenergy

pytestmark = pytest.mark.gpu
THREADS = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 16)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ks2(a, b):
    a, b = np.sort(a), np.sort(b)
    allv = np.concatenate([a, b])
    fa = np.searchsorted(a, allv, side="right") / a.size
    fb = np.searchsorted(b, allv, side="right") / b.size
    return np.abs(fa - fb).max()


def swarm_device(n, steps, seed, tracked=(True, False, False)):
    te, tg, tp = H.tables()
    mp = H.make_device(te, tg, tp, seed, caps=(4 * n + 1024, 1024, 4 * n + 1024), tracked=tracked)
    mp.init_particles(0, n, 1e6 * co.eV, loguniform=False)
    mp.init()
    say("device swarm n=%d steps=%d seed=%d" % (n, steps, seed))
    mp.run_steps(0.0, 1e-12, steps, compact_every=8, check_every=steps)
    return mp


def swarm_oracle(n, steps, seed, tracked=(True, False, False)):
    te, tg, tp = H.tables()
    ow = H.make_oracle(te, tg, tp, seed, caps=(4 * n + 1024, 1024, 4 * n + 1024), tracked=tracked,
                       nthreads=THREADS)
    ow.add(0, O.init_swarm(n, 0, seed, e1=1e6))
    ow.init()
    ow.run(0.0, 1e-12, steps)
    return ow


def spectra(p):
    E = kinenergy(0, p) / co.eV
    cz = -p[:, 2] / np.linalg.norm(p, axis=1)  # cosine to the accelerating direction (+z force)
    return E, cz


def test_spectra_ks_independent_seeds():
    n, steps = 200_000, 100
    mp = swarm_device(n, steps, seed=101)
    a = mp.download(0)
    ea, ca = spectra(a["p"])
    mp.close()
    say("oracle swarm n=%d steps=%d threads=%d" % (n, steps, THREADS))
    ow = swarm_oracle(n, steps, seed=202)
    b = ow.particles(0)
    eb, cb = spectra(b["p"])
    dE, dC = ks2(ea, eb), ks2(ca, cb)
    floor = 1.36 * math.sqrt(1 / ea.size + 1 / eb.size)
    say("KS energy %.4f angle %.4f (n = %d / %d, 95%% noise floor %.4f)" % (dE, dC, ea.size, eb.size, floor))
    assert dE <= 0.01 and dC <= 0.01
    # same seed: the same particles, so the spectra coincide far below the noise
    ow2 = swarm_oracle(20_000, 20, seed=303)
    mp2 = swarm_device(20_000, 20, seed=303)
    e1, _ = spectra(ow2.particles(0)["p"])
    e2, _ = spectra(mp2.download(0)["p"])
    mp2.close()
    assert ks2(e1, e2) < 2e-3


def _growth_and_energy(counts, energies, T):
    return math.log(counts[-1] / counts[0]) / T, energies[-1]


def test_growth_rate_and_mean_energy_32_seeds():
    n, steps, dt = 2000, 400, 1e-12
    dev_g, dev_e, cpu_g, cpu_e = [], [], [], []
    for k in range(32):
        mp = swarm_device(n, steps, seed=1000 + k)
        st = mp.stats(0)
        dev_g.append(math.log(st.nactive / n) / (steps * dt))
        dev_e.append(st.mean_energy / co.eV)
        mp.close()
        ow = swarm_oracle(n, steps, seed=5000 + k)
        s = ow.stats(0)
        cpu_g.append(math.log(s[0] / n) / (steps * dt))
        cpu_e.append(s[2] / co.eV)
    for name, a, b in (("growth rate (1/s)", dev_g, cpu_g), ("mean energy (eV)", dev_e, cpu_e)):
        a, b = np.array(a), np.array(b)
        sig = math.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
        say("%s: device %.5g +- %.2g, oracle %.5g +- %.2g, |diff| / sigma = %.2f"
            % (name, a.mean(), a.std(ddof=1) / math.sqrt(a.size), b.mean(),
               b.std(ddof=1) / math.sqrt(b.size), abs(a.mean() - b.mean()) / sig))
        assert abs(a.mean() - b.mean()) <= 2 * sig


def test_photon_birth_tally_matches_oracle():
    """Photons tallied at birth (not tracked): counts and 512-bin log spectra.  10 MeV
    primaries so that enough bremsstrahlung photons above the cut are born."""
    n, steps = 40_000, 200
    te, tg, tp = H.tables()
    tracked = (True, False, False)
    mp = H.make_device(te, tg, tp, 77, caps=(4 * n, 1024, 4 * n), tracked=tracked)
    mp.init_particles(0, n, 1e7 * co.eV)
    mp.init()
    mp.run_steps(0.0, 1e-12, steps, compact_every=8, check_every=steps)
    h, cnt, w = mp.tally(1)
    mp.close()
    ow = H.make_oracle(te, tg, tp, 88, caps=(4 * n, 1024, 4 * n), tracked=tracked, nthreads=THREADS,
                       tally={1: (512, 1e4 * co.eV, 1e7 * co.eV)})
    ow.add(0, O.init_swarm(n, 0, 88, e1=1e7))
    ow.init()
    ow.run(0.0, 1e-12, steps)
    ocnt = ow.w.pop[1].tally_n
    say("photon births: device %d oracle %d" % (cnt, ocnt))
    assert cnt > 50 and ocnt > 50
    assert abs(cnt - ocnt) <= 4 * math.sqrt(cnt + ocnt)
    assert 0.9 * cnt <= h.sum() <= cnt  # weights are 1; births above the last bin are not binned
