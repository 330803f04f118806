"""Pins of the CPU oracle against what fixes it independently of its own code: published
known-answer vectors, closed forms, textbook cross sections and their integrals,
independent data (NIST XCOM), conservation laws.  The reference has no tests or fixtures
of its own (test/runtests.jl is empty), so these are the oracle's anchors."""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O


@pytest.fixture(scope="module")
def co(O):
    return O.co()


# ------------------------------------------------------------------ Philox
def test_philox_known_answers(O):
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniforms_open_interval_and_uniform(O):
    lib = O.lib()
    assert lib.or_u01(0) == 0.5 * 2.0 ** -52
    assert lib.or_u01(2 ** 64 - 1) == (2.0 ** 52 - 0.5) * 2.0 ** -52 < 1.0
    g = O.Uniforms(seed=9, uid=3, step=1)
    u = np.array([g() for _ in range(200000)])
    from scipy.stats import kstest
    assert kstest(u, "uniform").pvalue > 1e-3
    assert 0.0 < u.min() and u.max() < 1.0


def test_child_uids_distinct(O):
    lib = O.lib()
    uids = {lib.or_child_uid(1, parent, 0, d, k) for parent in range(50) for d in range(40) for k in (0, 1)}
    assert len(uids) == 50 * 40 * 2


# ------------------------------------------------------------------ binary intervals
def test_precheb_closed_form(O, co):
    xmax = 1e3 * co.eV * 2 ** 18
    k = 32
    for m in range(-6, 18):  # interval edges 2^m keV
        E = 1e3 * co.eV * 2.0 ** m
        i, t = O.precheb(E)
        assert i == m + 15 and t[1] == -1.0 and t[2] == 1.0
        # just below the edge: previous interval, xi -> +1
        i2, t2 = O.precheb(np.nextafter(E, 0))
        assert i2 == m + 14 and 1.0 - t2[1] < 1e-14
        lo, hi = C.c_double(), C.c_double()
        O.lib().or_interval(k, xmax, i, C.byref(lo), C.byref(hi))
        assert lo.value == E and hi.value == 2 * E
    # interval 0 covers (0, xmax 2^-32) linearly
    i, t = O.precheb(xmax * 2.0 ** -33)
    assert i == 0 and t[1] == 0.0
    rng = np.random.default_rng(0)
    for xi in rng.uniform(-1, 1, 100):
        E = (1e3 * co.eV) * (1.5 + xi / 2.0)  # interval [1, 2) keV mapped to [-1, 1)
        i, t = O.precheb(E)
        assert i == 15
        assert abs(t[1] - xi) < 1e-14
        assert abs(t[2] - (2 * t[1] * t[1] - 1)) == 0.0


def test_chebfit_interpolates_at_nodes(O, co):
    # a quadratic is reproduced everywhere; a smooth function at the three nodes
    p = O.make_proc(O.ELASTIC, 7, coef=1.0)
    xmax = 1e3 * co.eV * 2 ** 18
    out = np.zeros((33, 3))
    O.lib().or_chebfit_proc(C.byref(p), 0, 32, xmax, 3, O.dptr(out))
    for i in (10, 16, 25, 31):
        lo, hi = xmax * 2.0 ** (i - 33), xmax * 2.0 ** (i - 32)
        for q in range(3):
            x = (lo + hi) / 2 + (hi - lo) / 2 * math.cos((2 * q + 1) * math.pi / 6)
            xi = (2 * x - (lo + hi)) / (hi - lo)
            val = out[i, 0] + out[i, 1] * xi + out[i, 2] * (2 * xi * xi - 1)
            want = O.lib().or_speed(0, x) * O.totalcs(p, x)
            assert abs(val - want) <= 1e-12 * abs(want)


# ------------------------------------------------------------------ cross sections
def kn_textbook(k, re):
    """Klein-Nishina per electron (Heitler's form), k = E / mc^2"""
    a = 1.0 + 2.0 * k
    return 2 * math.pi * re ** 2 * ((1 + k) / k ** 2 * (2 * (1 + k) / a - math.log(a) / k) +
                                    math.log(a) / (2 * k) - (1 + 3 * k) / a ** 2)


def test_klein_nishina(O, co):
    p = O.make_proc(O.KN_COMPTON, 1)
    assert abs(O.totalcs(p, 1e6 * co.eV) - 0.2112e-28) < 0.0005e-28  # 0.2112 b at 1 MeV
    for E in (5e4, 2e5, 1e6, 5e6, 5e7):
        k = E * co.eV / co.mc2
        assert abs(O.totalcs(p, E * co.eV) / kn_textbook(k, co.r_e) - 1) < 1e-9
    # Thomson limit of the textbook form
    assert abs(kn_textbook(1e-3, co.r_e) / (8 * math.pi / 3 * co.r_e ** 2) - 1) < 3e-3
    # Geant4's empirical Compton is Klein-Nishina times Z above ~1 MeV
    g = O.make_proc(O.COMPTON, 7)
    for E in (1e6, 1e7):
        assert abs(O.totalcs(g, E * co.eV) / (7 * O.totalcs(p, E * co.eV)) - 1) < 0.03


def _xcom(Z):
    path = os.path.join(os.path.dirname(O.HERE), "tests", "golden", "xcom-Z%d.txt" % Z)
    return np.loadtxt(path)


@pytest.mark.parametrize("Z", [7, 8])
def test_photon_cross_sections_against_xcom(O, co, Z):
    """Sandia photoelectric, G4 Compton and G4 pair cross sections vs NIST XCOM (barn/atom)."""
    x = _xcom(Z)
    barn = 1e-28
    pe, cp, bh = O.make_proc(O.PHOTO, Z), O.make_proc(O.COMPTON, Z), O.make_proc(O.BH, Z)
    for row in x:
        E = row[0] * 1e6 * co.eV
        if 2e4 <= row[0] * 1e6 <= 1e6:
            assert abs(O.totalcs(pe, E) / (row[3] * barn) - 1) < 0.15, row
        if 1e5 <= row[0] * 1e6:
            assert abs(O.totalcs(cp, E) / (row[2] * barn) - 1) < 0.05, row
        if 5e6 <= row[0] * 1e6:
            assert abs(O.totalcs(bh, E) / ((row[4] + row[5]) * barn) - 1) < 0.1, row


def test_pair_production_threshold(O, co):
    for Z in (7, 8, 18):
        p = O.make_proc(O.BH, Z)
        assert O.totalcs(p, 2 * co.mc2 * 0.999) == 0.0
        assert O.totalcs(p, 2 * co.mc2 * 1.01) > 0.0
        lim = 1.5e6 * co.eV
        a, b = O.totalcs(p, lim * (1 - 1e-12)), O.totalcs(p, lim * (1 + 1e-12))
        assert abs(a / b - 1) < 1e-9


@pytest.mark.parametrize("T", [2e4, 1e5, 1e6, 1e7])
def test_rbeb_total_is_integral_of_differential(O, co, T):
    for B, U, N in O.N2_ORBITALS + O.O2_ORBITALS:
        B, U = B * co.eV, U * co.eV
        p = O.make_proc(O.RBEB, 0, B=B, U=U, N=1)
        t = T * co.eV / B
        f = lambda s: O.lib().or_rbeb_dsdw((math.exp(s) - 1) * B, T * co.eV, B, U) * math.exp(s)
        smax = math.log((t + 1) / 2)
        I = sum(quad(f, a, b, limit=200, epsrel=1e-12)[0]
                for a, b in zip(np.linspace(0, smax, 9)[:-1], np.linspace(0, smax, 9)[1:]))
        assert abs(I / O.totalcs(p, T * co.eV) - 1) < 1e-8


def moller_dsde(eps, g):
    """Moller (G4 Physics Reference Manual) per atom / (2 pi re^2 Z / (beta^2 (g-1)))"""
    q = (2 * g - 1) / g ** 2
    return (g - 1) ** 2 / g ** 2 + 1 / eps * (1 / eps - q) + 1 / (1 - eps) * (1 / (1 - eps) - q)


def bhabha_dsde(eps, g):
    """Bhabha (G4 Physics Reference Manual): 1/(beta^2 eps^2) - B1/eps + B2 - B3 eps + B4 eps^2"""
    y = 1 / (g + 1)
    B1 = 2 - y * y
    B2 = (1 - 2 * y) * (3 + y * y)
    B4 = (1 - 2 * y) ** 3
    B3 = B4 + (1 - 2 * y) ** 2
    b2 = 1 - 1 / g ** 2
    return 1 / (b2 * eps ** 2) - B1 / eps + B2 - B3 * eps + B4 * eps ** 2


@pytest.mark.parametrize("E", [3e4, 2e5, 3e6, 5e7])
def test_moller_bhabha_totals(O, co, E):
    tcut = 1e4 * co.eV
    eng = E * co.eV
    g = 1 + eng / co.mc2
    b2 = 1 - 1 / g ** 2
    x = tcut / eng
    m = O.make_proc(O.MOLLER, 7, tcut=tcut)
    I = quad(lambda e: moller_dsde(e, g), x, 0.5, limit=200)[0]
    assert abs(O.totalcs(m, eng) / (2 * math.pi * co.r_e ** 2 * 7 * I / (b2 * (g - 1))) - 1) < 1e-8
    b = O.make_proc(O.BHABHA, 7, tcut=tcut)
    I = quad(lambda e: bhabha_dsde(e, g), x, 1.0, limit=200)[0]
    assert abs(O.totalcs(b, eng) / (2 * math.pi * co.r_e ** 2 * 7 * I / (g - 1)) - 1) < 1e-8


@pytest.mark.parametrize("E", [1e4, 1e5, 1e6, 1e7])
def test_elastic_total_is_integral_of_screened_rutherford(O, co, E):
    """sigma = 4 pi int_0^1 Z^2 re^2 (1 - beta^2 x) / (4 beta^4 gamma^2 (x + alpha)^2) dx,
    x = sin^2(theta/2), with Moliere screening alpha = hbar^2 / (4 p^2 a^2)."""
    for Z in (7, 8, 18):
        p = O.make_proc(O.ELASTIC, Z)
        K = E * co.eV + 1e-4 * co.eV
        a = 1.3413 * Z ** (-1 / 3) * co.a_0
        pm = math.sqrt(K * (K + 2 * co.mc2)) / co.c
        g = 1 + K / co.mc2
        beta = pm / (g * co.m_e * co.c)
        al = co.hbar ** 2 / (4 * pm ** 2 * a ** 2)
        # x = al (e^s - 1): smooth integrand in s
        f = lambda s: (1 - beta ** 2 * al * (math.exp(s) - 1)) / (al * math.exp(s)) ** 2 * al * math.exp(s)
        smax = math.log(1 + 1 / al)
        I = sum(quad(f, a, b, limit=200, epsrel=1e-12)[0]
                for a, b in zip(np.linspace(0, smax, 17)[:-1], np.linspace(0, smax, 17)[1:]))
        want = 4 * math.pi * Z ** 2 * co.r_e ** 2 / (4 * beta ** 4 * g ** 2) * I
        assert abs(O.totalcs(p, E * co.eV) / want - 1) < 1e-6


@pytest.mark.parametrize("E", [1e4, 1e5, 1e6, 1e7])
def test_annihilation_total_is_half_heitler_integral(O, co, E):
    g = 1 + E * co.eV / co.mc2
    sq = math.sqrt((g - 1) / (g + 1))
    S = lambda e: -(g + 1) ** 2 + (g * g + 4 * g + 1) / e - 1 / e ** 2
    I = quad(lambda e: S(e) + S(1 - e), (1 - sq) / 2, (1 + sq) / 2, limit=200)[0]
    p = O.make_proc(O.ANNIH, 7)
    want = 7 * 0.5 * math.pi * co.r_e ** 2 / ((g + 1) * (g * g - 1)) * I
    assert abs(O.totalcs(p, E * co.eV) / want - 1) < 1e-8


def test_sb_inverse_cdf_newton_vs_closed_form(O, co):
    """The oracle's Newton inversion (the reference's method) agrees with the exact inverse of
    the piecewise-quadratic cumulative."""
    from particulator_hip.physics import sb_inverse_cdf
    sbo = O.sb(7)
    logk = np.log(sbo.x)
    pc = np.arange(1000) / 999.0
    for j in (0, 10, 30, 56):
        T1 = math.exp(sbo.log_energy[j])
        q, _ = sb_inverse_cdf(logk, sbo.vals[j], pc, math.log(1e2 * co.eV / T1), math.log(5e7 * co.eV / T1))
        assert np.abs(q - sbo.data[j]).max() < 1e-9


# ------------------------------------------------------------------ samplers (KS vs textbook pdfs)
def _ks(samples, cdf_grid_x, cdf_grid_y):
    s = np.sort(samples)
    F = np.interp(s, cdf_grid_x, cdf_grid_y)
    n = s.size
    d = max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n))
    return d


def _numcdf(pdf, a, b, m=20001, log=False):
    x = np.geomspace(a, b, m) if log else np.linspace(a, b, m)
    f = np.array([pdf(v) for v in x])
    F = np.concatenate([[0], np.cumsum((f[1:] + f[:-1]) / 2 * np.diff(x))])
    return x, F / F[-1]


def _draws(fn, n, seed):
    g = O.Uniforms(seed=seed, uid=1, step=0)
    return np.array([fn(g) for _ in range(n)])


N_KS = 60000
KS_BAR = 2.0 / math.sqrt(N_KS)  # well above the 95% noise of one sample against a cdf


@pytest.mark.parametrize("T", [1e4, 1e6, 1e8])
def test_tsai_sampler_ks(O, co, T):
    eng = T * co.eV
    umax = 2 * (1 + eng / co.mc2)
    from scipy.stats import gamma
    # u = uu * a with uu ~ Gamma(2), a = 1.6 (w.p. 1/4) or 1.6/3, truncated at umax
    Fm = lambda u: 0.25 * gamma.cdf(u / 1.6, 2) + 0.75 * gamma.cdf(u * 3 / 1.6, 2)
    u = _draws(lambda g: O.lib().or_sample_tsai(eng, g.ref), N_KS, 1)
    u = np.sqrt((1 - u) * umax ** 2 / 2)
    x = np.linspace(0, umax, 20001)
    assert _ks(u, x, Fm(x) / Fm(umax)) < KS_BAR


@pytest.mark.parametrize("E", [1e4, 1e6])
def test_elastic_sampler_ks(O, co, E):
    K = E * co.eV
    a = 1.3413 * 7 ** (-1 / 3) * co.a_0
    pm = math.sqrt(K * (K + 2 * co.mc2)) / co.c
    g = 1 + K / co.mc2
    beta = pm / (g * co.m_e * co.c)
    al = co.hbar ** 2 / (4 * pm ** 2 * a ** 2)
    c = _draws(lambda g_: O.lib().or_sample_rel_sr(al, beta, g_.ref), N_KS, 2)
    xs = (1 - c) / 2
    F = lambda x: ((1 + al * beta ** 2) * (1 / al - 1 / (x + al)) - beta ** 2 * np.log((x + al) / al))
    grid = np.geomspace(1e-30, 1, 40001)
    assert _ks(xs, grid, F(grid) / F(1.0)) < KS_BAR


@pytest.mark.parametrize("T", [2e4, 1e6])
def test_rbeb_sampler_ks(O, co, T):
    B, U, N = O.N2_ORBITALS[1]
    B, U = B * co.eV, U * co.eV
    eng = T * co.eV
    w = _draws(lambda g: O.lib().or_rbeb_sample(eng, B, g.ref), N_KS, 3) / B
    t = eng / B
    x, F = _numcdf(lambda v: O.lib().or_rbeb_dsdw(v * B, eng, B, U), 1e-9, (t - 1) / 2, log=True)
    assert _ks(w, x, F) < KS_BAR


@pytest.mark.parametrize("E", [3e4, 1e6, 1e8])
def test_moller_sampler_ks(O, co, E):
    tcut, eng = 1e4 * co.eV, E * co.eV
    g = 1 + eng / co.mc2
    eps = _draws(lambda u: O.lib().or_moller_sample(tcut, eng, u.ref), N_KS, 4) / eng
    x, F = _numcdf(lambda e: moller_dsde(e, g), tcut / eng, 0.5, log=True)
    assert _ks(eps, x, F) < KS_BAR


@pytest.mark.parametrize("E", [3e4, 1e6, 1e8])
def test_bhabha_sampler_ks(O, co, E):
    """The reference's sampler (src/bhaba.jl:55-81) accepts with g1(e)/g1(e0) where
    g1 = B0 + B1 e + ... + B4 e^4 has only non-negative coefficients, so the ratio is >= 1 for
    every e >= e0 and the rejection step never rejects: e follows the 1/e^2 envelope on
    [e0, 1] exactly (unlike its total cross section, which integrates the G4 alternating-sign
    form).  Reproduced as is; pinned here against the envelope's closed-form CDF."""
    tcut, eng = 1e4 * co.eV, E * co.eV
    g = 1 + eng / co.mc2
    y = 1 / (g + 1)
    B1, B2 = 2 - y * y, (1 - 2 * y) * (3 + y * y)
    B4 = (1 - 2 * y) ** 3
    B3 = (1 - 2 * y) ** 2 + B4
    assert min(B1, B2, B3, B4) >= 0
    e0 = tcut / eng
    eps = _draws(lambda u: O.lib().or_bhabha_sample(tcut, eng, u.ref), N_KS, 5) / eng
    x = np.geomspace(e0, 1.0, 40001)
    F = (1 / e0 - 1 / x) / (1 / e0 - 1)
    assert _ks(eps, x, F) < KS_BAR


@pytest.mark.parametrize("E", [2e4, 5e5, 1e7])
def test_compton_sampler_ks(O, co, E):
    """Klein-Nishina: d sigma / d eps ~ (1/eps + eps)(1 - eps sin^2 / (1 + eps^2))."""
    eng = E * co.eV
    k = eng / co.mc2
    E1 = np.zeros(1)
    ct = np.zeros(1)

    def one(g):
        O.lib().or_compton_sample(eng, g.ref, O.dptr(E1), O.dptr(ct))
        return E1[0] / eng, ct[0]

    r = _draws(one, N_KS, 6)
    eps = r[:, 0]
    assert np.allclose(r[:, 1], 1 - (1 - eps) / (eps * k), atol=1e-12)  # Compton kinematics

    def pdf(e):
        t = (1 - e) / (e * k)
        s2 = t * (2 - t)
        return (1 / e + e) * (1 - e * s2 / (1 + e * e))

    x, F = _numcdf(pdf, 1 / (1 + 2 * k), 1.0)
    assert _ks(eps, x, F) < KS_BAR


@pytest.mark.parametrize("E", [1e4, 1e5, 1e6])
def test_sauter_sampler_ks(O, co, E):
    """Sauter: dsigma/dOmega ~ sin^2 / (1 - b cos)^4 [1 + g(g-1)(g-2)(1 - b cos)/2]."""
    ee = E * co.eV
    g = 1 + ee / co.mc2
    b = math.sqrt(1 - 1 / g ** 2)
    c = _draws(lambda u: O.lib().or_sauter_sample(ee, u.ref), N_KS, 7)
    pdf = lambda cs: (1 - cs ** 2) / (1 - b * cs) ** 4 * (1 + g * (g - 1) * (g - 2) * (1 - b * cs) / 2)
    x, F = _numcdf(pdf, -1.0, 1.0, m=200001)
    assert _ks(c, x, F) < KS_BAR


@pytest.mark.parametrize("E", [1e4, 1e6, 1e8])
def test_annihilation_sampler_ks(O, co, E):
    eng = E * co.eV
    g = 1 + eng / co.mc2
    sq = math.sqrt((g - 1) / (g + 1))
    S = lambda e: -(g + 1) ** 2 + (g * g + 4 * g + 1) / e - 1 / e ** 2
    eps = _draws(lambda u: O.lib().or_annih_sample(eng, u.ref), N_KS, 8)
    # photon a takes eps ~ S(eps) (Heitler's term for the photon that carries eps), photon b
    # takes 1 - eps; S(eps) + S(1 - eps) is the distribution of an unlabelled photon
    x, F = _numcdf(S, (1 - sq) / 2, (1 + sq) / 2, log=True)
    assert _ks(eps, x, F) < KS_BAR


@pytest.mark.parametrize("E", [3e6, 1e8])
def test_bethe_heitler_sampler_ks(O, co, E):
    """G4 Bethe-Heitler with screening: d sigma / d eps ~ (eps - 1/2)^2 (F1(delta) - FZ) +
    (F2(delta) - FZ)/2 on [eps_min, 1/2] mirrored (G4 Physics Reference Manual, 6.5)."""
    Z, eng = 7, E * co.eV
    e0 = co.mc2 / eng
    d0 = 136 * e0 / Z ** (1 / 3)
    FZ = 8 * math.log(Z) / 3 + (8 * 0 if eng <= 50e6 * co.eV else 0)
    if eng > 50e6 * co.eV:
        az2 = co.alpha ** 2
        FZ += 8 * (1 / (1 + az2) + 0.20206 - 0.0369 * az2 + 0.0083 * az2 ** 2 - 0.0020 * az2 ** 3) * az2
    dmax = math.exp((42.24 - FZ) / 8.368) - 0.952
    emin = max(e0, (1 - math.sqrt(1 - 4 * d0 / dmax)) / 2)
    f1 = lambda d: 42.038 - 8.29 * math.log(d + 0.958) if d > 1.4 else 42.184 - d * (7.444 - 1.623 * d)
    f2 = lambda d: 42.038 - 8.29 * math.log(d + 0.958) if d > 1.4 else 41.326 - d * (5.848 - 0.902 * d)
    ek, pk = np.zeros(1), np.zeros(1)

    def one(g):
        O.lib().or_bh_sample(Z, eng, g.ref, O.dptr(ek), O.dptr(pk))
        return (ek[0] + co.mc2) / eng

    eps = _draws(one, N_KS, 9)
    eps = np.minimum(eps, 1 - eps)

    def pdf(e):
        d = d0 / (e * (1 - e))
        return (e - 0.5) ** 2 * (f1(d) - FZ) + 0.5 * (f2(d) - FZ)

    x, F = _numcdf(pdf, emin, 0.5)
    assert _ks(eps, x, F) < KS_BAR


def _synthetic_sb(rows_equal):
    """An SB table with two energy rows whose quantiles are those of chi(log k) = 1 (a
    log-uniform photon spectrum on [log k0, 0]) and, for the second row, chi = 2 + log k."""
    ncum = 1000
    pc = np.arange(ncum) / (ncum - 1.0)
    k0 = -8.0
    a = k0 * (1 - pc)                      # uniform in log k
    c = 2.0 + np.linspace(k0, 0, 4001)     # chi = 2 + log k, cumulative by quadrature
    xx = np.linspace(k0, 0, 4001)
    Fc = np.concatenate([[0], np.cumsum((c[1:] + c[:-1]) / 2 * np.diff(xx))])
    b = np.interp(pc, Fc / Fc[-1], xx)
    data = np.stack([a, a if rows_equal else b])
    le = np.array([math.log(1e5 * 1.6021766208e-19), math.log(1e6 * 1.6021766208e-19)])
    return O.OracleSB.from_arrays(8, le, np.ones(2), data), a, b, pc, k0


def test_sb_sampler_ks(O, co):
    """With both rows equal the bilinear interpolation is exact: the sampled log(k/T) must
    be uniform on [k0, 0] (pins the pcum search, interpolation and exponent)."""
    sbt, a, b, pc, k0 = _synthetic_sb(True)
    T = math.exp(0.5 * (sbt.log_energy[0] + sbt.log_energy[1]))
    kk = _draws(lambda g: O.lib().or_sb_sample(C.byref(sbt.s), T, g.ref), N_KS, 10)
    x = np.linspace(k0, 0, 2001)
    assert _ks(np.log(kk), x, (x - k0) / -k0) < KS_BAR


def test_sb_bilinear_corner_order_of_the_reference(O, co):
    """src/seltzer.jl:116-119 weights u[i1,j2] by (x-x1)(y2-y) and u[i2,j1] by (x2-x)(y-y1):
    the off-diagonal corners are exchanged relative to bilinear interpolation.  At the upper
    row (y == y2) the reference therefore returns u[i2,j1](x2-x)/(x2-x1) + u[i2,j2](x-x1)/(x2-x1).
    The oracle reproduces the reference formula (and the kernels reproduce the oracle)."""
    sbt, a, b, pc, k0 = _synthetic_sb(False)
    T = math.exp(sbt.log_energy[1])
    assert math.log(T) == sbt.log_energy[1]
    rng = np.random.default_rng(3)
    for x in rng.random(200):
        got = math.log(O.lib().or_sb_sample(C.byref(sbt.s), T, O.Uniforms(inject=[x]).ref))
        i2 = int(np.searchsorted(pc, x, side="left"))
        x1, x2 = pc[i2 - 1], pc[i2]
        want = (a[i2] * (x2 - x) + b[i2] * (x - x1)) / (x2 - x1)
        assert abs(got - want) < 1e-12
        correct = (b[i2 - 1] * (x2 - x) + b[i2] * (x - x1)) / (x2 - x1)
        assert abs(got - correct) > 1e-6 or abs(a[i2] - b[i2 - 1]) < 1e-6


# ------------------------------------------------------------------ kinematics
def _state(sp, E, rng, co):
    st = O.State()
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    pn = O.lib().or_momentum_norm_from_kin(sp, E)
    st.p[:] = list(d * pn)
    st.w, st.active, st.uid = 1.0, 1, 77
    return st


@pytest.mark.parametrize("kind,sp", [(O.RBEB, 0), (O.MOLLER, 0), (O.BHABHA, 2), (O.SB, 0),
                                     (O.COMPTON, 1), (O.PHOTO, 1), (O.BH, 1), (O.ANNIH, 2),
                                     (O.ELASTIC, 0)])
def test_outcome_conservation(O, co, kind, sp):
    rng = np.random.default_rng(kind)
    B = 41.72 * co.eV
    p = O.make_proc(kind, 7, B=B, U=71.13 * co.eV, N=2, tcut=1e4 * co.eV)
    for _ in range(300):
        E = math.exp(rng.uniform(math.log(3e6 if kind == O.BH else 3e4), math.log(1e8))) * co.eV
        st = _state(sp, E, rng, co)
        out = O.Outcome()
        g = O.Uniforms(seed=1, uid=int(rng.integers(1 << 40)), step=0)
        err = O.lib().or_collide(C.byref(p), sp, C.byref(st), E, 1, 0, g.ref, C.byref(out))
        assert err == 0
        ke = lambda s, q: O.lib().or_kinenergy(s, out.st[q].p)
        p0 = np.array(st.p[:])
        ps = [np.array(out.st[q].p[:]) for q in range(3)]
        if kind == O.RBEB:
            assert abs(ke(0, 0) + ke(0, 1) + B - E) < 1e-9 * E
        elif kind in (O.MOLLER, O.BHABHA):
            assert abs(ke(sp, 0) + ke(0, 1) - E) < 1e-9 * E
        elif kind in (O.SB, O.COMPTON):
            assert np.abs(ps[0] + ps[1] - p0).max() <= 1e-12 * np.abs(p0).max()
            if kind == O.COMPTON:
                assert abs(ke(1, 0) + ke(0, 1) - E) < 1e-8 * E
        elif kind == O.PHOTO:
            assert abs(ke(0, 1) - (E - 403.0 * co.eV)) < 1e-9 * E
        elif kind == O.ANNIH:
            assert np.abs(ps[1] + ps[2] - p0).max() <= 1e-12 * np.abs(p0).max()
            assert abs(ke(1, 1) + ke(1, 2) - (E + 2 * co.mc2)) < 1e-8 * E
        elif kind == O.BH:
            assert abs(ke(0, 1) + ke(2, 2) - max(E - 2 * co.mc2, 0)) < 1e-9 * E
        elif kind == O.ELASTIC:
            assert abs(np.linalg.norm(ps[0]) / np.linalg.norm(p0) - 1) < 1e-14


# ------------------------------------------------------------------ push
def test_push_uniform_field_exact_momentum_and_energy_relation(O, co):
    """Uniform E along z, B = 0: p(t) = p0 + qE t exactly up to rounding, and
    z(t) - z0 = (W(t) - W0) / (qE) (work-energy theorem) to O(dt^2)."""
    Ez = -1e6
    f = O.field_homogeneous((0.0, 0.0, Ez))
    st = O.State()
    st.p[:] = [1e-22, -3e-23, 2e-22]
    p0 = np.array(st.p[:])
    W0 = O.lib().or_kinenergy(0, st.p) + co.mc2
    dt = 1e-13
    for _ in range(1000):
        O.lib().or_push(0, C.byref(f), C.byref(st), dt)
    F = -co.e * Ez
    assert abs(st.p[2] - (p0[2] + F * 1000 * dt)) < 1e-12 * abs(st.p[2])
    assert st.p[0] == p0[0] and st.p[1] == p0[1]
    W = O.lib().or_kinenergy(0, st.p) + co.mc2
    assert abs(st.x[2] - (W - W0) / F) < 1e-6 * abs(st.x[2])
    assert abs(st.t - 1000 * dt) < 1e-13 * 1000 * dt  # 1000 rounded additions


def test_push_magnetic_rotation(O, co):
    f = O.Field()
    f.e.kind = O.FIELD_HOMOGENEOUS
    f.b.kind = O.FIELD_HOMOGENEOUS
    f.b.v[:] = [0.0, 0.0, 0.1]
    st = O.State()
    st.p[:] = [3e-22, 0.0, 1e-22]
    n0 = np.linalg.norm(st.p[:])
    for _ in range(1000):
        O.lib().or_push(0, C.byref(f), C.byref(st), 1e-13)
    assert abs(np.linalg.norm(st.p[:]) / n0 - 1) < 1e-6
    assert st.p[2] == 1e-22


def test_push_photon_straight_line(O, co):
    f = O.field_homogeneous((0.0, 0.0, -1e6))
    st = O.State()
    st.p[:] = [1e-22, 2e-22, -2e-22]
    O.lib().or_push(1, C.byref(f), C.byref(st), 1e-9)
    d = np.array(st.x[:]) / (co.c * 1e-9)
    assert np.allclose(d, np.array([1, 2, -2]) / 3.0, rtol=1e-14)


# ------------------------------------------------------------------ step invariants
def test_step_is_thread_count_invariant(O):
    import helpers as H
    te, tg, tp = H.tables()
    res = []
    for nth in (1, 4):
        w = H.make_oracle(te, tg, tp, seed=3, nthreads=nth, caps=(1 << 16, 1 << 14, 1 << 14))
        w.add(0, O.init_swarm(3000, 0, 3, e1=1e5, e2=1e7, loguniform=True))
        w.init()
        w.run(0.0, 1e-12, 5)
        res.append((H.by_uid(w.particles(0)), w.counts(0)))
    assert np.array_equal(res[0][0]["uid"], res[1][0]["uid"])
    assert np.array_equal(res[0][0]["p"], res[1][0]["p"])
    assert np.array_equal(res[0][1], res[1][1])


def test_step_weight_and_cut_bookkeeping(O):
    import helpers as H
    te, tg, tp = H.tables()
    w = H.make_oracle(te, tg, tp, seed=4, caps=(1 << 16, 1 << 14, 1 << 14))
    w.add(0, O.init_swarm(2000, 0, 4, e1=1e4 * 1.2, e2=1e7, loguniform=True))
    w.init()
    w.run(0.0, 1e-12, 4)
    for sp in (0, 1, 2):
        a = w.particles(sp)
        if a.size:
            E = np.array([O.lib().or_kinenergy(sp, O.dptr(np.ascontiguousarray(q))) for q in a["p"]])
            assert E.min() >= H.CUT
            assert np.all(a["w"] == 1.0)
            # the step loop stops once trem <= eps(Float64) s (src/mixed_population.jl:65)
            assert np.all(np.abs(a["t"] - 4e-12) <= 2.3e-16)
    assert w.w.ratebound_violations == 0
