"""Continuum losses (src/continuum.jl): the restricted stopping-power formula against the
NIST ESTAR collision stopping power of air where the restriction is lifted, its Chebyshev
fit, the fit's evaluation (product vs oracle, bit for bit), and the oracle's push with the
friction term (work-energy over a short push)."""
import math

import numpy as np
import pytest

import oracle as O
from particulator_hip import air, co
from particulator_hip._lib import PT_ELECTRON, PT_POSITRON
from particulator_hip.continuum import ChebContinuumLoss, ContinuumLoss, air_loss
from particulator_hip.population import kinenergy

MEV_CM2_PER_G = 1.0 / (1e6 * co.eV) * 1e4 * 1e-3  # (J m^2 / kg) -> MeV cm^2 / g


def air_density(comp):
    molar = {"N2": 28.0134e-3, "O2": 31.9988e-3, "Ar": 39.948e-3}  # kg/mol
    return sum(comp[k] * molar[k] for k in comp) / co.N_A


def test_unrestricted_loss_is_the_collision_stopping_power_of_air():
    """With Tcut above T/2 the formula is the Berger-Seltzer collision stopping power;
    NIST ESTAR gives 1.661 MeV cm^2/g for electrons of 1 MeV in dry air (the density
    effect is negligible in a gas at 1 MeV)."""
    comp = air.composition()
    nel = 14 * comp["N2"] + 16 * comp["O2"] + 18 * comp["Ar"]
    cl = ContinuumLoss(nel, 85.7 * co.eV, 1.0)  # Tcut = 1 J: no restriction
    s = cl.energy_loss(PT_ELECTRON, 1e6 * co.eV) / air_density(comp) * MEV_CM2_PER_G
    assert abs(s - 1.661) <= 0.02 * 1.661
    # restricting to losses below 10 keV removes the hard collisions
    cr = ContinuumLoss(nel, 85.7 * co.eV, 1e4 * co.eV)
    for E in (1e5, 1e6, 1e7):
        le, lr = cl.energy_loss(PT_ELECTRON, E * co.eV), cr.energy_loss(PT_ELECTRON, E * co.eV)
        assert 0 < lr < le
        assert cr.energy_loss(PT_POSITRON, E * co.eV) > 0


def test_chebyshev_fit_of_the_loss():
    """5 terms per octave fit the loss to 0.1% (the density-effect kink near 20 MeV sits
    inside an interval and bounds the accuracy; 3 terms give 0.9% at 20 keV)"""
    comp = air.composition()
    cheb = air_loss(comp, 1e4 * co.eV, n=5)
    E = np.geomspace(2e4 * co.eV, 9e7 * co.eV, 4000)
    for sp in (PT_ELECTRON, PT_POSITRON):
        exact = cheb.cl.energy_loss(sp, E)
        rel = np.abs(cheb(sp, E) / exact - 1)
        assert rel.max() < 2e-3


def test_fit_evaluation_matches_oracle_bit_for_bit(O):
    cheb = air_loss(air.composition(), 1e4 * co.eV, n=5)
    E = np.geomspace(1e3 * co.eV, 2e8 * co.eV, 997)
    for sp, c in ((PT_ELECTRON, cheb.ec), (PT_POSITRON, cheb.pc)):
        mine = cheb(sp, E)
        ptr = O.dptr(np.ascontiguousarray(c))
        theirs = np.array([O.lib().or_chebeval_n(ptr, cheb.k, cheb.order, cheb.b.xmax, e) for e in E])
        assert np.array_equal(mine, theirs)


def test_oracle_push_loses_energy_at_the_fitted_rate(O):
    cheb = air_loss(air.composition(), 1e4 * co.eV, n=3)
    f = O.with_continuum_loss(O.field_homogeneous((0.0, 0.0, 0.0)), cheb.k, cheb.order, cheb.b.xmax,
                              cheb.ec, cheb.pc)
    for sp in (0, 2):
        st = O.init_swarm(64, sp, 3, e1=5e4, e2=5e6, loguniform=True)
        E0 = kinenergy(sp, st["p"])
        v = co.c * np.sqrt(1 - (co.electron_mc2 / (co.electron_mc2 + E0)) ** 2)
        dt = 1e-14
        out = st.copy()
        for q in range(out.size):
            s = O.State.from_buffer_copy(out[q].tobytes())
            O.lib().or_push(sp, O.C.byref(f), O.C.byref(s), dt)
            out[q] = np.frombuffer(bytes(s), dtype=O.STATE_DTYPE)[0]
        dE = kinenergy(sp, out["p"]) - E0
        expect = -cheb(sp, E0) * v * dt
        assert np.allclose(dE, expect, rtol=1e-4)
