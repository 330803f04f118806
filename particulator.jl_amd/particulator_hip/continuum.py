"""Continuum energy losses of electrons and positrons below a cut (src/continuum.jl): the
restricted collision stopping power as a friction force, and its Chebyshev fit that the
device evaluates inside the pusher's force."""
import math

import numpy as np

from . import constants as co
from ._lib import PT_ELECTRON, PT_POSITRON
from .tables import BinaryIntervals, chebfit

LOG10 = math.log(10.0)


def _x0x1(C):
    """src/continuum.jl:124-139"""
    if C < 10:
        return 1.6, 4.0
    if C < 10.5:
        return 1.7, 4.0
    if C < 11.0:
        return 1.8, 4.0
    if C < 11.5:
        return 1.9, 4.0
    if C < 12.25:
        return 2.0, 4.0
    if C < 13.804:
        return 2.0, 5.0
    return 0.326 * C - 2.5, 5.0


class ContinuumLoss:
    """ContinuumLoss(nel, I, Tcut), src/continuum.jl:6-15: nel electrons per m^3, mean
    excitation energy I (J), losses to secondaries below Tcut (J)."""

    def __init__(self, nel, I, Tcut):
        self.nel, self.I, self.Tcut = nel, I, Tcut

    def energy_loss(self, species, eng):
        """energy_loss(cl, s, eng), src/continuum.jl:67-105 (vectorised; J/m)"""
        nel, I, Tcut = self.nel, self.I, self.Tcut
        mc2, r_e = co.electron_mc2, co.r_e
        eng = np.asarray(eng, dtype=np.float64)
        tau = eng / mc2
        tauc = Tcut / mc2
        taumax = tau if species == PT_POSITRON else tau / 2
        g = 1.0 + tau
        b2 = 1.0 - 1.0 / g ** 2
        tup = np.minimum(tauc, taumax)
        with np.errstate(divide="ignore", invalid="ignore"):
            if species == PT_POSITRON:
                y = 1.0 / (2.0 + tau)
                F = (np.log(tau * tup)
                     - (tup ** 2 / tau) * (tau * 2 * tup - 3 * tup ** 2 * y / 2 - (tup - tup ** 3 / 3) * y ** 2
                                           - (tup ** 2 / 2 - tau * tup ** 3 / 3 + tup ** 4 / 4) * y ** 3))
            else:
                F = (-1 - b2 + np.log((tau - tup) * tup) + tau / (tau - tup)
                     + (tup ** 2 / 2 + (2 * tau + 1) * np.log(1 - tup / tau)) / g ** 2)
            x = np.log(g ** 2 * b2) / LOG10 / 2
        hnup = co.hbar * co.c * math.sqrt(4 * math.pi * nel * r_e)
        C = 1 + 2 * math.log(I / hnup)
        xa = C / LOG10 / 2
        m = 3
        x0, x1 = _x0x1(C)
        a = 2 * LOG10 * (xa - x) / (x1 - x0) ** m  # as written in the reference (depends on x)
        delta = np.where(x < x0, 0.0, np.where(x < x1, 2 * LOG10 * x - C + a * (x1 - x) ** m, 2 * LOG10 * x - C))
        return (2 * math.pi * r_e ** 2 * mc2 * nel / b2) * (np.log((2 * (g + 1)) / (I / mc2) ** 2) + F - delta)


class ChebContinuumLoss:
    """ChebContinuumLoss(cl, Tmax, n), src/continuum.jl:21-41: fits of the electron and
    positron losses with n Chebyshev terms on BinaryIntervals(k, 2^k Tcut), 2^k Tcut >= Tmax."""

    def __init__(self, cl, Tmax, n):
        self.cl = cl
        self.k = int(math.ceil(math.log2(Tmax / cl.Tcut)))
        self.b = BinaryIntervals(self.k, 2.0 ** self.k * cl.Tcut)
        self.order = n
        self.ec = np.ascontiguousarray(chebfit(lambda x: cl.energy_loss(PT_ELECTRON, x), self.b, n))
        self.pc = np.ascontiguousarray(chebfit(lambda x: cl.energy_loss(PT_POSITRON, x), self.b, n))

    def __call__(self, species, eng):
        """the fitted loss (J/m), evaluated as the device does"""
        c = self.ec if species == PT_ELECTRON else self.pc
        i, t = self.b.precheb(eng)
        i = np.minimum(i, self.k)
        xi = t[1]
        a = c[i]
        acc = a[..., 0] * 1.0
        if self.order > 1:
            acc = acc + a[..., 1] * xi
        tm2, tm1 = np.ones_like(xi), xi
        for j in range(2, self.order):
            tj = 2.0 * xi * tm1 - tm2
            acc = acc + a[..., j] * tj
            tm2, tm1 = tm1, tj
        return acc


def air_loss(comp, Tcut, I=85.7 * co.eV, Tmax=1e8 * co.eV, n=3):
    """ChebContinuumLoss of dry air (electron density of N2, O2, Ar; I = 85.7 eV, ICRU 37)"""
    nel = 14 * comp["N2"] + 16 * comp["O2"] + 18 * comp.get("Ar", 0.0)
    return ChebContinuumLoss(ContinuumLoss(nel, I, Tcut), Tmax, n)
