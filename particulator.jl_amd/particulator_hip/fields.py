"""Field functors, forcing and pushers of Particulator.jl (src/field.jl, src/pusher.jl).

Fields here are descriptions uploaded to the device; the RK2 push itself runs in the HIP
kernels (csrc/pt_device.h `push`).
"""
import numpy as np

from ._lib import (FIELD_CONFINED, FIELD_DOUBLELAYER, FIELD_HOMOGENEOUS, FIELD_NONE,
                   FIELD_STEP, FieldDesc, FieldFn)


class HomogeneousField:
    """src/field.jl:4-8"""

    def __init__(self, v):
        self.v = tuple(float(a) for a in v)

    def __call__(self, x, t=0.0):
        return np.array(self.v)

    def fn(self):
        f = FieldFn()
        f.kind = FIELD_HOMOGENEOUS
        f.v[:] = list(self.v)
        return f


class DoubleLayerField:
    """src/field.jl:10-16: v for z1 < z < z2, else 0"""

    def __init__(self, z1, z2, v):
        self.z1, self.z2, self.v = float(z1), float(z2), tuple(float(a) for a in v)

    def __call__(self, x, t=0.0):
        return np.array(self.v) if self.z1 < x[2] < self.z2 else np.zeros(3)

    def fn(self):
        f = FieldFn()
        f.kind = FIELD_DOUBLELAYER
        f.v[:] = list(self.v)
        f.z1, f.z2 = self.z1, self.z2
        return f


class StepField:
    """src/field.jl:22-28: v1 below z, v2 above"""

    def __init__(self, z, v1, v2):
        self.z, self.v1, self.v2 = float(z), tuple(map(float, v1)), tuple(map(float, v2))

    def __call__(self, x, t=0.0):
        return np.array(self.v1) if x[2] < self.z else np.array(self.v2)

    def fn(self):
        f = FieldFn()
        f.kind = FIELD_STEP
        f.v[:] = list(self.v1)
        f.v2[:] = list(self.v2)
        f.z1 = self.z
        return f


class ConfinedDoubleLayerField:
    """src/field.jl:35-52"""

    def __init__(self, sx, sy, sz, ez0):
        self.sx, self.sy, self.sz, self.ez0 = map(float, (sx, sy, sz, ez0))

    def fn(self):
        f = FieldFn()
        f.kind = FIELD_CONFINED
        f.sx, f.sy, f.sz, f.ez0 = self.sx, self.sy, self.sz, self.ez0
        return f


class NullField:
    def fn(self):
        f = FieldFn()
        f.kind = FIELD_NONE
        return f


class ElectromagneticField:
    """Forcing q e (E + v x B), src/field.jl:54-70"""

    def __init__(self, e, b=None):
        self.e = e
        self.b = b if b is not None else HomogeneousField((0.0, 0.0, 0.0))

    def desc(self):
        d = FieldDesc()
        d.e = self.e.fn()
        d.b = self.b.fn()
        return d


class NullForcing:
    """src/pusher.jl:14"""

    def desc(self):
        d = FieldDesc()
        d.e = NullField().fn()
        d.b = NullField().fn()
        return d


class CombinedForcing:
    """CombinedForcing(field, loss), src/pusher.jl:14-23: the electromagnetic field plus a
    ChebContinuumLoss friction (the device adds the two forces)."""

    def __init__(self, field, continuum=None):
        self.field = field
        self.continuum = continuum

    def desc(self):
        return self.field.desc()


class RK2Pusher:
    """Two-stage Runge-Kutta pusher (Ralston weights 1/4, 3/4), src/pusher.jl:37-63."""

    def __init__(self, forcing):
        self.forcing = forcing

    def desc(self):
        return self.forcing.desc()
