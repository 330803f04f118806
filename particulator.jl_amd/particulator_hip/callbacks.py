"""Callbacks of Particulator.jl (src/callback.jl).

`onstep` / `onoutput` callbacks run on the host between device steps.  Per-collision and
per-advance callbacks cannot call back into Python from a GPU kernel; the built-in ones
(CollisionCounter) are computed on the device instead.
"""
import numpy as np

from .population import SPECIES, roulette_, split_


class AbstractCallback:
    """src/callback.jl:8-35"""

    def onstep(self, mpopl, t=None):
        return True

    def onoutput(self, mpopl, t=None, i=None):
        return None

    def device_collision_counter(self):
        return False


class VoidCallback(AbstractCallback):
    pass


class CombinedCallback(AbstractCallback):
    """src/callback.jl:41-108: callbacks called in order; onstep is the AND of all."""

    def __init__(self, tpl):
        self.tpl = tuple(tpl)

    def install(self, mpopl):
        for c in self.tpl:
            if hasattr(c, "install"):
                c.install(mpopl)

    def __getitem__(self, i):
        return self.tpl[i]

    def onstep(self, mpopl, t=None):
        r = True
        for c in self.tpl:
            r = c.onstep(mpopl, t) and r
        return r

    def onoutput(self, mpopl, t=None, i=None):
        for c in self.tpl:
            c.onoutput(mpopl, t, i)


class CollisionCounter(AbstractCallback):
    """Counts collisions per process (src/callback.jl:118-141), read from device counters.
    `counts(mpopl)` returns {species: {process name: count}} including null collisions."""

    def device_collision_counter(self):
        return True

    def counts(self, mpopl):
        out = {}
        for name, p in mpopl.index.items():
            if not p.tracked or p.collisions is None:
                continue
            c = mpopl.collision_counts(p.species)
            d = {"NullCollision": int(c[0])}
            for j, proc in enumerate(p.collisions.proc):
                key = "%s[%d]" % (type(proc).__name__, j)
                d[key] = int(c[1 + j])
            out[name] = d
        return out


class WallCallback(AbstractCallback):
    """WallCallback{P}(coord, v, drop), src/callback.jl:146-184, evaluated on the device after
    every push.  `install(mpopl)` sets the wall (run_ calls it); `accum(mpopl)` returns the
    recorded states."""

    def __init__(self, species, coord, v, drop=True, capacity=1 << 20):
        self.species = SPECIES.get(species, species)
        self.coord, self.v, self.drop, self.capacity = coord, v, drop, capacity

    def install(self, mpopl):
        mpopl.set_wall(self.species, self.coord, self.v, self.drop, self.capacity)

    def accum(self, mpopl):
        rec, _ = mpopl.wall_records()
        return rec[rec["species"] == self.species]


class ParticleCountCallback(AbstractCallback):
    """Saves the weight of each species at every output, src/callback.jl:190-204."""

    def __init__(self, particles):
        self.p = list(particles)
        self.counts = []

    def onoutput(self, mpopl, t=None, i=None):
        self.counts.append([t] + [mpopl.get(s).weight() for s in self.p])


class RouletteCallback(AbstractCallback):
    """Russian roulette to at most m particles per species, src/callback.jl:210-226."""

    def __init__(self, m):
        self.m = m

    def onstep(self, mpopl, t=None):
        for name, popl in mpopl.index.items():
            if not popl.tracked:
                continue
            n = popl.nactives()
            if n > self.m:
                roulette_(self.m / n, popl)
            mpopl.compact(popl.species)
        return True


class SplitCallback(AbstractCallback):
    """Particle splitting below m * min_frac particles, src/callback.jl:231-247."""

    def __init__(self, species, m, min_frac):
        self.species, self.m, self.min_frac = species, m, min_frac

    def onstep(self, mpopl, t=None):
        popl = mpopl.get(self.species)
        n = popl.nactives()
        if 0 < n < self.m * self.min_frac:
            split_(self.m / n - 1.0, popl)
        mpopl.compact(popl.species)
        return True


class PopulationTargetCallback(AbstractCallback):
    """Early stop when a species' active count crosses n, src/callback.jl:253-267."""

    def __init__(self, species, n, ismax=True):
        self.species, self.n, self.ismax = species, n, ismax

    def onstep(self, mpopl, t=None):
        na = mpopl.get(self.species).nactives()
        return na < self.n if self.ismax else na > self.n
