"""run! (src/run.jl:1-40): the time loop over device steps."""
import logging

from .callbacks import VoidCallback

log = logging.getLogger("particulator_hip")


def _msg(mpopl, t, perc):
    """src/run.jl:31-40"""
    lines = ["%.2f%% complete" % perc, "time = %.3f ns" % (t / 1e-9), "# of particles:"]
    for name, p in mpopl.index.items():
        if p.tracked:
            lines.append("%15s => %-10d" % (name, p.nparticles()))
    lines.append("centroid locations:")
    for name, p in mpopl.index.items():
        if p.tracked:
            lines.append("%15s => %s" % (name, p.spread()[0]))
    return "\n".join(lines)


def run_(mpopl, pusher, tfinal, dt, callback=None, output_dt="default", verbosity=1,
         compact_every=1, t0=0.0):
    """run!(mpopl, pusher, tfinal, dt, callback; output_dt, verbosity).

    Each step is advance! then droplow! of every species (src/run.jl:7-9); the device
    applies the energy cut in the step and frees slots every `compact_every` steps
    (1 reproduces droplow! each step)."""
    callback = callback or VoidCallback()
    if output_dt == "default":
        output_dt = tfinal / 20
    if pusher is not mpopl.pusher:
        mpopl.set_pusher(pusher)
    if hasattr(callback, "install"):
        callback.install(mpopl)  # device-side parts (walls)
    t = t0
    nxt = t
    isave = 0
    istep = 0
    while t < tfinal:
        mpopl.advance(t + dt)
        istep += 1
        if compact_every and istep % compact_every == 0:
            mpopl.compact(-1)
        t += dt
        cont = callback.onstep(mpopl, t)
        if output_dt is not None and t >= nxt:
            nxt += output_dt
            isave += 1
            if verbosity > 0:
                log.info(_msg(mpopl, t, 100 * t / tfinal))
            callback.onoutput(mpopl, t, isave)
        if not cont:
            if verbosity > 0:
                log.info("Early stop due to onstep() callback")
            break
    return t
