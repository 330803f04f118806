#!/usr/bin/env python3
"""Throughput of the null-collision Monte Carlo step on one MI355X.

Workload (default): the coupled relativistic avalanche in dry air (N2:O2:Ar =
0.781:0.210:0.009, 1 atm, 273 K), 10^7 initial electrons with kinetic energy log-uniform in
100 keV - 10 MeV, isotropic, from the origin, uniform E = 1 MV/m along -z, B = 0;
electrons, photons and positrons tracked with a 10 keV cut, populations capped at 2^27,
compaction every 8 steps, FP64 state, dt = 1 ps.  A step is one pt_advance (push,
null-collision selection, sampling, secondary append and the secondaries' own passes).

Metric: particle-steps/s = sum over timed steps of the active particles at step start,
divided by the wall time of the timed loop (inputs resident in HBM, tables uploaded).

Prints one JSON line.  `roofline` prices the main advance kernel with the algorithmic HBM
bytes of its launches (153 B per particle-step + 81 B per appended secondary) over its
average launch duration measured with HIP events on the library's stream.  `cpu_baseline`
times the CPU restatement of the reference step (oracle/, OpenMP over the host cores) on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "particulator.jl_amd"))

HBM_PEAK_GBPS = 8000.0
BYTES_PER_PSTEP = 153.0
BYTES_PER_SECONDARY = 81.0
BYTES_PER_MOVE = 162.0

WORKLOADS = {
    "avalanche": dict(n=10_000_000, e1=1e5, e2=1e7, loguniform=True, efield=1e6, dt=1e-12,
                      cap=1 << 27, compact_every=8, tracked=(1, 1, 1),
                      name="coupled relativistic avalanche, 1e7 e- log-uniform 0.1-10 MeV, "
                           "dry air 1 atm 273 K, E=1 MV/m, dt=1 ps, e-/photon/e+ tracked, "
                           "cap 2^27, compaction every 8 steps"),
    "swarm": dict(n=10_000, e1=1e6, e2=None, loguniform=False, efield=1e6, dt=1e-12,
                  cap=1 << 20, compact_every=1, tracked=(1, 0, 0),
                  name="runaway swarm, 1e4 e- at 1 MeV, dry air 1 atm 273 K, E=1 MV/m, "
                       "dt=1 ps, photons tallied"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, seed, nsample, nsteps):
    """Oracle (C, OpenMP) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import helpers as H
    from particulator_hip import co
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    te, tg, tp = H.tables(efield=wl["efield"], dt=wl["dt"])
    caps = (nsample * 2, nsample, nsample)
    ow = H.make_oracle(te, tg, tp, seed, caps=caps, tracked=tuple(bool(x) for x in wl["tracked"]),
                       efield=(0.0, 0.0, -wl["efield"]), nthreads=threads)
    init = O.init_swarm(nsample, 0, seed, e1=wl["e1"], e2=wl["e2"], loguniform=wl["loguniform"])
    ow.add(0, init)
    ow.init()
    t, psteps = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(nsteps):
        psteps += sum(int(O.lib().or_nactives(O.C.byref(ow.w.pop[s]))) for s in range(3)
                      if ow.bufs[s] is not None)
        ow.advance(t + wl["dt"])
        t += wl["dt"]
    el = time.perf_counter() - t0
    return {"value": psteps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "%d e- of the same distribution, %d steps (%d particle-steps, %.1f s wall "
                      "on %d OpenMP threads), reference-semantics C restatement"
                      % (nsample, nsteps, psteps, el, threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="avalanche", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="override the initial particle count")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rounds", type=int, default=None, help="collision rounds per step (0: history mode)")
    ap.add_argument("--block", type=int, default=None, help="threads per workgroup of the step kernels")
    ap.add_argument("--grid-mult", type=int, default=None)
    ap.add_argument("--defer-mask", type=lambda v: int(v, 0), default=None,
                    help="sampler kinds queued in collision rounds (bit = pt_proc_kind)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (to scale the HBM "
                         "byte counters of a profiler) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import particulator_hip as P
    from particulator_hip import air, co

    wl = dict(WORKLOADS[args.workload])
    n = args.n or wl["n"]
    cut = 1e4 * co.eV
    t_setup = time.perf_counter()
    te, tg, tp = air.build_tables(air.composition(), wl["efield"], wl["dt"], cut)
    pops = {}
    for name, tb, tr in (("electron", te, wl["tracked"][0]), ("photon", tg, wl["tracked"][1]),
                         ("positron", tp, wl["tracked"][2])):
        pops[name] = P.Population(wl["cap"], [], tb, cut, tracked=bool(tr))
    mp = P.MultiPopulation(seed=args.seed, **pops)
    mp.set_pusher(P.RK2Pusher(P.ElectromagneticField(P.HomogeneousField((0.0, 0.0, -wl["efield"])))))
    mp.init_particles(P.PT_ELECTRON, n, wl["e1"] * co.eV, (wl["e2"] or wl["e1"]) * co.eV,
                      loguniform=wl["loguniform"])
    mp.init()
    launch = mp.get_launch()
    for k, v in (("rounds", args.rounds), ("block", args.block), ("grid_mult", args.grid_mult),
                 ("defer_mask", args.defer_mask)):
        if v is not None:
            launch[k] = v
    launch["compact_every"] = wl["compact_every"]
    mp.set_launch(**launch)
    log("setup %.1f s, launch %s" % (time.perf_counter() - t_setup, launch))
    dt = wl["dt"]
    t = mp.run_steps(0.0, dt, args.warmup, compact_every=wl["compact_every"], check_every=args.warmup or 1)
    mp.work(reset=True)
    mp.profile(True)
    check(mp)
    t0 = time.perf_counter()
    t = mp.run_steps(t, dt, args.steps, compact_every=wl["compact_every"], check_every=args.steps)
    check(mp)
    el = time.perf_counter() - t0
    w = mp.work(reset=False)
    counts = [mp.count(s) for s in range(3)]
    psteps = w.particle_steps
    value = psteps / el
    adv_avg_ms = w.advance_ms / max(w.advance_launches, 1)
    adv_bytes = (BYTES_PER_PSTEP * psteps + BYTES_PER_SECONDARY * w.main_secondaries) / max(w.advance_launches, 1)
    achieved = adv_bytes / (adv_avg_ms * 1e-3) / 1e9
    log("timed %d steps in %.3f s: %.4g particle-steps/s; advance %.3f ms/launch, tail %.3f ms, "
        "compact %.3f ms total; populations %s; events/pstep %.2f; secondaries %d"
        % (args.steps, el, value, adv_avg_ms, w.tail_ms / max(w.advance_launches, 1), w.compact_ms,
           counts, w.events / max(psteps, 1), w.secondaries))
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated initial particles from a seeded Philox stream)",
        "config": {"workload": wl["name"], "initial_particles": n, "seed": args.seed,
                   "population_end": counts, "events_per_particle_step": w.events / max(psteps, 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "k_advance", "avg_launch_ms": adv_avg_ms,
                     "algorithmic_bytes_per_launch": adv_bytes},
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(wl, args.seed, args.cpu_sample, args.cpu_steps)
        except Exception as e:  # the CPU leg must not hide the GPU figure
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out))
    mp.close()


def check(mp):
    rc, flags = mp.flags()
    if rc != 0:
        raise SystemExit("device error flags %s" % flags)


if __name__ == "__main__":
    main()
