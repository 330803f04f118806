"""The C host (tests/c_host/swarm.c) runs a runaway-electron swarm in dry air with tables it
built through the ABI's host entry points; the Python host, with tables from the same
builders (tables.native_table, SeltzerBerger(native=True)), gives the same swarm (the random
streams are keyed by seed, uid and step only): the same count, weight and maximum energy,
and the mean energy to rounding."""
import json
import subprocess
import sys

import pytest

import particulator_hip as P
from particulator_hip import air, co, tables
from particulator_hip.physics import SeltzerBerger
from test_c_host import SB_DIR, build_c_host

pytestmark = pytest.mark.gpu

CUT = 1e4 * co.eV


def python_host(n, steps, seed, efield=1e6, dt=1e-12):
    comp = air.composition()
    procs = [(c, SeltzerBerger(p.Z, native=True) if isinstance(p, SeltzerBerger) else p)
             for c, p in air.electron_processes(comp, CUT)]
    te = tables.native_table(procs, P.PT_ELECTRON, co.e * efield * dt, energy_cut=CUT)
    mp = P.MultiPopulation(seed=seed, tally_bins=64, electron=P.Population(4 * n + 4096, [], te, CUT))
    mp.set_pusher(P.RK2Pusher(P.ElectromagneticField(P.HomogeneousField((0.0, 0.0, -efield)))))
    mp.init_particles(P.PT_ELECTRON, n, 1e6 * co.eV)
    mp.init()
    mp.run_steps(0.0, dt, steps, compact_every=8, check_every=steps)
    s = mp.stats(P.PT_ELECTRON)
    out = {"nactive": int(s.nactive), "weight": s.weight, "mean_energy": s.mean_energy,
           "max_energy": s.max_energy}
    mp.close()
    return out


def test_c_host_matches_python_host(tmp_path):
    n, steps, seed = 10000, 300, 4242
    exe = build_c_host(tmp_path)
    print("C host: %d electrons, %d steps, seed %d" % (n, steps, seed), flush=True)
    sys.stdout.flush()
    r = subprocess.run([exe, SB_DIR, str(n), str(steps), str(seed)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout.strip().splitlines()[-1])
    print("python host", flush=True)
    p = python_host(n, steps, seed)
    print(c, p, flush=True)
    assert c["nactive"] > n  # the swarm grows at 1 MV/m
    # the particles are the same; the mean is a sum over slots, whose order follows the
    # atomic slot claims, so it agrees to rounding
    for k in ("nactive", "weight", "max_energy"):
        assert c[k] == p[k], k
    assert abs(c["mean_energy"] - p["mean_energy"]) <= 1e-12 * p["mean_energy"]
    # the same run as replays of a captured graph (pt_run_graph)
    print("C host, captured graph", flush=True)
    r = subprocess.run([exe, SB_DIR, str(n), str(steps), str(seed), "graph"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("nactive", "weight", "max_energy"):
        assert g[k] == c[k], k
    assert abs(g["mean_energy"] - c["mean_energy"]) <= 1e-12 * c["mean_energy"]
