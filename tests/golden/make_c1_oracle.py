"""Write c1_oracle.json: the oracle's runaway-electron swarm at full size and length (10^4
electrons at 1 MeV, isotropic, dry air, E = 1 MV/m along -z, 10 keV cut, dt = 1 ps, 2000
steps, photons tallied), 32 independent samples run as 32 swarms of one world: per-sample
electron count and mean kinetic energy at the end.  Run from the repository root:
    python tests/golden/make_c1_oracle.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "particulator.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import helpers as H  # noqa: E402
import oracle as O  # noqa: E402
from particulator_hip import air, co  # noqa: E402
from particulator_hip.population import kinenergy  # noqa: E402

NS, N, STEPS, DT, E, SEED = 32, 10_000, 2000, 1e-12, 1e6, 9071


def main():
    O.build()
    threads = os.cpu_count() or 1
    te, tg, tp = air.swarm_tables(air.composition(), [E] * NS, DT, H.CUT)
    ots = [H.oracle_table(te), H.oracle_table(tg), H.oracle_table(tp)]
    ow = O.OracleWorld(ots, SEED, (1 << 21, 1024, 1 << 16), (H.CUT,) * 3, tracked=(True, False, False),
                       nthreads=threads, fields=[O.field_homogeneous((0.0, 0.0, -E))] * NS)
    for i in range(NS):
        ow.add(0, O.init_swarm(N, 0, SEED, uid0=i * N, swarm=i, e1=1e6))
    ow.init()
    t = 0.0
    for k in range(STEPS // 100):
        t = ow.run(t, DT, 100)
        print("step %d" % ((k + 1) * 100), flush=True)
    a = ow.particles(0)
    cnt = np.bincount(a["swarm"], minlength=NS)
    esum = np.bincount(a["swarm"], weights=kinenergy(0, a["p"]), minlength=NS)
    out = {"workload": "runaway swarm: %d samples x %d e- at 1 MeV isotropic, dry air 1 atm 273 K, E = 1 MV/m "
                       "along -z, cut 10 keV, dt = 1 ps, %d steps, photons tallied" % (NS, N, STEPS),
           "seed": SEED, "threads": threads, "n0": N, "steps": STEPS, "dt": DT,
           "count": cnt.tolist(), "mean_energy_eV": (esum / cnt / co.eV).tolist()}
    with open(os.path.join(HERE, "c1_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
