"""The oracle's depth-dose bookkeeping (build-defined deposition: particles dropped below the
cut, secondaries born below it, binding energies) closes the energy balance where the
reference's kinematics conserve energy: electrons with elastic scattering, RBEB and Moller,
photons with Compton and photoabsorption, no field.  (Bremsstrahlung builds the electron
from p - p_gamma, src/seltzer.jl:88, and pair production creates rest mass, so both are left
out of the balance.)"""
import numpy as np

import oracle as O
from particulator_hip import co
from particulator_hip.population import kinenergy


def test_dose_tally_closes_the_energy_balance(O):
    comp = O.air_composition()
    cut = 1e4 * co.eV
    keep_e = (O.ELASTIC, O.RBEB, O.MOLLER)
    keep_g = (O.COMPTON, O.PHOTO)
    pe = [p for p in O.air_processes(O.ELECTRON, comp, cut) if p.kind in keep_e]
    pg = [p for p in O.air_processes(O.PHOTON, comp, cut) if p.kind in keep_g]
    te = O.OracleTable(pe, O.ELECTRON, 0.0)
    tg = O.OracleTable(pg, O.PHOTON, 0.0)
    tp = O.OracleTable(O.air_processes(O.POSITRON, comp, cut), O.POSITRON, 0.0)
    w = O.OracleWorld([te, tg, tp], 7, (1 << 16, 1 << 16, 1 << 16), (cut, cut, cut), nthreads=4,
                      field=O.field_homogeneous((0.0, 0.0, 0.0)), dose=(64, -0.02, 0.02))
    e0 = O.init_swarm(3000, O.ELECTRON, 7, uid0=0, e1=1.5e4, e2=1e5, loguniform=True)
    g0 = O.init_swarm(3000, O.PHOTON, 7, uid0=10_000, e1=1.5e4, e2=1e5, loguniform=True)
    w.add(O.ELECTRON, e0)
    w.add(O.PHOTON, g0)
    w.init()
    ein = kinenergy(0, e0["p"]).sum() + kinenergy(1, g0["p"]).sum()
    w.run(0.0, 1e-12, 200)
    eout = w.dose[-1] + sum(kinenergy(s, w.particles(s)["p"]).sum() for s in (0, 1))
    assert w.dose[-1] > 0.1 * ein  # a good part of the energy was deposited
    assert abs(eout - ein) <= 1e-9 * ein
    # the histogram (with under- and overflow) holds the total
    assert abs(w.dose[:-1].sum() - w.dose[-1]) <= 1e-9 * w.dose[-1]
