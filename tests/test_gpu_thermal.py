"""Thermal checks of the device Langevin path (the known answers that pin the
oracle's noise term, test_oracle_thermal.py), at FP64 and FP32.

* Two-site chain, E = S1.S2 (test/test_samplers.jl:120-227): 64 replicas of
  the chain stepped as one batch; after 1000 equilibration steps, 32 samples
  500 steps apart from every replica (2048 energies) against the exact
  energy distribution, atol 0.05, for :dipole and :SUN.  At FP64 replica 0
  also follows the oracle's trajectory with the same Philox draws.
* Paramagnet of kappa = 3/2 moments: <S.h^> = -kappa L(kappa |h| / kT).
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce
from test_oracle_thermal import (CHAIN, PARA, chain_distribution_error, paramagnet, paramagnet_expected,
                                 two_site_chain)

pytestmark = pytest.mark.gpu

R_CHAIN = 64


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("mode", ["dipole", "SUN"])
def test_chain_energy_distribution(mode, precision):
    announce(f"thermal two-site chain {mode} fp{precision} x{R_CHAIN}")
    systems = [two_site_chain(mode, precision=precision, seed=100 + r) for r in range(R_CHAIN)]
    batch = sh.ReplicaBatch(systems, precision=precision)
    integ = sh.Langevin(CHAIN["dt"], damping=CHAIN["damping"], kT=CHAIN["kT"])
    if precision == 64:
        # replica 0 follows the oracle with the same draws
        s0 = systems[0]
        o = O.Oracle(s0.spec())
        st = dict(S=s0.dipoles_flat()) if mode == "dipole" else dict(Z=s0.coherents_flat())
        _, E_ref = o.langevin_run(**st, dt=integ.dt, damping=integ.damping, kT=integ.kT,
                                  key=O.replica_key(batch.noise_seed, 0), step0=0, nsteps=200, every=20)
        E_dev = []
        for _ in range(10):
            sh.step_(batch, integ, nsteps=20)
            E_dev.append(sh.energy(batch)[0])
        np.testing.assert_allclose(E_dev, E_ref, atol=1e-10)
        sh.step_(batch, integ, nsteps=CHAIN["n_equilib"] - 200)
    else:
        sh.step_(batch, integ, nsteps=CHAIN["n_equilib"])
    n = CHAIN["n_samples"] // R_CHAIN
    E = []
    for _ in range(n):
        sh.step_(batch, integ, nsteps=CHAIN["n_decorr"])
        E.append(sh.energy(batch))
    E = np.concatenate(E)
    assert np.all(np.abs(E) <= 1 + 1e-5)
    err = chain_distribution_error(E)
    print(f"[thermal] chain {mode} fp{precision}: max bin error {err:.4f}", flush=True)
    assert err <= CHAIN["atol"], err


@pytest.mark.parametrize("precision", [64, 32])
def test_paramagnet_langevin_function(precision):
    p = PARA
    sys, h = paramagnet(precision=precision)
    announce(f"thermal paramagnet kappa=3/2 fp{precision} x{sys.nsites}")
    integ = sh.Langevin(p["dt"], damping=p["damping"], kT=p["kT"])
    S0 = sys.dipoles_flat()
    sh.step_(sys, integ, nsteps=p["n_equilib"])
    if precision == 64:
        S_ref, _ = O.Oracle(sys.spec()).langevin_run(S=S0, dt=integ.dt, damping=integ.damping, kT=integ.kT,
                                                     key=O.replica_key(sys.noise_seed, 0), step0=0,
                                                     nsteps=p["n_equilib"], every=0)
        assert np.max(np.abs(sys.dipoles_flat() - S_ref)) <= 1e-11
    E = []
    for _ in range(p["nsteps"] // p["every"]):
        sh.step_(sys, integ, nsteps=p["every"])
        E.append(sh.energy(sys))
    m = float(np.mean(E)) / (h * sys.nsites)
    print(f"[thermal] paramagnet fp{precision}: {m:.5f} vs {paramagnet_expected():.5f}", flush=True)
    assert abs(m - paramagnet_expected()) <= 0.01 * p["kappa"]
