"""GPU parity of inhomogeneous systems (per-site couplings, vacancies,
per-site fields and anisotropies) against the oracle, through the C ABI.
The oracle's inhomogeneous semantics are pinned in
test_oracle_inhomogeneous.py (direct sums, identity with homogeneous
systems, the reference's two-site chain)."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce, rel_err
from test_oracle_inhomogeneous import edited_cubic, reference_chain
from test_oracle_thermal import CHAIN, chain_distribution_error

pytestmark = pytest.mark.gpu

TOL = {64: 1e-11, 32: 2e-5}


def models(precision):
    t = edited_cubic(seed=11)
    t.precision = precision
    yield "cubic_edited", t
    p = sh.to_inhomogeneous(sh.workloads.pyrochlore_model(L=2, seed=12, precision=precision))
    sh.set_vacancy_at_(p, (1, 0, 1, 5))
    sh.set_exchange_at_(p, np.array([[0.4, 0.1, 0], [0.1, 0.4, 0], [0, 0, 0.9]]), (0, 0, 0, 0), (0, 0, 0, 1),
                        offset=(0, 0, 0))
    sh.randomize_spins_(p)
    yield "pyrochlore2_edited", p


@pytest.mark.parametrize("precision", [64, 32])
def test_inhomogeneous_dipole_parity(precision, layout):
    for name, s in models(precision):
        s.layout = layout
        announce(f"inhomogeneous {name} fp{precision} layout {layout}")
        o = O.Oracle(s.spec())
        S = s.dipoles_flat()
        assert rel_err(sh.energy_grad_dipoles(s), o.grad(S)) <= TOL[precision], name
        E_ref = o.energy(S)
        assert abs(sh.energy(s) - E_ref) <= TOL[precision] * max(1.0, abs(E_ref)) * 10, name
        rng = np.random.default_rng(4)
        integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
        for _ in range(2):
            nz = rng.standard_normal((s.nsites, 3))
            S = o.step_langevin(S=S, noise=nz, dt=integ.dt, damping=integ.damping, kT=integ.kT)
            sh.step_(s, integ, noise=nz)
        assert rel_err(s.dipoles_flat(), S, scale=1.5) <= TOL[precision] * 2, name
        vac = s.site_kappas() == 0
        assert np.all(s.dipoles_flat()[vac] == 0)
        tol = 1e-12 if precision == 64 else 1e-6
        S, _ = o.step_midpoint(S=s.dipoles_flat(), dt=0.02, tol=tol)
        sh.step_(s, sh.ImplicitMidpoint(0.02, tol=tol))
        assert rel_err(s.dipoles_flat(), S, scale=1.5) <= (TOL[precision] if precision == 64 else 5e-6), name


def test_inhomogeneous_copy_matches_homogeneous_device():
    """An inhomogeneous copy steps exactly like its homogeneous original
    (FP64, device against device, same Philox noise)."""
    h = sh.workloads.pyrochlore_model(L=2, seed=13, precision=64)
    t = sh.to_inhomogeneous(h)
    announce("inhomogeneous copy vs homogeneous, pyrochlore 2^3 fp64")
    integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
    sh.step_(h, integ, nsteps=5)
    sh.step_(t, integ, nsteps=5)
    assert rel_err(t.dipoles_flat(), h.dipoles_flat(), scale=1.5) <= 1e-12
    assert abs(sh.energy(t) - sh.energy(h)) <= 1e-11 * abs(sh.energy(h))


@pytest.mark.parametrize("precision", [64, 32])
def test_inhomogeneous_sun_and_ewald(precision):
    from test_oracle_general import pair_model
    s, _ = pair_model(N=3, seed=14, L=(3, 2, 2), precision=precision)
    t = sh.to_inhomogeneous(s)
    sh.set_vacancy_at_(t, (2, 1, 0, 0))
    Sx, Sy, Sz = sh.spin_matrices(1.0)
    sh.set_onsite_coupling_at_(t, 0.4 * Sx @ Sx, (0, 0, 1, 0))
    sh.set_field_at_(t, [0.0, 0.2, 0.1], (1, 0, 0, 0))
    sh.randomize_spins_(t)
    announce(f"inhomogeneous SU(3) general + vacancy fp{precision}")
    o = O.Oracle(t.spec())
    Z = t.coherents_flat()
    assert rel_err(sh.energy_grad_coherents(t), o.grad_coherents(Z)) <= TOL[precision]
    rng = np.random.default_rng(6)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
    nz = (rng.standard_normal((t.nsites, 3)) + 1j * rng.standard_normal((t.nsites, 3))) / np.sqrt(2)
    Z = o.step_langevin(Z=Z, noise=nz, dt=integ.dt, damping=integ.damping, kT=integ.kT)
    sh.step_(t, integ, noise=nz)
    assert rel_err(t.coherents_flat(), Z, scale=1.0) <= TOL[precision] * 2
    # Ewald on the mapped cell: every pair of the 16 pyrochlore sites, one vacancy
    e = sh.to_inhomogeneous(sh.workloads.pyrochlore_model(L=1, J1=0.05, J2=0, A=0, s=3.5, seed=15, dipolar=True,
                                                          precision=precision))
    sh.set_vacancy_at_(e, (0, 0, 0, 3))
    sh.randomize_spins_(e)
    announce(f"inhomogeneous Ewald pyrochlore 1^3 fp{precision}")
    oe = O.Oracle(e.spec())
    S = e.dipoles_flat()
    assert rel_err(sh.energy_grad_dipoles(e), oe.grad(S)) <= TOL[precision] * 3
    assert abs(sh.energy(e) - oe.energy(S)) <= TOL[precision] * 10 * max(1.0, abs(oe.energy(S)))


@pytest.mark.parametrize("mode", ["dipole", "SUN"])
def test_reference_chain_distribution(mode):
    """The two-site chain of test_samplers.jl:174-227 as the reference writes
    it (inhomogeneous, one bond), 64 replicas on the device, FP64."""
    R = 64
    systems = [reference_chain(mode, seed=200 + r) for r in range(R)]
    announce(f"inhomogeneous reference chain {mode} x{R}")
    batch = sh.ReplicaBatch(systems, precision=64)
    integ = sh.Langevin(CHAIN["dt"], damping=CHAIN["damping"], kT=CHAIN["kT"])
    sh.step_(batch, integ, nsteps=CHAIN["n_equilib"])
    E = []
    for _ in range(CHAIN["n_samples"] // R):
        sh.step_(batch, integ, nsteps=CHAIN["n_decorr"])
        E.append(sh.energy(batch))
    assert chain_distribution_error(np.concatenate(E)) <= CHAIN["atol"]


def test_inhomogeneous_local_sampler():
    t = edited_cubic(seed=16)
    announce("inhomogeneous LocalSampler fp64")
    m, nc, col = sh.local_coloring(t)
    o = O.Oracle(t.spec())
    S0 = t.dipoles_flat()
    key = O.replica_key(t.noise_seed, 0)
    S_ref, acc = o.local_sweeps(S0, m, col.reshape(-1), nc, [(2, 0.3, 1.0)], 0.5, key, t.step_counter, 2.0)
    sampler = sh.LocalSampler(kT=0.5, nsweeps=2.0, propose=sh.propose_delta(0.3))
    sh.step_(t, sampler)
    assert np.max(np.abs(t.dipoles_flat() - S_ref)) <= 1e-12 * 1.5
    assert np.all(t.dipoles_flat()[t.site_kappas() == 0] == 0)
