"""GPU parity of the :SUN (coherent-state) path against the FP64 oracle."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce, energy_scale, rel_err

pytestmark = pytest.mark.gpu

TOL = {64: 1e-11, 32: 2e-5}


def tri_model(precision, L1=6, L2=5, seed=3):
    return sh.workloads.triangular_sun_model(L1=L1, L2=L2, seed=seed, precision=precision)


def spin2_model(precision):
    """SU(5) (s = 2) on a diamond 1x1x2 cell: field, exchange matrix, dense
    onsite matrix, and a quadrupole coupling."""
    cryst = sh.diamond_crystal()
    s = sh.System(cryst, [(i, sh.Moment(s=2, g=2)) for i in range(8)], "SUN", dims=(1, 1, 2), seed=9,
                  precision=precision)
    sh.set_field_(s, [0.1, -0.2, 0.3])
    sh.set_exchange_(s, np.array([[0.5, 0.2, -0.4], [0.2, 0.5, -0.4], [0.4, 0.4, 1.5]]), sh.Bond(0, 1, (0, 0, 0)))
    Sx, Sy, Sz = sh.spin_matrices(2)
    for i in range(8):
        sh.set_onsite_coupling_(s, 0.2 * (Sx @ Sx @ Sx @ Sx + Sy @ Sy @ Sy @ Sy + Sz @ Sz @ Sz @ Sz), i)
    sh.set_pair_biquad_(s, 0.1 * np.diag([1.2, 1, 1.1, 1, 1.3]), sh.Bond(0, 0, (0, 0, 1)))
    sh.randomize_spins_(s)
    return s


def general_model(precision, N=3, seed=40):
    """A general pair operator (scalar, bilinear, biquadratic and tensor
    terms) on one bond of an SU(N) cubic lattice, plus a field."""
    from test_oracle_general import pair_model
    s, _ = pair_model(N=N, seed=seed, L=(4, 3, 2), precision=precision)
    sh.set_field_(s, [0.1, -0.2, 0.3])
    sh.randomize_spins_(s)
    return s


def cases(precision):
    yield "tri6x5", tri_model(precision)
    yield "tri1x3", tri_model(precision, L1=1, L2=3, seed=5)
    yield "su5_diamond", spin2_model(precision)
    yield "su3_general", general_model(precision)
    yield "su4_general", general_model(precision, N=4, seed=41)


@pytest.mark.parametrize("precision", [64, 32])
def test_grad_energy_sun(precision):
    for name, s in cases(precision):
        announce(f"sun grad/energy {name} fp{precision}")
        Z = s.coherents_flat()
        o = O.Oracle(s.spec())
        H_dev = sh.energy_grad_coherents(s)
        H_ref = o.grad_coherents(Z)
        assert rel_err(H_dev, H_ref) <= TOL[precision], name
        E_dev, E_ref = sh.energy(s), o.energy(Z=Z)
        assert abs(E_dev - E_ref) <= TOL[precision] * 0.1 * energy_scale(s) + 1e-13, (name, E_dev, E_ref)
        # dipoles reported for a coherent state are expected_spin(Z)
        assert rel_err(s.dipoles_flat(), o.expected_spin(Z), scale=1.0) <= TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_langevin_sun_host_noise(precision):
    for name, s in cases(precision):
        announce(f"sun langevin host-noise {name} fp{precision}")
        o = O.Oracle(s.spec())
        rng = np.random.default_rng(5)
        integ = sh.Langevin(0.02, damping=0.1, kT=0.7)
        Z = s.coherents_flat()
        for _ in range(3):
            noise = (rng.standard_normal((s.nsites, s.N)) + 1j * rng.standard_normal((s.nsites, s.N))) / np.sqrt(2)
            Z = o.step_langevin(Z=Z, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
            sh.step_(s, integ, noise=noise)
        Zd = s.coherents_flat()
        assert rel_err(Zd, Z, scale=1.0) <= 3 * TOL[precision], name
        assert np.max(np.abs((np.abs(Zd) ** 2).sum(1) - 1.0)) <= (1e-12 if precision == 64 else 1e-6)


def test_langevin_sun_philox():
    s = tri_model(64, 8, 8, seed=4)
    announce("sun langevin philox tri8x8 fp64 x10")
    o = O.Oracle(s.spec())
    integ = sh.Langevin(0.02, damping=0.1, kT=0.4)
    Z = s.coherents_flat()
    key = O.replica_key(s.noise_seed, 0)
    for k in range(10):
        n = O.philox_normals(key, k, s.nsites, 2 * s.N) / np.sqrt(2)
        noise = n[:, 0::2] + 1j * n[:, 1::2]
        Z = o.step_langevin(Z=Z, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
    sh.step_(s, integ, nsteps=10)
    assert rel_err(s.coherents_flat(), Z, scale=1.0) <= 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_midpoint_sun(precision):
    for name, s in cases(precision):
        announce(f"sun midpoint {name} fp{precision}")
        o = O.Oracle(s.spec())
        tol = 1e-12 if precision == 64 else 1e-6
        integ = sh.ImplicitMidpoint(0.02, damping=0.05, tol=tol)
        Z = s.coherents_flat()
        its_ref = []
        for _ in range(3):
            Z, it = o.step_midpoint(Z=Z, dt=integ.dt, damping=0.05, tol=tol)
            its_ref.append(it)
        its_dev = [int(sh.step_(s, integ)[0]) for _ in range(3)]
        err = rel_err(s.coherents_flat(), Z, scale=1.0)
        if precision == 64:
            assert err <= 1e-11, name
            assert all(abs(a - b) <= 1 for a, b in zip(its_dev, its_ref)), (its_dev, its_ref)
        else:
            assert err <= 5e-6, (name, err)


def test_sun_batch_uneven():
    sizes = [(8, 8), (5, 7), (12, 3)]
    systems = [tri_model(32, a, b, seed=30 + r) for r, (a, b) in enumerate(sizes)]
    announce(f"sun batch uneven {sizes} fp32")
    batch = sh.ReplicaBatch(systems, precision=32)
    kTs = np.array([0.1, 0.5, 2.0])
    integ = sh.Langevin(0.02, damping=0.1, kT=0.0)
    refs = [s.coherents_flat() for s in systems]
    sh.step_(batch, integ, nsteps=4, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        o = O.Oracle(s.spec())
        Z = refs[r]
        key = O.replica_key(batch.noise_seed, r)
        for k in range(4):
            n = O.philox_normals(key, k, s.nsites, 2 * s.N) / np.sqrt(2)
            Z = o.step_langevin(Z=Z, noise=n[:, 0::2] + 1j * n[:, 1::2], dt=integ.dt, damping=integ.damping,
                                kT=kTs[r])
        assert rel_err(s.coherents_flat(), Z, scale=1.0) <= 1e-4
