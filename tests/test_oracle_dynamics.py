"""Pins the oracle's energy gradients and integrators (CPU only).

Finite differences of the energy, closed-form precession of the spherical
midpoint method with the reference's handedness (test/test_chirality.jl),
energy conservation (test/test_energy_consistency.jl), norm invariance and
kappa time rescaling (test/test_rescaling.jl), the SU(3) thermal mean
energy (test/test_samplers.jl:6-15), non-convergence, and the Philox noise
stream (Random123 known answers).
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O


def rich_dipole_system(seed=1):
    """Diamond, s = 3/2: field, anisotropic exchange, k = 2, 4 Stevens,
    quadrupole coupling, scalar biquadratic and Ewald."""
    cryst = sh.diamond_crystal()
    s = sh.System(cryst, [(i, sh.Moment(s=1.5, g=2)) for i in range(8)], "dipole", dims=(1, 1, 2), seed=seed)
    sh.set_field_(s, [0.1, 0.2, 0.3])
    J = np.array([[0.5, 0.2, -0.4], [0.2, 0.5, -0.4], [0.4, 0.4, 1.5]])
    sh.set_exchange_(s, J, sh.Bond(0, 1, (0, 0, 0)))
    sh.set_exchange_(s, 0.3, sh.Bond(0, 2, (0, 0, 0)), biquad=0.2)
    Om = sh.stevens_matrices(1.5)
    for i in range(8):
        sh.set_onsite_coupling_(s, 0.3 * Om[2][0] + 0.2 * Om[2][4], i)
    sh.set_pair_biquad_(s, 0.1 * np.diag([1.2, 1, 1.1, 1, 1.3]), sh.Bond(0, 0, (0, 0, 1)))
    sh.enable_dipole_dipole_(s, 0.5)
    sh.randomize_spins_(s)
    return s


def test_dipole_gradient_finite_differences():
    s = rich_dipole_system()
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    g = o.grad(S)
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in rng.choice(s.nsites, 6, replace=False):
        fd = np.zeros(3)
        for a in range(3):
            Sp, Sm = S.copy(), S.copy()
            Sp[k, a] += h
            Sm[k, a] -= h
            fd[a] = (o.energy(Sp) - o.energy(Sm)) / (2 * h)
        u = S[k] / np.linalg.norm(S[k])
        # the Stevens gradient holds |S|^2 fixed: compare perpendicular parts
        np.testing.assert_allclose(g[k] - (g[k] @ u) * u, fd - (fd @ u) * u, atol=1e-6)


def test_sun_gradient_finite_differences():
    """H Z = dE/dZbar: the directional derivative of E(Z) along dZ is
    2 Re <dZ, HZ> (Interactions.jl:594-595)."""
    s = sh.workloads.triangular_sun_model(L1=3, L2=4, seed=3, precision=64)
    sh.set_field_(s, [0.2, 0.0, 0.3])
    sh.randomize_spins_(s)
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    H = o.grad_coherents(Z)
    rng = np.random.default_rng(1)
    for _ in range(3):
        dZ = rng.standard_normal(Z.shape) + 1j * rng.standard_normal(Z.shape)
        h = 1e-6
        fd = (o.energy(Z=Z + h * dZ) - o.energy(Z=Z - h * dZ)) / (2 * h)
        assert abs(fd - 2 * np.real(np.vdot(dZ, H))) < 1e-6 * max(1, abs(fd))


def single_spin(mode, B=(0, 0, 1.0)):
    cryst = sh.cubic_crystal(1.0)
    s = sh.System(cryst, [(0, sh.Moment(s=1, g=2))], mode)
    sh.set_field_(s, B)
    sh.set_dipole_(s, [1 / np.sqrt(2), 0, 1 / np.sqrt(2)], (0, 0, 0, 0))
    return s


def test_midpoint_precession_closed_form():
    """Spin perpendicular to its field: one spherical-midpoint step rotates it
    by exactly 2 asin(dt |dE/dS| / 2) about the field."""
    s = single_spin("dipole", B=(0, 0, 1.0))
    o = O.Oracle(s.spec())
    S0 = np.array([[1.0, 0.0, 0.0]])
    dt = 0.3
    S1, it = o.step_midpoint(S=S0, dt=dt, tol=1e-15)
    field = 2.0  # |g' B|
    phi = np.arccos(np.clip(S0[0] @ S1[0], -1, 1))
    assert abs(phi - 2 * np.arcsin(dt * field / 2)) < 1e-12
    assert abs(S1[0, 2]) < 1e-15


@pytest.mark.parametrize("mode", ["dipole", "SUN"])
def test_precession_handedness(mode):
    """B.(S0 x S) > 0 after 5 midpoint steps (test/test_chirality.jl:1-26,
    stated there for magnetic moments -g S)."""
    s = single_spin(mode)
    o = O.Oracle(s.spec())
    ic = np.array([1 / np.sqrt(2), 0, 1 / np.sqrt(2)])
    if mode == "dipole":
        S = s.dipoles_flat()
        for _ in range(5):
            S, _ = o.step_midpoint(S=S, dt=0.05)
        S1 = S[0]
    else:
        Z = s.coherents_flat()
        for _ in range(5):
            Z, _ = o.step_midpoint(Z=Z, dt=0.05)
        S1 = o.expected_spin(Z)[0]
    assert np.array([0, 0, 1.0]) @ np.cross(ic, S1) > 0


def test_midpoint_energy_conservation():
    s = rich_dipole_system(seed=2)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    E = []
    for _ in range(150):
        S, it = o.step_midpoint(S=S, dt=0.002)
        assert it > 0
        E.append(o.energy(S))
    assert (max(E) - min(E)) / np.sqrt(s.nsites) < 1e-3


def test_norms_preserved():
    s = rich_dipole_system(seed=3)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    rng = np.random.default_rng(0)
    for _ in range(20):
        S = o.step_langevin(S=S, noise=rng.standard_normal(S.shape), dt=0.01, damping=0.1, kT=0.5)
    np.testing.assert_allclose(np.linalg.norm(S, axis=1), 1.5, atol=1e-13)
    t = sh.workloads.triangular_sun_model(L1=3, L2=3, seed=4, precision=64)
    ot = O.Oracle(t.spec())
    Z = t.coherents_flat()
    for _ in range(20):
        nz = (rng.standard_normal(Z.shape) + 1j * rng.standard_normal(Z.shape)) / np.sqrt(2)
        Z = ot.step_langevin(Z=Z, noise=nz, dt=0.01, damping=0.1, kT=0.5)
    np.testing.assert_allclose(np.sum(np.abs(Z) ** 2, axis=1), 1.0, atol=1e-13)


@pytest.mark.parametrize("order,power", [("exchange", 1), ("quartic", 3)])
def test_kappa_time_rescaling(order, power):
    """Scaling |S| by kappa and dt by kappa^-p maps trajectories onto each
    other (test/test_rescaling.jl:61-89): p = 1 for exchange (quadratic in
    S), p = 3 for a quartic anisotropy."""
    def traj(kappa, dt):
        cryst = sh.diamond_crystal()
        s = sh.System(cryst, [(i, sh.Moment(s=1.5, g=2)) for i in range(8)], "dipole", dims=(1, 1, 1), seed=0)
        if order == "exchange":
            sh.set_exchange_(s, np.array([[0.5, 0.2, -0.4], [0.2, 0.5, -0.4], [0.4, 0.4, 1.5]]),
                             sh.Bond(0, 1, (0, 0, 0)))
        else:
            st = sh.StevensExpansion(np.zeros(1), np.zeros(5), np.array([0, 0, 0, 0.01, 0.02, 0, 0, 0, 0.03]),
                                     np.zeros(13))
            for i in range(8):
                sh.set_stevens_expansion_(s, st, i)
        sh.set_spin_rescaling_(s, np.full(8, kappa))
        sh.randomize_spins_(s)
        o = O.Oracle(s.spec())
        S = s.dipoles_flat()
        for _ in range(50):
            S, _ = o.step_midpoint(S=S, dt=dt)
        return S

    dt = 0.005
    k = 2.0
    np.testing.assert_allclose(traj(1.0, dt), traj(k, dt / k ** power) / k, atol=1e-9)


def su3_mean_energy(kT, D):
    a = D / kT
    return D * (2 - (2 + 2 * a + a ** 2) * np.exp(-a)) / (a * (1 - (1 + a) * np.exp(-a)))


@pytest.mark.slow
@pytest.mark.parametrize("kT", [0.125, 0.5])
def test_su3_thermal_mean_energy(kT):
    """Langevin sampling of Lambda = D Sz^2 for SU(3) reproduces the exact
    mean energy (test/test_samplers.jl:6-15, rtol 0.1).  Pins the noise
    magnitude sqrt(2 dt lambda kT) and the complex-noise convention."""
    D, L, dt, damping = 1.0, 20, 0.01, 1.0
    cryst = sh.Crystal(sh.lattice_vectors(1, 1, 1, 90, 90, 90), [[0, 0, 0]])
    s = sh.System(cryst, [(0, sh.Moment(s=1, g=2))], "SUN", dims=(L, 1, 1), seed=0)
    sh.set_onsite_coupling_(s, lambda S: D * S[2] @ S[2], 0)
    sh.randomize_spins_(s)
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    rng = np.random.default_rng(7)
    E = []
    for k in range(int(10 / dt) + int(50 / dt)):
        nz = (rng.standard_normal(Z.shape) + 1j * rng.standard_normal(Z.shape)) / np.sqrt(2)
        Z = o.step_langevin(Z=Z, noise=nz, dt=dt, damping=damping, kT=kT)
        if k >= int(10 / dt):
            E.append(o.energy(Z=Z) / L)
    assert abs(np.mean(E) - su3_mean_energy(kT, D)) < 0.1 * su3_mean_energy(kT, D)


def test_midpoint_failure_returns_minus_one():
    s = rich_dipole_system(seed=5)
    o = O.Oracle(s.spec())
    _, it = o.step_midpoint(S=s.dipoles_flat(), dt=0.05, max_iters=1)
    assert it == -1


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    assert O.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_normals_statistics():
    x = O.philox_normals(12345, 7, 200000, 3).ravel()
    assert abs(x.mean()) < 0.01
    assert abs(x.var() - 1) < 0.01
    assert abs(np.mean(x ** 4) - 3) < 0.05
