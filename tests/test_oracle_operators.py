"""Pins the oracle's operator algebra and onsite anisotropy (CPU only).

Checks are independent of the formulas under test: commutation relations,
finite differences, coherent-state identities, and the SU(N) <-> dipole
equivalence of renormalised Stevens anisotropies (test/test_rescaling.jl:
111-137 in the reference).
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O


@pytest.mark.parametrize("N", [2, 3, 4, 5, 7])
def test_spin_matrices_algebra(N):
    s = (N - 1) / 2
    Sx, Sy, Sz = O.spin_matrices(N)
    np.testing.assert_allclose(Sx @ Sy - Sy @ Sx, 1j * Sz, atol=1e-13)
    np.testing.assert_allclose(Sy @ Sz - Sz @ Sy, 1j * Sx, atol=1e-13)
    np.testing.assert_allclose(Sx @ Sx + Sy @ Sy + Sz @ Sz, s * (s + 1) * np.eye(N), atol=1e-12)
    # host (numpy) and oracle (C) constructions agree
    for a, b in zip((Sx, Sy, Sz), sh.spin_matrices(s)):
        np.testing.assert_allclose(a, b, atol=1e-14)


@pytest.mark.parametrize("N", [3, 4, 5, 7])
def test_quadrupole_operators(N):
    """O[2,q] from the oracle equal the host Stevens operators, and their
    coherent-state expectations are rcs_2(s) times the classical quadrupole."""
    s = (N - 1) / 2
    Q = O.stevens_quadrupoles(N)
    Oh = sh.stevens_matrices(s)[2]
    for a in range(5):
        np.testing.assert_allclose(Q[a], Oh[a], atol=1e-12)
    rng = np.random.default_rng(N)
    for _ in range(3):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        Z = sh.ket_from_dipole(n, N)
        ev = np.array([np.vdot(Z, Q[a] @ Z).real for a in range(5)])
        x, y, z = s * n
        classical = np.array([x * x - y * y, z * x, 3 * z * z - (x * x + y * y + z * z), z * y, 2 * x * y])
        np.testing.assert_allclose(ev, sh.rcs_factors(s)[2] * classical, atol=1e-12)


def test_expected_spin_of_coherent_state():
    for N in (2, 3, 5, 7):
        s = (N - 1) / 2
        n = np.array([0.3, -0.5, 0.8])
        n /= np.linalg.norm(n)
        Z = sh.ket_from_dipole(n, N)[None, :]
        np.testing.assert_allclose(O.expected_spin(N, Z)[0], s * n, atol=1e-13)


def test_sz2_stevens_coefficients():
    """-0.1 Sz^2 for s = 1: c0 = -0.1*X/3 with X = s(s+1) = 2, c2[q=0] =
    -0.1/3 * rcs_2(1) = -1/60 (OnsiteCoupling.jl:46-55)."""
    s = sh.workloads.cubic_model(L=1, D=-0.1)
    c = s.onsite[0].packed()
    assert abs(c[0] - (-1 / 15)) < 1e-15
    np.testing.assert_allclose(c[1:6], [0, 0, -1 / 60, 0, 0], atol=1e-15)
    # classical energy of S along z equals the quantum energy of |m=1>: -0.1
    E, g = O.stevens_energy_grad(np.array([0, 0, 1.0]), c, 2)
    assert abs(E - (-0.1)) < 1e-15


@pytest.mark.parametrize("N", [3, 5, 7])
def test_stevens_expansion_is_complete(N):
    """Any Hermitian N x N matrix (N <= 7) is recovered from its Stevens
    coefficients (Stevens.jl:144-159)."""
    rng = np.random.default_rng(N)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    A = A + A.conj().T
    c = sh.matrix_to_stevens_coefficients(A)
    Om = sh.stevens_matrices((N - 1) / 2)
    R = sum(c[k][m] * Om[k][m] for k in range(7) for m in range(2 * k + 1))
    np.testing.assert_allclose(R, A, atol=1e-10)


def _random_stevens(rng, kmax=6):
    c = np.zeros(28)
    c[0] = rng.standard_normal()
    c[1:6] = rng.standard_normal(5)
    if kmax >= 4:
        c[6:15] = rng.standard_normal(9)
    if kmax >= 6:
        c[15:28] = rng.standard_normal(13)
    return c


@pytest.mark.parametrize("kmax", [2, 4, 6])
def test_stevens_gradient_finite_differences(kmax):
    """Perpendicular part of the closed-form gradient matches central
    differences of the energy (test/test_operators.jl:263-298)."""
    rng = np.random.default_rng(kmax)
    c = _random_stevens(rng, kmax)
    for _ in range(5):
        S = rng.standard_normal(3)
        S /= np.linalg.norm(S)
        _, g = O.stevens_energy_grad(S, c, kmax)
        h = 1e-6
        fd = np.array([(O.stevens_energy_grad(S + h * e, c, kmax)[0] - O.stevens_energy_grad(S - h * e, c, kmax)[0])
                       / (2 * h) for e in np.eye(3)])
        g_perp = g - (g @ S) * S
        fd_perp = fd - (fd @ S) * S
        np.testing.assert_allclose(g_perp, fd_perp, atol=1e-7 * max(1, np.abs(fd).max()))


@pytest.mark.parametrize("s", [2.0, 3.0])
def test_anisotropy_sun_dipole_equivalence(s):
    """E of an SU(N) coherent state with onsite Lambda = sum c O equals the
    :dipole energy with rcs-renormalised classical Stevens coefficients
    (test/test_rescaling.jl:111-137).  Pins the Stevens matrices, the
    coefficient extraction, the renormalisation factors and the classical
    closed form together."""
    N = int(2 * s + 1)
    rng = np.random.default_rng(int(s))
    Om = sh.stevens_matrices(s)
    Lam = sum(rng.standard_normal() * Om[k][m] for k in (2, 4, 6) if k < N for m in range(2 * k + 1))
    cryst = sh.cubic_crystal()
    sd = sh.System(cryst, [(0, sh.Moment(s=s, g=2))], "dipole")
    sh.set_onsite_coupling_(sd, Lam, 0)
    su = sh.System(cryst, [(0, sh.Moment(s=s, g=2))], "SUN")
    sh.set_onsite_coupling_(su, Lam, 0)
    od, ou = O.Oracle(sd.spec()), O.Oracle(su.spec())
    for _ in range(4):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        Z = sh.ket_from_dipole(n, N)[None, :]
        E_sun = ou.energy(Z=Z)
        E_dip = od.energy(S=(s * n)[None, :])
        assert abs(E_sun - E_dip) < 1e-10 * max(1, abs(E_sun))
