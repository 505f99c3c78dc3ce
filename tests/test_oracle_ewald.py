"""Pins the oracle's Ewald dipole-dipole sum (CPU only).

Known answers and identities from the reference's test/test_ewald.jl:
single dipole in a cubic box (-1/6), FFT field equals the pairwise sum,
E = 1/2 sum S.dE, surface term M'NM/2V, and resize invariance.
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O


def single_dipole(dims, demag=0.0):
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=1))], "dipole", dims=dims)
    sh.enable_dipole_dipole_(s, 1.0, demag=demag)
    return s


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4)])
def test_single_dipole_energy(dims):
    s = single_dipole(dims)
    o = O.Oracle(s.spec())
    assert abs(o.energy(s.dipoles_flat()) / s.nsites - (-1 / 6)) < 1e-13


def triclinic(demag):
    rng = np.random.default_rng(0)
    cryst = sh.Crystal(sh.lattice_vectors(1.1, 0.9, 0.8, 92, 85, 95),
                       [[0, 0.01, 0.01], [0.1, 0.01, 0.01], [0.6, 0.4, 0.5]])
    mom = [(0, sh.Moment(s=1, g=rng.random((3, 3)))), (1, sh.Moment(s=1.5, g=rng.random((3, 3)))),
           (2, sh.Moment(s=2, g=rng.random((3, 3))))]
    s = sh.System(cryst, mom, "dipole", dims=(2, 3, 1), seed=1)
    sh.enable_dipole_dipole_(s, 1.0, demag=demag)
    sh.randomize_spins_(s)
    return s


def test_fft_field_equals_pairwise_sum():
    R = np.random.default_rng(3).standard_normal((3, 3))
    s = triclinic(R.T @ R)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    np.testing.assert_allclose(o.grad(S), o.ewald_grad_bruteforce(S), atol=1e-12)
    # energy as a sum over pairs
    assert abs(o.energy(S) - 0.5 * np.sum(S * o.ewald_grad_bruteforce(S))) < 1e-12


def test_surface_term():
    """Only the k = 0 term depends on demag: E(N) - E(0) = M'NM/(2V)."""
    R = np.random.default_rng(4).standard_normal((3, 3))
    N = R.T @ R
    s1, s0 = triclinic(N), triclinic(np.zeros((3, 3)))
    S = s1.dipoles_flat()
    E1 = O.Oracle(s1.spec()).energy(S)
    E0 = O.Oracle(s0.spec()).energy(S)
    g = s1.gs
    mom = np.array([-(g[i] @ S[k]) for i in range(3) for k in range(i * 6, (i + 1) * 6)])
    M = mom.sum(0)
    V = abs(np.linalg.det(s1.crystal.latvecs)) * np.prod(s1.dims)
    assert abs((E1 - E0) - M @ N @ M / (2 * V)) < 1e-12


def test_energy_per_site_resize_invariant():
    s = triclinic(np.eye(3) / 3)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    e1 = o.energy(S) / s.nsites
    big = sh.System(s.crystal, [(i, sh.Moment(s=s.spins[i], g=s.gs[i])) for i in range(3)], "dipole",
                    dims=(4, 3, 2))
    sh.enable_dipole_dipole_(big, 1.0, demag=np.eye(3) / 3)
    Sv = S.reshape(3, 1, 3, 2, 3)  # [i][c3][c2][c1][xyz] of the 2x3x1 system
    Sb = np.tile(Sv, (1, 2, 1, 2, 1)).reshape(-1, 3)
    e2 = O.Oracle(big.spec()).energy(Sb) / big.nsites
    assert abs(e1 - e2) < 1e-12


def test_ewald_tensor_symmetry():
    """A[off, i, j] = A[-off, j, i]' (the reference leaves this unverified,
    src/System/Ewald.jl:149); it makes every F[A]_k Hermitian."""
    s = triclinic(np.eye(3) / 3)
    A = O.ewald_A(s.spec(), s.dims)  # [j][i][cell][3][3]
    L1, L2, L3 = s.dims
    for j in range(3):
        for i in range(3):
            for c in range(L1 * L2 * L3):
                c1, c2, c3 = c % L1, (c // L1) % L2, c // (L1 * L2)
                m = (-c1) % L1 + L1 * ((-c2) % L2 + L2 * ((-c3) % L3))
                np.testing.assert_allclose(A[j, i, c], A[i, j, m].T, atol=1e-14)


def smooth_cutoff(x, x0=0.5):
    """1 for x <= x0, 0 for x >= 1, infinitely differentiable between."""
    t = np.clip((x - x0) / (1 - x0), 0, 1)

    def f(u):
        return np.where(u > 0, np.exp(-1 / np.maximum(u, 1e-300)), 0.0)

    return f(1 - t) / (f(1 - t) + f(t))


def direct_dipole_sum(sys, S, Rc):
    """Energy and dE/dS of the periodic dipole lattice by a direct sum over
    image dipoles, weighted by a smooth spherical cutoff of radius Rc.

    The kernel is that of modify_exchange_with_truncated_dipole_dipole!
    (Ewald.jl:279): mu_a.T(r) mu_b with T(r) = (I - 3 rhat rhat') / (4 pi r^3),
    mu = -g S.  A spherical cutoff sums the conditionally convergent lattice
    sum in the shape of a sphere in vacuum, whose limit is the Ewald sum with
    demag = 1/3; the smooth cutoff removes the lattice-point discreteness so
    the limit is reached at moderate Rc.  Nothing here is shared with the
    oracle's Ewald split (no erfc, no reciprocal sum, no FFT)."""
    cr = sys.crystal
    L1, L2, L3 = sys.dims
    X, mu, g = [], [], []
    for i in range(cr.natoms):
        for c3 in range(L3):
            for c2 in range(L2):
                for c1 in range(L1):
                    k = c1 + L1 * (c2 + L2 * (c3 + L3 * i))
                    X.append(cr.latvecs @ (np.array([c1, c2, c3]) + cr.positions[i]))
                    mu.append(-sys.gs[i] @ S[k])
                    g.append(sys.gs[i])
    X, mu = np.array(X), np.array(mu)
    SL = cr.latvecs * np.array(sys.dims)[None, :]
    diam = np.max(np.linalg.norm(X[:, None] - X[None], axis=-1))
    inv = np.linalg.inv(SL)
    nmax = [int(np.ceil((Rc + diam) * np.linalg.norm(inv[k]))) + 1 for k in range(3)]
    n = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in nmax], indexing="ij"), -1).reshape(-1, 3)
    R = n @ SL.T
    R = R[np.linalg.norm(R, axis=1) < Rc + diam + 1]
    E, grad = 0.0, np.zeros_like(mu)
    for a in range(len(X)):
        r = X[None, :, :] - X[a][None, None, :] + R[:, None, :]
        d = np.linalg.norm(r, axis=-1)
        mask = d > 1e-12
        dd = np.where(mask, d, 1.0)
        w = np.where(mask, smooth_cutoff(dd / Rc), 0.0)
        rh = r / dd[..., None]
        mb = mu[None, :, :]
        Tm = (mb - 3 * rh * np.sum(rh * mb, -1, keepdims=True)) / (4 * np.pi * dd[..., None] ** 3) * w[..., None]
        f = Tm.sum(axis=(0, 1))
        E += 0.5 * mu[a] @ f
        grad[a] = -g[a].T @ f
    return E, grad


def triclinic_dims(dims, demag=np.eye(3) / 3):
    s = triclinic(demag)
    big = sh.System(s.crystal, [(i, sh.Moment(s=s.spins[i], g=s.gs[i])) for i in range(3)], "dipole", dims=dims,
                    seed=5)
    sh.enable_dipole_dipole_(big, 1.0, demag=demag)
    sh.randomize_spins_(big)
    return big


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 1)])
def test_triclinic_ewald_equals_direct_sum(dims):
    """The random triclinic 3-atom cell of test_ewald.jl:33-57, compared
    against an independent direct dipole sum (there: Ewalder.jl, absent here)
    with demag = 1/3: energy and gradient."""
    s = triclinic_dims(dims)
    S = s.dipoles_flat()
    o = O.Oracle(s.spec())
    E_dir, g_dir = direct_dipole_sum(s, S, Rc=20.0)
    E = o.energy(S)
    # the cutoff sum reaches its limit to a few 1e-7 at Rc = 20 (about 20 cells)
    assert abs(E - E_dir) <= 2e-6 * abs(E), (E, E_dir)
    np.testing.assert_allclose(o.grad(S), g_dir, atol=2e-6 * np.max(np.abs(g_dir)))


def test_direct_sum_catches_planted_displacement_mistakes():
    """Fractional instead of Cartesian displacements, or the two atoms of a
    pair swapped (c + p_i - p_j), make a different but still periodic
    crystal: the identities of the tests above all still hold, the direct
    sum does not.  An axis of length 3 is needed for the swap: with axes of
    length 1 or 2 every offset equals its own negative."""
    s = triclinic_dims((2, 3, 1))
    S = s.dipoles_flat()
    E_dir, _ = direct_dipole_sum(s, S, Rc=20.0)
    for m in (O.mutation.EWALD_FRACTIONAL, O.mutation.EWALD_SWAPPED):
        with O.mutation(m):
            E = O.Oracle(s.spec()).energy(S)
        assert abs(E - E_dir) > 1e-3 * abs(E_dir), (m, E, E_dir)


def test_ewald_tensor_independent_of_splitting_width():
    """The real/reciprocal split of Ewald.jl:61-152 at sigma = cbrt(V)/3 and
    cbrt(V)/2.5 gives the same tensor: a dropped or mis-scaled term on either
    side would make it depend on sigma."""
    s = triclinic_dims((2, 3, 1), demag=np.diag([0.2, 0.3, 0.5]))
    A3 = O.ewald_A(s.spec(), s.dims)
    A25 = O.ewald_A(s.spec(), s.dims, sigma_div=2.5)
    assert np.max(np.abs(A3 - A25)) <= 1e-10 * np.max(np.abs(A3))
