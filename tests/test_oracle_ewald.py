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
