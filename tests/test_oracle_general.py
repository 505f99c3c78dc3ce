"""General SU(N) pair couplings (set_pair_coupling!, the TensorDecomposition
part; src/System/PairExchange.jl:71-111,274-312, Interactions.jl:371-377,
496-503, 659-666) in the host decomposition and the oracle (CPU only).

Known answers: the decomposition reconstructs the operator; the classical
SU(N) energy of a pair coupling is its expectation value in the product of
the two (normalised) coherent states, <Zi (x) Zj| op |Zi (x) Zj>, summed over
bonds - computed here directly with Kronecker products, independently of
the oracle's term-by-term restatement; H Z is dE/dZbar (directional
derivatives); local_energy_change equals energy differences.
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O


def random_hermitian(rng, n):
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (M + M.conj().T) / 2


def pair_model(N=3, seed=0, extract_parts=True, L=(4, 3, 1), op=None, precision=64):
    """SU(N) chain of a cubic lattice with one general pair operator on
    Bond(0, 0, (1, 0, 0)) (that bond only) and an onsite term."""
    rng = np.random.default_rng(seed)
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=(N - 1) / 2, g=2))], "SUN", dims=L, seed=seed + 1,
                  precision=precision)
    op = random_hermitian(rng, N * N) if op is None else op
    sh.set_pair_coupling_(s, op, sh.Bond(0, 0, (1, 0, 0)), extract_parts=extract_parts)
    Sx, Sy, Sz = sh.spin_matrices((N - 1) / 2)
    if N > 2:
        sh.set_onsite_coupling_(s, 0.3 * Sz @ Sz + 0.1 * Sx, 0)
    sh.randomize_spins_(s)
    return s, op


def product_state_energy(s, op):
    """sum over cells of <Zc (x) Zc+x| op |Zc (x) Zc+x> with normalised kets."""
    Z = s.coherents_flat()
    L1, L2, L3 = s.dims
    E = 0.0
    for c3 in range(L3):
        for c2 in range(L2):
            for c1 in range(L1):
                a = c1 + L1 * (c2 + L2 * c3)
                b = (c1 + 1) % L1 + L1 * (c2 + L2 * c3)
                psi = np.kron(Z[a] / np.linalg.norm(Z[a]), Z[b] / np.linalg.norm(Z[b]))
                E += np.real(np.vdot(psi, op @ psi))
    return E


@pytest.mark.parametrize("N", [3, 4, 5])
def test_decomposition_reconstructs_operator(N):
    rng = np.random.default_rng(N)
    op = random_hermitian(rng, N * N)
    sc, J, B, G = sh.decompose_pair_coupling(op, N, N)
    S = sh.spin_matrices((N - 1) / 2)
    O2 = sh.stevens_matrices((N - 1) / 2)[2]
    rec = sc * np.eye(N * N, dtype=complex)
    rec += sum(J[a, b] * np.kron(S[a], S[b]) for a in range(3) for b in range(3))
    if B is not None:
        rec += sum(B[p, q] * np.kron(O2[p], O2[q]) for p in range(5) for q in range(5))
    rec += sum(np.kron(A, Bm) for A, Bm in G)
    np.testing.assert_allclose(rec, op, atol=1e-12)
    for A, Bm in G:
        np.testing.assert_allclose(A, A.conj().T, atol=1e-14)
        np.testing.assert_allclose(Bm, Bm.conj().T, atol=1e-14)
    # without extraction everything but the scalar is general
    sc2, J2, B2, G2 = sh.decompose_pair_coupling(op, N, N, extract_parts=False)
    assert np.all(J2 == 0) and B2 is None
    np.testing.assert_allclose(sc2 * np.eye(N * N) + sum(np.kron(A, Bm) for A, Bm in G2), op, atol=1e-12)


@pytest.mark.parametrize("N,extract", [(3, True), (3, False), (4, True), (5, True)])
def test_energy_is_product_state_expectation(N, extract):
    s, op = pair_model(N=N, seed=10 + N, extract_parts=extract)
    base, _ = pair_model(N=N, seed=10 + N, op=np.zeros((N * N, N * N)))
    base.set_state_flat(coh=s.coherents_flat())
    Z = s.coherents_flat()
    E = O.Oracle(s.spec()).energy(Z=Z)
    E0 = O.Oracle(base.spec()).energy(Z=Z)
    assert abs((E - E0) - product_state_energy(s, op)) < 1e-11 * max(1.0, abs(E))


def test_general_part_is_present_and_reversed():
    s, op = pair_model(N=3, seed=3)
    sp = s.spec()
    assert sp["gen_start"] is not None and sp["gen_start"][-1] > 0
    b = sh.Bond(0, 0, (1, 0, 0))
    G, Gr = s.pairs[b][3], s.pairs[b.reverse()][3]
    assert len(G) == len(Gr) and all(np.array_equal(A, Br) and np.array_equal(B, Ar)
                                     for (A, B), (Ar, Br) in zip(G, Gr))


@pytest.mark.parametrize("N", [3, 5])
def test_general_gradient_directional_derivatives(N):
    s, _ = pair_model(N=N, seed=20 + N)
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    H = o.grad_coherents(Z)
    rng = np.random.default_rng(2)
    for _ in range(3):
        dZ = rng.standard_normal(Z.shape) + 1j * rng.standard_normal(Z.shape)
        h = 1e-6
        fd = (o.energy(Z=Z + h * dZ) - o.energy(Z=Z - h * dZ)) / (2 * h)
        assert abs(fd - 2 * np.real(np.vdot(dZ, H))) < 1e-6 * max(1, abs(fd))


def test_general_local_energy_change():
    s, _ = pair_model(N=4, seed=31)
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    rng = np.random.default_rng(5)
    for site in rng.choice(s.nsites, 4, replace=False):
        zn = rng.standard_normal(s.N) + 1j * rng.standard_normal(s.N)
        zn /= np.linalg.norm(zn)
        Z2 = Z.copy()
        Z2[site] = zn
        dE = o.local_energy_change_sun(Z, site, zn)
        assert abs(dE - (o.energy(Z=Z2) - o.energy(Z=Z))) < 1e-11


def test_dipole_mode_rejects_general_remainder():
    rng = np.random.default_rng(7)
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=2))], "dipole", dims=(3, 1, 1))
    with pytest.raises(ValueError, match="dipole mode"):
        sh.set_pair_coupling_(s, random_hermitian(rng, 9), sh.Bond(0, 0, (1, 0, 0)))
    # a bilinear plus biquadratic operator is allowed
    Si, Sj = sh.to_product_space(sh.spin_matrices(1), sh.spin_matrices(1))
    dot = sum(a @ b for a, b in zip(Si, Sj))
    sh.set_pair_coupling_(s, lambda Si, Sj: 0.5 * dot + 0.2 * dot @ dot, sh.Bond(0, 0, (1, 0, 0)))
    assert s.spec()["gen_start"] is None
