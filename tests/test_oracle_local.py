"""Pins the oracle's Metropolis restatement (CPU only).

local_energy_change (Interactions.jl:311-380) against differences of the
already pinned total energy, with every interaction kind including Ewald;
the self-interaction error; the accumulated dE, dS against the states; and
the sampled distribution of the colour-class passes against closed forms (a
paramagnet's Langevin function, the periodic Ising chain, which a fixed
sweep order with flips fails) and against the
reference's random-site order (Samplers.jl:133-153).
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O
from helpers import coloring_is_proper, langevin_function, site_coloring
from test_oracle_dynamics import rich_dipole_system

UNIFORM, FLIP, DELTA = 0, 1, 2


def proposals_for(S0, rng, kappa):
    """A uniform, a flipped and a perturbed dipole for one site."""
    n = rng.standard_normal(3)
    yield kappa * n / np.linalg.norm(n)
    yield -S0
    d = S0 + 0.3 * kappa * rng.standard_normal(3)
    yield kappa * d / np.linalg.norm(d)


@pytest.mark.parametrize("which", ["rich", "pyrochlore"])
def test_local_energy_change_is_energy_difference(which):
    if which == "rich":
        s = rich_dipole_system(seed=4)
    else:
        s = sh.workloads.pyrochlore_model(L=2, seed=5)
        sh.set_field_(s, [0.05, -0.1, 0.2])
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    E0 = o.energy(S)
    kap = np.repeat(s.spec()["kappa"], s.ncells)
    rng = np.random.default_rng(7)
    scale = max(abs(E0), 1.0)
    for site in rng.choice(s.nsites, 8, replace=False):
        for Sn in proposals_for(S[site], rng, kap[site]):
            S2 = S.copy()
            S2[site] = Sn
            dE = o.local_energy_change(S, site, Sn)
            assert abs(dE - (o.energy(S2) - E0)) <= 1e-11 * scale, (which, site)


def test_self_interaction_is_rejected():
    s = sh.System(sh.cubic_crystal(), [(0, sh.Moment(s=1, g=2))], "dipole", dims=(1, 2, 2), seed=0)
    sh.set_exchange_(s, 1.0, sh.Bond(0, 0, (1, 0, 0)))
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    assert np.isnan(o.local_energy_change(S, 0, -S[0]))
    m, nc, col = (1, 1, 1), 1, np.zeros(1, np.int32)
    with pytest.raises(ValueError, match="self-interaction"):
        o.local_sweeps(S, m, col, nc, [(FLIP, 0, 1)], 1.0, 0, 0, 1)


def test_accumulated_changes_match_states():
    s = sh.workloads.pyrochlore_model(L=2, seed=2)
    spec = s.spec()
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    assert coloring_is_proper(spec, s.dims, m, col)
    S = s.dipoles_flat()
    props = [(FLIP, 0, 0.3), (DELTA, 0.5, 0.5), (UNIFORM, 0, 0.2)]
    S2, acc = o.local_sweeps(S, m, col.reshape(-1), nc, props, 0.7, 123, 5, 3)
    assert np.sum(np.any(S2 != S, axis=1)) > s.nsites // 4
    np.testing.assert_allclose(acc[0], o.energy(S2) - o.energy(S), atol=1e-9)
    np.testing.assert_allclose(acc[1:], S2.sum(0) - S.sum(0), atol=1e-10)
    # kT = 0 never raises the energy; a fractional sweep tries about that share
    S3, acc3 = o.local_sweeps(S, m, col.reshape(-1), nc, props, 0.0, 9, 0, 1)
    assert acc3[0] <= 0 and o.energy(S3) <= o.energy(S) + 1e-12
    S4, _ = o.local_sweeps(S, m, col.reshape(-1), nc, [(FLIP, 0, 1)], 1e9, 9, 0, 0.25)
    moved = np.sum(np.any(S4 != S, axis=1)) / s.nsites
    assert 0.15 < moved < 0.35


def test_sequential_trials_with_ewald():
    """Sequential trials (the device's form when the Ewald term is on): the
    accumulated changes are the changes of the Ewald-inclusive energy and of
    the summed dipoles; kT = 0 never raises the energy; the draws pick sites
    uniformly (every site of 16 is tried over 40 sweeps)."""
    s = sh.workloads.pyrochlore_model(L=1, seed=21, dipolar=True)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    props = [(FLIP, 0, 0.3), (DELTA, 0.4, 0.5), (UNIFORM, 0, 0.2)]
    S2, acc = o.local_sequential(S, props, 0.5, 77, 3, 4.0)
    assert np.sum(np.any(S2 != S, axis=1)) > s.nsites // 4
    np.testing.assert_allclose(acc[0], o.energy(S2) - o.energy(S), atol=1e-10)
    np.testing.assert_allclose(acc[1:], S2.sum(0) - S.sum(0), atol=1e-12)
    S3, acc3 = o.local_sequential(S, props, 0.0, 5, 0, 3.0)
    assert acc3[0] <= 0 and o.energy(S3) <= o.energy(S) + 1e-12
    S4, _ = o.local_sequential(S, [(UNIFORM, 0, 1)], 1e9, 8, 0, 40.0)
    assert np.all(np.any(S4 != S, axis=1))


def chain_mean(o, S, m, col, nc, props, kT, key, nsweeps, observe, burn=200):
    vals = []
    for q in range(nsweeps):
        S, _ = o.local_sweeps(S, m, col, nc, props, kT, key, q, 1)
        if q >= burn:
            vals.append(observe(S))
    vals = np.asarray(vals)
    blocks = vals[: len(vals) // 20 * 20].reshape(20, -1).mean(1)
    return vals.mean(), blocks.std(ddof=1) / np.sqrt(len(blocks))


def test_paramagnet_matches_langevin_function():
    """Independent moments in a field: <S.h>/kappa = -L(kappa |h| / kT)."""
    s = sh.workloads.cubic_model(L=4, J=0.0, D=0.0, seed=1)
    sh.set_field_(s, [0.0, 0.3, 0.4])
    spec = s.spec()
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    kap = spec["kappa"][0]
    h = np.asarray(spec["g"][0]).reshape(3, 3).T @ np.asarray(spec["field"][0])
    kT = 0.6
    hn = h / np.linalg.norm(h)
    mean, err = chain_mean(o, s.dipoles_flat(), m, col.reshape(-1), nc, [(UNIFORM, 0, 1)], kT, 77, 4000,
                           lambda S: np.mean(S @ hn))
    exact = -kap * langevin_function(kap * np.linalg.norm(h) / kT)
    assert abs(mean - exact) < max(4 * err, 2e-3), (mean, exact, err)


def ising_chains(L, dims=None, precision=64):
    """Orthorhombic cell (1, 2, 3): the only bond of length 1 runs along x,
    so J = -1 on it makes ferromagnetic chains."""
    cryst = sh.Crystal(sh.lattice_vectors(1, 2, 3, 90, 90, 90), [[0, 0, 0]])
    s = sh.System(cryst, [(0, sh.Moment(s=1, g=2))], "dipole", dims=dims or (L, 2, 2), seed=0, precision=precision)
    sh.set_exchange_(s, -1.0, sh.Bond(0, 0, (1, 0, 0)))
    return s


def test_ising_chain_with_flips():
    """Dipoles along +-z with flip proposals form periodic Ising chains:
    <s_i s_i+1> = (t + t^(L-1)) / (1 + t^L), t = tanh(K), K = |J| kappa^2 / kT."""
    L = 16
    s = ising_chains(L)
    spec = s.spec()
    kap = spec["kappa"][0]
    rng = np.random.default_rng(3)
    S = np.zeros((s.nsites, 3))
    S[:, 2] = kap * rng.choice([-1.0, 1.0], s.nsites)
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    kT = 1.0
    mean, err = chain_mean(o, S, m, col.reshape(-1), nc, [(FLIP, 0, 1)], kT, 5, 6000,
                           lambda X: o.energy(X) / s.nsites)
    t = np.tanh(kap**2 / kT)
    exact = -kap**2 * (t + t ** (L - 1)) / (1 + t**L)
    assert abs(mean - exact) < max(4 * err, 5e-3), (mean, exact, err)


def test_colour_passes_and_random_site_orders_sample_alike():
    """Mean energy of a frustrated pyrochlore from the colour-class passes and
    from the reference's random-site trials agree within their errors."""
    s = sh.workloads.pyrochlore_model(L=2, seed=6)
    spec = s.spec()
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    props = [(DELTA, 0.4, 1.0)]
    kT = 0.8
    e_cb, err_cb = chain_mean(o, s.dipoles_flat(), m, col.reshape(-1), nc, props, kT, 11, 3000,
                              lambda X: o.energy(X) / s.nsites)
    S = s.dipoles_flat()
    vals = []
    for q in range(3000):
        S, _ = o.local_random_sites(S, props, kT, 1000 + q, 1.0)
        if q >= 200:
            vals.append(o.energy(S) / s.nsites)
    vals = np.asarray(vals)
    blocks = vals[: len(vals) // 20 * 20].reshape(20, -1).mean(1)
    e_rs, err_rs = vals.mean(), blocks.std(ddof=1) / np.sqrt(20)
    assert abs(e_cb - e_rs) < 4 * np.hypot(err_cb, err_rs) + 1e-3, (e_cb, err_cb, e_rs, err_rs)


# --------------------------------------------------------------------------
# :SUN mode


def sun_biquad_system(seed=3):
    """Triangular SU(3) with J1-J3, D Sz^2, a field and a biquadratic J1 bond."""
    s = sh.workloads.triangular_sun_model(L1=3, L2=4, seed=seed, precision=64)
    sh.set_field_(s, [0.1, -0.2, 0.3])
    sh.set_exchange_(s, -0.24, sh.Bond(0, 0, (1, 0, 0)), biquad=0.15)
    sh.randomize_spins_(s)
    return s


def kets_for(Z0, rng, kappa):
    """A random, a time-reversed and a perturbed coherent state of one site."""
    N = len(Z0)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    yield np.sqrt(kappa) * z / np.linalg.norm(z)
    par = np.array([1 - 2 * (j % 2) for j in range(N)])
    yield (par * np.conj(Z0))[::-1]
    d = Z0 + 0.3 * np.sqrt(kappa) * (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / np.sqrt(2)
    yield np.sqrt(kappa) * d / np.linalg.norm(d)


def test_sun_local_energy_change_is_energy_difference():
    s = sun_biquad_system()
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    E0 = o.energy(Z=Z)
    kap = s.spec()["kappa"][0]
    rng = np.random.default_rng(9)
    for site in rng.choice(s.nsites, 6, replace=False):
        for Zn in kets_for(Z[site], rng, kap):
            Z2 = Z.copy()
            Z2[site] = Zn
            dE = o.local_energy_change_sun(Z, site, Zn)
            assert abs(dE - (o.energy(Z=Z2) - E0)) <= 1e-11 * max(abs(E0), 1.0), site


def sun_dipolar_system(seed=5, dims=(2, 2, 3)):
    """SU(3) on a cubic lattice with J1, an easy axis and the Ewald term acting
    on expected dipoles (axis of length 3: offsets differ from their
    negatives)."""
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=2))], "SUN", dims=dims, seed=seed)
    sh.set_exchange_(s, 0.3, sh.Bond(0, 0, (1, 0, 0)))
    sh.set_onsite_coupling_(s, lambda S: -0.2 * S[2] @ S[2], 0)
    sh.enable_dipole_dipole_(s, 0.5)
    sh.randomize_spins_(s)
    return s


def test_sun_ewald_local_energy_change_and_sequential_trials():
    """:SUN with Ewald: local_energy_change equals differences of the pinned
    total energy; sequential trials accumulate the Ewald-inclusive energy
    change and the change of the summed expected dipoles; kT = 0 never raises
    the energy; every site of 12 is tried over 40 sweeps."""
    s = sun_dipolar_system()
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    E0 = o.energy(Z=Z)
    kap = s.spec()["kappa"][0]
    rng = np.random.default_rng(10)
    for site in rng.choice(s.nsites, 5, replace=False):
        for Zn in kets_for(Z[site], rng, kap):
            Z2 = Z.copy()
            Z2[site] = Zn
            assert abs(o.local_energy_change_sun(Z, site, Zn) - (o.energy(Z=Z2) - E0)) <= 1e-11 * max(abs(E0), 1)
    props = [(FLIP, 0, 0.3), (DELTA, 0.4, 0.5), (UNIFORM, 0, 0.2)]
    Z2, acc = o.local_sequential_sun(Z, props, 0.5, 77, 3, 4.0)
    assert np.sum(np.any(np.abs(Z2 - Z) > 1e-12, axis=1)) > s.nsites // 4
    np.testing.assert_allclose(acc[0], o.energy(Z=Z2) - E0, atol=1e-10)
    np.testing.assert_allclose(acc[1:], o.expected_spin(Z2).sum(0) - o.expected_spin(Z).sum(0), atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(Z2, axis=1) ** 2, kap, rtol=1e-13)
    Z3, acc3 = o.local_sequential_sun(Z, props, 0.0, 5, 0, 3.0)
    assert acc3[0] <= 0 and o.energy(Z=Z3) <= E0 + 1e-12
    Z4, _ = o.local_sequential_sun(Z, [(UNIFORM, 0, 1)], 1e9, 8, 0, 40.0)
    assert np.all(np.any(np.abs(Z4 - Z) > 1e-12, axis=1))


def test_sun_flip_reverses_expected_spin():
    """Time reversal (flip_ket) negates <S>: one pass at a huge kT accepts every
    flip it tries, so every site is either unchanged or reversed."""
    s = sun_biquad_system(seed=4)
    spec = s.spec()
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    Z = s.coherents_flat()
    S0 = o.expected_spin(Z)
    Z2, acc = o.local_sweeps_sun(Z, m, col.reshape(-1), nc, [(FLIP, 0, 1)], 1e12, 3, 0, 0.5)
    S2 = o.expected_spin(Z2)
    same = np.all(np.abs(S2 - S0) < 1e-12, axis=1)
    flipped = np.all(np.abs(S2 + S0) < 1e-12, axis=1)
    assert np.all(same | flipped) and 0.2 < flipped.mean() < 0.8
    np.testing.assert_allclose(acc[1:], S2.sum(0) - S0.sum(0), atol=1e-10)
    np.testing.assert_allclose(np.linalg.norm(Z2, axis=1) ** 2, spec["kappa"][0], rtol=1e-13)


def su5_mean_energy(kT, D):
    """test/test_samplers.jl:11-15: SU(5) with Lambda = D (Sz^2 - Sz^4 / 5)."""
    a = 4 * D / (5 * kT)
    return 4 * D * (np.exp(-a) * (-a * (a * (a * (a + 4) + 12) + 24) - 24) + 24) / (
        5 * a * (np.exp(-a) * (-a * (a * (a + 3) + 6) - 6) + 6))


def sun_anisotropy_chain(N, D, L=20):
    cryst = sh.Crystal(sh.lattice_vectors(1, 1, 1, 90, 90, 90), [[0, 0, 0]])
    s = sh.System(cryst, [(0, sh.Moment(s=(N - 1) / 2, g=2))], "SUN", dims=(L, 1, 1), seed=0)
    if N == 3:
        sh.set_onsite_coupling_(s, lambda S: D * S[2] @ S[2], 0)
    else:
        sh.set_onsite_coupling_(s, lambda S: D * (S[2] @ S[2] - 0.2 * S[2] @ S[2] @ S[2] @ S[2]), 0)
    sh.randomize_spins_(s)
    return s


@pytest.mark.parametrize("N,kT", [(3, 0.125), (3, 0.5), (5, 0.25)])
def test_sun_local_sampler_mean_energy(N, kT):
    """The closed-form mean energies of test/test_samplers.jl:6-15 (written for
    Langevin sampling) from Metropolis trials with uniform proposals."""
    from test_oracle_dynamics import su3_mean_energy
    D = 1.0
    s = sun_anisotropy_chain(N, D)
    spec = s.spec()
    o = O.Oracle(spec)
    m, nc, col = site_coloring(spec, s.dims)
    Z = s.coherents_flat()
    vals = []
    for q in range(6000):
        Z, _ = o.local_sweeps_sun(Z, m, col.reshape(-1), nc, [(UNIFORM, 0, 1)], kT, 21, q, 1.0)
        if q >= 200:
            vals.append(o.energy(Z=Z) / s.nsites)
    vals = np.asarray(vals)
    blocks = vals[: len(vals) // 20 * 20].reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(20)
    exact = su3_mean_energy(kT, D) if N == 3 else su5_mean_energy(kT, D)
    assert abs(vals.mean() - exact) < max(4 * err, 0.01 * abs(exact)), (vals.mean(), exact, err)
