"""Inhomogeneous systems (to_inhomogeneous, set_exchange_at!,
set_onsite_coupling_at!, set_vacancy_at!, set_field_at!;
src/System/Interactions.jl:80-100,227-260, PairExchange.jl:390-500,
OnsiteCoupling.jl:206-225) in the host mirror and the oracle (CPU only).

Checks: an inhomogeneous copy of a homogeneous system has the same energy,
gradient and trajectories (the supercell-as-cell mapping keeps every site,
bond image and, for Ewald, every position); energies of edited systems equal
direct sums written here with numpy (bonds, vacancies, per-site fields and
anisotropies); the two-site chain of test_samplers.jl:174-186 written as the
reference writes it (one bond, offset (-1,0,0)) has E = S1.S2 and samples
the exact energy distribution.
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from test_oracle_thermal import CHAIN, chain_distribution_error


def homog_models():
    s = sh.workloads.pyrochlore_model(L=2, seed=3)
    sh.set_field_(s, [0.1, -0.2, 0.05])
    yield "pyrochlore2", s
    s = sh.workloads.cubic_model(dims=(3, 2, 4), seed=4)
    Om = sh.stevens_matrices(1.0)
    sh.set_onsite_coupling_(s, 0.2 * Om[2][0] + 0.1 * Om[2][1], 0)
    yield "cubic_ragged", s
    s = sh.workloads.pyrochlore_model(L=1, J1=0.05, J2=0, A=0, s=3.5, seed=5, dipolar=True)
    yield "pyro1_ewald", s
    from test_oracle_general import pair_model
    s, _ = pair_model(N=3, seed=6, L=(3, 2, 2))
    yield "su3_general", s


@pytest.mark.parametrize("name,s", list(homog_models()), ids=lambda v: v if isinstance(v, str) else "")
def test_inhomogeneous_copy_is_identical(name, s):
    t = sh.to_inhomogeneous(s)
    assert not sh.is_homogeneous(t) and t.params == {}
    o, ot = O.Oracle(s.spec()), O.Oracle(t.spec())
    rng = np.random.default_rng(1)
    if s.N == 0:
        S = s.dipoles_flat()
        np.testing.assert_allclose(ot.energy(S), o.energy(S), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(ot.grad(S), o.grad(S), rtol=0, atol=1e-12)
        nz = rng.standard_normal((s.nsites, 3))
        np.testing.assert_allclose(ot.step_langevin(S=S, noise=nz, dt=0.02, damping=0.1, kT=0.3),
                                   o.step_langevin(S=S, noise=nz, dt=0.02, damping=0.1, kT=0.3), atol=1e-12)
        a, ia = ot.step_midpoint(S=S, dt=0.02)
        b, ib = o.step_midpoint(S=S, dt=0.02)
        np.testing.assert_allclose(a, b, atol=1e-12)
        assert abs(ia - ib) <= 1
    else:
        Z = s.coherents_flat()
        np.testing.assert_allclose(ot.energy(Z=Z), o.energy(Z=Z), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(ot.grad_coherents(Z), o.grad_coherents(Z), atol=1e-12)
        nz = (rng.standard_normal((s.nsites, s.N)) + 1j * rng.standard_normal((s.nsites, s.N))) / np.sqrt(2)
        np.testing.assert_allclose(ot.step_langevin(Z=Z, noise=nz, dt=0.02, damping=0.1, kT=0.3),
                                   o.step_langevin(Z=Z, noise=nz, dt=0.02, damping=0.1, kT=0.3), atol=1e-12)


def edited_cubic(seed=7, L=(4, 3, 2), s_=1.5):
    """Cubic Heisenberg J = 1 with D Sz^2 and a field, then per-site edits:
    two vacancies, one stronger bond, one site's own anisotropy and field."""
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=s_, g=2))], "dipole", dims=L, seed=seed)
    sh.set_exchange_(s, 1.0, sh.Bond(0, 0, (1, 0, 0)))
    sh.set_onsite_coupling_(s, lambda S: -0.1 * S[2] @ S[2], 0)
    sh.set_field_(s, [0.0, 0.0, 0.2])
    t = sh.to_inhomogeneous(s)
    sh.set_vacancy_at_(t, (1, 1, 0, 0))
    sh.set_vacancy_at_(t, (3, 0, 1, 0))
    # L1 = 4: (1,0,0) is not 4x shorter than (-3,0,0), so the offset is explicit
    sh.set_exchange_at_(t, 2.5, (0, 0, 0, 0), (1, 0, 0, 0), offset=(1, 0, 0))
    sh.set_onsite_coupling_at_(t, lambda S: 0.3 * S[0] @ S[0], (2, 2, 1, 0))
    sh.set_field_at_(t, [0.3, 0.0, -0.1], (0, 2, 1, 0))
    sh.randomize_spins_(t)
    return t


def direct_energy(t):
    """Bonds of the cubic lattice (each once), onsite and Zeeman terms with the
    edits above, summed site by site."""
    L = t.dims
    S = t.dipoles_flat()
    kap = t.site_kappas()
    vac = kap == 0
    idx = lambda c1, c2, c3: c1 + L[0] * (c2 + L[1] * c3)  # noqa: E731
    E = 0.0
    # classical Stevens renormalisation of an O[2,q] term: rcs_2 = 1 - 1/(2s)
    rcs2 = 1 - 1 / (2 * t.spins[0])
    for c3 in range(L[2]):
        for c2 in range(L[1]):
            for c1 in range(L[0]):
                k = idx(c1, c2, c3)
                if vac[k]:
                    continue
                for n in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    j = idx((c1 + n[0]) % L[0], (c2 + n[1]) % L[1], (c3 + n[2]) % L[2])
                    if vac[j]:
                        continue
                    J = 2.5 if (k, j) in ((idx(0, 0, 0), idx(1, 0, 0)),) else 1.0
                    E += J * S[k] @ S[j]
                x, y, z = S[k]
                X = S[k] @ S[k]
                s_ = t.spins[0]
                if k == idx(2, 2, 1):
                    # 0.3 Sx^2 = 0.3 [s(s+1)/3 - O20/6 + O22/2]; classical O's, rank 2 scaled by rcs2
                    E += 0.3 * (s_ * (s_ + 1) / 3 + rcs2 * (x * x - X / 3))
                else:
                    # -0.1 Sz^2 = -0.1 [s(s+1)/3 + O20/3]
                    E += -0.1 * (s_ * (s_ + 1) / 3 + rcs2 * (z * z - X / 3))
                B = np.array([0.3, 0.0, -0.1]) if k == idx(0, 2, 1) else np.array([0.0, 0.0, 0.2])
                E += B @ (2 * S[k])
    return E


def test_edited_energy_matches_direct_sum():
    t = edited_cubic()
    S = t.dipoles_flat()
    assert np.all(S[t.site_kappas() == 0] == 0)
    E = O.Oracle(t.spec()).energy(S)
    assert abs(E - direct_energy(t)) < 1e-11 * max(1, abs(E))


def test_edited_gradient_finite_differences():
    t = edited_cubic(seed=8)
    o = O.Oracle(t.spec())
    S = t.dipoles_flat()
    g = o.grad(S)
    # a vacancy keeps only its Zeeman term g'B (set_energy_grad_dipoles! adds
    # it at every site); no bond reaches it and its own state is zero
    np.testing.assert_allclose(g[t.site_kappas() == 0], [[0, 0, 0.4]] * 2, atol=1e-15)
    h = 1e-6
    for k in range(t.nsites):
        if t.site_kappas()[k] == 0:
            continue
        u = S[k] / np.linalg.norm(S[k])
        fd = np.zeros(3)
        for a in range(3):
            Sp, Sm = S.copy(), S.copy()
            Sp[k, a] += h
            Sm[k, a] -= h
            fd[a] = (o.energy(Sp) - o.energy(Sm)) / (2 * h)
        np.testing.assert_allclose(g[k] - (g[k] @ u) * u, fd - (fd @ u) * u, atol=1e-6)


def test_vacancy_stays_empty_under_dynamics():
    t = edited_cubic(seed=9)
    o = O.Oracle(t.spec())
    S = t.dipoles_flat()
    rng = np.random.default_rng(3)
    for _ in range(3):
        S = o.step_langevin(S=S, noise=rng.standard_normal((t.nsites, 3)), dt=0.02, damping=0.1, kT=0.4)
    S, _ = o.step_midpoint(S=S, dt=0.02)
    vac = t.site_kappas() == 0
    assert np.all(S[vac] == 0)
    np.testing.assert_allclose(np.linalg.norm(S[~vac], axis=1), 1.5, atol=1e-12)


def test_sites_to_internal_bond_rules():
    s = sh.to_inhomogeneous(sh.workloads.cubic_model(dims=(2, 1, 1), seed=1))
    # two sites two cells apart in a 2-cell chain: +1 and -1 are equally short
    with pytest.raises(ValueError, match="Ambiguous offset"):
        sh.sites_to_internal_bond(s, (0, 0, 0, 0), (1, 0, 0, 0))
    assert sh.sites_to_internal_bond(s, (0, 0, 0, 0), (1, 0, 0, 0), offset=(-1, 0, 0)).n == (-1, 0, 0)
    with pytest.raises(ValueError, match="not compatible"):
        sh.sites_to_internal_bond(s, (0, 0, 0, 0), (1, 0, 0, 0), offset=(0, 0, 0))


def reference_chain(mode, precision=64, seed=0):
    """test_samplers.jl:174-186 as written there: dims (2,1,1), to_inhomogeneous,
    kappa 1 (dipole, s = 1) or 2 (SU(2), s = 1/2), one exchange J = 1 between
    the two sites with offset (-1, 0, 0)."""
    s_ = 1.0 if mode == "dipole" else 0.5
    sys = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=s_, g=2))], mode, dims=(2, 1, 1), seed=seed,
                    precision=precision)
    sys = sh.to_inhomogeneous(sys)
    sys.inhomog.kappa[:] = 1.0 if mode == "dipole" else 2.0
    sh.set_exchange_at_(sys, 1.0, (0, 0, 0, 0), (1, 0, 0, 0), offset=(-1, 0, 0))
    sh.randomize_spins_(sys)
    return sys


@pytest.mark.parametrize("mode", ["dipole", "SUN"])
def test_reference_chain_energy_and_distribution(mode):
    sys = reference_chain(mode)
    assert sys.spec()["nbonds"] == 2  # one bond, both directions
    o = O.Oracle(sys.spec())
    S = sys.dipoles_flat()
    st = dict(S=S) if mode == "dipole" else dict(Z=sys.coherents_flat())
    E = o.energy(**st)
    assert abs(E - S[0] @ S[1]) < 1e-13
    key = O.replica_key(sys.noise_seed, 0)
    kw = dict(dt=CHAIN["dt"], damping=CHAIN["damping"], kT=CHAIN["kT"], key=key)
    state, _ = o.langevin_run(**st, **kw, step0=0, nsteps=CHAIN["n_equilib"], every=0)
    st = dict(S=state) if mode == "dipole" else dict(Z=state)
    _, Es = o.langevin_run(**st, **kw, step0=CHAIN["n_equilib"], nsteps=CHAIN["n_samples"] * CHAIN["n_decorr"],
                           every=CHAIN["n_decorr"])
    assert chain_distribution_error(Es) <= CHAIN["atol"]
