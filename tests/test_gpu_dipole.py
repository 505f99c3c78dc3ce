"""GPU parity of the :dipole path against the FP64 oracle.

Bars: FP64 <= 1e-11 relative (per step and over short trajectories); FP32
<= 2e-5 relative on fields and states, energy within 1e-6 of the energy
scale.  All calls go through libsunnyhip's C ABI.
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce, energy_scale, rel_err

pytestmark = pytest.mark.gpu

TOL = {64: 1e-11, 32: 2e-5}


def models(precision):
    yield "cubic4", sh.workloads.cubic_model(L=4, seed=3, precision=precision)
    yield "cubic_ragged", sh.workloads.cubic_model(dims=(3, 5, 2), seed=4, precision=precision)
    yield "pyro2", sh.workloads.pyrochlore_model(L=2, seed=5, precision=precision)
    s = sh.workloads.pyrochlore_model(dims=(1, 2, 3), seed=6, precision=precision)
    yield "pyro_ragged", s
    # bonds longer than the box: ghost images wrap more than once
    s = sh.workloads.cubic_model(dims=(1, 2, 3), seed=8, precision=precision)
    sh.set_exchange_(s, 0.3, sh.Bond(0, 0, (2, 1, 0)), propagate=[sh.Bond(0, 0, (2, 1, 0))])
    sh.set_exchange_(s, -0.2, sh.Bond(0, 0, (0, 3, 4)), propagate=[sh.Bond(0, 0, (0, 3, 4))])
    sh.randomize_spins_(s)
    yield "cubic_long_bonds", s
    # the large-s mode: polynomial anisotropy up to order 6, exact (Si.Sj)^2
    from test_oracle_uncorrected import _poly
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=2, g=2))], "dipole_uncorrected", dims=(2, 3, 2), seed=9,
                  precision=precision)
    sh.set_exchange_(s, 0.7, sh.Bond(0, 0, (1, 0, 0)), biquad=0.3)
    sh.set_onsite_coupling_(s, _poly, 0)
    sh.randomize_spins_(s)
    yield "cubic_uncorrected", s


def aniso_model(precision):
    """Diamond, s = 3: bilinear exchange matrix, field, k = 2, 4, 6 Stevens
    terms and a quadrupole coupling (the shared.jl interaction set)."""
    cryst = sh.diamond_crystal()
    s = sh.System(cryst, [(i, sh.Moment(s=3, g=2)) for i in range(8)], "dipole", dims=(2, 2, 2), seed=7,
                  precision=precision)
    sh.set_field_(s, [0.0, 0.2, 0.3])
    J = np.array([[0.5, 0.2, -0.4], [0.2, 0.5, -0.4], [0.4, 0.4, 1.5]])
    sh.set_exchange_(s, J, sh.Bond(0, 1, (0, 0, 0)))
    O_ = sh.stevens_matrices(3)
    for i in range(8):
        sh.set_onsite_coupling_(s, 0.01 * ((1 / 20) * O_[4][4] + (1 / 4) * O_[4][0]) + 2e-4 * O_[6][6]
                                + 0.05 * O_[2][2], i)
    B = np.array([[1.2, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, 0, 1.1, 0, -1.4], [0, 0, 0, 1, 0], [0, 0, 1.4, 0, 1.3]])
    sh.set_pair_biquad_(s, 0.1 * B, sh.Bond(0, 0, (0, 0, 1)))
    sh.set_exchange_(s, 0.3, sh.Bond(0, 2, (0, 0, 0)), biquad=0.2)
    sh.randomize_spins_(s)
    return s


@pytest.mark.parametrize("precision", [64, 32])
def test_grad_and_energy(precision, layout):
    cases = list(models(precision)) + [("diamond_aniso", aniso_model(precision))]
    for name, s in cases:
        s.layout = layout
        announce(f"grad/energy {name} fp{precision} layout {layout}")
        S = s.dipoles_flat()
        o = O.Oracle(s.spec())
        g_dev = sh.energy_grad_dipoles(s)
        g_ref = o.grad(S)
        assert rel_err(g_dev, g_ref) <= TOL[precision], name
        E_dev, E_ref = sh.energy(s), o.energy(S)
        assert abs(E_dev - E_ref) <= TOL[precision] * 0.1 * energy_scale(s) + 1e-13, (name, E_dev, E_ref)


@pytest.mark.parametrize("precision", [64, 32])
def test_langevin_host_noise(precision, layout):
    for name, s in list(models(precision)) + [("diamond_aniso", aniso_model(precision))]:
        s.layout = layout
        announce(f"langevin host-noise {name} fp{precision} layout {layout}")
        o = O.Oracle(s.spec())
        rng = np.random.default_rng(11)
        integ = sh.Langevin(0.02, damping=0.1, kT=0.5)
        S = s.dipoles_flat()
        for _ in range(3):
            noise = rng.standard_normal((s.nsites, 3))
            S = o.step_langevin(S=S, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
            sh.step_(s, integ, noise=noise)
        S_dev = s.dipoles_flat()
        kap = np.repeat(s.kappas, s.ncells)
        assert rel_err(S_dev, S, scale=kap.max()) <= 3 * TOL[precision], name
        assert np.max(np.abs(np.linalg.norm(S_dev, axis=1) - kap)) <= 1e-12 * kap.max() + (
            1e-6 * kap.max() if precision == 32 else 0)


def test_langevin_device_philox_trajectory():
    """Device Philox noise reproduced on the CPU: 20 FP64 steps agree."""
    s = sh.workloads.pyrochlore_model(L=2, seed=12, precision=64)
    announce("langevin philox pyro2 fp64 x20")
    o = O.Oracle(s.spec())
    integ = sh.Langevin(0.025, damping=0.2, kT=0.3)
    S = s.dipoles_flat()
    key = O.replica_key(s.noise_seed, 0)
    for k in range(20):
        noise = O.philox_normals(key, s.step_counter + k, s.nsites, 3)
        S = o.step_langevin(S=S, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
    sh.step_(s, integ, nsteps=20)
    assert rel_err(s.dipoles_flat(), S, scale=1.5) <= 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_midpoint(precision, layout):
    for name, s in list(models(precision)) + [("diamond_aniso", aniso_model(precision))]:
        s.layout = layout
        announce(f"midpoint {name} fp{precision} layout {layout}")
        o = O.Oracle(s.spec())
        damping = 0.1 if name.startswith("cubic") else 0.0
        tol = 1e-12 if precision == 64 else 1e-6
        integ = sh.ImplicitMidpoint(0.02, damping=damping, tol=tol)
        S = s.dipoles_flat()
        its_ref = []
        for _ in range(3):
            S, it = o.step_midpoint(S=S, dt=integ.dt, damping=damping, tol=tol)
            its_ref.append(it)
        its_dev = [int(sh.step_(s, integ)[0]) for _ in range(3)]
        kap = np.repeat(s.kappas, s.ncells)
        err = rel_err(s.dipoles_flat(), S, scale=kap.max())
        if precision == 64:
            assert err <= 1e-11, name
            assert all(abs(a - b) <= 1 for a, b in zip(its_dev, its_ref)), (its_dev, its_ref)
        else:
            assert err <= 5e-6, (name, err)


def test_midpoint_nonconvergence_message():
    s = sh.workloads.cubic_model(L=4, seed=2, precision=64)
    announce("midpoint nonconvergence cubic4")
    S0 = s.dipoles_flat()
    with pytest.raises(sh.ConvergenceError, match="Spherical midpoint method failed to converge"):
        sh.step_(s, sh.ImplicitMidpoint(0.05), max_iters=1)
    np.testing.assert_array_equal(s.dipoles_flat(), S0)


@pytest.mark.parametrize("precision", [64, 32])
def test_replica_batch_uneven(precision, layout):
    """Replicas of different dims in one launch, each matching its own oracle."""
    dims = [(4, 4, 4), (3, 2, 5), (6, 1, 2)]
    systems = [sh.workloads.cubic_model(dims=d, seed=20 + r, precision=precision) for r, d in enumerate(dims)]
    announce(f"replica batch cubic {dims} fp{precision} layout {layout}")
    batch = sh.ReplicaBatch(systems, precision=precision, layout=layout)
    kTs = np.array([0.1, 0.5, 1.0])
    integ = sh.Langevin(0.03, damping=0.1, kT=0.0)
    refs = [s.dipoles_flat() for s in systems]
    sh.step_(batch, integ, nsteps=5, kTs=kTs)
    E_dev = sh.energy(batch)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        o = O.Oracle(s.spec())
        S = refs[r]
        key = O.replica_key(batch.noise_seed, r)
        for k in range(5):
            noise = O.philox_normals(key, k, s.nsites, 3)
            S = o.step_langevin(S=S, noise=noise, dt=integ.dt, damping=integ.damping, kT=kTs[r])
        assert rel_err(s.dipoles_flat(), S, scale=1.0) <= (1e-10 if precision == 64 else 1e-4)
        assert abs(E_dev[r] - o.energy(S)) <= (1e-10 if precision == 64 else 1e-4) * energy_scale(s)


def test_layout_selection(monkeypatch):
    """The automatic plan picks the LDS-tiled sweep for :dipole models and
    the ghost-padded gather when asked, in :SUN mode, and for one small
    replica stepped by the cooperative midpoint."""
    c = sh.workloads.pyrochlore_model(L=4, seed=1, precision=32)
    assert c._device().query("cooperative") == 1 and c._device().query("layout") == 1
    monkeypatch.setenv("SNH_COOP", "0")
    s = sh.workloads.pyrochlore_model(L=4, seed=1, precision=32)
    h = s._device()
    assert h.query("layout") == 0 and h.query("tile_T1") >= 1 and h.query("tile_work") >= 1
    assert h.query("tile_lds") > 0
    g = sh.workloads.pyrochlore_model(L=4, seed=1, precision=32)
    g.layout = 1
    assert g._device().query("layout") == 1
    np.testing.assert_allclose(sh.energy(s), sh.energy(g), rtol=1e-6)
    u = sh.workloads.triangular_sun_model(L1=4, L2=4)
    assert u._device().query("layout") == 1


def test_tiled_bench_shape_step():
    """One FP32 Heun step of the 32^3 pyrochlore replica with the automatic
    (benchmark) tile plan against the ghost-padded gather on the same state
    and noise, plus the energy after the step against the FP64 oracle."""
    a = sh.workloads.pyrochlore_model(L=32, seed=3, precision=32)
    b = sh.clone_system(a)
    b.layout = 1
    announce(f"bench-shape tile plan T1={a._device().query('tile_T1')} T2={a._device().query('tile_T2')}")
    integ = sh.Langevin(0.025, damping=0.2, kT=0.3)
    noise = np.random.default_rng(4).standard_normal((a.nsites, 3))
    sh.step_(a, integ, noise=noise)
    sh.step_(b, integ, noise=noise)
    assert rel_err(a.dipoles_flat(), b.dipoles_flat(), scale=1.5) <= 2e-5
    o = O.Oracle(a.spec())
    E_ref = o.energy(a.dipoles_flat())
    assert abs(sh.energy(a) - E_ref) <= 1e-6 * energy_scale(a)


def test_autotune_keeps_results():
    """The empirical autotuner times the cost model's best tile plans and
    keeps the fastest; every plan computes the same per-site arithmetic, so
    the stepped states stay bitwise equal to those of the default plan."""
    dims = [(8, 8, 8), (6, 4, 10), (5, 5, 5)]

    def batch():
        systems = [sh.workloads.cubic_model(dims=d, seed=40 + r, precision=32) for r, d in enumerate(dims)]
        return sh.ReplicaBatch(systems, precision=32)

    announce(f"autotune cubic {dims} fp32")
    a, b = batch(), batch()
    hb = b._device()
    assert hb.query("tile_candidates") >= 1 and hb.query("tile_tuned") == -1
    ms = hb.autotune(candidates=6, launches=2)
    assert ms > 0
    assert 0 <= hb.query("tile_tuned") < min(6, hb.query("tile_candidates"))
    assert 0 <= hb.query("tile_tuned_correct") < min(6, hb.query("tile_candidates"))
    integ = sh.Langevin(0.03, damping=0.1, kT=0.4)
    kTs = np.array([0.2, 0.5, 0.8])
    sh.step_(a, integ, nsteps=5, kTs=kTs)
    sh.step_(b, integ, nsteps=5, kTs=kTs)
    da = np.empty((a._device().total, 3))
    db = np.empty_like(da)
    a._device().call("snh_get_dipoles", sh._lib.dptr(da))
    hb.call("snh_get_dipoles", sh._lib.dptr(db))
    assert np.array_equal(da, db)
    np.testing.assert_allclose(sh.energy(a), sh.energy(b), rtol=1e-12)
    # the midpoint after autotuning (which allocates scratch buffers of its
    # own) finds every buffer it needs
    mp = sh.ImplicitMidpoint(0.03, tol=1e-5)
    sh.step_(a, mp, nsteps=2)
    sh.step_(b, mp, nsteps=2)
    a._device().call("snh_get_dipoles", sh._lib.dptr(da))
    hb.call("snh_get_dipoles", sh._lib.dptr(db))
    np.testing.assert_allclose(da, db, rtol=0, atol=1e-6)


def test_cooperative_midpoint_matches_launches(monkeypatch):
    """One :dipole replica of 512 < sites <= 8192 without Ewald steps the
    midpoint in one cooperative launch, a grid barrier per fixed-point
    iteration: iteration counts and states equal, bitwise, the launched
    sweeps on the same ghost layout, and the oracle's midpoint steps; a step
    that does not converge raises the reference's error and keeps the state
    of the previous step."""
    def make(coop, seed=71):
        monkeypatch.setenv("SNH_COOP", "1" if coop else "0")
        s = sh.workloads.cubic_model(L=16, J=1.0, D=-0.1, seed=seed, precision=64)
        s.layout = 1
        assert s._device().query("cooperative") == int(coop) and s._device().query("layout") == 1
        return s

    announce("cooperative midpoint cubic 16^3 fp64")
    a, b = make(True), make(False)
    S0 = a.dipoles_flat()
    for dt, n in ((0.05, 6), (0.02, 3)):
        mp = sh.ImplicitMidpoint(dt, damping=0.1)
        ia = sh.step_(a, mp, nsteps=n)
        ib = sh.step_(b, mp, nsteps=n)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(a.dipoles_flat(), b.dipoles_flat())
    o = O.Oracle(a.spec())
    S = S0
    for dt, n in ((0.05, 6), (0.02, 3)):
        for _ in range(n):
            S, _ = o.step_midpoint(S=S, dt=dt, damping=0.1)
    assert rel_err(a.dipoles_flat(), S, scale=1.0) <= 1e-11
    # a pyrochlore replica of 8192 sites (32 workgroups, 18 bonds per site)
    monkeypatch.setenv("SNH_COOP", "1")
    p = sh.workloads.pyrochlore_model(L=8, seed=72, precision=64)
    assert p._device().query("cooperative") == 1
    P0 = p.dipoles_flat()
    sh.step_(p, sh.ImplicitMidpoint(0.03, damping=0.05), nsteps=3)
    S = P0
    op = O.Oracle(p.spec())
    for _ in range(3):
        S, _ = op.step_midpoint(S=S, dt=0.03, damping=0.05)
    assert rel_err(p.dipoles_flat(), S, scale=1.5) <= 1e-11
    # non-convergence: the error of Integrators.jl and the state of the last good step
    c = make(True, seed=73)
    before = c.dipoles_flat()
    with pytest.raises(RuntimeError, match="failed to converge"):
        sh.step_(c, sh.ImplicitMidpoint(0.05, damping=0.1), nsteps=2, max_iters=2)
    np.testing.assert_array_equal(c.dipoles_flat(), before)


def test_persistent_small_replicas_match_launches(monkeypatch):
    """Small :dipole replicas on the ghost layout step in one persistent
    workgroup per replica; the results equal the per-stage launches (same
    per-site arithmetic) and the oracle."""
    dims = [(6, 6, 6), (4, 5, 3), (8, 2, 4)]

    def batch(persist):
        monkeypatch.setenv("SNH_PERSIST", "1" if persist else "0")
        systems = [sh.workloads.cubic_model(dims=d, seed=50 + r, precision=64) for r, d in enumerate(dims)]
        b = sh.ReplicaBatch(systems, precision=64, layout=1)
        assert b._device().query("persistent") == int(persist)
        return b, systems

    announce(f"persistent vs launches cubic {dims} fp64")
    (a, sa), (b, sb) = batch(True), batch(False)
    integ = sh.Langevin(0.03, damping=0.1, kT=0.4)
    kTs = np.array([0.2, 0.5, 0.8])
    refs = [s.dipoles_flat() for s in sa]
    sh.step_(a, integ, nsteps=7, kTs=kTs)
    sh.step_(b, integ, nsteps=7, kTs=kTs)
    a.sync_to_systems()
    b.sync_to_systems()
    for r in range(3):
        assert rel_err(sa[r].dipoles_flat(), sb[r].dipoles_flat(), scale=1.0) <= 1e-13
        o = O.Oracle(sa[r].spec())
        S = refs[r]
        key = O.replica_key(a.noise_seed, r)
        for k in range(7):
            S = o.step_langevin(S=S, noise=O.philox_normals(key, k, sa[r].nsites, 3), dt=integ.dt,
                                damping=integ.damping, kT=kTs[r])
        assert rel_err(sa[r].dipoles_flat(), S, scale=1.0) <= 1e-11
    # midpoint on one replica: every step iterated to convergence on the device
    monkeypatch.setenv("SNH_PERSIST", "1")
    p = sh.workloads.cubic_model(L=6, seed=61, precision=64)
    assert p._device().query("persistent") == 1
    monkeypatch.setenv("SNH_PERSIST", "0")
    q = sh.workloads.cubic_model(L=6, seed=61, precision=64)
    assert q._device().query("persistent") == 0
    mp = sh.ImplicitMidpoint(0.03, damping=0.05)
    its_p = sh.step_(p, mp, nsteps=5)
    its_q = sh.step_(q, mp, nsteps=5)
    assert abs(int(its_p[0]) - int(its_q[0])) <= 1
    assert rel_err(p.dipoles_flat(), q.dipoles_flat(), scale=1.0) <= 1e-12
    o = O.Oracle(p.spec())
    S = sh.workloads.cubic_model(L=6, seed=61, precision=64).dipoles_flat()
    for _ in range(5):
        S, _ = o.step_midpoint(S=S, dt=0.03, damping=0.05)
    assert rel_err(p.dipoles_flat(), S, scale=1.0) <= 1e-11
