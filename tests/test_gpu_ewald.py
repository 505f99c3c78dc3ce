"""GPU Ewald dipole-dipole path: known answers and parity with the oracle."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce, rel_err

pytestmark = pytest.mark.gpu


def single_dipole(dims, precision=64):
    cryst = sh.cubic_crystal(1.0)
    s = sh.System(cryst, [(0, sh.Moment(s=1, g=1))], "dipole", dims=dims, precision=precision)
    sh.enable_dipole_dipole_(s, 1.0, demag=0.0)
    return s


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (8, 8, 8), (16, 16, 16), (32, 8, 8)])
def test_single_dipole_cubic_box(dims):
    """E per site = -1/6 for a lattice of parallel dipoles, demag = 0
    (test/test_ewald.jl:16-31): the register line transforms of every length
    the larger grids use (2 ... 32) and the device tensor at 16^3."""
    announce(f"ewald single dipole {dims}")
    s = single_dipole(dims)
    assert abs(sh.energy_per_site(s) - (-1 / 6)) < 1e-12
    s32 = single_dipole(dims, precision=32)
    assert abs(sh.energy_per_site(s32) - (-1 / 6)) < 1e-6


@pytest.mark.parametrize("dims", [(24, 24, 24), (40, 6, 6), (64, 4, 2)])
def test_single_dipole_large_grids(dims):
    """Grids whose transform does not fit one workgroup's LDS, or with an
    axis longer than 32, take the hipFFT path: E per site = -1/6 for parallel
    dipoles, demag = 0 (test/test_ewald.jl:16-31)."""
    announce(f"ewald single dipole {dims} (library FFT)")
    s = single_dipole(dims)
    assert abs(sh.energy_per_site(s) - (-1 / 6)) < 1e-12
    assert s._device().query("ewald_fft_lib") == 1
    s32 = single_dipole(dims, precision=32)
    assert abs(sh.energy_per_site(s32) - (-1 / 6)) < 1e-6


def random_triclinic(precision):
    """Three atoms in a skewed cell with random g tensors and demag
    (test/test_ewald.jl:33-57)."""
    rng = np.random.default_rng(0)
    cryst = sh.Crystal(sh.lattice_vectors(1.1, 0.9, 0.8, 92, 85, 95),
                       [[0, 0.01, 0.01], [0.1, 0.01, 0.01], [0.6, 0.4, 0.5]])
    mom = [(0, sh.Moment(s=1, g=rng.random((3, 3)))), (1, sh.Moment(s=1.5, g=rng.random((3, 3)))),
           (2, sh.Moment(s=2, g=rng.random((3, 3))))]
    s = sh.System(cryst, mom, "dipole", dims=(2, 3, 1), seed=1, precision=precision)
    R = rng.standard_normal((3, 3))
    sh.enable_dipole_dipole_(s, 1.0, demag=R.T @ R)
    sh.randomize_spins_(s)
    return s


@pytest.mark.parametrize("precision", [64, 32])
def test_ewald_grad_energy_parity(precision):
    tol = 1e-11 if precision == 64 else 3e-5
    for name, s in [("triclinic", random_triclinic(precision)),
                    ("pyro2_dipolar", sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=2,
                                                                    dipolar=True, precision=precision))]:
        announce(f"ewald parity {name} fp{precision}")
        S = s.dipoles_flat()
        o = O.Oracle(s.spec())
        g_ref = o.grad(S)
        assert rel_err(sh.energy_grad_dipoles(s), g_ref) <= tol, name
        E_ref = o.energy(S)
        assert abs(sh.energy(s) - E_ref) <= tol * max(1.0, np.abs(g_ref).max() * np.abs(S).max() * s.nsites), name


def test_ewald_langevin_and_midpoint_fp64():
    s = sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=8, dipolar=True, precision=64)
    announce("ewald langevin+midpoint pyro2 fp64")
    o = O.Oracle(s.spec())
    rng = np.random.default_rng(1)
    S = s.dipoles_flat()
    integ = sh.Langevin(0.01, damping=0.1, kT=0.2)
    for _ in range(2):
        noise = rng.standard_normal((s.nsites, 3))
        S = o.step_langevin(S=S, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
        sh.step_(s, integ, noise=noise)
    assert rel_err(s.dipoles_flat(), S, scale=3.5) <= 1e-11
    mp = sh.ImplicitMidpoint(0.01)
    S, it = o.step_midpoint(S=S, dt=0.01)
    it_dev = sh.step_(s, mp)[0]
    assert rel_err(s.dipoles_flat(), S, scale=3.5) <= 1e-11
    assert abs(int(it_dev) - it) <= 1


def test_ewald_sun():
    """Ewald acting on expected dipoles of SU(3) coherent states."""
    cryst = sh.cubic_crystal(1.0)
    s = sh.System(cryst, [(0, sh.Moment(s=1, g=2))], "SUN", dims=(2, 2, 3), seed=3, precision=64)
    sh.set_exchange_(s, 0.3, sh.Bond(0, 0, (1, 0, 0)))
    sh.enable_dipole_dipole_(s, 0.5)
    sh.randomize_spins_(s)
    announce("ewald sun cubic 2x2x3 fp64")
    Z = s.coherents_flat()
    o = O.Oracle(s.spec())
    assert rel_err(sh.energy_grad_coherents(s), o.grad_coherents(Z)) <= 1e-11
    assert abs(sh.energy(s) - o.energy(Z=Z)) <= 1e-11
    # Langevin: both Heun stages take the Ewald potential of expected dipoles
    rng = np.random.default_rng(8)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
    for _ in range(2):
        noise = (rng.standard_normal((s.nsites, s.N)) + 1j * rng.standard_normal((s.nsites, s.N))) / np.sqrt(2)
        Z = o.step_langevin(Z=Z, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
        sh.step_(s, integ, noise=noise)
    assert rel_err(s.coherents_flat(), Z, scale=1.0) <= 1e-11


@pytest.mark.parametrize("R,model", [(9, "pyro2"), (20, "cubic444"), (40, "cubic234")])
def test_ewald_batch_fp32(R, model):
    """Replica batches of 8 or more share the transformed tensor and run the
    per-k contraction on the matrix cores (1 or 2 column tiles of 16
    replicas, a partial last tile, two replica groups at R = 40); every
    replica's energy and Langevin step match its own FP64 oracle."""
    def build(r):
        if model == "pyro2":
            return sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=60 + r, dipolar=True,
                                                 precision=32)
        s = single_dipole((4, 4, 4) if model == "cubic444" else (2, 3, 4), precision=32)
        sh.set_exchange_(s, 0.2, sh.Bond(0, 0, (1, 0, 0)))
        s.rng = np.random.default_rng(70 + r)
        sh.randomize_spins_(s)
        return s

    systems = [build(r) for r in range(R)]
    announce(f"ewald batch {model} R={R} fp32")
    batch = sh.ReplicaBatch(systems, precision=32)
    E_dev = sh.energy(batch)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.0)
    kTs = np.linspace(0.1, 1.0, R)
    refs = [s.dipoles_flat() for s in systems]
    sh.step_(batch, integ, nsteps=2, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        o = O.Oracle(s.spec())
        S = refs[r]
        scale = max(1.0, np.abs(o.grad(S)).max() * np.abs(S).max() * s.nsites)
        assert abs(E_dev[r] - o.energy(S)) <= 3e-5 * scale, (r, E_dev[r], o.energy(S))
        key = O.replica_key(batch.noise_seed, r)
        for k in range(2):
            S = o.step_langevin(S=S, noise=O.philox_normals(key, k, s.nsites, 3), dt=integ.dt,
                                damping=integ.damping, kT=kTs[r])
        assert rel_err(s.dipoles_flat(), S, scale=1.0) <= 1e-4, r


def test_triclinic_demag_third_matches_direct_sum():
    """The triclinic cell with demag = 1/3 on the device against the oracle
    and against the independent direct dipole sum that pins the oracle."""
    from test_oracle_ewald import direct_dipole_sum, triclinic_dims

    for dims in [(1, 1, 1), (2, 3, 1), (3, 2, 2)]:
        announce(f"ewald triclinic demag 1/3 {dims} fp64")
        s = triclinic_dims(dims)
        S = s.dipoles_flat()
        o = O.Oracle(s.spec())
        E = sh.energy(s)
        assert abs(E - o.energy(S)) <= 1e-11 * abs(E)
        assert rel_err(sh.energy_grad_dipoles(s), o.grad(S)) <= 1e-11
        if s.nsites <= 18:  # where the cutoff sum converges at Rc = 20
            E_dir, _ = direct_dipole_sum(s, S, Rc=20.0)
            assert abs(E - E_dir) <= 2e-6 * abs(E)


@pytest.mark.parametrize("L", [4, 8, 16])
def test_dipolar_pyrochlore_sizes(L):
    """Dipolar pyrochlore (16 sublattices) at L^3 cells: device tensor, line
    transforms of length L and the per-k contraction against the oracle, FP64
    to 1e-11 and FP32 to 3e-5; at 16^3 this is the Ewald workload's size."""
    kw = dict(L=L, J1=0.05, J2=0, A=0, s=3.5, seed=40 + L, dipolar=True)
    s64 = sh.workloads.pyrochlore_model(**kw, precision=64)
    S = s64.dipoles_flat()
    announce(f"ewald pyrochlore {L}^3 fp64 + fp32")
    o = O.Oracle(s64.spec())
    g_ref, E_ref = o.grad(S), o.energy(S)
    scale = np.abs(g_ref).max() * np.abs(S).max() * s64.nsites
    assert rel_err(sh.energy_grad_dipoles(s64), g_ref) <= 1e-11
    assert abs(sh.energy(s64) - E_ref) <= 1e-11 * scale
    s32 = sh.workloads.pyrochlore_model(**kw, precision=32)
    s32.set_state_flat(dip=S)
    assert rel_err(sh.energy_grad_dipoles(s32), g_ref) <= 3e-5
    assert abs(sh.energy(s32) - E_ref) <= 3e-5 * scale


def test_ewald_batch_pyrochlore8_mfma():
    """Sixteen dipolar pyrochlore 8^3 replicas: one column tile of the matrix
    core contraction; energies and one Langevin step of every replica, three
    of them against their own FP64 oracle."""
    R = 16
    systems = [sh.workloads.pyrochlore_model(L=8, J1=0.05, J2=0, A=0, s=3.5, seed=80 + r, dipolar=True,
                                             precision=32) for r in range(R)]
    announce(f"ewald batch pyrochlore 8^3 R={R} fp32")
    refs = [s.dipoles_flat() for s in systems]
    batch = sh.ReplicaBatch(systems, precision=32)
    E_dev = sh.energy(batch)
    integ = sh.Langevin(0.01, damping=0.1, kT=0.2)
    sh.step_(batch, integ, nsteps=1)
    batch.sync_to_systems()
    A = O.ewald_A(systems[0].spec(), systems[0].dims)
    for r in (0, 7, 15):
        s = systems[r]
        o = O.Oracle(s.spec(), A=A)
        S = refs[r]
        g = o.grad(S)
        scale = np.abs(g).max() * np.abs(S).max() * s.nsites
        assert abs(E_dev[r] - o.energy(S)) <= 3e-5 * scale, r
        key = O.replica_key(batch.noise_seed, r)
        S = o.step_langevin(S=S, noise=O.philox_normals(key, 0, s.nsites, 3), dt=integ.dt, damping=integ.damping,
                            kT=integ.kT)
        assert rel_err(s.dipoles_flat(), S, scale=3.5) <= 1e-4, r


@pytest.mark.parametrize("precision", [64, 32])
def test_library_fft_path_matches_oracle(precision, monkeypatch):
    """SNH_EWALD_FFT=lib sends small grids through the hipFFT path (gather,
    batched 3D transforms, plain spectra, scatter): energies, gradients and
    Langevin steps against the oracle, a :SUN system, and an FP32 batch of 9
    replicas on the matrix-core contraction over plain spectra."""
    monkeypatch.setenv("SNH_EWALD_FFT", "lib")
    tol = 1e-11 if precision == 64 else 3e-5
    models = [("triclinic", random_triclinic(precision)),
              ("pyro3_dipolar", sh.workloads.pyrochlore_model(dims=(3, 2, 4), J1=0.05, J2=0, A=0, s=3.5, seed=5,
                                                              dipolar=True, precision=precision))]
    for name, s in models:
        announce(f"ewald library FFT {name} fp{precision}")
        assert s._device().query("ewald_fft_lib") == 1
        S = s.dipoles_flat()
        o = O.Oracle(s.spec())
        g_ref = o.grad(S)
        assert rel_err(sh.energy_grad_dipoles(s), g_ref) <= tol, name
        scale = max(1.0, np.abs(g_ref).max() * np.abs(S).max() * s.nsites)
        assert abs(sh.energy(s) - o.energy(S)) <= tol * scale, name
        rng = np.random.default_rng(3)
        nz = rng.standard_normal((s.nsites, 3))
        S = o.step_langevin(S=S, noise=nz, dt=0.01, damping=0.1, kT=0.2)
        sh.step_(s, sh.Langevin(0.01, damping=0.1, kT=0.2), noise=nz)
        assert rel_err(s.dipoles_flat(), S, scale=3.5) <= tol * 3, name
    cryst = sh.cubic_crystal(1.0)
    u = sh.System(cryst, [(0, sh.Moment(s=1, g=2))], "SUN", dims=(2, 2, 3), seed=3, precision=precision)
    sh.set_exchange_(u, 0.3, sh.Bond(0, 0, (1, 0, 0)))
    sh.enable_dipole_dipole_(u, 0.5)
    sh.randomize_spins_(u)
    announce(f"ewald library FFT SU(3) fp{precision}")
    Z = u.coherents_flat()
    o = O.Oracle(u.spec())
    assert rel_err(sh.energy_grad_coherents(u), o.grad_coherents(Z)) <= tol
    assert abs(sh.energy(u) - o.energy(Z=Z)) <= tol * 10
    if precision == 32:
        R = 9
        systems = [sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=60 + r, dipolar=True,
                                                 precision=32) for r in range(R)]
        announce("ewald library FFT batch pyro2 R=9 fp32")
        batch = sh.ReplicaBatch(systems, precision=32)
        assert batch._device().query("ewald_fft_lib") == 1
        E_dev = sh.energy(batch)
        for r in (0, 4, 8):
            S = systems[r].dipoles_flat()
            o = O.Oracle(systems[r].spec())
            scale = max(1.0, np.abs(o.grad(S)).max() * np.abs(S).max() * systems[r].nsites)
            assert abs(E_dev[r] - o.energy(S)) <= 3e-5 * scale, r


def test_library_fft_matches_workgroup_fft():
    """Dipolar pyrochlore 8^3 both ways at FP64 (one handle per path):
    energies and gradients agree to rounding."""
    import os
    kw = dict(L=8, J1=0.05, J2=0, A=0, s=3.5, seed=17, dipolar=True, precision=64)
    a = sh.workloads.pyrochlore_model(**kw)
    announce("ewald pyrochlore 8^3 library vs workgroup FFT fp64")
    Ea, ga = sh.energy(a), sh.energy_grad_dipoles(a)
    assert a._device().query("ewald_fft_lib") == 0
    os.environ["SNH_EWALD_FFT"] = "lib"
    try:
        b = sh.workloads.pyrochlore_model(**kw)
        Eb, gb = sh.energy(b), sh.energy_grad_dipoles(b)
        assert b._device().query("ewald_fft_lib") == 1
    finally:
        del os.environ["SNH_EWALD_FFT"]
    assert abs(Ea - Eb) <= 1e-12 * abs(Ea)
    assert rel_err(gb, ga) <= 1e-12


def test_ewald_batch_mixed_dims():
    """Replicas of different dims: each run of consecutive equal dims gets its
    own grid, tensor and spectra.  FP64: energies and two Philox Langevin
    steps of every replica against its own oracle; FP32: nine 2x2x1
    replicas (matrix-core contraction) beside two 1x2x3 ones."""
    dims = [(2, 2, 1), (2, 2, 1), (3, 2, 1), (2, 2, 1)]
    kw = dict(J1=0.05, J2=0, A=0, s=3.5, dipolar=True)
    systems = [sh.workloads.pyrochlore_model(dims=d, seed=110 + r, precision=64, **kw) for r, d in enumerate(dims)]
    announce("ewald batch of mixed dims fp64")
    refs = [s.dipoles_flat() for s in systems]
    batch = sh.ReplicaBatch(systems, precision=64)
    E_dev = sh.energy(batch)
    integ = sh.Langevin(0.01, damping=0.1, kT=0.2)
    sh.step_(batch, integ, nsteps=2)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        o = O.Oracle(s.spec())
        S = refs[r]
        scale = max(1.0, np.abs(o.grad(S)).max() * np.abs(S).max() * s.nsites)
        assert abs(E_dev[r] - o.energy(S)) <= 1e-11 * scale, r
        key = O.replica_key(batch.noise_seed, r)
        for k in range(2):
            S = o.step_langevin(S=S, noise=O.philox_normals(key, k, s.nsites, 3), dt=integ.dt, damping=integ.damping,
                                kT=integ.kT)
        assert rel_err(s.dipoles_flat(), S, scale=3.5) <= 1e-11, r
    dims32 = [(2, 2, 1)] * 9 + [(1, 2, 3)] * 2
    systems = [sh.workloads.pyrochlore_model(dims=d, seed=130 + r, precision=32, **kw) for r, d in enumerate(dims32)]
    announce("ewald batch of mixed dims fp32 (9 + 2 replicas)")
    batch = sh.ReplicaBatch(systems, precision=32)
    E_dev = sh.energy(batch)
    for r in (0, 8, 9, 10):
        s = systems[r]
        o = O.Oracle(s.spec())
        S = s.dipoles_flat()
        scale = max(1.0, np.abs(o.grad(S)).max() * np.abs(S).max() * s.nsites)
        assert abs(E_dev[r] - o.energy(S)) <= 3e-5 * scale, r


def test_nonsymmetric_demag_matches_oracle():
    """A non-symmetric demagnetization tensor (the reference takes any 3x3):
    gradient and energy against the oracle (FP64)."""
    s = random_triclinic(64)
    rng = np.random.default_rng(5)
    Rm = rng.standard_normal((3, 3))
    K = rng.standard_normal((3, 3))
    sh.enable_dipole_dipole_(s, 1.0, demag=Rm.T @ Rm + 0.3 * (K - K.T))
    announce("ewald non-symmetric demag fp64")
    S = s.dipoles_flat()
    o = O.Oracle(s.spec())
    g_ref = o.grad(S)
    assert rel_err(sh.energy_grad_dipoles(s), g_ref) <= 1e-11
    E_ref = o.energy(S)
    assert abs(sh.energy(s) - E_ref) <= 1e-11 * max(1.0, np.abs(g_ref).max() * np.abs(S).max() * s.nsites)


@pytest.mark.parametrize("precision", [64, 32])
def test_slab_path_matches_field_transforms(precision, monkeypatch):
    """Few replicas take the slab path (per-plane 2D transforms, the third
    axis and the contraction per in-plane wavevector); one to four replicas
    of the dipolar pyrochlore (a 3 x 4 x 2 grid: odd, even and length-2
    axes) and the triclinic three-atom cell agree with the per-field
    transforms and with the oracle."""
    tol = 1e-12 if precision == 64 else 3e-5

    def build(mode, systems):
        monkeypatch.setenv("SNH_EWALD_SLAB", mode)
        b = sh.ReplicaBatch(systems, precision=precision)
        b._device()  # the handle reads SNH_EWALD_SLAB when it is created
        return b

    for R in (1, 2, 4):
        def systems():
            return [sh.workloads.pyrochlore_model(dims=(3, 4, 2), J1=0.05, J2=0, A=0, s=3.5, seed=70 + r,
                                                  dipolar=True, precision=precision) for r in range(R)]
        announce(f"ewald slab path pyrochlore 3x4x2, {R} replicas, fp{precision}")
        a = build("2", systems())
        b = build("0", systems())
        assert a._device().query("ewald_slab") == 1 and b._device().query("ewald_slab") == 0
        Ea, Eb = sh.energy(a), sh.energy(b)
        np.testing.assert_allclose(Ea, Eb, rtol=tol, atol=0)
        ga, gb = sh.energy_grad_dipoles(a), sh.energy_grad_dipoles(b)
        assert rel_err(ga, gb) <= tol
        if precision == 64:
            for r, s in enumerate(systems()):
                o = O.Oracle(s.spec())
                assert abs(Ea[r] - o.energy(s.dipoles_flat())) <= 1e-11 * abs(Ea[r]), r
    announce(f"ewald slab path triclinic fp{precision}")
    monkeypatch.setenv("SNH_EWALD_SLAB", "1")
    s = random_triclinic(precision)
    assert s._device().query("ewald_slab") == 1
    S = s.dipoles_flat()
    o = O.Oracle(s.spec())
    assert rel_err(sh.energy_grad_dipoles(s), o.grad(S)) <= (1e-11 if precision == 64 else 3e-5)
