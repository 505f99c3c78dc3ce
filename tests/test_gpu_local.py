"""GPU parity of the device LocalSampler against the oracle.

The device runs colour-class passes of Metropolis trials with Philox draws;
the oracle's restatement (oracle/sunny_oracle.c: so_local_sweeps, whose
local_energy_change and sampling are pinned in test_oracle_local.py) makes
the same trials with the same draws, so FP64 states agree to rounding
(<= 1e-12 of kappa) and FP32 states to 2e-5 of kappa.  Known answers at
sizes the oracle would take minutes for: a paramagnet's Langevin function
and periodic Ising chains with flip proposals.  All calls go through
libsunnyhip's C ABI.
"""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O

from helpers import announce, coloring_is_proper, energy_scale, langevin_function
from test_gpu_dipole import aniso_model
from test_oracle_local import ising_chains, su5_mean_energy, sun_anisotropy_chain, sun_biquad_system
from test_oracle_dynamics import su3_mean_energy

pytestmark = pytest.mark.gpu

TOL = {64: 1e-12, 32: 2e-5}

PROPOSALS = {
    "uniform": sh.propose_uniform,
    "flip": sh.propose_flip,
    "delta": sh.propose_delta(0.3),
    "mix": sh.mix_proposals(0.4, sh.propose_flip, 0.6, sh.propose_delta(0.2)),
}


def models(precision):
    yield "cubic4", sh.workloads.cubic_model(L=4, seed=3, precision=precision)
    yield "cubic_ragged", sh.workloads.cubic_model(dims=(3, 5, 2), seed=4, precision=precision)
    yield "pyro2", sh.workloads.pyrochlore_model(L=2, seed=5, precision=precision)
    yield "pyro_ragged", sh.workloads.pyrochlore_model(dims=(1, 2, 3), seed=6, precision=precision)
    s = sh.workloads.cubic_model(dims=(2, 2, 3), seed=8, precision=precision)
    sh.set_exchange_(s, 0.3, sh.Bond(0, 0, (2, 1, 0)), propagate=[sh.Bond(0, 0, (2, 1, 0))])
    sh.set_exchange_(s, -0.2, sh.Bond(0, 0, (0, 3, 4)), propagate=[sh.Bond(0, 0, (0, 3, 4))])
    sh.randomize_spins_(s)
    yield "cubic_long_bonds", s
    yield "diamond_aniso", aniso_model(precision)


def oracle_props(p):
    return [(k, m, w) for k, m, w in p.terms]


def run_oracle(s, S, kT, prop, key, sweep0, nsweeps):
    m, nc, col = sh.local_coloring(s)
    return O.Oracle(s.spec()).local_sweeps(S, m, col.reshape(-1), nc, oracle_props(prop), kT, key, sweep0, nsweeps)


@pytest.mark.parametrize("precision", [64, 32])
def test_sweeps_match_oracle(precision, layout):
    for name, s in models(precision):
        s.layout = layout
        kap = np.repeat(s.kappas, s.ncells)
        for pname, prop in PROPOSALS.items():
            if precision == 32 and pname != "flip":
                # FP32 draws and proposals differ from the oracle's by rounding, and a
                # trial near its acceptance threshold could then go the other way
                continue
            announce(f"local sweeps {name} {pname} fp{precision} layout {layout}")
            S0 = s.dipoles_flat()
            E0 = sh.energy(s)
            key = O.replica_key(s.noise_seed, 0)
            sweep0 = s.step_counter
            nsweeps = 1.5 if pname == "mix" else 2.0
            S_ref, acc = run_oracle(s, S0, 0.7, prop, key, sweep0, nsweeps)
            sampler = sh.LocalSampler(kT=0.7, nsweeps=nsweeps, propose=prop)
            sh.step_(s, sampler)
            S_dev = s.dipoles_flat()
            assert np.max(np.abs(S_dev - S_ref)) <= TOL[precision] * kap.max(), (name, pname)
            assert np.sum(np.any(np.abs(S_dev - S0) > 1e-9, axis=1)) > 0, (name, pname)
            scale = energy_scale(s)
            assert abs(sampler.dE - acc[0]) <= 10 * TOL[precision] * scale, (name, pname, sampler.dE, acc[0])
            np.testing.assert_allclose(sampler.dS, acc[1:], atol=10 * TOL[precision] * s.nsites * kap.max())
            # the accumulated change is the change of the energy
            assert abs(sh.energy(s) - E0 - sampler.dE) <= 10 * TOL[precision] * scale, (name, pname)


def test_zero_temperature_descends():
    s = sh.workloads.pyrochlore_model(L=2, seed=9, precision=64)
    announce("local sweeps kT=0 pyro2")
    E = [sh.energy(s)]
    sampler = sh.LocalSampler(kT=0.0, propose=sh.propose_delta(0.5))
    key = O.replica_key(s.noise_seed, 0)
    S_ref, _ = run_oracle(s, s.dipoles_flat(), 0.0, sampler.propose, key, s.step_counter, 3.0)
    for _ in range(3):
        sh.step_(s, sampler)
        E.append(sh.energy(s))
    assert all(b <= a + 1e-12 for a, b in zip(E, E[1:])), E
    assert np.max(np.abs(s.dipoles_flat() - S_ref)) <= 1e-12 * 1.5


def test_replica_batch_matches_oracle():
    dims = [(2, 2, 2), (2, 2, 4), (4, 2, 2)]
    systems = [sh.workloads.pyrochlore_model(dims=d, seed=30 + r, precision=64) for r, d in enumerate(dims)]
    announce(f"local sweeps replica batch pyrochlore {dims}")
    batch = sh.ReplicaBatch(systems, precision=64, layout=1)
    kTs = np.array([0.2, 0.6, 1.5])
    refs = [s.dipoles_flat() for s in systems]
    prop = sh.mix_proposals(0.5, sh.propose_uniform, 0.5, sh.propose_delta(0.25))
    sampler = sh.LocalSampler(kT=0.0, nsweeps=2, propose=prop)
    m, nc, col = sh.local_coloring(batch)
    for d in dims:
        assert coloring_is_proper(systems[0].spec(), d, m, col)
    sh.step_(batch, sampler, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        o = O.Oracle(s.spec())
        S, acc = o.local_sweeps(refs[r], m, col.reshape(-1), nc, oracle_props(prop), kTs[r],
                                O.replica_key(batch.noise_seed, r), 0, 2.0)
        assert np.max(np.abs(s.dipoles_flat() - S)) <= 1e-12 * 1.5, r
        assert abs(sampler.dE[r] - acc[0]) <= 1e-9 * energy_scale(s), r


def test_colorings_are_proper():
    for name, s in models(64):
        announce(f"local colouring {name}")
        m, nc, col = sh.local_coloring(s)
        assert coloring_is_proper(s.spec(), s.dims, m, col), name
        assert 2 <= nc <= 4 * s.crystal.natoms + 8, (name, nc)


def test_langevin_after_sweeps_sees_updated_images():
    """Ghost images written by accepted trials feed the next Langevin step."""
    s = sh.workloads.cubic_model(L=4, seed=13, precision=64)
    s.layout = 1
    announce("local sweeps then langevin, ghost layout")
    sh.step_(s, sh.LocalSampler(kT=0.5, propose=sh.propose_flip))
    S = s.dipoles_flat()
    o = O.Oracle(s.spec())
    noise = np.random.default_rng(2).standard_normal((s.nsites, 3))
    integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
    S_ref = o.step_langevin(S=S, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
    sh.step_(s, integ, noise=noise)
    assert np.max(np.abs(s.dipoles_flat() - S_ref)) <= 1e-11


@pytest.mark.parametrize("precision", [64, 32])
def test_ewald_sequential_trials_match_oracle(precision):
    """With the Ewald term the device tries one site at a time per replica, in
    chunks of 64 trials per launch (each trial sees the potential of the
    chunk's start plus the moves accepted since), and adds a chunk's moves to
    the FP64 potential everywhere after it; the oracle's so_local_sequential
    makes the same trials with the same draws, updating after every move.
    pyro321 (96 sites, 192 trials) spans three chunks, the last a partial one."""
    cases = [("pyro1", sh.workloads.pyrochlore_model(L=1, seed=23, precision=precision, dipolar=True)),
             ("pyro211", sh.workloads.pyrochlore_model(dims=(2, 1, 1), seed=24, precision=precision, dipolar=True)),
             # an axis of length 3: offsets differ from their negatives, so the
             # atom order and offset sign of the potential update are pinned
             ("pyro321", sh.workloads.pyrochlore_model(dims=(3, 2, 1), seed=25, precision=precision, dipolar=True)),
             # 288 sites, 576 trials in nine chunks: most trials take the
             # parallel evaluation's exact corrections, a few are made again
             ("pyro332", sh.workloads.pyrochlore_model(dims=(3, 3, 2), seed=27, precision=precision, dipolar=True))]
    for name, s in cases:
        kap = np.repeat(s.kappas, s.ncells)
        for pname in (["uniform", "mix"] if precision == 64 else ["flip"]):
            prop = PROPOSALS[pname]
            announce(f"local sequential ewald {name} {pname} fp{precision}")
            S0 = s.dipoles_flat()
            E0 = sh.energy(s)
            key = O.replica_key(s.noise_seed, 0)
            S_ref, acc = O.Oracle(s.spec()).local_sequential(S0, oracle_props(prop), 0.6, key, s.step_counter, 2.0)
            sampler = sh.LocalSampler(kT=0.6, nsweeps=2.0, propose=prop)
            sh.step_(s, sampler)
            S_dev = s.dipoles_flat()
            assert np.max(np.abs(S_dev - S_ref)) <= TOL[precision] * kap.max(), (name, pname)
            assert np.sum(np.any(np.abs(S_dev - S0) > 1e-9, axis=1)) > 0, (name, pname)
            scale = energy_scale(s)
            assert abs(sampler.dE - acc[0]) <= 10 * TOL[precision] * scale, (name, pname, sampler.dE, acc[0])
            assert abs(sh.energy(s) - E0 - sampler.dE) <= 10 * TOL[precision] * scale, (name, pname)


@pytest.mark.parametrize("precision", [64, 32])
def test_sun_ewald_sequential_trials_match_oracle(precision):
    """:SUN with Ewald: the device's sequential trials (k_sun_mc_ewald_chunk)
    against the oracle's so_local_sequential_sun with the same draws."""
    from test_oracle_local import sun_dipolar_system
    s = sun_dipolar_system(seed=26)
    s.precision = precision
    kap = s.spec()["kappa"][0]
    for pname in (["uniform", "mix"] if precision == 64 else ["flip"]):
        prop = PROPOSALS[pname]
        announce(f"local sequential ewald SU(3) cubic 2x2x3 {pname} fp{precision}")
        Z0 = s.coherents_flat()
        E0 = sh.energy(s)
        key = O.replica_key(s.noise_seed, 0)
        Z_ref, acc = O.Oracle(s.spec()).local_sequential_sun(Z0, oracle_props(prop), 0.6, key, s.step_counter, 2.0)
        sampler = sh.LocalSampler(kT=0.6, nsweeps=2.0, propose=prop)
        sh.step_(s, sampler)
        Z_dev = s.coherents_flat()
        assert np.max(np.abs(Z_dev - Z_ref)) <= TOL[precision] * np.sqrt(kap) * 4, pname
        assert np.sum(np.any(np.abs(Z_dev - Z0) > 1e-9, axis=1)) > 0, pname
        scale = energy_scale(s)
        assert abs(sampler.dE - acc[0]) <= 10 * TOL[precision] * scale, (pname, sampler.dE, acc[0])
        assert abs(sh.energy(s) - E0 - sampler.dE) <= 10 * TOL[precision] * scale, pname


def test_ewald_sequential_batch_matches_oracle():
    systems = [sh.workloads.pyrochlore_model(dims=(2, 2, 1), seed=40 + r, precision=64, dipolar=True)
               for r in range(2)]
    announce("local sequential ewald replica batch")
    batch = sh.ReplicaBatch(systems, precision=64)
    kTs = np.array([0.3, 1.2])
    refs = [s.dipoles_flat() for s in systems]
    prop = sh.mix_proposals(0.5, sh.propose_uniform, 0.5, sh.propose_delta(0.25))
    sampler = sh.LocalSampler(kT=0.0, nsweeps=1.5, propose=prop)
    sh.step_(batch, sampler, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        S, acc = O.Oracle(s.spec()).local_sequential(refs[r], oracle_props(prop), kTs[r],
                                                     O.replica_key(batch.noise_seed, r), 0, 1.5)
        assert np.max(np.abs(s.dipoles_flat() - S)) <= 1e-12 * 1.5, r
        assert abs(sampler.dE[r] - acc[0]) <= 1e-9 * energy_scale(s), r


@pytest.mark.parametrize("fft", ["workgroup", "lib"])
def test_ewald_sequential_batch_fft_refresh(fft, monkeypatch):
    """A batch of at least four replicas refreshes the potential of its
    sequential trials by one FFT evaluation after every eighth chunk
    (instead of adding every accepted move at every site; the chunks between
    add the moves since the refresh at their trial sites), written in FP64
    by the inverse transforms or converted from the library path's output:
    states and energy changes against per-replica oracles over eleven chunks
    each (a refresh, then three more chunks)."""
    if fft == "lib":
        monkeypatch.setenv("SNH_EWALD_FFT", "lib")
    systems = [sh.workloads.pyrochlore_model(dims=(3, 2, 1), seed=150 + r, precision=64, dipolar=True)
               for r in range(4)]
    announce(f"local sequential ewald replica batch of four, FFT refresh ({fft})")
    batch = sh.ReplicaBatch(systems, precision=64)
    assert batch._device().query("ewald_fft_lib") == (fft == "lib")
    refs = [s.dipoles_flat() for s in systems]
    kTs = np.array([0.2, 0.5, 1.0, 3.0])
    prop = sh.mix_proposals(0.5, sh.propose_uniform, 0.5, sh.propose_delta(0.25))
    sampler = sh.LocalSampler(kT=0.0, nsweeps=7.0, propose=prop)
    sh.step_(batch, sampler, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        S, acc = O.Oracle(s.spec()).local_sequential(refs[r], oracle_props(prop), kTs[r],
                                                     O.replica_key(batch.noise_seed, r), 0, 7.0)
        assert np.max(np.abs(s.dipoles_flat() - S)) <= 1e-12 * 1.5, r
        assert abs(sampler.dE[r] - acc[0]) <= 1e-9 * energy_scale(s), r


def test_sun_ewald_sequential_batch_fft_refresh():
    """:SUN with Ewald in a batch of four replicas (k_sun_mc_ewald_chunk with
    the FFT refresh after every eighth chunk, the chunks between adding the
    moves since the refresh at their trial sites): kets and energy changes
    against per-replica oracles over ten chunks each."""
    from test_oracle_local import sun_dipolar_system
    systems = [sun_dipolar_system(seed=160 + r, dims=(4, 4, 3)) for r in range(4)]
    announce("local sequential ewald SU(3) replica batch of four, FFT refresh")
    refs = [s.coherents_flat() for s in systems]
    batch = sh.ReplicaBatch(systems, precision=64)
    kTs = np.array([0.2, 0.5, 1.0, 2.0])
    prop = PROPOSALS["mix"]
    sampler = sh.LocalSampler(kT=0.0, nsweeps=13.0, propose=prop)
    sh.step_(batch, sampler, kTs=kTs)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        kap = s.spec()["kappa"][0]
        Z_ref, acc = O.Oracle(s.spec()).local_sequential_sun(refs[r], oracle_props(prop), kTs[r],
                                                             O.replica_key(batch.noise_seed, r), 0, 13.0)
        assert np.max(np.abs(s.coherents_flat() - Z_ref)) <= TOL[64] * np.sqrt(kap) * 4, r
        assert abs(sampler.dE[r] - acc[0]) <= 10 * TOL[64] * energy_scale(s), r


def test_ewald_sequential_batch_mixed_dims():
    """Sequential trials with the Ewald term on replicas of different dims
    (each run of equal dims has its own real-space tensor)."""
    dims = [(2, 2, 1), (1, 2, 3), (1, 2, 3)]
    systems = [sh.workloads.pyrochlore_model(dims=d, seed=140 + r, precision=64, dipolar=True)
               for r, d in enumerate(dims)]
    announce("local sequential ewald replica batch of mixed dims")
    batch = sh.ReplicaBatch(systems, precision=64)
    refs = [s.dipoles_flat() for s in systems]
    prop = sh.mix_proposals(0.5, sh.propose_uniform, 0.5, sh.propose_delta(0.25))
    sampler = sh.LocalSampler(kT=0.5, nsweeps=1.0, propose=prop)
    sh.step_(batch, sampler)
    batch.sync_to_systems()
    for r, s in enumerate(systems):
        S, acc = O.Oracle(s.spec()).local_sequential(refs[r], oracle_props(prop), 0.5,
                                                     O.replica_key(batch.noise_seed, r), 0, 1.0)
        assert np.max(np.abs(s.dipoles_flat() - S)) <= 1e-12 * 1.5, r
        assert abs(sampler.dE[r] - acc[0]) <= 1e-9 * energy_scale(s), r


def test_unsupported_and_invalid():
    announce("local sampler errors")
    selfint = sh.workloads.cubic_model(dims=(1, 2, 2), seed=3, precision=64)
    with pytest.raises(ValueError, match="self-interaction"):
        sh.step_(selfint, sh.LocalSampler(kT=1.0))
    s = sh.workloads.cubic_model(L=2, seed=4, precision=64)
    bad = sh.LocalSampler(kT=1.0)
    bad.propose = sh.Proposal([(3, 0.0, 1.0)], "bad")
    with pytest.raises(ValueError, match="kind"):
        sh.step_(s, bad)
    with pytest.raises(ValueError, match="kT"):
        sh.step_(s, sh.LocalSampler(kT=-1.0))
    # the sequential trials keep the Ewald tensor's blocks as upper triangles
    d = sh.workloads.pyrochlore_model(L=1, seed=5, precision=64, dipolar=True)
    sh.enable_dipole_dipole_(d, 0.6745817653324668, np.array([[0.4, 0.1, 0.0], [0.0, 0.3, 0.0], [0.0, 0.0, 0.3]]))
    with pytest.raises(Exception, match="symmetric demagnetization"):
        sh.step_(d, sh.LocalSampler(kT=1.0))


@pytest.mark.parametrize("precision", [32, 64])
def test_paramagnet_langevin_function(precision):
    s = sh.workloads.cubic_model(L=16, J=0.0, D=0.0, seed=5, precision=precision)
    sh.set_field_(s, [0.0, 0.3, 0.4])
    announce(f"local sampler paramagnet 16^3 fp{precision}")
    spec = s.spec()
    kap = spec["kappa"][0]
    h = np.asarray(spec["g"][0]).reshape(3, 3).T @ np.asarray(spec["field"][0])
    hn = h / np.linalg.norm(h)
    kT = 0.6
    sampler = sh.LocalSampler(kT=kT, propose=sh.propose_uniform)
    sh.step_(s, sampler, nsteps=20)
    vals = []
    for _ in range(200):
        sh.step_(s, sampler)
        vals.append(np.mean(s.dipoles_flat() @ hn))
    blocks = np.asarray(vals).reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(20)
    exact = -kap * langevin_function(kap * np.linalg.norm(h) / kT)
    assert abs(np.mean(vals) - exact) < max(4 * err, 1e-3), (np.mean(vals), exact, err)


def test_ising_chains_with_flips():
    L = 16
    s = ising_chains(L, dims=(L, 8, 8), precision=32)
    announce("local sampler Ising chains 16x8x8 fp32")
    kap = s.spec()["kappa"][0]
    rng = np.random.default_rng(4)
    S = np.zeros((s.nsites, 3))
    S[:, 2] = kap * rng.choice([-1.0, 1.0], s.nsites)
    s.set_state_flat(dip=S)
    kT = 1.0
    sampler = sh.LocalSampler(kT=kT, propose=sh.propose_flip)
    sh.step_(s, sampler, nsteps=200)
    vals = []
    for _ in range(1000):
        sh.step_(s, sampler)
        vals.append(sh.energy(s) / s.nsites)
    vals = np.asarray(vals)
    blocks = vals.reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(20)
    t = np.tanh(kap**2 / kT)
    exact = -kap**2 * (t + t ** (L - 1)) / (1 + t**L)
    assert abs(vals.mean() - exact) < max(4 * err, 3e-3), (vals.mean(), exact, err)


def sun_models(precision):
    yield "tri_sun3_biquad", sun_biquad_system(seed=5)
    from test_gpu_sun import general_model
    yield "su3_general", general_model(precision, seed=42)
    s5 = sh.workloads.triangular_sun_model(L1=3, L2=3, s=2.0, D=0.7, seed=6, precision=precision)
    sh.set_field_(s5, [0.0, 0.3, -0.1])
    sh.randomize_spins_(s5)
    yield "tri_sun5", s5


@pytest.mark.parametrize("precision", [64, 32])
def test_sun_sweeps_match_oracle(precision):
    for name, s in sun_models(precision):
        s.precision = precision
        for pname, prop in PROPOSALS.items():
            if precision == 32 and pname != "flip":
                continue
            announce(f"local sweeps {name} {pname} fp{precision}")
            Z0 = s.coherents_flat()
            key = O.replica_key(s.noise_seed, 0)
            m, nc, col = sh.local_coloring(s)
            o = O.Oracle(s.spec())
            nsweeps = 1.5 if pname == "mix" else 2.0
            Z_ref, acc = o.local_sweeps_sun(Z0, m, col.reshape(-1), nc, oracle_props(prop), 0.7, key,
                                            s.step_counter, nsweeps)
            sampler = sh.LocalSampler(kT=0.7, nsweeps=nsweeps, propose=prop)
            sh.step_(s, sampler)
            Z_dev = s.coherents_flat()
            kap = s.spec()["kappa"][0]
            assert np.max(np.abs(Z_dev - Z_ref)) <= TOL[precision] * np.sqrt(kap) * 10, (name, pname)
            assert np.sum(np.any(np.abs(Z_dev - Z0) > 1e-9, axis=1)) > 0, (name, pname)
            assert abs(sampler.dE - acc[0]) <= 10 * TOL[precision] * energy_scale(s), (name, pname)


@pytest.mark.parametrize("N,kT", [(3, 0.125), (3, 0.5), (5, 0.25)])
def test_sun_mean_energy_closed_form(N, kT):
    """test/test_samplers.jl:6-15 closed forms from device Metropolis trials."""
    announce(f"local sampler SU({N}) anisotropy kT={kT}")
    D = 1.0
    s = sun_anisotropy_chain(N, D, L=512)
    s.precision = 32
    sampler = sh.LocalSampler(kT=kT, propose=sh.propose_uniform)
    sh.step_(s, sampler, nsteps=50)
    vals = []
    for _ in range(400):
        sh.step_(s, sampler)
        vals.append(sh.energy(s) / s.nsites)
    vals = np.asarray(vals)
    blocks = vals.reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(20)
    exact = su3_mean_energy(kT, D) if N == 3 else su5_mean_energy(kT, D)
    assert abs(vals.mean() - exact) < max(4 * err, 0.01 * abs(exact)), (vals.mean(), exact, err)


def test_sun_langevin_after_sweeps_sees_updated_kets():
    """A bilinear-only SU(3) model carries the expected-dipole planes from one
    Langevin call to the next; Metropolis sweeps in between rewrite the kets,
    so the next Langevin step must start from the swept state's dipoles."""
    s = sh.workloads.triangular_sun_model(L1=6, L2=5, seed=31, precision=64)
    announce("sun langevin, local sweeps, langevin with host noise")
    integ = sh.Langevin(0.02, damping=0.1, kT=0.3)
    sh.step_(s, integ, nsteps=3)  # device noise: the dipole planes become current
    sh.step_(s, sh.LocalSampler(kT=0.5, nsweeps=1.0, propose=sh.propose_uniform))
    Z = s.coherents_flat()
    rng = np.random.default_rng(4)
    noise = (rng.standard_normal((s.nsites, s.N)) + 1j * rng.standard_normal((s.nsites, s.N))) / np.sqrt(2)
    Z_ref = O.Oracle(s.spec()).step_langevin(Z=Z, noise=noise, dt=integ.dt, damping=integ.damping, kT=integ.kT)
    sh.step_(s, integ, noise=noise)
    assert np.max(np.abs(s.coherents_flat() - Z_ref)) <= 1e-11


def test_ewald_sequential_many_bonds():
    """More bonds per site than the chunk kernel keeps neighbour ids for (56:
    exchange to the fifth shell of a simple cubic lattice, bond lists padded
    to a multiple of the gather width) with the Ewald term: sequential trials
    and their energy change against the oracle."""
    s = sh.workloads.cubic_model(L=6, J=1.0, D=-0.05, seed=31, precision=64)
    cryst = s.crystal
    for d, J in ((2 ** 0.5, 0.3), (3 ** 0.5, -0.2), (2.0, 0.1), (5 ** 0.5, 0.05)):
        sh.set_exchange_(s, J, sh.workloads.bond_at_distance(cryst, d))
    sh.enable_dipole_dipole_(s, 0.5, demag=1 / 3)
    announce("local sequential ewald cubic 6^3, 56 bonds per site")
    prop = PROPOSALS["mix"]
    S0 = s.dipoles_flat()
    E0 = sh.energy(s)
    key = O.replica_key(s.noise_seed, 0)
    S_ref, acc = O.Oracle(s.spec()).local_sequential(S0, oracle_props(prop), 0.6, key, s.step_counter, 1.5)
    sampler = sh.LocalSampler(kT=0.6, nsweeps=1.5, propose=prop)
    sh.step_(s, sampler)
    assert np.max(np.abs(s.dipoles_flat() - S_ref)) <= 1e-12
    scale = energy_scale(s)
    assert abs(sampler.dE - acc[0]) <= 1e-9 * scale
    assert abs(sh.energy(s) - E0 - sampler.dE) <= 1e-9 * scale
