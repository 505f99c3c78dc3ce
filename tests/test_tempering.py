"""Parallel tempering host logic (ParallelTempering.jl:39-95) over a replica
batch: exchange acceptance, rank alternation, and (on the GPU) a short run."""
import math

import numpy as np
import pytest

import sunnyhip as sh
import sunnyhip.tempering as tp

from helpers import announce


def make(kTs, seed=0):
    s = sh.workloads.cubic_model(L=3, seed=1, precision=64)
    return tp.ParallelTempering(s, sh.Langevin(0.02, damping=0.1, kT=0.0), kTs, seed=seed)


def test_equal_energies_always_exchange():
    pt = make([0.1, 0.2, 0.4, 0.8])
    pt.replica_exchange_(0, energies=np.zeros(4))
    assert list(pt.system_ids) == [1, 0, 3, 2]
    pt.replica_exchange_(1, energies=np.zeros(4))
    assert list(pt.system_ids) == [1, 3, 0, 2]
    assert list(pt.n_exch) == [1, 1, 1, 0] and list(pt.n_accept) == [1, 1, 1, 0]
    # every replica keeps exactly one temperature
    assert sorted(pt.replica_kTs()) == [0.1, 0.2, 0.4, 0.8]
    assert pt.replica_kTs()[pt.system_ids[1]] == 0.2


def test_acceptance_probability():
    """ln P = (E1 - E2)(1/kT1 - 1/kT2): a swap that lowers the weight by e is
    accepted with probability 1/e; a favourable one always."""
    pt = make([1.0, 2.0], seed=7)
    n, acc = 20000, 0
    for _ in range(n):
        pt.system_ids[:] = [0, 1]
        pt.replica_exchange_(0, energies=np.array([-3.0, -1.0]))  # ln P = (-3 + 1)(1 - 1/2) = -1
        acc += pt.system_ids[0] == 1
    p = math.exp(-1)
    assert abs(acc / n - p) < 4 * math.sqrt(p * (1 - p) / n)
    for _ in range(50):
        # the colder rank holds the higher energy: the swap is always taken
        pt.system_ids[:] = [0, 1]
        pt.replica_exchange_(0, energies=np.array([-1.0, -3.0]))  # ln P = +1
        assert pt.system_ids[0] == 1


def test_rounds_alternate(monkeypatch):
    pt = make([0.1, 0.2, 0.4])
    calls, starts = [], []
    monkeypatch.setattr(tp, "step_", lambda b, integ, nsteps, kTs: calls.append((nsteps, tuple(kTs))))
    monkeypatch.setattr(pt, "replica_exchange_", lambda s: starts.append(s))
    pt.step_ensemble_(25, 10)
    assert [c[0] for c in calls] == [10, 10, 10] and starts == [1, 0, 1]


@pytest.mark.gpu
def test_tempering_run_orders_energies():
    kTs = [0.05, 0.2, 0.8, 3.2]
    pt = make(kTs, seed=3)
    announce("parallel tempering cubic 3^3 fp64, 4 ranks")
    means = np.zeros(4)
    for _ in range(20):
        pt.step_ensemble_(40, 10)
        E = sh.energy(pt.batch)
        means += E[pt.system_ids] / 20
    assert sorted(pt.system_ids) == [0, 1, 2, 3]
    assert np.all(np.diff(means) > 0), means
    assert np.all(pt.n_exch[:3] > 0)


def oracle_replay(systems, kT_sched, key_seed, nsteps, exch_interval, seed, langevin):
    """step_ensemble! (ParallelTempering.jl:79-95) restated on the oracle:
    each replica id steps with its own Philox stream at its rank's
    temperature, energies from the oracle, exchanges of neighbouring ranks
    with the same acceptance draws."""
    from oracle import oracle as O
    R = len(systems)
    orc = [O.Oracle(s.spec()) for s in systems]
    S = [s.dipoles_flat() for s in systems]
    ids = np.arange(R)
    n_acc, n_ex = np.zeros(R, np.int64), np.zeros(R, np.int64)
    rng = np.random.default_rng(seed)
    counter = 0
    E = np.array([orc[r].energy(S[r]) for r in range(R)])
    for i in range(1, -(-nsteps // exch_interval) + 1):
        for rank in range(R):
            r = ids[rank]
            key = O.replica_key(key_seed, int(r))
            if langevin is not None:
                S[r], _ = orc[r].langevin_run(S=S[r], dt=langevin.dt, damping=langevin.damping, kT=kT_sched[rank],
                                              key=key, step0=counter, nsteps=exch_interval, every=0)
            else:
                m, nc, col = systems[0]._pt_coloring
                for k in range(exch_interval):
                    S[r], acc = orc[r].local_sweeps(S[r], m, col.reshape(-1), nc, [(2, 0.3, 1.0)], kT_sched[rank],
                                                    key, counter + k, 1.0)
                    E[r] += acc[0]
        counter += exch_interval
        if langevin is not None:
            E = np.array([orc[r].energy(S[r]) for r in range(R)])
        for rank in range(i % 2, R - 1, 2):
            E1, E2 = E[ids[rank]], E[ids[rank + 1]]
            lnP = (E1 / kT_sched[rank] - E2 / kT_sched[rank]) + (E2 / kT_sched[rank + 1] - E1 / kT_sched[rank + 1])
            if lnP >= 0 or rng.random() < math.exp(lnP):
                ids[rank], ids[rank + 1] = ids[rank + 1], ids[rank]
                n_acc[rank] += 1
            n_ex[rank] += 1
    return S, ids, n_acc, n_ex, E


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["langevin", "local"])
def test_tempering_replays_on_oracle(kind):
    """The device ensemble (Langevin or LocalSampler, FP64) against the
    oracle's replay with the same draws: labels, counters and states."""
    kTs = [0.1, 0.25, 0.6, 1.5, 4.0]
    base = sh.workloads.pyrochlore_model(L=1, seed=21, precision=64)
    systems = [sh.clone_system(base, seed=300 + r) for r in range(len(kTs))]
    for s in systems:
        sh.randomize_spins_(s)
    init = [sh.clone_system(s) for s in systems]
    announce(f"parallel tempering replay {kind}, pyrochlore 1^3 fp64, 5 ranks")
    if kind == "langevin":
        sampler = sh.Langevin(0.02, damping=0.2, kT=0.0)
    else:
        sampler = sh.LocalSampler(kT=1.0, propose=sh.propose_delta(0.3))
    pt = tp.ParallelTempering(base, sampler, kTs, systems=systems, precision=64, seed=9)
    if kind == "local":
        init[0]._pt_coloring = sh.local_coloring(pt.batch)
    pt.step_ensemble_(60, 10)
    pt.batch.sync_to_systems()
    S, ids, n_acc, n_ex, E = oracle_replay(init, kTs, pt.batch.noise_seed, 60, 10, 9,
                                           sampler if kind == "langevin" else None)
    assert list(pt.system_ids) == list(ids)
    assert list(pt.n_exch) == list(n_ex) and list(pt.n_accept) == list(n_acc)
    assert n_acc.sum() > 0
    for r, s in enumerate(pt.systems):
        assert np.max(np.abs(s.dipoles_flat() - S[r])) <= 1e-10, r
    if kind == "local":
        np.testing.assert_allclose(pt.E, E, atol=1e-9)
        np.testing.assert_allclose(pt.E, sh.energy(pt.batch), atol=1e-9)
