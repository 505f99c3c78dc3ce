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
