"""Captured HIP graphs of Langevin step chunks replay exactly the eager
launches: a handle stepped in long calls (graph chunks) and a twin stepped
one call per step (eager launches) stay bitwise equal, across replays
separated by other work on the stream (energy evaluations)."""

import numpy as np
import pytest

import sunnyhip as sh

from helpers import announce

pytestmark = pytest.mark.gpu


def twins(kind):
    def one():
        if kind == "tile_fp32_batch":
            systems = [sh.workloads.cubic_model(dims=d, seed=20 + r, precision=32)
                       for r, d in enumerate([(8, 8, 8), (6, 4, 10), (5, 5, 5)])]
            return sh.ReplicaBatch(systems, precision=32)
        if kind == "ghost_fp64":
            return sh.workloads.pyrochlore_model(L=3, seed=4, precision=64)
        if kind == "sun3_batch":
            systems = [sh.workloads.triangular_sun_model(L1=L, seed=30 + r, precision=32)
                       for r, L in enumerate([9, 16, 5])]
            return sh.ReplicaBatch(systems, precision=32)
        if kind == "ewald_fp32":
            return sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=6, dipolar=True, precision=32)
        raise ValueError(kind)
    return one(), one()


def state(obj):
    h = obj._device()
    d = np.empty((h.total, 3))
    h.call("snh_get_dipoles", sh._lib.dptr(d))
    return d


@pytest.mark.parametrize("kind", ["tile_fp32_batch", "ghost_fp64", "sun3_batch", "ewald_fp32"])
def test_graph_replays_match_eager(kind, monkeypatch):
    announce(f"graph vs eager {kind}")
    # graphs are opt-in, read at handle creation; they replay the per-stage
    # launches, which small ghost-layout replicas would otherwise replace by
    # the persistent kernel
    monkeypatch.setenv("SNH_PERSIST", "0")
    monkeypatch.setenv("SNH_GRAPH_STEPS", "8")
    a = twins(kind)[0]
    a._device()
    monkeypatch.setenv("SNH_GRAPH_STEPS", "0")
    b = twins(kind)[1]
    b._device()
    R = getattr(a, "nreplicas", 1)
    kTs = np.linspace(0.2, 0.9, R)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.5)
    for block in range(3):
        # a: one call of 21 steps -> captured chunks of graph_steps plus eager rest
        sh.step_(a, integ, nsteps=21, kTs=kTs if R > 1 else None)
        for _ in range(21):
            sh.step_(b, integ, nsteps=1, kTs=kTs if R > 1 else None)
        Ea, Eb = sh.energy(a), sh.energy(b)
        assert np.array_equal(state(a), state(b)), (kind, block)
        assert np.array_equal(np.atleast_1d(Ea), np.atleast_1d(Eb)), (kind, block)
    h = a._device()
    assert h.query("graph_steps") == 8
    assert h.query("graph_live") == 1
    assert b._device().query("graph_steps") == 0 and b._device().query("graph_live") == 0
