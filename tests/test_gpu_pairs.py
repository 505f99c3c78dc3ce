"""Tiled Langevin batches stepped as two replica groups on two streams (one
group's corrector beside the other's predictor) equal, bitwise, the same
batch stepped with one stream; the state stays consistent with later work
on the handle's stream (energies, midpoint steps)."""
import numpy as np
import pytest

import sunnyhip as sh

from helpers import announce

pytestmark = pytest.mark.gpu


def batch(precision):
    dims = [(8, 8, 8), (6, 4, 10), (5, 5, 5), (4, 8, 6)]
    systems = [sh.workloads.cubic_model(dims=d, seed=40 + r, precision=precision) for r, d in enumerate(dims)]
    return sh.ReplicaBatch(systems, precision=precision, layout=0)


def state(obj):
    h = obj._device()
    d = np.empty((h.total, 3))
    h.call("snh_get_dipoles", sh._lib.dptr(d))
    return d


@pytest.mark.parametrize("precision", [32, 64])
def test_paired_groups_match_one_stream(precision, monkeypatch):
    announce(f"paired replica groups fp{precision}")
    if precision == 64:
        # the cost model leaves FP64 batches on the ghost layout (too few
        # resident waves); an explicit tile request is honoured
        monkeypatch.setenv("SNH_TILE_T1", "4")
        monkeypatch.setenv("SNH_TILE_T2", "2")
    monkeypatch.setenv("SNH_PAIR", "2")
    a = batch(precision)
    ha = a._device()
    monkeypatch.setenv("SNH_PAIR", "0")
    b = batch(precision)
    hb = b._device()
    assert ha.query("layout") == 0 and hb.query("layout") == 0
    assert ha.query("pair_on") == 1 and hb.query("pair_on") == 0
    kTs = np.linspace(0.2, 0.9, a.nreplicas)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.5)
    for _ in range(3):
        sh.step_(a, integ, nsteps=7, kTs=kTs)
        sh.step_(b, integ, nsteps=7, kTs=kTs)
        np.testing.assert_array_equal(sh.energy(a), sh.energy(b))
    np.testing.assert_array_equal(state(a), state(b))
    # host-supplied noise, one step per call
    rng = np.random.default_rng(1)
    for _ in range(2):
        noise = rng.standard_normal((ha.total, 3))
        sh.step_(a, integ, noise=noise, kTs=kTs)
        sh.step_(b, integ, noise=noise, kTs=kTs)
    np.testing.assert_array_equal(state(a), state(b))
    # a midpoint step on the handle's stream after paired Langevin steps
    sh.step_(a, integ, nsteps=5, kTs=kTs)
    sh.step_(b, integ, nsteps=5, kTs=kTs)
    mid = sh.ImplicitMidpoint(0.02, damping=0.1, tol=1e-6 if precision == 32 else 1e-12)
    sh.step_(a, mid)
    sh.step_(b, mid)
    np.testing.assert_array_equal(state(a), state(b))
