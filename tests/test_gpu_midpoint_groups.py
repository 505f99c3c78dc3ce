"""Implicit-midpoint steps queued in groups with a fixed iteration budget and
one host check (undone, retried with a larger budget, then rerun step by step
when a replica needs more iterations) equal, bitwise, steps checked on the
host one by one."""
import numpy as np
import pytest

import sunnyhip as sh

from helpers import announce

pytestmark = pytest.mark.gpu


def twins(kind, precision):
    def one():
        if kind == "cubic":
            return sh.workloads.cubic_model(L=6, seed=3, precision=precision)
        if kind == "pyro_batch":
            systems = [sh.workloads.pyrochlore_model(dims=d, seed=10 + r, precision=precision)
                       for r, d in enumerate([(2, 2, 2), (3, 2, 2)])]
            return sh.ReplicaBatch(systems, precision=precision)
        if kind == "sun3":
            return sh.workloads.triangular_sun_model(L1=5, L2=4, seed=4, precision=precision)
        raise ValueError(kind)
    return one(), one()


def state(obj):
    h = obj._device()
    if obj.N == 0:
        d = np.empty((h.total, 3))
        h.call("snh_get_dipoles", sh._lib.dptr(d))
    else:
        d = np.empty((h.total, obj.N, 2))
        h.call("snh_get_coherents", sh._lib.dptr(d))
    return d


@pytest.mark.parametrize("kind,precision", [("cubic", 64), ("pyro_batch", 32), ("pyro_batch", 64), ("sun3", 64)])
def test_budgeted_groups_match_checked_steps(kind, precision, monkeypatch):
    announce(f"midpoint groups {kind} fp{precision}")
    monkeypatch.setenv("SNH_PERSIST", "0")
    a, b = twins(kind, precision)
    monkeypatch.setenv("SNH_MID_SYNC", "0")
    a._device()
    monkeypatch.setenv("SNH_MID_SYNC", "1")
    b._device()
    tol = 1e-6 if precision == 32 else 1e-12
    # a larger dt in the second call needs more iterations than the budget
    # learned in the first: that group is undone and rerun step by step
    for dt, n in ((0.02, 9), (0.06, 7), (0.06, 5)):
        integ = sh.ImplicitMidpoint(dt, damping=0.1, tol=tol)
        ia = sh.step_(a, integ, nsteps=n)
        ib = sh.step_(b, integ, nsteps=n)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(state(a), state(b))
    assert a._device().query("mid_reruns") >= 1 and b._device().query("mid_reruns") == 0
    assert b._device().query("mid_fallbacks") == 0


def test_budgeted_groups_fall_back_to_checked_steps(monkeypatch):
    """A tenfold dt jump overruns both the learned budget and its retry (two
    more iterations): the group runs with per-step checks, bitwise equal to
    a handle that checks every step."""
    announce("midpoint groups fallback")
    monkeypatch.setenv("SNH_PERSIST", "0")
    a, b = twins("cubic", 64)
    monkeypatch.setenv("SNH_MID_SYNC", "0")
    a._device()
    monkeypatch.setenv("SNH_MID_SYNC", "1")
    b._device()
    for dt, n in ((0.005, 3), (0.05, 4)):
        integ = sh.ImplicitMidpoint(dt, damping=0.0, tol=1e-12)
        ia = sh.step_(a, integ, nsteps=n)
        ib = sh.step_(b, integ, nsteps=n)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(state(a), state(b))
    assert a._device().query("mid_fallbacks") >= 1


def test_budgeted_groups_report_nonconvergence(monkeypatch):
    announce("midpoint groups nonconvergence")
    monkeypatch.setenv("SNH_PERSIST", "0")
    s = sh.workloads.cubic_model(L=6, seed=5, precision=64)
    sh.step_(s, sh.ImplicitMidpoint(0.01, tol=1e-12), nsteps=3)
    S0 = s.dipoles_flat()
    with pytest.raises(sh.ConvergenceError, match="Spherical midpoint method failed to converge"):
        sh.step_(s, sh.ImplicitMidpoint(0.05, tol=1e-12), nsteps=4, max_iters=2)
    # the failing step (the first) left the state where it started
    np.testing.assert_array_equal(s.dipoles_flat(), S0)
