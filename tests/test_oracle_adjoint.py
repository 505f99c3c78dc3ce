"""Pins the checkers of the midpoint adjoint (CPU only): the forward-mode
dual-number restatement against the C oracle and central differences, the
structure-factor loss and its gradient, and the labeled-parameter API
(couplings linear in the parameter values, src/System/Interactions.jl:1-42)."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O
from oracle.adjoint_ref import MidpointAD


def _run_oracle(J, D, S0, nsteps, dt=0.05, L=4, seed=3):
    s = sh.workloads.cubic_model(L=L, J=J, D=D, seed=seed, params=True)
    o = O.Oracle(s.spec())
    S = S0.copy()
    for _ in range(nsteps):
        S, it = o.step_midpoint(S=S, dt=dt, damping=0.0, tol=1e-13)
        assert it > 0
    return S


def test_dual_restatement_matches_oracle_and_fd():
    s = sh.workloads.cubic_model(L=4, J=0.9, D=-0.08, seed=3, params=True)
    S0 = s.dipoles_flat()
    ad = MidpointAD(s.spec(), [sh.param_spec(s, "J"), sh.param_spec(s, "D")])
    Sd = ad.trajectory(S0, 0.05, 20)
    np.testing.assert_allclose(Sd.v, _run_oracle(0.9, -0.08, S0, 20), atol=1e-13)
    h = 1e-5
    fdJ = (_run_oracle(0.9 + h, -0.08, S0, 20) - _run_oracle(0.9 - h, -0.08, S0, 20)) / (2 * h)
    fdD = (_run_oracle(0.9, -0.08 + h, S0, 20) - _run_oracle(0.9, -0.08 - h, S0, 20)) / (2 * h)
    assert np.abs(Sd.d[..., 0] - fdJ).max() <= 1e-7 * np.abs(fdJ).max()
    assert np.abs(Sd.d[..., 1] - fdD).max() <= 1e-7 * np.abs(fdD).max()


def test_dual_initial_state_direction():
    s = sh.workloads.cubic_model(L=4, J=0.9, D=-0.08, seed=4, params=True)
    S0 = s.dipoles_flat()
    v = np.random.default_rng(1).standard_normal(S0.shape)
    v -= (v * S0).sum(1, keepdims=True) * S0  # tangent to the spheres
    ad = MidpointAD(s.spec(), [])
    Sd = ad.trajectory(S0, 0.05, 10, dS0=v)
    h = 1e-6
    o = O.Oracle(s.spec())

    def run(S):
        for _ in range(10):
            S, _ = o.step_midpoint(S=S, dt=0.05, damping=0.0, tol=1e-13)
        return S
    fd = (run(S0 + h * v) - run(S0 - h * v)) / (2 * h)
    assert np.abs(Sd.d[..., -1] - fd).max() <= 1e-6 * np.abs(fd).max()


def _dipolar(J, D, seed=6):
    s = sh.workloads.cubic_model(dims=(3, 2, 2), J=J, D=D, seed=seed, params=True)
    sh.enable_dipole_dipole_(s, 0.4, demag=0.2)
    return s


def test_dual_restatement_with_ewald():
    """The Ewald field enters the dual restatement as a dense linear operator:
    values equal the C oracle's midpoint steps (which run its Ewald sum), and
    parameter tangents equal central differences."""
    s = _dipolar(0.9, -0.08)
    S0 = s.dipoles_flat()
    ad = MidpointAD(s.spec(), [sh.param_spec(s, "J"), sh.param_spec(s, "D")])
    Sd = ad.trajectory(S0, 0.05, 12)

    def run(J, D):
        o = O.Oracle(_dipolar(J, D).spec())
        S = S0.copy()
        for _ in range(12):
            S, _ = o.step_midpoint(S=S, dt=0.05, damping=0.0, tol=1e-13)
        return S
    np.testing.assert_allclose(Sd.v, run(0.9, -0.08), atol=1e-13)
    h = 1e-5
    fdJ = (run(0.9 + h, -0.08) - run(0.9 - h, -0.08)) / (2 * h)
    fdD = (run(0.9, -0.08 + h) - run(0.9, -0.08 - h)) / (2 * h)
    assert np.abs(Sd.d[..., 0] - fdJ).max() <= 1e-7 * np.abs(fdJ).max()
    assert np.abs(Sd.d[..., 1] - fdD).max() <= 1e-7 * np.abs(fdD).max()


def _nonlinear(J, D, B, seed=9):
    """Cubic 3x2x2 with J1 (label J), a biquadratic J1 bond (label B), an easy
    axis D Sz^2 (label D) and fixed Stevens terms of order 4 and 6."""
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=3, g=2))], "dipole", dims=(3, 2, 2), seed=seed)
    sh.set_exchange_(s, 1.0, sh.Bond(0, 0, (1, 0, 0)), param=("J", J))
    sh.set_exchange_(s, 0.0, sh.Bond(0, 0, (0, 1, 0)), biquad=1.0, param=("B", B))
    O4 = sh.stevens_matrices(3.0)
    sh.set_onsite_coupling_(s, lambda S: S[2] @ S[2], 0, param=("D", D))
    sh.set_onsite_coupling_(s, 0.01 * O4[4][0] + 0.004 * O4[6][3], 0)
    sh.randomize_spins_(s)
    return s


def test_dual_restatement_nonlinear_terms():
    """Biquadratic bonds and Stevens order 4 and 6 enter the dual restatement
    through the oracle's gradient: values equal the oracle's midpoint steps,
    tangents along J, D, B and the initial state equal central differences."""
    s = _nonlinear(0.9, -0.08, 0.05)
    S0 = s.dipoles_flat()
    u = np.random.default_rng(2).standard_normal(S0.shape)
    u -= (u * S0).sum(1, keepdims=True) * S0 / 9  # tangent to the spheres (|S| = 3)
    ad = MidpointAD(s.spec(), [sh.param_spec(s, k) for k in ("J", "D", "B")])
    Sd = ad.trajectory(S0, 0.05, 8, dS0=u)

    def run(J, D, B, S=S0):
        o = O.Oracle(_nonlinear(J, D, B).spec())
        for _ in range(8):
            S, _ = o.step_midpoint(S=S, dt=0.05, damping=0.0, tol=1e-13)
        return S
    np.testing.assert_allclose(Sd.v, run(0.9, -0.08, 0.05), atol=1e-13)
    h = 1e-5
    for t, (dJ, dD, dB) in enumerate([(h, 0, 0), (0, h, 0), (0, 0, h)]):
        fd = (run(0.9 + dJ, -0.08 + dD, 0.05 + dB) - run(0.9 - dJ, -0.08 - dD, 0.05 - dB)) / (2 * h)
        assert np.abs(Sd.d[..., t] - fd).max() <= 1e-6 * np.abs(fd).max(), t
    fd = (run(0.9, -0.08, 0.05, S0 + 1e-6 * u) - run(0.9, -0.08, 0.05, S0 - 1e-6 * u)) / 2e-6
    assert np.abs(Sd.d[..., 3] - fd).max() <= 1e-6 * np.abs(fd).max()


def test_structure_factor_direct_sum():
    dims = (8, 4, 4)
    rng = np.random.default_rng(0)
    S = rng.standard_normal((128, 3))
    Sq, _ = sh.structure_factor(S, dims, nq=4)
    c = np.arange(128)
    r = np.stack([c % 8, (c // 8) % 4, c // 32], 1)
    for m in [(0, 0, 0), (1, 2, 3), (3, 0, 1)]:
        q = 2 * np.pi * np.array(m) / 4
        F = (np.exp(1j * r @ q)[:, None] * S).sum(0)
        ref = (np.abs(F) ** 2).sum() / 128
        assert abs(Sq[m[2], m[1], m[0]] - ref) < 1e-10


def test_structure_factor_loss_gradient():
    dims = (8, 8, 8)
    rng = np.random.default_rng(2)
    S = rng.standard_normal((512, 3))
    T = rng.random((4, 4, 4))
    L, g = sh.structure_factor_loss(S, dims, T, nq=4)
    for _ in range(4):
        v = rng.standard_normal(S.shape)
        h = 1e-6
        fd = (sh.structure_factor_loss(S + h * v, dims, T, nq=4)[0] -
              sh.structure_factor_loss(S - h * v, dims, T, nq=4)[0]) / (2 * h)
        assert abs((g * v).sum() - fd) <= 1e-7 * max(1.0, abs(fd))


def test_labeled_parameters_are_linear():
    s = sh.workloads.cubic_model(L=2, J=0.9, D=-0.08, seed=1, params=True)
    assert sh.get_params(s, ["J", "D"]) == [0.9, -0.08]
    b = sh.Bond(0, 0, (1, 0, 0))
    np.testing.assert_allclose(s.pairs[b][1], 0.9 * np.eye(3))
    c_D = s.onsite[0].packed().copy()
    sh.set_params_(s, ["J", "D"], [1.0, -0.1])
    np.testing.assert_allclose(s.pairs[b][1], np.eye(3))
    np.testing.assert_allclose(s.onsite[0].packed(), c_D * (-0.1 / -0.08), atol=1e-15)
    # the unit spec is the derivative of the couplings
    uJ, uD = sh.param_spec(s, "J"), sh.param_spec(s, "D")
    full = s.spec()
    np.testing.assert_allclose(full["bilin"], 1.0 * uJ["bilin"] - 0.1 * uD["bilin"], atol=1e-15)
    np.testing.assert_allclose(full["stevens"], 1.0 * uJ["stevens"] - 0.1 * uD["stevens"], atol=1e-15)
    with pytest.raises(KeyError):
        sh.set_params_(s, ["K"], [1.0])
