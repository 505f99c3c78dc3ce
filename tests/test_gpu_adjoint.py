"""GPU parity of the midpoint adjoint (snh_midpoint_record / _adjoint).

Parameter gradients and the initial-state gradient of a structure-factor
loss after a deterministic ImplicitMidpoint trajectory (J, D at 0.9, -0.08,
target S(q) from 1.0, -0.1) are checked
 - against forward-mode dual numbers through the CPU restatement on an 8^3
   copy over the full 500-step horizon with checkpoints every 50 (and on a
   short run with an odd checkpoint interval), and
 - against central finite differences (h = 1e-4, bar 1e-4 relative) of the
   device forward trajectory on the 64^3 lattice, over 100 steps.
From random initial spins the dynamics is chaotic: the exact derivative of
the 500-step map grows to ~1e8, far outside the linear range of a 1e-4
finite difference, so the 64^3 difference check stops at 100 steps (at 200
steps h = 1e-4 already departs from h = 1e-6 by 13%)."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle.adjoint_ref import MidpointAD

from helpers import announce

pytestmark = pytest.mark.gpu


def _loss_fn(dims, target, nq):
    return lambda S: sh.structure_factor_loss(S, dims, target, nq)


def test_adjoint_matches_forward_mode():
    L, nq, nsteps, dt = 8, 4, 30, 0.05
    ref = sh.workloads.cubic_model(L=L, J=1.0, D=-0.1, seed=7, params=True)
    S0 = ref.dipoles_flat()
    sh.step_(ref, sh.ImplicitMidpoint(dt), nsteps=nsteps)
    target, _ = sh.structure_factor(ref.dipoles_flat(), (L, L, L), nq)
    s = sh.workloads.cubic_model(L=L, J=0.9, D=-0.08, seed=7, params=True)
    announce(f"adjoint cubic {L}^3, {nsteps} steps, checkpoints every 7")
    loss, grad, dS0 = sh.midpoint_gradient(s, sh.ImplicitMidpoint(dt), nsteps, ["J", "D"],
                                           _loss_fn((L, L, L), target, nq), checkpoint_every=7)
    v = np.random.default_rng(3).standard_normal(S0.shape)
    ad = MidpointAD(s.spec(), [sh.param_spec(s, "J"), sh.param_spec(s, "D")])
    Sd = ad.trajectory(S0, dt, nsteps, dS0=v)
    L_ref, g_final = sh.structure_factor_loss(Sd.v, (L, L, L), target, nq)
    assert abs(loss - L_ref) <= 1e-12 * max(1.0, abs(L_ref))
    # chain rule through the dual tangents: dL/dtheta = g_final . dS_final/dtheta
    fwd = np.einsum("sa,sap->p", g_final, Sd.d)
    np.testing.assert_allclose(grad, fwd[:2], rtol=1e-9, atol=1e-14)
    assert abs((dS0 * v).sum() - fwd[2]) <= 1e-9 * max(1e-12, abs(fwd[2]))
    # the handle ends at the final state
    np.testing.assert_allclose(s.dipoles_flat(), Sd.v, atol=1e-12)


def test_adjoint_with_ewald_matches_forward_mode():
    """The Ewald term in the adjoint (the field at Shat and the Ewald part of
    M w, both from the device FFT potential) against forward-mode duals
    through the dual restatement (whose Ewald operator is the oracle's
    pairwise field), on a dipolar cubic 3x2x2 lattice."""
    from test_oracle_adjoint import _dipolar
    nsteps, dt = 12, 0.05
    s = _dipolar(0.9, -0.08, seed=8)
    s.layout = 1
    S0 = s.dipoles_flat()
    v = np.random.default_rng(5).standard_normal(S0.shape)
    announce(f"adjoint dipolar cubic 3x2x2 (Ewald), {nsteps} steps, checkpoints every 5")
    loss, grad, dS0 = sh.midpoint_gradient(s, sh.ImplicitMidpoint(dt), nsteps, ["J", "D"],
                                           lambda S: (float((S * v).sum()), v), checkpoint_every=5)
    u = np.random.default_rng(6).standard_normal(S0.shape)
    ad = MidpointAD(s.spec(), [sh.param_spec(s, "J"), sh.param_spec(s, "D")])
    Sd = ad.trajectory(S0, dt, nsteps, dS0=u)
    assert abs(loss - float((Sd.v * v).sum())) <= 1e-12 * max(1.0, abs(loss))
    fwd = np.einsum("sa,sap->p", v, Sd.d)
    np.testing.assert_allclose(grad, fwd[:2], rtol=1e-9, atol=1e-13)
    assert abs((dS0 * u).sum() - fwd[2]) <= 1e-9 * max(1e-12, abs(fwd[2]))
    np.testing.assert_allclose(s.dipoles_flat(), Sd.v, atol=1e-12)


def test_adjoint_nonlinear_terms_match_forward_mode():
    """Biquadratic bonds and Stevens anisotropy of order 4 and 6 in the
    adjoint (their derivative along w by dual numbers on the device) against
    forward-mode duals of the same trajectory, with gradients along the
    labels J, D and B (a biquadratic coupling)."""
    from test_oracle_adjoint import _nonlinear
    nsteps, dt = 10, 0.05
    s = _nonlinear(0.9, -0.08, 0.05, seed=12)
    s.layout = 1
    S0 = s.dipoles_flat()
    v = np.random.default_rng(7).standard_normal(S0.shape)
    announce(f"adjoint cubic 3x2x2 biquadratic + Stevens 4/6, {nsteps} steps, checkpoints every 4")
    loss, grad, dS0 = sh.midpoint_gradient(s, sh.ImplicitMidpoint(dt), nsteps, ["J", "D", "B"],
                                           lambda S: (float((S * v).sum()), v), checkpoint_every=4)
    u = np.random.default_rng(8).standard_normal(S0.shape)
    ad = MidpointAD(s.spec(), [sh.param_spec(s, k) for k in ("J", "D", "B")])
    Sd = ad.trajectory(S0, dt, nsteps, dS0=u)
    assert abs(loss - float((Sd.v * v).sum())) <= 1e-12 * max(1.0, abs(loss))
    fwd = np.einsum("sa,sap->p", v, Sd.d)
    np.testing.assert_allclose(grad, fwd[:3], rtol=1e-8, atol=1e-12)
    assert abs((dS0 * u).sum() - fwd[3]) <= 1e-8 * max(1e-12, abs(fwd[3]))


def test_adjoint_500_steps_forward_mode_8cubed():
    L, nq, nsteps, dt = 8, 4, 500, 0.05
    ref = sh.workloads.cubic_model(L=L, J=1.0, D=-0.1, seed=5, params=True)
    S0 = ref.dipoles_flat()
    sh.step_(ref, sh.ImplicitMidpoint(dt), nsteps=nsteps)
    target, _ = sh.structure_factor(ref.dipoles_flat(), (L, L, L), nq)
    s = sh.workloads.cubic_model(L=L, J=0.9, D=-0.08, seed=5, params=True)
    announce(f"adjoint cubic {L}^3, {nsteps} steps, checkpoints every 50")
    loss, grad, _ = sh.midpoint_gradient(s, sh.ImplicitMidpoint(dt), nsteps, ["J", "D"],
                                         _loss_fn((L, L, L), target, nq), checkpoint_every=50)
    ad = MidpointAD(s.spec(), [sh.param_spec(s, "J"), sh.param_spec(s, "D")])
    Sd = ad.trajectory(S0, dt, nsteps)
    _, g_final = sh.structure_factor_loss(Sd.v, (L, L, L), target, nq)
    fwd = np.einsum("sa,sap->p", g_final, Sd.d)
    announce(f"loss {loss:.6e} adjoint {grad} forward-mode {fwd}")
    np.testing.assert_allclose(grad, fwd, rtol=1e-4)


def test_adjoint_cubic64_finite_differences():
    L, nq, nsteps, dt = 64, 8, 100, 0.05
    dims = (L, L, L)
    ref = sh.workloads.cubic_model(L=L, J=1.0, D=-0.1, seed=11, params=True)
    sh.step_(ref, sh.ImplicitMidpoint(dt), nsteps=nsteps)
    target, _ = sh.structure_factor(ref.dipoles_flat(), dims, nq)
    announce(f"adjoint cubic 64^3, {nsteps} steps, checkpoints every 50")
    s = sh.workloads.cubic_model(L=L, J=0.9, D=-0.08, seed=11, params=True)
    loss, grad, _ = sh.midpoint_gradient(s, sh.ImplicitMidpoint(dt), nsteps, ["J", "D"],
                                         _loss_fn(dims, target, nq), checkpoint_every=50)

    def loss_at(J, D):
        t = sh.workloads.cubic_model(L=L, J=J, D=D, seed=11, params=True)
        sh.step_(t, sh.ImplicitMidpoint(dt), nsteps=nsteps)
        return sh.structure_factor_loss(t.dipoles_flat(), dims, target, nq)[0]
    h = 1e-4
    fd = np.array([(loss_at(0.9 + h, -0.08) - loss_at(0.9 - h, -0.08)) / (2 * h),
                   (loss_at(0.9, -0.08 + h) - loss_at(0.9, -0.08 - h)) / (2 * h)])
    announce(f"loss {loss:.6e} grad {grad} fd {fd}")
    np.testing.assert_allclose(grad, fd, rtol=1e-4)
