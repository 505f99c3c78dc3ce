"""Gradients through deterministic implicit-midpoint trajectories.

Sunny has no reverse rule for `step!`; its fitting code shows the pattern a
pullback follows (src/Fitting.jl:123-152, SystemTangent over the labeled
parameters, src/System/Types.jl:143-146).  `midpoint_gradient` is that
pullback for a trajectory of `step!(sys, ImplicitMidpoint(dt))` (damping 0):
the forward pass runs on the device with checkpoints
(snh_midpoint_record), the loss and its gradient are evaluated on the host,
and the reverse sweep runs on the device (snh_midpoint_adjoint), returning
dL/dparams for the requested labels and dL/dS at the initial state.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .system import param_spec


def structure_factor(S, dims, nq=8):
    """Equal-time structure factor S(q) = (1/Ns) sum_a |sum_r e^{i q.r} S_r^a|^2
    on the nq^3 grid q = 2 pi m / nq (one site per cell, positions = cells;
    every L divisible by nq).  Returns (Sq [nq,nq,nq] indexed [m3,m2,m1], F)."""
    L1, L2, L3 = dims
    if L1 % nq or L2 % nq or L3 % nq:
        raise ValueError("dims must be multiples of nq")
    A = np.asarray(S, float).reshape(L3, L2, L1, 3)
    # e^{i q.r} depends on r mod nq only: fold the lattice onto one nq^3 block
    B = A.reshape(L3 // nq, nq, L2 // nq, nq, L1 // nq, nq, 3).sum(axis=(0, 2, 4))
    F = np.fft.ifftn(B, axes=(0, 1, 2)) * nq ** 3
    Ns = L1 * L2 * L3
    return (np.abs(F) ** 2).sum(-1) / Ns, F


def structure_factor_loss(S, dims, target, nq=8):
    """L = mean_q (S(q) - target(q))^2 and dL/dS ([sites][3], Julia order)."""
    Sq, F = structure_factor(S, dims, nq)
    Rq = Sq - target
    L = float(np.mean(Rq ** 2))
    L1, L2, L3 = dims
    Ns = L1 * L2 * L3
    Nq = nq ** 3
    # dL/dS_r^a = 4/(Nq Ns) Re sum_q R(q) conj(F^a(q)) e^{i q.r}
    G = np.fft.fftn(Rq[..., None] * F.conj(), axes=(0, 1, 2))
    # fftn gives sum_m X[m] e^{-2 pi i m.a/nq}; the sign flips with a -> -a
    idx = [(-np.arange(nq)) % nq] * 3
    G = G[np.ix_(*idx)].real * (4.0 / (Nq * Ns))
    full = np.tile(G, (L3 // nq, L2 // nq, L1 // nq, 1))
    return L, full.reshape(Ns, 3)


def midpoint_gradient(sys, integ, nsteps, labels, loss_fn, checkpoint_every=50, max_iters=100):
    """Runs `nsteps` of step!(sys, integ) (ImplicitMidpoint, damping 0) from the
    current state, evaluates (L, dL/dS_final) = loss_fn(S_final), and returns
    (L, dL/dparams for `labels`, dL/dS_initial).  The system ends at the
    final state."""
    if getattr(integ, "damping", 0.0) != 0.0:
        raise ValueError("the midpoint adjoint needs a deterministic trajectory (damping 0)")
    h = sys._device()
    h.call("snh_midpoint_record", float(integ.dt), float(integ.tol), int(max_iters), int(nsteps),
           int(checkpoint_every))
    sys._dev_dirty = True
    sys.step_counter += nsteps
    S_final = sys.dipoles_flat()
    L, dLdS = loss_fn(S_final)
    specs = [param_spec(sys, lab) for lab in labels]
    packed = (_lib.SnhSpec * max(1, len(specs)))(*[_lib.pack_spec(sp) for sp in specs])
    grad = np.zeros(max(1, len(specs)))
    dLdS0 = np.zeros((sys.nsites, 3))
    h.call("snh_midpoint_adjoint", len(specs), ctypes.cast(packed, ctypes.POINTER(_lib.SnhSpec)),
           _lib.dptr(np.ascontiguousarray(dLdS, float)), _lib.dptr(grad), _lib.dptr(dLdS0))
    return L, grad[:len(specs)], dLdS0
