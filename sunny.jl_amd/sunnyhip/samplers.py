"""Metropolis sampling of single spins (dipoles or SU(N) coherent states) on the device.

Host mirror of Sunny's `LocalSampler` and its proposal functions
(src/MonteCarlo/Samplers.jl:1-153) over snh_step_local.  A proposal is a
mixture of up to four device proposals; `propose_uniform`, `propose_flip`,
`propose_delta(magnitude)` and `mix_proposals(w1, p1, w2, p2, ...)` build
them as the reference's functions and `@mix_proposals` macro do.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib

KIND_UNIFORM, KIND_FLIP, KIND_DELTA = 0, 1, 2


class Proposal:
    """Mixture of device proposals: terms [(kind, magnitude, weight)]."""

    def __init__(self, terms, name):
        self.terms = [(int(k), float(m), float(w)) for k, m, w in terms]
        self.name = name

    def __repr__(self):
        return self.name


# randspin (System.jl:444-447)
propose_uniform = Proposal([(KIND_UNIFORM, 0.0, 1.0)], "propose_uniform")
# flip (System.jl:440-442)
propose_flip = Proposal([(KIND_FLIP, 0.0, 1.0)], "propose_flip")


def propose_delta(magnitude):
    """S + |S| magnitude randn, normalized (perturbed_spin, System.jl:454-460);
    an infinite magnitude is propose_uniform."""
    magnitude = float(magnitude)
    if not magnitude >= 0:
        raise ValueError("magnitude must be >= 0")
    return Proposal([(KIND_DELTA, magnitude, 1.0)], f"propose_delta({magnitude})")


def mix_proposals(*terms):
    """@mix_proposals weight1 propose1 weight2 propose2 ... (Samplers.jl:67-100)."""
    if len(terms) % 2:
        raise ValueError("Alternate weights and proposal functions.")
    out = []
    for w, p in zip(terms[0::2], terms[1::2]):
        tot = sum(t[2] for t in p.terms)
        out += [(k, m, float(w) * wi / tot) for k, m, wi in p.terms]
    if len(out) > 4:
        raise ValueError("the device mixes at most four proposals")
    if not sum(t[2] for t in out) > 0:
        raise ValueError("weights must not all be zero")
    name = " ".join(f"{w} {p!r}" for w, p in zip(terms[0::2], terms[1::2]))
    return Proposal(out, f"@mix_proposals {name}")


class LocalSampler:
    """LocalSampler(; kT, nsweeps=1.0, propose=propose_uniform) (Samplers.jl:102-131).

    `dE` and `dS` accumulate the energy and net dipole changes of accepted
    trials (for a ReplicaBatch, per replica)."""

    def __init__(self, *, kT, nsweeps=1.0, propose=propose_uniform):
        self.kT = float(kT)
        self.nsweeps = float(nsweeps)
        if not isinstance(propose, Proposal):
            raise TypeError("propose must be propose_uniform, propose_flip, propose_delta(m) or a mix_proposals")
        self.propose = propose
        self.dE = 0.0
        self.dS = np.zeros(3)

    def copy(self):
        return LocalSampler(kT=self.kT, nsweeps=self.nsweeps, propose=self.propose)

    def __repr__(self):
        return f"LocalSampler(kT={self.kT}, nsweeps={self.nsweeps}, propose={self.propose!r})"


def local_coloring(sys):
    """(m, ncolors, color[na][m3][m2][m1]) of the device's Metropolis sweeps."""
    h = sys._device()
    m = np.zeros(3, np.int32)
    nc = _lib.C.c_int32()
    h.call("snh_local_coloring", _lib.iptr(m), _lib.C.byref(nc), None)
    # sublattices of the device cell (an inhomogeneous system has one per site)
    na = int(h._keep["na"])
    color = np.zeros(na * int(np.prod(m)), np.int32)
    h.call("snh_local_coloring", None, None, _lib.iptr(color))
    return tuple(int(x) for x in m), nc.value, color.reshape(na, m[2], m[1], m[0])


def step_local_(sys, sampler: LocalSampler, *, nsteps=1, kTs=None):
    """step!(sys, sampler) repeated nsteps times (Samplers.jl:133-153) on the
    device; sys is a System or a ReplicaBatch (kTs: per-replica temperatures).
    Returns the [R][4] sums (dE, dSx, dSy, dSz) over accepted trials."""
    h = sys._device()
    R = getattr(sys, "nreplicas", 1)
    kT = np.full(R, sampler.kT) if kTs is None else np.ascontiguousarray(kTs, float).reshape(R)
    kinds = np.array([t[0] for t in sampler.propose.terms], np.int32)
    mags = np.array([t[1] for t in sampler.propose.terms], float)
    wts = np.array([t[2] for t in sampler.propose.terms], float)
    per = max(1, math.ceil(sampler.nsweeps)) if sampler.nsweeps > 0 else 0
    total = np.zeros((R, 4))
    for _ in range(int(nsteps)):
        out = np.zeros((R, 4))
        try:
            h.call("snh_step_local", _lib.dptr(kT), len(kinds), _lib.iptr(kinds), _lib.dptr(mags), _lib.dptr(wts),
                   sys.noise_seed, sys.step_counter, sampler.nsweeps, _lib.dptr(out))
        finally:
            sys._dev_dirty = True
        sys.step_counter += per
        total += out
    if R == 1 and not hasattr(sys, "nreplicas"):
        sampler.dE += float(total[0, 0])
        sampler.dS = sampler.dS + total[0, 1:]
    else:
        if np.ndim(sampler.dE) == 0:
            sampler.dE = np.full(R, float(sampler.dE))
            sampler.dS = np.tile(np.asarray(sampler.dS, float).reshape(-1)[:3], (R, 1))
        sampler.dE = sampler.dE + total[:, 0]
        sampler.dS = sampler.dS + total[:, 1:]
    return total
