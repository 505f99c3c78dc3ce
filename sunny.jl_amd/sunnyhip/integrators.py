"""`Langevin`, `ImplicitMidpoint`, `step_` (Sunny's `step!`), energies.

`step_` also runs a `LocalSampler` (samplers.py).

Constructors validate exactly as src/Integrators.jl:56-135; `step_` runs the
stochastic Heun or implicit-midpoint step on the device
(src/Integrators.jl:313-452).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._lib import ConvergenceError  # noqa: F401  (re-exported)
from .samplers import LocalSampler, step_local_
from .system import System


class Langevin:
    """Langevin(dt; damping, kT): stochastic Heun with renormalisation."""

    def __init__(self, dt=math.nan, *, damping=None, kT, lam=None):
        if lam is not None:
            damping = damping if damping is not None else lam
        if damping is None:
            raise ValueError("`damping` parameter required")
        if damping == 0:
            raise ValueError("Use ImplicitMidpoint instead for energy-conserving dynamics")
        if dt <= 0:
            raise ValueError("Select positive dt")
        if kT < 0:
            raise ValueError("Select nonnegative kT")
        if damping <= 0:
            raise ValueError("Select positive damping")
        self.dt, self.damping, self.kT = float(dt), float(damping), float(kT)

    def __repr__(self):
        return f"Langevin({self.dt}; damping={self.damping}, kT={self.kT})"


class ImplicitMidpoint:
    """ImplicitMidpoint(dt; damping=0, kT=0, tol=1e-12): spherical (dipole) or
    Schroedinger (SU(N)) midpoint method."""

    def __init__(self, dt=math.nan, *, damping=0.0, kT=0.0, tol=1e-12):
        if dt <= 0:
            raise ValueError("Select positive dt")
        if kT < 0:
            raise ValueError("Select nonnegative kT")
        if damping < 0:
            raise ValueError("Select nonnegative damping")
        if kT != 0:
            raise ValueError("ImplicitMidpoint with a Langevin thermostat is not currently supported.")
        self.dt, self.damping, self.kT, self.tol = float(dt), float(damping), float(kT), float(tol)

    def __repr__(self):
        return f"ImplicitMidpoint({self.dt}; tol={self.tol})"


def _check_dt(integrator):
    if math.isnan(integrator.dt):
        raise ValueError("Set integration timestep `dt`.")


def step_(sys, integrator, *, max_iters=100, noise=None, nsteps=1, kTs=None):
    """Advance `sys` (a System or a ReplicaBatch) by `nsteps` steps.

    noise: raw standard-normal draws for one Langevin step, laid out like the
    state (the draws of randn!(rng, xi) before scaling); None uses the device
    Philox stream.  kTs: per-replica temperatures for a ReplicaBatch.
    Returns the midpoint iteration counts of the last step (ImplicitMidpoint),
    or the [R][4] sums (dE, dS) of accepted trials (LocalSampler)."""
    if isinstance(integrator, LocalSampler):
        return step_local_(sys, integrator, nsteps=nsteps, kTs=kTs)
    _check_dt(integrator)
    h = sys._device()
    R = getattr(sys, "nreplicas", 1)
    if isinstance(integrator, Langevin):
        kT = np.full(R, integrator.kT) if kTs is None else np.ascontiguousarray(kTs, float).reshape(R)
        nz = None
        if noise is not None:
            if nsteps != 1:
                raise ValueError("host-supplied noise covers exactly one step")
            if sys.N > 0 and np.iscomplexobj(noise):
                nz = np.ascontiguousarray(np.stack([noise.real, noise.imag], -1), float)
            else:
                nz = np.ascontiguousarray(noise, float)
            if nz.size != h.total * (3 if sys.N == 0 else 2 * sys.N):
                raise ValueError("noise buffer has the wrong size")
        h.call("snh_step_langevin", integrator.dt, integrator.damping, _lib.dptr(kT), _lib.dptr(nz),
               sys.noise_seed, sys.step_counter, int(nsteps))
        sys.step_counter += int(nsteps)
        sys._dev_dirty = True
        return None
    if isinstance(integrator, ImplicitMidpoint):
        iters = np.zeros(R, np.int32)
        try:
            h.call("snh_step_midpoint", integrator.dt, integrator.damping, integrator.tol, int(max_iters),
                   int(nsteps), _lib.iptr(iters))
        finally:
            sys._dev_dirty = True
        return iters
    raise TypeError(f"unsupported integrator {integrator!r}")


def energy(sys):
    """energy(sys): FP64 total energy (per replica for a ReplicaBatch)."""
    h = sys._device()
    R = getattr(sys, "nreplicas", 1)
    E = np.zeros(R)
    h.call("snh_energy", _lib.dptr(E))
    return E if isinstance(sys, ReplicaBatch) else float(E[0])


def energy_per_site(sys):
    return energy(sys) / sys.nsites


def energy_grad_dipoles(sys: System):
    """dE/dS at every site, flat [site][xyz] (set_energy_grad_dipoles!)."""
    h = sys._device()
    out = np.zeros((h.total, 3))
    h.call("snh_energy_grad", _lib.dptr(out))
    return out


def energy_grad_coherents(sys: System):
    """H Z at every site, flat [site][N] complex (set_energy_grad_coherents!)."""
    h = sys._device()
    out = np.zeros((h.total, sys.N, 2))
    h.call("snh_energy_grad", _lib.dptr(out))
    return out[..., 0] + 1j * out[..., 1]


class ReplicaBatch:
    """Many replicas of one model (same couplings; dims may differ without
    Ewald) advanced together by one set of kernel launches.

    This is the device counterpart of stepping clones of a System under
    `Threads.@threads` (src/MonteCarlo/ParallelTempering.jl:80-95)."""

    def __init__(self, systems, precision=32, layout=0):
        self.systems = list(systems)
        s0 = self.systems[0]
        self.N = s0.N
        self.nreplicas = len(self.systems)
        self.precision = int(precision)
        self.noise_seed = s0.noise_seed
        self.step_counter = 0
        ref = s0.spec(precision=self.precision)
        for s in self.systems[1:]:
            sp = s.spec(precision=self.precision)
            for k in ("stevens", "kmax", "Lambda", "bond", "bilin", "biquad", "kappa", "g", "field"):
                if not np.array_equal(sp[k], ref[k]):
                    raise ValueError("replicas of a ReplicaBatch must share their couplings")
        self._spec = s0.spec(precision=self.precision, dims_list=[s.dims for s in self.systems])
        self._spec["layout"] = int(layout)
        self._handle = None
        self._host_dirty = True
        self._dev_dirty = False
        self.nsites = sum(s.nsites for s in self.systems)

    def _device(self):
        if self._handle is None:
            self._handle = _lib.Handle(self._spec)
        if self._host_dirty:
            if self.N == 0:
                x = np.concatenate([s.dipoles_flat() for s in self.systems])
                self._handle.call("snh_set_dipoles", _lib.dptr(np.ascontiguousarray(x)))
            else:
                z = np.concatenate([s.coherents_flat() for s in self.systems])
                x = np.ascontiguousarray(np.stack([z.real, z.imag], -1))
                self._handle.call("snh_set_coherents", _lib.dptr(x))
            self._host_dirty = False
        return self._handle

    def sync_to_systems(self):
        """Copy device states back into the member Systems."""
        if not self._dev_dirty:
            return
        h = self._handle
        d = np.empty((h.total, 3))
        h.call("snh_get_dipoles", _lib.dptr(d))
        z = None
        if self.N > 0:
            zz = np.empty((h.total, self.N, 2))
            h.call("snh_get_coherents", _lib.dptr(zz))
            z = zz[..., 0] + 1j * zz[..., 1]
        o = 0
        for s in self.systems:
            n = s.nsites
            if self.N == 0:
                s.set_state_flat(dip=d[o:o + n])
            else:
                s.set_state_flat(coh=z[o:o + n])
            o += n
        self._dev_dirty = False

    def synchronize(self):
        if self._handle is not None:
            self._handle.call("snh_synchronize")
