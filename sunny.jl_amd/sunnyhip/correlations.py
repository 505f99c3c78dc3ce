"""Sampled dynamical correlations on the device.

Host mirror of Sunny's `SampledCorrelations` / `add_sample!`
(src/SampledCorrelations/SampledCorrelations.jl, CorrelationSampling.jl)
for one System (:dipole, or :SUN through the expected spins): the
trajectory, the lattice and time transforms,
and the running mean of every correlation run on the GPU
(snh_corr_init / snh_corr_add_sample / snh_corr_data).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .system import System

# ssf_custom's correlation pairs (MeasureSpec.jl:117), 0-based
SSF_CORR_PAIRS = [(2, 2), (1, 2), (0, 2), (1, 1), (0, 1), (0, 0)]


def adjusted_dt_and_downsampling_factor(dt, n_omega, omega_max):
    """(dt, measperiod) such that the sampled trajectory resolves omega_max
    with n_omega non-negative energies exactly (SampledCorrelations.jl:152-173)."""
    if not math.pi / dt > omega_max:
        raise ValueError("Desired `ωmax` not possible with specified `dt`. Choose smaller `dt` value.")
    n_all = 2 * (int(n_omega) - 1)
    d_omega = omega_max / (n_omega - 1)
    measperiod = math.ceil(math.pi / (dt * omega_max))
    return 2 * math.pi / (d_omega * measperiod * n_all), measperiod


def dipole_observables(sys: System, apply_g=True):
    """all_dipole_observables(sys; apply_g) (MeasureSpec.jl:53-64) for a
    homogeneous system: [3][na][3], observable a of atom i is row a of g_i
    (or of the identity)."""
    na = sys.crystal.natoms
    obs = np.zeros((3, na, 3))
    for i in range(na):
        g = np.asarray(sys.gs[i], float).reshape(3, 3) if apply_g else np.eye(3)
        for a in range(3):
            obs[a, i] = g[a]
    return obs


class SampledCorrelations:
    """SampledCorrelations(sys; energies, dt, measure=ssf observables)
    (SampledCorrelations.jl:194-268), accumulating on the device of `sys`.
    `observables` [nobs][na][3] and `corr_pairs` default to the spin
    structure factor's (ssf_custom: rows of g, six pairs).  The trajectory
    integrator is ImplicitMidpoint(dt) with damping 0 and tolerance `tol`."""

    def __init__(self, sys: System, *, energies, dt, observables=None, corr_pairs=None, window="cosine",
                 tol=1e-12, apply_g=True):
        energies = np.asarray(energies, float)
        n_omega = len(energies)
        omega_max = energies[-1]
        if not (energies[0] == 0 and omega_max > 0):
            raise ValueError("`energies` must be a range from 0 to a positive value")
        d = np.diff(energies)
        if not np.allclose(d, d[0]):
            raise ValueError("`energies` must be equally spaced.")
        if window not in ("cosine", "rectangular"):
            raise ValueError("window must be :cosine or :rectangular")
        self.dt, self.measperiod = adjusted_dt_and_downsampling_factor(dt, n_omega, omega_max)
        self.d_omega = omega_max / (n_omega - 1)
        self.n_all_omega = 2 * (n_omega - 1)
        self.observables = np.ascontiguousarray(dipole_observables(sys, apply_g) if observables is None
                                                else np.asarray(observables, float))
        self.corr_pairs = list(SSF_CORR_PAIRS if corr_pairs is None else corr_pairs)
        self.window = window
        self.tol = tol
        if getattr(sys, "inhomog", None) is not None:
            # the device holds an inhomogeneous system as one supercell-sized
            # cell, which has no lattice to transform over
            raise NotImplementedError("sampled correlations of an inhomogeneous system are not supported here")
        self.dims = tuple(int(x) for x in sys.dims)
        self.na = sys.crystal.natoms
        self.sys = sys
        h = sys._device()
        pairs = np.ascontiguousarray(np.asarray(self.corr_pairs, np.int32).reshape(-1))
        h.call("snh_corr_init", self.observables.shape[0], _lib.dptr(self.observables.reshape(-1)),
               len(self.corr_pairs), _lib.iptr(pairs), self.n_all_omega, self.measperiod,
               1 if window == "cosine" else 0)

    @property
    def nsamples(self):
        n = _lib.C.c_int64()
        self.sys._device().call("snh_corr_data", None, _lib.C.byref(n))
        return n.value

    @property
    def data(self):
        """sc.data as complex [ncorr, na, na, L1, L2, L3, n_all_omega]."""
        nc, na = len(self.corr_pairs), self.na
        L1, L2, L3 = self.dims
        buf = np.zeros(2 * nc * na * na * L1 * L2 * L3 * self.n_all_omega)
        self.sys._device().call("snh_corr_data", _lib.dptr(buf), None)
        z = buf[0::2] + 1j * buf[1::2]
        # column-major (Julia) order: reverse the axes of a C-order view
        return z.reshape(self.n_all_omega, L3, L2, L1, na, na, nc).transpose(6, 5, 4, 3, 2, 1, 0)

    def omegas(self):
        """Energies of the n_all_omega slots: fftfreq(n, n dω) (CorrelationUtils.jl:34)."""
        n = self.n_all_omega
        return np.fft.fftfreq(n, 1.0 / (n * self.d_omega))


def add_sample_(sc: SampledCorrelations, sys: System, max_iters=100):
    """add_sample!(sc, sys) (CorrelationSampling.jl:154-175): integrates a
    trajectory from the current state of `sys` (which ends at its last
    state) and accumulates its correlations."""
    if sys is not sc.sys:
        raise ValueError("the correlations accumulate on the device of the System they were created for")
    h = sys._device()
    h.call("snh_corr_add_sample", float(sc.dt), float(sc.tol), int(max_iters))
    sys._dev_dirty = True
    sys.step_counter += sc.measperiod * (sc.n_all_omega // 2)
