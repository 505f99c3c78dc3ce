"""ctypes front end of the FP64 CPU oracle (oracle/sunny_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.  It takes the same
coupling-table dictionary that the product passes to snh_create
(System.spec()), and host spin arrays in Julia memory order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libsunny_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SoSpec(C.Structure):
    _fields_ = [("L", C.c_int * 3), ("na", C.c_int), ("N", C.c_int), ("kappa", _dp), ("g", _dp), ("field", _dp),
                ("stevens", _dp), ("kmax", _ip), ("Lambda", _dp), ("nb", C.c_int), ("bond", _ip), ("bilin", _dp),
                ("biquad", _dp), ("has_biquad", _ip), ("scalar", _dp), ("ewald_A", _dp), ("gen_start", _ip),
                ("gen_A", _dp), ("gen_B", _dp)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sp = C.POINTER(SoSpec)
        L.so_stevens_energy_grad.argtypes = [_dp, _dp, C.c_int, _dp, _dp]
        L.so_spin_matrices.argtypes = [C.c_int, _dp, _dp, _dp]
        L.so_stevens_quadrupoles.argtypes = [C.c_int, _dp]
        L.so_expected_spin.argtypes = [C.c_int, C.c_long, _dp, _dp]
        L.so_grad_dipoles.argtypes = [sp, _dp, _dp]
        L.so_grad_coherents.argtypes = [sp, _dp, _dp]
        L.so_energy.argtypes = [sp, _dp, _dp]
        L.so_energy.restype = C.c_double
        L.so_step_langevin_dipole.argtypes = [sp, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.so_step_langevin_sun.argtypes = [sp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.so_step_midpoint_dipole.argtypes = [sp, _dp, C.c_double, C.c_double, C.c_double, C.c_int]
        L.so_step_midpoint_sun.argtypes = [sp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int]
        L.so_ewald_precompute.argtypes = [_dp, C.c_int, _dp, C.POINTER(C.c_int), _dp, C.c_double, _dp]
        L.so_ewald_energy.argtypes = [sp, _dp]
        L.so_ewald_energy.restype = C.c_double
        L.so_ewald_grad_bruteforce.argtypes = [sp, _dp, _dp]
        L.so_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.so_philox_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_long, C.c_int, _dp]
        L.so_langevin_run.argtypes = [sp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                      C.c_long, C.c_long, _dp]
        L.so_ewald_precompute_sigma.argtypes = [_dp, C.c_int, _dp, C.POINTER(C.c_int), _dp, C.c_double,
                                                C.c_double, _dp]
        L.so_set_mutation.argtypes = [C.c_int]
        L.so_local_energy_change.argtypes = [sp, _dp, C.c_long, _dp]
        L.so_local_energy_change.restype = C.c_double
        L.so_local_sweeps.argtypes = [sp, _dp, _ip, _ip, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, C.c_uint64,
                                      C.c_uint64, C.c_double, _dp]
        L.so_local_energy_change_sun.argtypes = [sp, _dp, _dp, C.c_long, _dp]
        L.so_local_energy_change_sun.restype = C.c_double
        L.so_local_sweeps_sun.argtypes = [sp, _dp, _dp, _ip, _ip, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double,
                                          C.c_uint64, C.c_uint64, C.c_double, _dp]
        L.so_local_random_sites.argtypes = [sp, _dp, C.c_int, _ip, _dp, _dp, C.c_double, C.c_uint64, C.c_double,
                                            _dp]
        L.so_local_sequential.argtypes = [sp, _dp, C.c_int, _ip, _dp, _dp, C.c_double, C.c_uint64, C.c_uint64,
                                          C.c_double, _dp]
        L.so_local_sequential_sun.argtypes = [sp, _dp, _dp, C.c_int, _ip, _dp, _dp, C.c_double, C.c_uint64,
                                              C.c_uint64, C.c_double, _dp]
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def ewald_A(spec, dims, sigma_div=3.0):
    """Real-space Ewald tensor A[j][i][cell][3][3] (Ewald.jl:61-152), with the
    splitting width sigma = cbrt(V) / sigma_div (the reference's is 3)."""
    na = spec["na"]
    nc = int(np.prod(dims))
    A = np.zeros((na, na, nc, 3, 3))
    L = (C.c_int * 3)(*dims)
    lv = np.ascontiguousarray(spec["latvecs"], float)
    pos = np.ascontiguousarray(spec["positions"], float)
    dm = np.ascontiguousarray(spec["demag"], float)
    if sigma_div == 3.0:
        lib().so_ewald_precompute(_d(lv), na, _d(pos), L, _d(dm), spec["mu0_muB2"], _d(A))
    else:
        lib().so_ewald_precompute_sigma(_d(lv), na, _d(pos), L, _d(dm), spec["mu0_muB2"], float(sigma_div), _d(A))
    return A


class mutation:
    """Context manager that plants one of the oracle's known mistakes (see
    so_set_mutation) for a test showing that a check catches it."""

    NOISE_SCALE, CORRECTOR_NOISE, DAMPING_KAPPA, EWALD_FRACTIONAL, EWALD_SWAPPED = 1, 2, 3, 4, 5

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().so_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().so_set_mutation(0)
        return False


class Oracle:
    """One System (one replica) of a spec dict, with its tables kept alive."""

    def __init__(self, spec, dims=None, A=None):
        self.spec = spec
        self.dims = tuple(spec["dims"][0]) if dims is None else tuple(dims)
        self.N = spec["N"]
        self.ns = int(np.prod(self.dims)) * spec["na"]
        self.A = None
        if spec["ewald"]:
            self.A = ewald_A(spec, self.dims) if A is None else A
        s = SoSpec()
        s.L[:] = list(self.dims)
        s.na, s.N = spec["na"], spec["N"]
        s.kappa, s.g, s.field = _d(spec["kappa"]), _d(spec["g"]), _d(spec["field"])
        s.stevens, s.kmax, s.Lambda = _d(spec["stevens"]), _i(spec["kmax"]), _d(spec["Lambda"])
        s.nb = spec["nbonds"]
        s.bond, s.bilin, s.biquad = _i(spec["bond"]), _d(spec["bilin"]), _d(spec["biquad"])
        s.has_biquad, s.scalar = _i(spec["has_biquad"]), _d(spec["scalar"])
        s.ewald_A = _d(self.A)
        s.gen_start, s.gen_A, s.gen_B = _i(spec.get("gen_start")), _d(spec.get("gen_A")), _d(spec.get("gen_B"))
        self.s = s

    def _p(self):
        return C.byref(self.s)

    def grad(self, S):
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3)
        out = np.zeros((self.ns, 3))
        lib().so_grad_dipoles(self._p(), _d(S), _d(out))
        return out

    def grad_coherents(self, Z):
        Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float)
        out = np.zeros_like(Zr)
        lib().so_grad_coherents(self._p(), _d(Zr), _d(out))
        return out[..., 0] + 1j * out[..., 1]

    def energy(self, S=None, Z=None):
        if Z is not None:
            S = self.expected_spin(Z)
            Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float)
        else:
            Zr = None
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3)
        return lib().so_energy(self._p(), _d(S), _d(Zr))

    def local_energy_change(self, S, site, Snew):
        """local_energy_change(sys, site, state) for a :dipole state."""
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3)
        sn = np.ascontiguousarray(Snew, float).reshape(3)
        return lib().so_local_energy_change(self._p(), _d(S), int(site), _d(sn))

    @staticmethod
    def _props(proposals):
        """[(kind, magnitude, weight), ...] -> kind, mag, cumulative probability arrays."""
        kind = np.array([p[0] for p in proposals], np.int32)
        mag = np.array([p[1] for p in proposals], float)
        w = np.array([p[2] for p in proposals], float)
        return kind, mag, np.cumsum(w / w.sum())

    def local_sweeps(self, S, m, color, ncolors, proposals, kT, key, sweep0, nsweeps):
        """LocalSampler sweeps in colour-class passes; returns (S, [dE, dSx, dSy, dSz])."""
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
        kind, mag, cum = self._props(proposals)
        mm = np.ascontiguousarray(m, np.int32)
        col = np.ascontiguousarray(color, np.int32)
        acc = np.zeros(4)
        rc = lib().so_local_sweeps(self._p(), _d(S), _i(mm), _i(col), int(ncolors), len(kind), _i(kind), _d(mag),
                                   _d(cum), float(kT), int(key) % 2**64, int(sweep0), float(nsweeps), _d(acc))
        if rc:
            raise ValueError("Energy delta for self-interaction not supported")
        return S, acc

    def _ket_arrays(self, Z):
        Zr = np.ascontiguousarray(np.stack([np.asarray(Z).real, np.asarray(Z).imag], -1), float).reshape(
            self.ns, self.N, 2)
        return Zr

    def local_energy_change_sun(self, Z, site, Znew):
        """local_energy_change(sys, site, state) for a :SUN state Z [sites][N]."""
        Zr = self._ket_arrays(Z)
        S = self.expected_spin(Z)
        zn = np.ascontiguousarray(np.stack([np.asarray(Znew).real, np.asarray(Znew).imag], -1), float)
        return lib().so_local_energy_change_sun(self._p(), _d(Zr), _d(S), int(site), _d(zn))

    def local_sweeps_sun(self, Z, m, color, ncolors, proposals, kT, key, sweep0, nsweeps):
        """:SUN LocalSampler passes; returns (Z, [dE, dSx, dSy, dSz])."""
        Zr = self._ket_arrays(Z).copy()
        S = self.expected_spin(Z)
        kind, mag, cum = self._props(proposals)
        mm = np.ascontiguousarray(m, np.int32)
        col = np.ascontiguousarray(color, np.int32)
        acc = np.zeros(4)
        rc = lib().so_local_sweeps_sun(self._p(), _d(Zr), _d(S), _i(mm), _i(col), int(ncolors), len(kind), _i(kind),
                                       _d(mag), _d(cum), float(kT), int(key) % 2**64, int(sweep0), float(nsweeps),
                                       _d(acc))
        if rc:
            raise ValueError("Energy delta for self-interaction not supported")
        return Zr[..., 0] + 1j * Zr[..., 1], acc

    def local_random_sites(self, S, proposals, kT, seed, nsweeps):
        """LocalSampler trials at random sites (the reference's order)."""
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
        kind, mag, cum = self._props(proposals)
        acc = np.zeros(4)
        rc = lib().so_local_random_sites(self._p(), _d(S), len(kind), _i(kind), _d(mag), _d(cum), float(kT),
                                         int(seed), float(nsweeps), _d(acc))
        if rc:
            raise ValueError("Energy delta for self-interaction not supported")
        return S, acc

    def local_sequential(self, S, proposals, kT, key, sweep0, nsweeps):
        """Sequential LocalSampler trials with Philox draws (sunny_oracle.c
        so_local_sequential: the device's form when the Ewald term is on)."""
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
        kind, mag, cum = self._props(proposals)
        acc = np.zeros(4)
        rc = lib().so_local_sequential(self._p(), _d(S), len(kind), _i(kind), _d(mag), _d(cum), float(kT),
                                       int(key) % 2**64, int(sweep0), float(nsweeps), _d(acc))
        if rc:
            raise ValueError("Energy delta for self-interaction not supported")
        return S, acc

    def local_sequential_sun(self, Z, proposals, kT, key, sweep0, nsweeps):
        """:SUN form of local_sequential; returns (Z, [dE, dSx, dSy, dSz])."""
        Zr = self._ket_arrays(Z).copy()
        S = self.expected_spin(Z)
        kind, mag, cum = self._props(proposals)
        acc = np.zeros(4)
        rc = lib().so_local_sequential_sun(self._p(), _d(Zr), _d(S), len(kind), _i(kind), _d(mag), _d(cum),
                                           float(kT), int(key) % 2**64, int(sweep0), float(nsweeps), _d(acc))
        if rc:
            raise ValueError("Energy delta for self-interaction not supported")
        return Zr[..., 0] + 1j * Zr[..., 1], acc

    def expected_spin(self, Z):
        Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float)
        S = np.zeros((self.ns, 3))
        lib().so_expected_spin(self.N, self.ns, _d(Zr), _d(S))
        return S

    def step_langevin(self, S=None, Z=None, noise=None, dt=0.05, damping=0.1, kT=0.0):
        if self.N == 0:
            S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
            nz = None if noise is None else np.ascontiguousarray(noise, float)
            lib().so_step_langevin_dipole(self._p(), _d(S), _d(nz), dt, damping, kT)
            return S
        Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float).copy()
        Sd = np.zeros((self.ns, 3))
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(np.stack([noise.real, noise.imag], -1), float)
        lib().so_step_langevin_sun(self._p(), _d(Zr), _d(Sd), _d(nz), dt, damping, kT)
        return Zr[..., 0] + 1j * Zr[..., 1]

    def step_midpoint(self, S=None, Z=None, dt=0.05, damping=0.0, tol=1e-12, max_iters=100):
        if self.N == 0:
            S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
            it = lib().so_step_midpoint_dipole(self._p(), _d(S), dt, damping, tol, max_iters)
            return S, it
        Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float).copy()
        Sd = np.zeros((self.ns, 3))
        it = lib().so_step_midpoint_sun(self._p(), _d(Zr), _d(Sd), dt, damping, tol, max_iters)
        return Zr[..., 0] + 1j * Zr[..., 1], it

    def langevin_run(self, S=None, Z=None, dt=0.02, damping=0.1, kT=0.1, key=0, step0=0, nsteps=1, every=1):
        """nsteps Langevin steps with the device's Philox noise stream; returns
        the final state and the energies after every `every` steps."""
        E = np.zeros(max(nsteps // every, 0) if every > 0 else 0)
        if self.N == 0:
            S = np.ascontiguousarray(S, float).reshape(self.ns, 3).copy()
            lib().so_langevin_run(self._p(), _d(S), None, dt, damping, kT, key, step0, nsteps, every, _d(E))
            return S, E
        Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float).copy()
        Sd = self.expected_spin(Z)
        lib().so_langevin_run(self._p(), _d(Sd), _d(Zr), dt, damping, kT, key, step0, nsteps, every, _d(E))
        return Zr[..., 0] + 1j * Zr[..., 1], E

    def ewald_energy(self, S):
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3)
        return lib().so_ewald_energy(self._p(), _d(S))

    def ewald_grad_bruteforce(self, S):
        S = np.ascontiguousarray(S, float).reshape(self.ns, 3)
        out = np.zeros((self.ns, 3))
        lib().so_ewald_grad_bruteforce(self._p(), _d(S), _d(out))
        return out


def expected_spin(N, Z):
    """expected_spin of every row of Z ([sites][N] complex)."""
    Z = np.atleast_2d(Z)
    Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], -1), float)
    S = np.zeros((Z.shape[0], 3))
    lib().so_expected_spin(N, Z.shape[0], _d(Zr), _d(S))
    return S


def stevens_energy_grad(S, c, kmax):
    S = np.ascontiguousarray(S, float)
    c = np.ascontiguousarray(c, float)
    E = C.c_double()
    g = np.zeros(3)
    lib().so_stevens_energy_grad(_d(S), _d(c), int(kmax), C.byref(E), _d(g))
    return E.value, g


def spin_matrices(N):
    out = [np.zeros((N, N, 2)) for _ in range(3)]
    lib().so_spin_matrices(N, *[_d(o) for o in out])
    return [o[..., 0] + 1j * o[..., 1] for o in out]


def stevens_quadrupoles(N):
    out = np.zeros((5, N, N, 2))
    lib().so_stevens_quadrupoles(N, _d(out))
    return out[..., 0] + 1j * out[..., 1]


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().so_philox4x32(c, k, o)
    return list(o)


def philox_normals(seed, step, nsites, per_site):
    out = np.zeros((nsites, per_site))
    lib().so_philox_normals(seed, step, nsites, per_site, _d(out))
    return out


def replica_key(seed, r):
    """Per-replica Philox key used by the device noise (seed + r * golden)."""
    return (seed + r * 0x9E3779B97F4A7C15) % (1 << 64)
