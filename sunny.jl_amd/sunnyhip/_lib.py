"""ctypes binding of libsunnyhip.so (declarations: include/sunny_hip.h).

The library is the only compute path.  Loading fails loudly when it has not
been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import glob
import hashlib
import os

import numpy as np

# SUNNYHIP_LIB selects another build of the library (A/B timing of variants).
LIB_PATH = os.environ.get("SUNNYHIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          "libsunnyhip.so")

SNH_OK = 0
SNH_ERR_INVALID = 1
SNH_ERR_HIP = 2
SNH_ERR_NOCONVERGE = 3
SNH_ERR_UNSUPPORTED = 4

EXPORTS = [
    "snh_create", "snh_destroy", "snh_last_error", "snh_build_id", "snh_num_sites", "snh_replica_sites",
    "snh_set_dipoles", "snh_get_dipoles", "snh_set_coherents", "snh_get_coherents",
    "snh_step_langevin", "snh_step_midpoint", "snh_energy", "snh_energy_grad", "snh_synchronize",
    "snh_enable_timing", "snh_kernel_time", "snh_copy_calibration", "snh_query",
    "snh_midpoint_record", "snh_midpoint_adjoint", "snh_autotune", "snh_corr_init", "snh_corr_add_sample",
    "snh_corr_data", "snh_step_local", "snh_local_coloring", "snh_energy_noalloc",
]


class SnhSpec(C.Structure):
    _fields_ = [
        ("na", C.c_int32), ("N", C.c_int32), ("precision", C.c_int32),
        ("kappa", C.POINTER(C.c_double)), ("g", C.POINTER(C.c_double)), ("field", C.POINTER(C.c_double)),
        ("stevens", C.POINTER(C.c_double)), ("kmax", C.POINTER(C.c_int32)), ("Lambda", C.POINTER(C.c_double)),
        ("nbonds", C.c_int32), ("bond", C.POINTER(C.c_int32)), ("bilin", C.POINTER(C.c_double)),
        ("biquad", C.POINTER(C.c_double)), ("has_biquad", C.POINTER(C.c_int32)), ("scalar", C.POINTER(C.c_double)),
        ("ewald", C.c_int32), ("mu0_muB2", C.c_double), ("demag", C.c_double * 9), ("latvecs", C.c_double * 9),
        ("positions", C.POINTER(C.c_double)), ("nreplicas", C.c_int32), ("dims", C.POINTER(C.c_int32)),
        ("layout", C.c_int32), ("gen_start", C.POINTER(C.c_int32)), ("gen_A", C.POINTER(C.c_double)),
        ("gen_B", C.POINTER(C.c_double)),
    ]


class HipError(RuntimeError):
    pass


class ConvergenceError(RuntimeError):
    pass


_lib = None


def source_build_id():
    """The build id of the sources beside this package, as the Makefile
    computes it (SHA-256 over csrc/*.hip, *.h, *.cpp and ../include/*.h in
    sorted path order, first 16 hex digits); None when they are absent."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rel = [os.path.relpath(f, root) for pat in ("csrc/*.hip", "csrc/*.h", "csrc/*.cpp", "../include/*.h")
           for f in glob.glob(os.path.join(root, pat))]
    if not rel:
        return None
    h = hashlib.sha256()
    for r in sorted(rel):
        with open(os.path.join(root, r), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def lib():
    """Load libsunnyhip.so once; raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libsunnyhip.so not found at {LIB_PATH}; build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.snh_create.argtypes = [C.POINTER(SnhSpec), C.POINTER(vp)]
        L.snh_destroy.argtypes = [vp]
        L.snh_last_error.argtypes = [vp]
        L.snh_last_error.restype = C.c_char_p
        L.snh_num_sites.argtypes = [vp, C.POINTER(C.c_int64)]
        L.snh_replica_sites.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
        for f in ("snh_set_dipoles", "snh_get_dipoles", "snh_set_coherents", "snh_get_coherents",
                  "snh_energy", "snh_energy_grad"):
            getattr(L, f).argtypes = [vp, dp]
        L.snh_energy_noalloc.argtypes = [vp, vp]  # device pointer
        L.snh_step_langevin.argtypes = [vp, C.c_double, C.c_double, dp, dp, C.c_uint64, C.c_uint64, C.c_int32]
        L.snh_step_midpoint.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                        C.POINTER(C.c_int32)]
        L.snh_synchronize.argtypes = [vp]
        L.snh_enable_timing.argtypes = [vp, C.c_int32]
        L.snh_kernel_time.argtypes = [vp, C.c_char_p, dp, C.POINTER(C.c_int64)]
        L.snh_copy_calibration.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
        L.snh_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        L.snh_autotune.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
        L.snh_corr_init.argtypes = [vp, C.c_int32, dp, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                    C.c_int32]
        L.snh_corr_add_sample.argtypes = [vp, C.c_double, C.c_double, C.c_int32]
        L.snh_corr_data.argtypes = [vp, dp, C.POINTER(C.c_int64)]
        ip = C.POINTER(C.c_int32)
        L.snh_step_local.argtypes = [vp, dp, C.c_int32, ip, dp, dp, C.c_uint64, C.c_uint64, C.c_double, dp]
        L.snh_local_coloring.argtypes = [vp, ip, ip, ip]
        L.snh_midpoint_record.argtypes = [vp, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32]
        L.snh_midpoint_adjoint.argtypes = [vp, C.c_int32, C.POINTER(SnhSpec), dp, dp, dp]
        L.snh_build_id.argtypes = []
        L.snh_build_id.restype = C.c_char_p
        for f in EXPORTS:
            if f not in ("snh_last_error", "snh_build_id"):
                getattr(L, f).restype = C.c_int
        # a library built from other sources than those beside it is refused
        # (SUNNYHIP_LIB, for timing variants, skips the check)
        want = source_build_id()
        have = L.snh_build_id().decode()
        if want is not None and not os.environ.get("SUNNYHIP_LIB") and have != want:
            raise ImportError(f"{LIB_PATH} was built from other sources (build id {have}, sources {want}); "
                              "rebuild it (python -c 'import __graft_entry__ as g; g.build()')")
        _lib = L
    return _lib


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def check(rc, handle=None):
    if rc == SNH_OK:
        return
    msg = lib().snh_last_error(handle).decode()
    if rc == SNH_ERR_NOCONVERGE:
        raise ConvergenceError(msg)
    if rc == SNH_ERR_HIP:
        raise HipError(msg)
    raise ValueError(msg)


def pack_spec(spec: dict) -> SnhSpec:
    """snh_spec pointing into the arrays of `spec` (which must outlive it)."""
    s = SnhSpec()
    s.na, s.N, s.precision = spec["na"], spec["N"], spec["precision"]
    s.kappa, s.g, s.field = dptr(spec["kappa"]), dptr(spec["g"]), dptr(spec["field"])
    s.stevens, s.kmax, s.Lambda = dptr(spec["stevens"]), iptr(spec["kmax"]), dptr(spec["Lambda"])
    s.nbonds = spec["nbonds"]
    s.bond, s.bilin, s.biquad = iptr(spec["bond"]), dptr(spec["bilin"]), dptr(spec["biquad"])
    s.has_biquad, s.scalar = iptr(spec["has_biquad"]), dptr(spec["scalar"])
    s.ewald = spec["ewald"]
    s.mu0_muB2 = spec["mu0_muB2"]
    s.demag[:] = list(spec["demag"])
    s.latvecs[:] = list(spec["latvecs"])
    s.positions = dptr(spec["positions"])
    s.nreplicas = spec["nreplicas"]
    s.dims = iptr(spec["dims"])
    s.layout = spec.get("layout", 0)
    s.gen_start = iptr(spec.get("gen_start"))
    s.gen_A, s.gen_B = dptr(spec.get("gen_A")), dptr(spec.get("gen_B"))
    return s


class Handle:
    """Owns one snh_system*; keeps the spec arrays alive for its lifetime."""

    def __init__(self, spec: dict):
        L = lib()
        self._keep = spec
        s = pack_spec(spec)
        h = C.c_void_p()
        rc = L.snh_create(C.byref(s), C.byref(h))
        if rc != SNH_OK:
            check(rc, None)
        self.h = h
        n = C.c_int64()
        check(L.snh_num_sites(self.h, C.byref(n)), self.h)
        self.total = n.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.snh_destroy(self.h)
            self.h = None

    def call(self, name, *args):
        check(getattr(lib(), name)(self.h, *args), self.h)

    def query(self, key):
        """Integer fact about the device plan (snh_query), e.g. "layout"."""
        v = C.c_int64()
        self.call("snh_query", key.encode(), C.byref(v))
        return v.value

    def autotune(self, candidates=8, launches=3):
        """Times the cost model's best tile plans on the device and keeps the
        fastest (snh_autotune); returns its ms per predictor + corrector pair
        (0 when the handle does not use the tiled sweep)."""
        ms = C.c_double()
        self.call("snh_autotune", int(candidates), int(launches), C.byref(ms))
        return ms.value
