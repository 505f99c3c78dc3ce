"""snh_energy_noalloc (the entry point that allocates nothing) against
snh_energy on the same handles, through the C ABI: two calls in a row, the
second on a smaller system, so that stale partial sums of the first call in
the shared static scratch would show.  Device buffers for the output come
from the HIP runtime the library itself uses (hipMalloc here, in the test,
outside the entry point)."""
import ctypes as C

import numpy as np
import pytest

import sunnyhip as sh
from sunnyhip import _lib

from helpers import announce

pytestmark = pytest.mark.gpu


def hip():
    _lib.lib()  # libsunnyhip loads the runtime; this returns the same library
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def noalloc_energy(obj):
    """Energies through snh_energy_noalloc into a fresh device buffer."""
    H = hip()
    h = obj._device()
    R = getattr(obj, "nreplicas", 1)
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), 8 * R) == 0
    try:
        assert H.hipMemset(p, 0xFF, 8 * R) == 0  # NaN pattern: every entry must be written
        h.call("snh_energy_noalloc", p)
        h.call("snh_synchronize")
        out = np.empty(R)
        assert H.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8 * R, 2) == 0  # device to host
    finally:
        H.hipFree(p)
    return out


def cases():
    big = sh.ReplicaBatch([sh.workloads.pyrochlore_model(L=4, seed=90 + r, precision=32) for r in range(3)],
                          precision=32)
    small = sh.workloads.cubic_model(dims=(3, 2, 2), seed=93, precision=32)
    yield "fp32 tiled batch -> small cubic", big, small
    big = sh.workloads.pyrochlore_model(L=3, seed=94, precision=64)
    big.layout = 1
    small = sh.workloads.triangular_sun_model(L1=3, L2=2, seed=95, precision=64)
    yield "fp64 ghost pyrochlore -> small SU(3)", big, small
    big = sh.workloads.pyrochlore_model(L=4, J1=0.05, J2=0, A=0, s=3.5, seed=96, dipolar=True, precision=64)
    small = sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=97, dipolar=True, precision=32)
    yield "ewald pyrochlore 4^3 -> 2^3", big, small


def test_noalloc_energy_matches_energy():
    for name, big, small in cases():
        announce(f"noalloc energy: {name}")
        for obj in (big, small):
            E_ref = np.atleast_1d(sh.energy(obj))
            E = noalloc_energy(obj)
            np.testing.assert_array_equal(E, E_ref)
        # the state changes: the static scratch serves the new state too
        sh.step_(big, sh.Langevin(0.02, damping=0.1, kT=0.2), nsteps=3)
        np.testing.assert_array_equal(noalloc_energy(big), np.atleast_1d(sh.energy(big)))


def test_noalloc_energy_rejects_null_output():
    s = sh.workloads.cubic_model(dims=(2, 2, 2), seed=98, precision=64)
    with pytest.raises(ValueError, match="null energy buffer"):
        s._device().call("snh_energy_noalloc", None)


def test_noalloc_energy_refuses_library_fft_grids():
    """An Ewald grid transformed by hipFFT (an axis longer than 32) is refused
    before any launch: the library's plans are not covered by the no-allocation
    promise."""
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=1))], "dipole", dims=(40, 2, 2), precision=64)
    sh.enable_dipole_dipole_(s, 1.0, demag=0.0)
    assert s._device().query("ewald_fft_lib") == 1
    with pytest.raises(ValueError, match="hipFFT"):
        noalloc_energy(s)
