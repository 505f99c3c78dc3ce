"""Host-side setup: lattices, bonds, integrator validation, C-ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import sunnyhip as sh
from sunnyhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pyrochlore_shells():
    a = 10.0
    c = sh.pyrochlore_crystal(a)
    assert c.natoms == 16
    for i in range(16):
        d = sorted(c.distance(b) for b in c.bonds_within(a * 0.62, 1) if b.i == i)
        assert np.allclose(d[:6], a * np.sqrt(2) / 4) and np.allclose(d[6:18], a * np.sqrt(6) / 4)
        assert len(d) == 18
    axes = sh.pyrochlore_local_axes(c)
    for n in axes:
        assert np.allclose(np.sort(np.abs(n)), np.full(3, 1 / np.sqrt(3)))
    s = sh.workloads.pyrochlore_model(L=1)
    assert s.spec()["nbonds"] == 16 * 18


def test_triangular_bonds_by_vector():
    """J1, J2, J3 select 6 bonds each; the c-axis bond (6.75 A, between J1
    and J2 in length) is not picked up."""
    s = sh.workloads.triangular_sun_model(L1=4, L2=4)
    sp = s.spec()
    assert sp["nbonds"] == 18
    lengths = sorted({round(s.crystal.distance(b), 6) for b in s.pairs})
    assert lengths == [4.0, round(4 * np.sqrt(3), 6), 8.0]


def test_reverse_bonds_carry_transpose():
    s = sh.System(sh.diamond_crystal(), [(i, sh.Moment(s=1, g=2)) for i in range(8)], "dipole")
    J = np.arange(9.0).reshape(3, 3)
    b = sh.Bond(0, 1, (0, 0, 0))
    sh.set_exchange_(s, J, b)
    np.testing.assert_array_equal(s.pairs[b][1], J)
    np.testing.assert_array_equal(s.pairs[b.reverse()][1], J.T)
    assert sh.bond_parity(b) != sh.bond_parity(b.reverse())


def test_integrator_validation_messages():
    with pytest.raises(ValueError, match="`damping` parameter required"):
        sh.Langevin(0.1, kT=0.1)
    with pytest.raises(ValueError, match="Use ImplicitMidpoint instead"):
        sh.Langevin(0.1, damping=0, kT=0.1)
    with pytest.raises(ValueError, match="Select positive dt"):
        sh.Langevin(-1, damping=0.1, kT=0.1)
    with pytest.raises(ValueError, match="Select nonnegative kT"):
        sh.Langevin(0.1, damping=0.1, kT=-1)
    with pytest.raises(ValueError, match="not currently supported"):
        sh.ImplicitMidpoint(0.1, kT=0.5)
    with pytest.raises(ValueError, match="Select nonnegative damping"):
        sh.ImplicitMidpoint(0.1, damping=-1)


def test_dipoles_view_matches_julia_layout():
    s = sh.workloads.pyrochlore_model(dims=(2, 3, 4), seed=1)
    flat = s.dipoles_flat()
    v = s.dipoles
    for (c1, c2, c3, i) in [(0, 0, 0, 0), (1, 2, 3, 15), (1, 0, 2, 7)]:
        np.testing.assert_array_equal(v[c1, c2, c3, i], flat[c1 + 2 * (c2 + 3 * (c3 + 4 * i))])


def test_onsite_spin_half_rejected():
    s = sh.System(sh.cubic_crystal(), [(0, sh.Moment(s=0.5, g=2))], "dipole")
    with pytest.raises(ValueError, match="trivial for quantum spin 1/2"):
        sh.set_onsite_coupling_(s, lambda S: S[2] @ S[2], 0)


def test_cabi_exports_every_declared_symbol():
    """libsunnyhip.so loads without a GPU and exports every entry point that
    include/sunny_hip.h declares."""
    header = open(os.path.join(ROOT, "include", "sunny_hip.h")).read()
    declared = set(re.findall(r"\b(snh_\w+)\s*\(", header))
    assert len(declared) >= 15
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTS)


def test_create_without_gpu_fails_cleanly():
    """Without a device snh_create returns an error code instead of crashing."""
    s = sh.workloads.cubic_model(L=2)
    try:
        h = _lib.Handle(s.spec())
    except (_lib.HipError, ValueError) as e:
        assert "device" in str(e) or "hip" in str(e).lower()
        return
    del h
    pytest.skip("a GPU is present")


def test_library_built_from_these_sources():
    """The loaded library carries the build id of the sources beside it
    (the loader refuses a stale build)."""
    from sunnyhip import _lib
    want = _lib.source_build_id()
    assert want is not None and len(want) == 16
    assert _lib.lib().snh_build_id().decode() == want
