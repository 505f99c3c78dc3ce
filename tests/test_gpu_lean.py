"""The tiled Langevin and midpoint kernels have a lean form (no Ewald
potential, device noise, onsite terms of rank <= 2; isotropic bonds for the
midpoint iteration) beside the general one.  A handle created with
SNH_LEAN=0 (general forms only) and one created without it take Langevin
steps to bitwise-equal states.  The general midpoint iteration evaluates
every bond as a 3x3 product; the lean one takes an isotropic bond's single
coupling (one fused multiply-add per component, as the ghost-layout kernels
do), so midpoint states agree to rounding there, and bitwise where the lean
form does not apply (anisotropic bonds)."""
import numpy as np
import pytest

import sunnyhip as sh

from helpers import announce

pytestmark = pytest.mark.gpu


def iso_batch(precision):
    dims = [(4, 4, 4), (3, 4, 5)]
    systems = [sh.workloads.pyrochlore_model(dims=d, seed=60 + r, precision=precision) for r, d in enumerate(dims)]
    return sh.ReplicaBatch(systems, precision=precision, layout=0)


def aniso_system(precision):
    """Diamond with a bilinear exchange matrix, a field and a rank-2 onsite
    term only: the lean predictor and corrector with the general bond path."""
    s = sh.System(sh.diamond_crystal(), [(i, sh.Moment(s=3, g=2)) for i in range(8)], "dipole", dims=(6, 6, 4),
                  seed=8, precision=precision)
    sh.set_field_(s, [0.0, 0.2, 0.3])
    J = np.array([[0.5, 0.2, -0.4], [0.2, 0.5, -0.4], [0.4, 0.4, 1.5]])
    sh.set_exchange_(s, J, sh.Bond(0, 1, (0, 0, 0)))
    O_ = sh.stevens_matrices(3)
    for i in range(8):
        sh.set_onsite_coupling_(s, 0.05 * O_[2][2] - 0.02 * O_[2][0], i)
    sh.randomize_spins_(s)
    s.layout = 0
    return s


def state(obj):
    h = obj._device()
    d = np.empty((h.total, 3))
    h.call("snh_get_dipoles", sh._lib.dptr(d))
    return d


@pytest.mark.parametrize("kind", ["iso", "aniso"])
@pytest.mark.parametrize("precision", [32, 64])
def test_lean_forms_match_general(kind, precision, monkeypatch):
    announce(f"lean vs general tiled kernels {kind} fp{precision}")
    make = (lambda: iso_batch(precision)) if kind == "iso" else (lambda: aniso_system(precision))
    monkeypatch.setenv("SNH_PAIR", "0")
    monkeypatch.setenv("SNH_COOP", "0")  # keep the one-replica model on the tiled layout
    if precision == 64:
        # small FP64 models default to the ghost layout: request a tile plan
        # (16 pyrochlore sublattices x 64 lanes need two sites per thread to
        # stay within 512 FP64 threads)
        monkeypatch.setenv("SNH_TILE_SP", "2" if kind == "iso" else "1")
    monkeypatch.setenv("SNH_LEAN", "0")
    a = make()
    ha = a._device()
    monkeypatch.delenv("SNH_LEAN")
    b = make()
    hb = b._device()
    if ha.query("layout") != 0 or hb.query("layout") != 0:
        pytest.skip("the model does not fit the tiled layout here")
    integ = sh.Langevin(0.02, damping=0.1, kT=0.5)
    for _ in range(2):
        sh.step_(a, integ, nsteps=6)
        sh.step_(b, integ, nsteps=6)
        np.testing.assert_array_equal(state(a), state(b))
    # the FP64 energy sums per workgroup of the predictor's plan, and a lean
    # handle's default plan (one site per thread) groups the sites differently
    np.testing.assert_allclose(sh.energy(a), sh.energy(b), rtol=1e-15, atol=0)
    mid = sh.ImplicitMidpoint(0.02, damping=0.1, tol=1e-6 if precision == 32 else 1e-12)
    ia = sh.step_(a, mid, nsteps=4)
    ib = sh.step_(b, mid, nsteps=4)
    np.testing.assert_array_equal(ia, ib)
    if kind == "aniso":
        np.testing.assert_array_equal(state(a), state(b))
    else:
        # |S| = 3/2: a few ulps of FP32 (FP64) per component after 4 steps
        tol = 2e-6 if precision == 32 else 1e-14
        np.testing.assert_allclose(state(a), state(b), rtol=0, atol=tol)
