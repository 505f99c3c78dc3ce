"""`:dipole_uncorrected` mode, the large-s limit (src/System/System.jl:33-93,
OnsiteCoupling.jl:1-44, PairExchange.jl:322-340) in the host mirror (CPU
only): onsite couplings are polynomials of the classical spin
(spin_matrices(inf)) expanded in classical Stevens polynomials on |S| = kappa
without renormalisation, and the scalar biquadratic coupling is exactly
(Si.Sj)^2.  Pinned by the oracle's energy against the polynomials evaluated
directly at random spins."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O


def _poly(S):
    Sx, Sy, Sz = S
    return -0.1 * Sz @ Sz + 0.01 * (Sx @ Sx) @ (Sx @ Sx) + 0.002 * Sx ** 2 * Sy ** 2 * Sz ** 2 - 0.03 * Sx @ Sy


def _direct(S):
    x, y, z = S[:, 0], S[:, 1], S[:, 2]
    return -0.1 * z * z + 0.01 * x ** 4 + 0.002 * x * x * y * y * z * z - 0.03 * x * y


@pytest.mark.parametrize("s_", [0.5, 2.0, 3.5])
def test_onsite_polynomial_energy(s_):
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=s_, g=2))], "dipole_uncorrected", dims=(2, 2, 1), seed=1)
    sh.set_onsite_coupling_(s, _poly, 0)
    sh.randomize_spins_(s)
    S = s.dipoles_flat()
    np.testing.assert_allclose(np.linalg.norm(S, axis=1), s_, rtol=1e-14)
    E = O.Oracle(s.spec()).energy(S)
    assert abs(E - _direct(S).sum()) <= 1e-12 * max(1.0, abs(E))


def test_biquadratic_is_exact_square():
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1.5, g=2))], "dipole_uncorrected", dims=(3, 2, 2),
                  seed=2)
    sh.set_exchange_(s, 0.7, sh.Bond(0, 0, (1, 0, 0)), biquad=0.3)
    sh.randomize_spins_(s)
    S = s.dipoles_flat()
    L = s.dims
    E = 0.0
    for c3 in range(L[2]):
        for c2 in range(L[1]):
            for c1 in range(L[0]):
                k = c1 + L[0] * (c2 + L[1] * c3)
                for n in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    j = (c1 + n[0]) % L[0] + L[0] * ((c2 + n[1]) % L[1] + L[1] * ((c3 + n[2]) % L[2]))
                    d = S[k] @ S[j]
                    E += 0.7 * d + 0.3 * d * d
    Eo = O.Oracle(s.spec()).energy(S)
    assert abs(Eo - E) <= 1e-12 * abs(E)


def test_mode_errors():
    u = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=2, g=2))], "dipole_uncorrected", dims=(2, 1, 1))
    d = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=2, g=2))], "dipole", dims=(2, 1, 1))
    with pytest.raises(ValueError, match="requires a symbolic operator"):
        sh.set_onsite_coupling_(u, np.eye(5), 0)
    with pytest.raises(ValueError, match="Symbolic operator only valid"):
        sh.set_onsite_coupling_(d, sh.spin_matrices(np.inf)[2] ** 2, 0)
    with pytest.raises(ValueError, match="time-reversal"):
        sh.set_onsite_coupling_(u, lambda S: S[2], 0)
    with pytest.raises(ValueError, match="General coupling not currently supported"):
        sh.set_pair_coupling_(u, lambda Si, Sj: Si[0] @ Sj[0], sh.Bond(0, 0, (1, 0, 0)))
    h = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=0.5, g=2))], "dipole", dims=(2, 1, 1))
    with pytest.raises(ValueError, match="Consider mode :dipole_uncorrected"):
        sh.set_onsite_coupling_(h, lambda S: S[2] @ S[2], 0)


def test_uncorrected_differs_from_dipole_by_renormalisation():
    """The same quadratic anisotropy in the two dipole modes: :dipole
    renormalises the rank-2 coefficients by rcs_2(s) = 1 - 1/(2s) and its
    constant by s(s+1)/3 (quantum S^2), the large-s mode not."""
    s_ = 2.0
    a = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=s_, g=2))], "dipole", dims=(1, 1, 1))
    b = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=s_, g=2))], "dipole_uncorrected", dims=(1, 1, 1))
    sh.set_onsite_coupling_(a, lambda S: S[2] @ S[2], 0)
    sh.set_onsite_coupling_(b, lambda S: S[2] @ S[2], 0)
    S = np.array([[0.6, 0.0, 0.8]]) * s_
    Ea, Eb = O.Oracle(a.spec()).energy(S), O.Oracle(b.spec()).energy(S)
    X = s_ * s_
    assert abs(Eb - S[0, 2] ** 2) < 1e-13
    assert abs(Ea - (s_ * (s_ + 1) / 3 + (1 - 1 / (2 * s_)) * (S[0, 2] ** 2 - X / 3))) < 1e-13
