"""Gradient-only callers on the device: suggest_timestep (Integrators.jl:187-245)
against closed forms, minimize_energy! (Optimization.jl:83-166) against the
oracle energy of the known ground state."""
import math

import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O
from sunnyhip.optimize import minimize_energy_, suggest_timestep_aux

from helpers import announce

pytestmark = pytest.mark.gpu


def test_suggest_timestep_single_spin_in_field():
    """One spin (s = 1, g = 2) in a field of 0.5 along z: |dE/dS| = 1, so
    dt = sqrt(tol / (1 + (kappa lambda)^2 + (lambda kT)^2))."""
    announce("suggest_timestep single spin")
    s = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=2))], "dipole", dims=(1, 1, 1), seed=0,
                  precision=64)
    sh.set_field_(s, [0.0, 0.0, 0.5])
    sh.randomize_spins_(s)
    tol = 1e-4
    assert abs(suggest_timestep_aux(s, sh.ImplicitMidpoint(0.01), tol=tol) - math.sqrt(tol)) < 1e-15
    b = suggest_timestep_aux(s, sh.Langevin(0.01, damping=0.5, kT=2.0), tol=tol)
    assert abs(b - math.sqrt(tol / (1 + 0.25 + 1.0))) < 1e-15
    with pytest.raises(ValueError):
        z = sh.System(sh.cubic_crystal(1.0), [(0, sh.Moment(s=1, g=2))], "dipole", dims=(1, 1, 1), precision=64)
        suggest_timestep_aux(z, sh.Langevin(0.01, damping=0.1, kT=0.0), tol=tol)


def test_minimize_energy_finds_neel_state():
    """Cubic antiferromagnet J = +1 with easy axis -0.1 Sz^2 on 4^3: the
    minimum is the staggered state along z; its energy from the oracle."""
    s = sh.workloads.cubic_model(L=4, J=1.0, D=-0.1, seed=3, precision=64)
    announce("minimize_energy cubic AFM 4^3")
    L = 4
    c = np.arange(L ** 3)
    stag = (-1.0) ** ((c % L) + (c // L) % L + c // (L * L))
    neel = np.zeros((L ** 3, 3))
    neel[:, 2] = stag
    E_neel = O.Oracle(s.spec()).energy(neel)
    # start near the Neel state (tilted by random angles of ~0.3 rad)
    rng = np.random.default_rng(4)
    x0 = neel + 0.3 * rng.standard_normal(neel.shape)
    s.set_state_flat(dip=x0 / np.linalg.norm(x0, axis=1, keepdims=True))
    res = minimize_energy_(s, maxiters=2000)
    assert res.converged, res
    assert abs(res.energy - E_neel) <= 1e-9 * abs(E_neel), (res.energy, E_neel)
    S = s.dipoles_flat()
    assert np.allclose(np.abs(S[:, 2]), 1.0, atol=1e-5)
    assert np.allclose(np.sign(S[:, 2]) * stag, np.sign(S[0, 2]) * stag[0])


def test_minimize_energy_sun_finds_polarised_state():
    """SU(3) (s = 1) triangular lattice, easy-axis D Sz^2 (D = -1), field
    B = 0.3 along z (g = 2) and ferromagnetic J1 = -0.2: the energy is minimal
    on every site and bond at once in the m = -1 state, so the ground energy is
    N (D - 2 B) + nbonds J1 = N (D - 2 B + 3 J1)."""
    s = sh.workloads.triangular_sun_model(L1=4, L2=3, J1=-0.2, J2=0.0, J3=0.0, D=-1.0, seed=11, precision=64)
    sh.set_field_(s, [0.0, 0.0, 0.3])
    sh.randomize_spins_(s)
    announce("minimize_energy SU(3) triangular 4x3")
    res = minimize_energy_(s, maxiters=2000)
    E_exact = s.nsites * (-1.0 - 2 * 0.3 + 3 * -0.2)
    assert res.converged, res
    assert abs(res.energy - E_exact) <= 1e-8 * abs(E_exact), (res.energy, E_exact)
    # the oracle agrees with the device energy of the final state
    o = O.Oracle(s.spec())
    assert abs(o.energy(Z=s.coherents_flat()) - res.energy) <= 1e-10 * abs(E_exact)
