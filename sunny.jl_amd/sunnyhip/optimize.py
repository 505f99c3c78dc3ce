"""Callers of the device energy gradient alone.

`suggest_timestep` (src/Integrators.jl:164-245) and `minimize_energy!`
(src/Optimization.jl:83-166) evaluate dE/dS (and E) on the GPU through the
same C ABI as the integrators; their control logic is host-side, as in the
reference.
"""
from __future__ import annotations

import math

import numpy as np

from .integrators import energy, energy_grad_coherents, energy_grad_dipoles
from .system import System


def suggest_timestep_aux(sys: System, integrator, *, tol):
    """dt bound sqrt(tol / (drift_rms^2 + (damping kT)^2)) from the rms drift
    of the current configuration (Integrators.jl:187-245)."""
    lam = integrator.damping
    kT = getattr(integrator, "kT", 0.0)
    if sys.N == 0:
        g = energy_grad_dipoles(sys)
        kap = sys.site_kappas()
        acc = float(np.sum((1 + (kap * lam) ** 2) * np.sum(g * g, axis=1)))
    else:
        g = energy_grad_coherents(sys)
        acc = float(np.sum((1 + lam ** 2) * np.sum(np.abs(g) ** 2, axis=1)))
    drift_rms = math.sqrt(acc / sys.nsites)
    if drift_rms == 0:
        raise ValueError("Cannot suggest a timestep without an energy scale!")
    return math.sqrt(tol / (drift_rms ** 2 + (lam * kT) ** 2))


def suggest_timestep(sys: System, integrator, *, tol):
    """Prints (and returns) the suggested dt, as suggest_timestep (Integrators.jl:164-185)."""
    bound = suggest_timestep_aux(sys, integrator, tol=tol)
    msg = f"Consider dt ≈ {bound:.4g} for this spin configuration at tol = {tol:.4g}."
    dt = integrator.dt
    if not math.isnan(dt):
        if dt <= bound / 2:
            msg += f"\nCurrent value dt = {dt:.4g} seems small! Increasing it will make the simulation faster."
        elif dt >= 2 * bound:
            msg += f"\nCurrent value dt = {dt:.4g} seems LARGE! Decreasing it will improve accuracy."
        else:
            msg += f" Current value is dt = {dt:.4g}."
    print(msg)
    return bound


class MinimizationResult:
    def __init__(self, converged, niters, energy):
        self.converged, self.niters, self.energy = converged, niters, energy

    def __repr__(self):
        state = "Converged" if self.converged else "Non-converged"
        return f"MinimizationResult({state} in {self.niters} iterations, E = {self.energy:.12g})"


def _manifold_cg(x, fg, g_abstol, maxiters):
    """Nonlinear conjugate gradients on a product of unit spheres (rows of x,
    real or complex): tangent projection, retraction by normalisation,
    Polak-Ribiere+ with restarts, backtracking Armijo line search.  fg(x)
    returns (E, G, g): G the projected real gradient (dE = Re sum conj(dx) G)
    and g the reference's gradient residual for the stopping rule."""
    def proj(x, v):  # rows of x are unit vectors (retraction by normalisation)
        return v - x * np.sum(np.conj(x) * v, axis=1, keepdims=True)

    def dot(a, b):
        return float(np.sum(np.real(np.conj(a) * b)))

    f, G, g = fg(x)
    d = -G
    alpha = 1.0 / max(np.abs(G).max(), 1e-300)
    converged = False
    it = 0
    for it in range(1, maxiters + 1):
        if np.abs(g).max() <= g_abstol:
            converged = True
            it -= 1
            break
        slope = dot(G, d)
        if slope >= 0:  # not a descent direction: restart along -G
            d = -G
            slope = dot(G, d)
        a = min(alpha * 2.0, 10.0 / max(np.abs(d).max(), 1e-300))
        while True:
            xn = x + a * d
            xn = xn / np.linalg.norm(xn, axis=1, keepdims=True)
            fn, Gn, gn = fg(xn)
            if fn <= f + 1e-4 * a * slope or a < 1e-20:
                break
            a *= 0.5
        dx = np.abs(xn - x).max()
        Gt = proj(xn, G)  # the old gradient transported by projection
        beta = max(0.0, dot(Gn, Gn - Gt) / max(dot(G, G), 1e-300))
        d = -Gn + beta * proj(xn, d)
        x, f, G, g, alpha = xn, fn, Gn, gn, a
        if dx <= 1e-12:
            converged = True
            break
    return x, converged, it


def minimize_energy_(sys: System, maxiters=1000, jitter=1e-8, seed=0):
    """minimize_energy!(sys; maxiters, jitter) (Optimization.jl:83-166):
    unit spins x on the product of spheres (dipoles, or kets for :SUN),
    the energy and its gradient from the device (kappa dE/dS, or
    sqrt(kappa) dE/dZbar), nonlinear conjugate gradients on the manifold,
    stopping when the largest gradient component falls below 1e-12 x the
    characteristic energy scale (Interactions.jl:674-683) or the largest
    spin change below 1e-12."""
    if sys.N != 0:
        return _minimize_sun(sys, maxiters, jitter, seed)
    rng = np.random.default_rng(seed)
    kap = sys.site_kappas()[:, None]
    x = sys.dipoles_flat()
    if jitter:  # perturb_spins!(sys, jitter)
        x = x + jitter * kap * rng.standard_normal(x.shape)
    n = np.linalg.norm(x, axis=1, keepdims=True)
    x = np.where(n > 0, x / np.where(n > 0, n, 1), np.array([1.0, 1.0, 1.0]) / math.sqrt(3))

    def load(x):
        sys.set_state_flat(dip=kap * x)

    def proj(x, g):
        return g - x * np.sum(x * g, axis=1, keepdims=True)

    def fg(x):
        load(x)
        g = proj(x, kap * energy_grad_dipoles(sys))
        return energy(sys), g, g

    load(x)
    scale = float(np.linalg.norm(kap * energy_grad_dipoles(sys)) / math.sqrt(sys.nsites))
    x, converged, it = _manifold_cg(x, fg, 1e-12 * max(scale, 1e-300), maxiters)
    load(x)
    return MinimizationResult(converged, it, energy(sys))


def _minimize_sun(sys: System, maxiters, jitter, seed):
    """:SUN branch of minimize_energy!: x = Z / |Z| per site, Z = sqrt(kappa) x;
    the reference's gradient is sqrt(kappa) H Z (H Z = dE/dZbar), and the real
    gradient of E(x) is twice that, projected on the tangent space."""
    rng = np.random.default_rng(seed)
    N = sys.N
    rk = np.sqrt(sys.site_kappas())[:, None]
    Z = sys.coherents_flat()
    if jitter:  # perturb_spins!(sys, jitter): Z + jitter sqrt(kappa) randn(ComplexF64)
        Z = Z + jitter * rk * (rng.standard_normal(Z.shape) + 1j * rng.standard_normal(Z.shape)) / math.sqrt(2)
    n = np.linalg.norm(Z, axis=1, keepdims=True)
    x = np.where(n > 0, Z / np.where(n > 0, n, 1), np.ones(N) / math.sqrt(N))

    def load(x):
        sys.set_state_flat(coh=rk * x)

    def proj(x, v):
        return v - x * np.sum(np.conj(x) * v, axis=1, keepdims=True)

    def fg(x):
        load(x)
        g = proj(x, rk * energy_grad_coherents(sys))
        return energy(sys), 2 * g, g

    load(x)
    scale = float(np.linalg.norm(rk * energy_grad_coherents(sys)) / math.sqrt(sys.nsites))
    x, converged, it = _manifold_cg(x, fg, 1e-12 * max(scale, 1e-300), maxiters)
    load(x)
    return MinimizationResult(converged, it, energy(sys))
