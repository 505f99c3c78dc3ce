"""Spin operators and Stevens expansions (host-side setup of onsite couplings).

Restates src/Operators/Spin.jl:1-14 (spin matrices), src/Operators/Stevens.jl
:12-86 (Stevens operators as spin polynomials) and :144-159 (coefficients by
trace projection), and src/System/OnsiteCoupling.jl:46-55 (renormalisation
factors for :dipole mode).
"""
from __future__ import annotations

import numpy as np

# Coefficients of the scalar biquadratic in the quadrupole basis,
# src/System/Types.jl:40.
SCALAR_BIQUAD_METRIC = np.array([1 / 2, 2, 1 / 6, 2, 1 / 2])


def spin_matrices(s: float):
    """(Sx, Sy, Sz) in the spin-s irrep, basis m = s..-s (Spin.jl:1-14)."""
    N = int(round(2 * s + 1))
    if abs(N - (2 * s + 1)) > 1e-12:
        raise ValueError("Spin `s` must be half-integer.")
    j = np.arange(1, N)
    off = np.sqrt(2 * (s + 1) * j - j * (j + 1)) / 2
    Sx = np.diag(off, 1) + np.diag(off, -1)
    Sy = np.diag(-1j * off, 1) + np.diag(1j * off, -1)
    Sz = np.diag(s - np.arange(N))
    return Sx.astype(complex), Sy.astype(complex), Sz.astype(complex)


def _stevens_polynomials(J, k):
    Jx, Jy, Jz = J
    Id = np.eye(Jx.shape[0], dtype=complex)
    X = Jx @ Jx + Jy @ Jy + Jz @ Jz
    Jp = Jx + 1j * Jy
    Jm = Jx - 1j * Jy
    mp = np.linalg.matrix_power
    A = [0.5 * (mp(Jp, m) + mp(Jm, m)) for m in range(k, 0, -1)] + [Id] + \
        [(1 / 2j) * (mp(Jp, m) - mp(Jm, m)) for m in range(1, k + 1)]
    Z = Jz
    if k == 0:
        B = [Id]
    elif k == 1:
        B = [Z, Id]
    elif k == 2:
        B = [3 * Z @ Z - X, Z, Id]
    elif k == 3:
        B = [5 * mp(Z, 3) - (3 * X - Id) @ Z, 5 * Z @ Z - X - Id / 2, Z, Id]
    elif k == 4:
        B = [35 * mp(Z, 4) - (30 * X - 25 * Id) @ Z @ Z + (3 * X @ X - 6 * X),
             7 * mp(Z, 3) - (3 * X + Id) @ Z, 7 * Z @ Z - (X + 5 * Id), Z, Id]
    elif k == 5:
        B = [63 * mp(Z, 5) - (70 * X - 105 * Id) @ mp(Z, 3) + (15 * X @ X - 50 * X + 12 * Id) @ Z,
             21 * mp(Z, 4) - 14 * X @ Z @ Z + (X @ X - X + 1.5 * Id),
             3 * mp(Z, 3) - (X + 6 * Id) @ Z, 9 * Z @ Z - (X + 16.5 * Id), Z, Id]
    elif k == 6:
        B = [231 * mp(Z, 6) - (315 * X - 735 * Id) @ mp(Z, 4) + (105 * X @ X - 525 * X + 294 * Id) @ Z @ Z
             - (5 * mp(X, 3) - 40 * X @ X + 60 * X),
             33 * mp(Z, 5) - (30 * X - 15 * Id) @ mp(Z, 3) + (5 * X @ X - 10 * X + 12 * Id) @ Z,
             33 * mp(Z, 4) - (18 * X + 123 * Id) @ Z @ Z + (X @ X + 10 * X + 102 * Id),
             11 * mp(Z, 3) - (3 * X + 59 * Id) @ Z, 11 * Z @ Z - (X + 38 * Id), Z, Id]
    else:
        raise ValueError(f"Stevens operators for k > 6 are currently unsupported, received k={k}.")
    B = list(reversed(B)) + B[1:]
    return [(a @ b + b @ a) / 2 for a, b in zip(A, B)]


def stevens_matrices(s: float):
    """O[k][q-index] for k = 0..6, q descending k..-k (Stevens.jl:80-86)."""
    J = spin_matrices(s)
    N = J[0].shape[0]
    out = {}
    for k in range(7):
        if k >= N:
            out[k] = [np.zeros((N, N), dtype=complex) for _ in range(2 * k + 1)]
        else:
            out[k] = _stevens_polynomials(J, k)
    return out


def matrix_to_stevens_coefficients(A: np.ndarray):
    """c[k] (q descending) with A = sum c[k][q] O[k][q] (Stevens.jl:144-159)."""
    A = np.asarray(A, dtype=complex)
    N = A.shape[0]
    O = stevens_matrices((N - 1) / 2)
    c = {}
    for k in range(7):
        if k >= N:
            c[k] = np.zeros(2 * k + 1)
            continue
        vals = []
        for Oq in O[k]:
            v = np.trace(Oq.conj().T @ A) / np.trace(Oq.conj().T @ Oq)
            assert abs(v.imag) < 1e-12
            vals.append(0.0 if abs(v) < 1e-12 else v.real)
        c[k] = np.array(vals)
    return c


def rcs_factors(s: float):
    """Renormalisation factors lambda[k] (OnsiteCoupling.jl:46-55)."""
    return [1.0, 1.0, 1 - 0.5 / s, 1 - 1.5 / s + 0.5 / s**2,
            1 - 3 / s + 2.75 / s**2 - 0.75 / s**3,
            1 - 5 / s + 8.75 / s**2 - 6.25 / s**3 + 1.5 / s**4,
            1 - 7.5 / s + 21.25 / s**2 - 28.125 / s**3 + 17.125 / s**4 - 3.75 / s**5]


class StevensExpansion:
    """c0, c2, c4, c6 (q descending) and kmax (Types.jl:4-21)."""

    def __init__(self, c0, c2, c4, c6):
        blocks = []
        for ck in (c0, c2, c4, c6):
            ck = np.asarray(ck, dtype=float)
            blocks.append(np.zeros_like(ck) if np.linalg.norm(ck) < 1e-12 else ck)
        self.c0, self.c2, self.c4, self.c6 = blocks
        self.kmax = max(2 * bool(np.any(self.c2)), 4 * bool(np.any(self.c4)), 6 * bool(np.any(self.c6)))

    @classmethod
    def from_coefficients(cls, c):
        if any(np.any(np.asarray(c[k]) != 0) for k in (1, 3, 5)):
            raise ValueError("Single-ion anisotropy must be time-reversal invariant.")
        return cls(c[0], c[2], c[4], c[6])

    @classmethod
    def zero(cls):
        return cls(np.zeros(1), np.zeros(5), np.zeros(9), np.zeros(13))

    def __add__(self, other):
        return StevensExpansion(self.c0 + other.c0, self.c2 + other.c2, self.c4 + other.c4, self.c6 + other.c6)

    def scaled(self, f):
        return StevensExpansion(self.c0 * f, self.c2 * f, self.c4 * f, self.c6 * f)

    def packed(self) -> np.ndarray:
        return np.concatenate([self.c0, self.c2, self.c4, self.c6])
