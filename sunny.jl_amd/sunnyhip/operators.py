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


class SpinPolynomial:
    """Polynomial in commuting classical spin components: the large-s spin
    operators that spin_matrices(Inf) gives for :dipole_uncorrected mode
    (Spin.jl:16-35, Symbolic.jl).  terms: {(a, b, c): coefficient} of
    Sx^a Sy^b Sz^c; + - * ** and @ (the same as *: the components commute)."""

    def __init__(self, terms):
        self.terms = {k: float(v) for k, v in terms.items() if v != 0}

    @staticmethod
    def _lift(o):
        if isinstance(o, SpinPolynomial):
            return o
        if np.ndim(o) == 0 and np.isreal(o):
            return SpinPolynomial({(0, 0, 0): float(np.real(o))})
        raise TypeError("spin polynomials combine with real numbers and spin polynomials only")

    def __add__(self, o):
        o = self._lift(o)
        t = dict(self.terms)
        for k, v in o.terms.items():
            t[k] = t.get(k, 0.0) + v
        return SpinPolynomial(t)

    __radd__ = __add__

    def __neg__(self):
        return SpinPolynomial({k: -v for k, v in self.terms.items()})

    def __sub__(self, o):
        return self + (-self._lift(o))

    def __rsub__(self, o):
        return self._lift(o) - self

    def __mul__(self, o):
        o = self._lift(o)
        t = {}
        for k1, v1 in self.terms.items():
            for k2, v2 in o.terms.items():
                k = (k1[0] + k2[0], k1[1] + k2[1], k1[2] + k2[2])
                t[k] = t.get(k, 0.0) + v1 * v2
        return SpinPolynomial(t)

    __rmul__ = __mul__
    __matmul__ = __mul__
    __rmatmul__ = __mul__

    def __pow__(self, n):
        out = SpinPolynomial({(0, 0, 0): 1.0})
        for _ in range(int(n)):
            out = out * self
        return out

    def degree(self):
        return max((sum(k) for k in self.terms), default=0)

    def __call__(self, x, y, z):
        """Value at classical spin components (arrays)."""
        out = np.zeros(np.broadcast(x, y, z).shape)
        for (a, b, c), v in self.terms.items():
            out = out + v * x ** a * y ** b * z ** c
        return out


def spin_matrices(s: float):
    """(Sx, Sy, Sz) in the spin-s irrep, basis m = s..-s (Spin.jl:1-14);
    s = inf gives the large-s operators of :dipole_uncorrected mode, commuting
    classical components (SpinPolynomial)."""
    if np.isinf(s):
        return (SpinPolynomial({(1, 0, 0): 1.0}), SpinPolynomial({(0, 1, 0): 1.0}),
                SpinPolynomial({(0, 0, 1): 1.0}))
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


def classical_stevens_basis(S):
    """Values [n, 28] of the classical Stevens polynomials (k = 0, 2, 4, 6; q
    descending) at spins S [n, 3], with X = |S|^2: the polynomials whose
    coefficients the classical anisotropy energy multiplies
    (OnsiteCoupling.jl:316-425)."""
    S = np.asarray(S, float)
    x, y, z = S[:, 0], S[:, 1], S[:, 2]
    X = x * x + y * y + z * z
    jp = x + 1j * y
    P = [np.ones_like(x) + 0j] + [jp ** m for m in range(1, 7)]

    def row(k, A):
        # A[m]: the z-polynomial of |q| = m; q = k..1 (Re Jp^q), 0, -1..-k (Im Jp^|q|)
        return [(P[q] * A[q]).real for q in range(k, 0, -1)] + [A[0]] + \
            [(P[q] * A[q]).imag for q in range(1, k + 1)]
    b = [np.ones_like(x)]
    b += row(2, [3 * z ** 2 - X, z, np.ones_like(x)])
    b += row(4, [35 * z ** 4 - 30 * X * z ** 2 + 3 * X ** 2, 7 * z ** 3 - 3 * X * z, 7 * z ** 2 - X, z,
                 np.ones_like(x)])
    b += row(6, [231 * z ** 6 - 315 * X * z ** 4 + 105 * X ** 2 * z ** 2 - 5 * X ** 3,
                 33 * z ** 5 - 30 * X * z ** 3 + 5 * X ** 2 * z, 33 * z ** 4 - 18 * X * z ** 2 + X ** 2,
                 11 * z ** 3 - 3 * X * z, 11 * z ** 2 - X, z, np.ones_like(x)])
    return np.stack(b, 1)


def polynomial_to_stevens_coefficients(p, kappa: float):
    """Classical Stevens coefficients {k: [2k+1]} of a spin polynomial on the
    sphere |S| = kappa (operator_to_stevens_coefficients(p, S^2 = kappa^2),
    Symbolic.jl:94-103, without renormalisation): an even polynomial of
    degree <= 6 restricted to the sphere is exactly such an expansion, found
    by least squares on fixed sample points."""
    p = SpinPolynomial._lift(p)
    if any(sum(k) % 2 for k in p.terms):
        raise ValueError("Single-ion anisotropy must be time-reversal invariant.")
    if p.degree() > 6:
        raise ValueError("Stevens operators for k > 6 are currently unsupported.")
    u = np.random.default_rng(0).standard_normal((400, 3))
    S = kappa * u / np.linalg.norm(u, axis=1, keepdims=True)
    B = classical_stevens_basis(S)
    v = p(S[:, 0], S[:, 1], S[:, 2])
    c, *_ = np.linalg.lstsq(B, v, rcond=None)
    if np.abs(B @ c - v).max() > 1e-9 * max(1.0, np.abs(v).max()):
        raise ValueError("spin polynomial is not a Stevens expansion of rank <= 6")
    c[np.abs(c) < 1e-13 * max(1.0, np.abs(c).max())] = 0.0
    out = {k: np.zeros(2 * k + 1) for k in range(7)}
    out[0], out[2], out[4], out[6] = c[:1], c[1:6], c[6:15], c[15:28]
    return out


def hermitian_basis(N: int):
    """An orthonormal basis of the N x N Hermitian matrices under the real
    inner product tr(A B): the diagonal units, then (E_ab + E_ba)/sqrt 2 and
    i(E_ba - E_ab)/sqrt 2 for a < b."""
    out = []
    for a in range(N):
        E = np.zeros((N, N), complex)
        E[a, a] = 1
        out.append(E)
    r = 1 / np.sqrt(2)
    for a in range(N):
        for b in range(a + 1, N):
            E = np.zeros((N, N), complex)
            E[a, b] = E[b, a] = r
            out.append(E)
            E = np.zeros((N, N), complex)
            E[a, b], E[b, a] = -1j * r, 1j * r
            out.append(E)
    return out


def tensor_expansion(op, N1: int, N2: int, tol=1e-12):
    """op = sum_k A_k (x) B_k with Hermitian A_k (N1 x N1) and B_k (N2 x N2):
    the coefficients of a Hermitian op in the product of Hermitian bases are
    real, and their singular value decomposition gives the terms (the
    construction of svd_tensor_expansion, src/Operators/TensorOperators.jl:77-113;
    each singular value is carried by A_k)."""
    op = np.asarray(op, complex)
    if np.max(np.abs(op - op.conj().T), initial=0) > 1e-12:
        raise ValueError("Detected non-Hermitian operator")
    H1, H2 = hermitian_basis(N1), hermitian_basis(N2)
    # tr((H1_m (x) H2_n) op) = sum_{a,b,i,j} H1_m[b,a] H2_n[j,i] op[(a,i),(b,j)]
    T4 = op.reshape(N1, N2, N1, N2)
    Cm = np.einsum("mba,nji,aibj->mn", np.array(H1), np.array(H2), T4)
    if np.max(np.abs(Cm.imag), initial=0) > 1e-10:
        raise ValueError("expansion coefficients of a Hermitian operator must be real")
    U, S, Vt = np.linalg.svd(Cm.real)
    terms = []
    for k, sig in enumerate(S):
        if sig < tol:
            break
        A = sig * np.tensordot(U[:, k], np.array(H1), axes=1)
        B = np.tensordot(Vt[k], np.array(H2), axes=1)
        terms.append((A, B))
    return terms


def decompose_pair_coupling(op, N1: int, N2: int, extract_parts=True):
    """Splits a pair operator on C^N1 (x) C^N2 into a scalar, a bilinear 3x3
    part over the spin operators, a biquadratic 5x5 part over the Stevens
    quadrupoles O[2,q] (q descending; None when absent or N < 3) and the
    general remainder as Hermitian tensor terms (decompose_general_coupling,
    src/System/PairExchange.jl:71-111)."""
    op = np.asarray(op, complex).copy()
    dim = N1 * N2
    if op.shape != (dim, dim):
        raise ValueError("pair operator has the wrong size")
    scalar = float(np.real(np.trace(op)) / dim)
    op = op - scalar * np.eye(dim)
    bilin = np.zeros((3, 3))
    biquad = None
    if extract_parts:
        S1, S2 = spin_matrices((N1 - 1) / 2), spin_matrices((N2 - 1) / 2)
        for a in range(3):
            for b in range(3):
                v = np.kron(S1[a], S2[b])
                J = np.trace(v.conj().T @ op) / np.trace(v.conj().T @ v)
                if abs(J.imag) > 1e-12:
                    raise ValueError("non-Hermitian bilinear part")
                bilin[a, b] = J.real
                op = op - v * J.real
        if N1 > 2 and N2 > 2:
            O1, O2 = stevens_matrices((N1 - 1) / 2)[2], stevens_matrices((N2 - 1) / 2)[2]
            bq = np.zeros((5, 5))
            for a in range(5):
                for b in range(5):
                    v = np.kron(O1[a], O2[b])
                    J = np.trace(v.conj().T @ op) / np.trace(v.conj().T @ v)
                    if abs(J.imag) > 1e-12:
                        raise ValueError("non-Hermitian biquadratic part")
                    bq[a, b] = J.real
                    op = op - v * J.real
            if np.max(np.abs(bq)) > 1e-12:
                biquad = bq
    return scalar, bilin, biquad, tensor_expansion(op, N1, N2)
