"""Forward-mode derivatives of the deterministic spherical-midpoint trajectory.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the library): a
vectorised numpy restatement of step!(::System{0}, ::ImplicitMidpoint) at
damping 0 (Sunny.jl src/Integrators.jl:379-413, fast_isapprox :361-371) whose
arithmetic runs on dual numbers: every quantity carries its value and its
tangents along P parameter directions (and optionally along the initial
state).  This is forward-mode automatic differentiation of the same discrete
algorithm, so it pins the device adjoint (a reverse sweep derived by implicit
differentiation) by an independent route, as ForwardDiff through the CPU
reference would.

Couplings supported: bilinear exchange (any 3x3 J per directed bond),
Zeeman, onsite anisotropy of order <= 2 (gradient 2 Q2 S with |S|^2 held
constant, OnsiteCoupling.jl:316-425), the Ewald dipole-dipole term
(Ewald.jl:185-209), which is linear in S: its field is applied as the dense
matrix whose columns are the oracle's pairwise Ewald field
(so_ewald_grad_bruteforce) of unit dipoles; and the nonlinear terms,
biquadratic bonds and Stevens anisotropy of order 4 and 6: their field is
the C oracle's gradient of a spec holding only those terms, a polynomial of
degree <= 5 in the spins, so its directional derivative along each tangent
is the 7-point central difference (exact for polynomials of degree <= 6, so
only rounding remains), and its derivative along a parameter is the same
gradient of the parameter's unit couplings (couplings are linear in the
parameters).  The couplings are linear in the
parameters (ModelParam, src/System/Types.jl:76-86): the tangent of the
couplings along parameter p is that parameter's unit coupling spec.
"""
from __future__ import annotations

import numpy as np


class Dual:
    """Value v [...] and tangents d [..., P]."""

    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v = np.asarray(v, float)
        self.d = np.asarray(d, float)

    @staticmethod
    def const(v, P):
        v = np.asarray(v, float)
        return Dual(v, np.zeros(v.shape + (P,)))

    def __add__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v + o.v, self.d + o.d)
        return Dual(self.v + o, self.d)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v - o.v, self.d - o.d)
        return Dual(self.v - o, self.d)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v * o.v, self.d * o.v[..., None] + self.v[..., None] * o.d)
        o = np.asarray(o, float)
        return Dual(self.v * o, self.d * o[..., None] if o.ndim else self.d * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v / o.v, (self.d * o.v[..., None] - self.v[..., None] * o.d) / (o.v ** 2)[..., None])
        o = np.asarray(o, float)
        return Dual(self.v / o, self.d / (o[..., None] if o.ndim else o))

    def __getitem__(self, k):
        # the tangent axis is last: index the value axes only
        k = k if isinstance(k, tuple) else (k,)
        return Dual(self.v[k], self.d[k + (slice(None),)])

    def sum(self, axis):
        return Dual(self.v.sum(axis), self.d.sum(axis))

    def sqrt(self):
        r = np.sqrt(self.v)
        return Dual(r, self.d / (2 * r)[..., None])


def _stack3(x, y, z):
    return Dual(np.stack([x.v, y.v, z.v], -1), np.stack([x.d, y.d, z.d], -2))


def _cross(a, b):
    ax, ay, az = a[..., 0], a[..., 1], a[..., 2]
    bx, by, bz = b[..., 0], b[..., 1], b[..., 2]
    return _stack3(ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx)


def _matvec(M, s):
    """M [..., 3, 3] (Dual) times s [..., 3] (Dual)."""
    rows = [(M[..., r, 0] * s[..., 0] + M[..., r, 1] * s[..., 1] + M[..., r, 2] * s[..., 2]) for r in range(3)]
    return _stack3(*rows)


def _normalize(s, kappa):
    n = (s[..., 0] * s[..., 0] + s[..., 1] * s[..., 1] + s[..., 2] * s[..., 2]).sqrt()
    f = Dual(kappa / n.v, (-kappa / n.v ** 2)[..., None] * n.d)
    return s * Dual(np.repeat(f.v[..., None], 3, -1), np.repeat(f.d[..., None, :], 3, -2))


def _nonlinear_part(spec, sites):
    """The spec's biquadratic bonds and its Stevens terms on the atoms
    `sites` (those the system's anisotropy makes nonlinear), nothing else."""
    s = dict(spec)
    s["bilin"] = np.zeros_like(np.asarray(spec["bilin"], float))
    s["scalar"] = np.zeros_like(np.asarray(spec["scalar"], float))
    s["field"] = np.zeros_like(np.asarray(spec["field"], float))
    st = np.array(spec["stevens"], float)
    st[~sites] = 0
    s["stevens"] = st
    s["kmax"] = np.where(sites, np.asarray(spec["kmax"]), 0).astype(np.int32)
    s["ewald"] = 0
    return s


_FD7 = ((1, 45.0), (2, -9.0), (3, 1.0))  # f'(0) = sum_k w_k (f(k h) - f(-k h)) / (60 h)


def _q2(c2):
    return np.array([[2 * c2[0], 2 * c2[4], c2[1]], [2 * c2[4], -2 * c2[0], c2[3]], [c2[1], c2[3], 6 * c2[2]]])


class MidpointAD:
    """Cubic-layout restatement for one replica: sites ordered as Sunny's
    Array{Vec3,4}(L1,L2,L3,na), bonds gathered per target site."""

    def __init__(self, spec, param_specs):
        assert spec["N"] == 0 and spec["nreplicas"] == 1
        L1, L2, L3 = (int(x) for x in spec["dims"][0])
        na = int(spec["na"])
        self.P = len(param_specs)
        self.ncell = L1 * L2 * L3
        self.Ns = self.ncell * na
        self.kappa = np.repeat(np.asarray(spec["kappa"], float), self.ncell)  # per site
        c = np.arange(self.ncell)
        c1, c2, c3 = c % L1, (c // L1) % L2, c // (L1 * L2)
        nb = int(spec["nbonds"])
        # per directed bond: site list of sources (i cells) and their neighbours
        self.src, self.nbr, self.J = [], [], []
        for b in range(nb):
            i, j, n1, n2, n3, _ = (int(x) for x in spec["bond"][b])
            self.src.append(c + self.ncell * i)
            self.nbr.append(((c1 + n1) % L1) + L1 * (((c2 + n2) % L2) + L2 * ((c3 + n3) % L3)) + self.ncell * j)
            Jv = np.asarray(spec["bilin"][b], float).reshape(3, 3)
            Jd = np.stack([np.asarray(ps["bilin"][b], float).reshape(3, 3) for ps in param_specs], -1) \
                if self.P else np.zeros((3, 3, 0))
            self.J.append(Dual(Jv, Jd))
        # onsite: 2 Q2 per site; Zeeman g^T B per site
        Qv = np.zeros((self.Ns, 3, 3))
        Qd = np.zeros((self.Ns, 3, 3, self.P))
        zee = np.zeros((self.Ns, 3))
        for i in range(na):
            sl = slice(self.ncell * i, self.ncell * (i + 1))
            Qv[sl] = _q2(np.asarray(spec["stevens"][i], float)[1:6])
            for p, ps in enumerate(param_specs):
                Qd[sl, :, :, p] = _q2(np.asarray(ps["stevens"][i], float)[1:6])
            g = np.asarray(spec["g"][i], float).reshape(3, 3)
            zee[sl] = g.T @ np.asarray(spec["field"][i], float)
        self.Q = Dual(Qv, Qd)
        self.zee = zee
        # nonlinear terms through the oracle; their sites drop out of Q
        self.nl = self.nl_p = None
        kmax = np.asarray(spec["kmax"])
        nl_sites = kmax > 2
        if np.asarray(spec["has_biquad"]).any() or nl_sites.any():
            from .oracle import Oracle
            self.nl = Oracle(_nonlinear_part(spec, nl_sites))
            self.nl_p = [Oracle(_nonlinear_part(dict(spec, bilin=ps["bilin"], biquad=ps["biquad"],
                                                     has_biquad=ps["has_biquad"], stevens=ps["stevens"],
                                                     kmax=ps["kmax"]), nl_sites))
                         for ps in param_specs]
            for i in range(na):
                if nl_sites[i]:
                    sl = slice(self.ncell * i, self.ncell * (i + 1))
                    self.Q.v[sl] = 0
                    self.Q.d[sl] = 0
        self.K = None  # Ewald field operator [3 Ns, 3 Ns] (parameter-independent)
        if spec["ewald"]:
            from .oracle import Oracle
            o = Oracle(spec)
            n3 = 3 * self.Ns
            self.K = np.zeros((n3, n3))
            for k in range(n3):
                e = np.zeros(n3)
                e[k] = 1.0
                self.K[:, k] = o.ewald_grad_bruteforce(e.reshape(self.Ns, 3)).reshape(-1)

    def field(self, s):
        """dE/dS (Dual [Ns, 3])."""
        g = _matvec(self.Q, s) + self.zee
        acc_v = g.v.copy()
        acc_d = g.d.copy()
        for src, nbr, J in zip(self.src, self.nbr, self.J):
            sj = s[nbr]
            Jb = Dual(np.broadcast_to(J.v, (len(src), 3, 3)), np.broadcast_to(J.d, (len(src), 3, 3, self.P)))
            t = _matvec(Jb, sj)
            np.add.at(acc_v, src, t.v)
            np.add.at(acc_d, src, t.d)
        if self.K is not None:
            acc_v += (self.K @ s.v.reshape(-1)).reshape(self.Ns, 3)
            acc_d += (self.K @ s.d.reshape(3 * self.Ns, -1)).reshape(acc_d.shape)
        if self.nl is not None:
            acc_v += self.nl.grad(s.v)
            for t in range(acc_d.shape[-1]):
                D = s.d[..., t]
                m = np.abs(D).max()
                if m > 0:
                    h = 0.3 / m
                    fd = sum(w * (self.nl.grad(s.v + k * h * D) - self.nl.grad(s.v - k * h * D)) for k, w in _FD7)
                    acc_d[..., t] += fd / (60 * h)
                if t < len(self.nl_p):  # the parameter's own nonlinear couplings
                    acc_d[..., t] += self.nl_p[t].grad(s.v)
        return Dual(acc_v, acc_d)

    def step(self, S, dt, tol, max_iters=100):
        """One spherical-midpoint step (damping 0) on Dual state S [Ns, 3];
        returns (S_next, iterations)."""
        atol2 = (tol * np.sqrt(self.Ns)) ** 2
        S1 = S
        kap = self.kappa
        for it in range(1, max_iters + 1):
            hat = _normalize((S + S1) * 0.5, kap)
            S2 = S - _cross(hat, self.field(hat)) * dt
            r = ((S1.v - S2.v) ** 2).sum()
            S1 = S2
            if r <= atol2:
                return _normalize(S2, kap), it
        raise RuntimeError("midpoint did not converge")

    def trajectory(self, S0, dt, nsteps, tol=1e-12, dS0=None):
        """Final state (Dual): tangents along the P parameters, plus along
        the initial-state direction dS0 when given (then P+1 tangents)."""
        P = self.P + (dS0 is not None)
        d = np.zeros(S0.shape + (P,))
        if dS0 is not None:
            d[..., -1] = dS0
            # the couplings carry no tangent along the state direction
            self.J = [Dual(J.v, np.concatenate([J.d, np.zeros((3, 3, 1))], -1)) for J in self.J]
            self.Q = Dual(self.Q.v, np.concatenate([self.Q.d, np.zeros(self.Q.v.shape + (1,))], -1))
            self.P = P
        S = Dual(np.asarray(S0, float), d)
        for _ in range(nsteps):
            S, _ = self.step(S, dt, tol)
        return S
