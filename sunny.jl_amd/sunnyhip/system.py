"""Host-side mirror of Sunny's `System` and its coupling setters.

Names follow Sunny (src/System/System.jl, src/System/PairExchange.jl,
src/System/OnsiteCoupling.jl, src/System/Interactions.jl) with Julia's `!`
spelled `_`: `set_exchange!` -> `set_exchange_`, `step!` -> `step_`.
Couplings are assembled here on the host (setup); every energy, gradient and
time step runs in libsunnyhip on the GPU through `sunnyhip._lib`.

Spin state layout matches Julia's column-major `Array{Vec3,4}(L1,L2,L3,na)`:
the flat buffer is `[site][xyz]` with site = c1 + L1*(c2 + L2*(c3 + L3*i)),
and `sys.dipoles` is a view indexed `[c1, c2, c3, i, :]` (0-based).
"""
from __future__ import annotations

import copy

import numpy as np

from . import _lib
from .crystal import Bond, Crystal, bond_parity
from .operators import (SCALAR_BIQUAD_METRIC, SpinPolynomial, StevensExpansion, decompose_pair_coupling,
                        matrix_to_stevens_coefficients, polynomial_to_stevens_coefficients, rcs_factors,
                        spin_matrices)

VACUUM_PERMEABILITY = 0.6745817653324668  # mu0 muB^2 in meV*A^3, src/Units.jl:91


class Moment:
    """Spin label s and g-tensor of one atom (src/System/Moment.jl)."""

    def __init__(self, s, g):
        self.s = float(s)
        g = np.asarray(g, dtype=float)
        self.g = g * np.eye(3) if g.ndim == 0 else g.reshape(3, 3)


def _julia_view(flat, dims, na, width):
    L1, L2, L3 = dims
    return flat.reshape(na, L3, L2, L1, width).transpose(3, 2, 1, 0, 4)


class System:
    """A supercell of `crystal` with per-sublattice moments and couplings.

    mode: "dipole", "dipole_uncorrected" (the large-s limit: no
    renormalisation of onsite and biquadratic couplings, which it takes as
    polynomials of the classical spin, spin_matrices(inf)) or "SUN"
    (src/System/System.jl:33-93).  precision selects
    the device kernels: 64 (plain FP64) or 32 (FP32 arithmetic with FP64
    renormalisation, residual and energy sums).
    """

    def __init__(self, crystal: Crystal, moments, mode: str, dims=(1, 1, 1), seed=None, precision=64):
        if mode not in ("dipole", "dipole_uncorrected", "SUN"):
            raise ValueError("Mode must be `:SUN`, `:dipole`, or `:dipole_uncorrected`")
        self.crystal = crystal
        self.mode = mode
        self.dims = tuple(int(d) for d in dims)
        na = crystal.natoms
        self.na = na
        mom = dict(moments)
        if sorted(mom) != list(range(na)):
            raise ValueError("a Moment is required for every atom (0-based)")
        self.spins = np.array([mom[i].s for i in range(na)])
        self.Ns_atom = np.rint(2 * self.spins + 1).astype(int)
        self.gs = np.array([mom[i].g for i in range(na)])
        if mode == "SUN":
            if len(set(self.Ns_atom)) != 1:
                raise ValueError("Currently all spins S must be equal in SU(N) mode")
            self.N = int(self.Ns_atom[0])
            self.kappas = np.ones(na)
        else:
            self.N = 0
            self.kappas = self.spins.copy()
        self.extfield = np.zeros((na, 3))
        self.onsite = [StevensExpansion.zero() if self.N == 0 else np.zeros((self.N, self.N), complex)
                       for _ in range(na)]
        # Bond -> [scalar, bilin 3x3, biquad 5x5 or None, general [(A, B), ...] or None],
        # accumulated from params
        self.pairs = {}
        # label -> ModelParam: couplings are linear in the parameter values
        # (src/System/Types.jl:76-86); unlabeled couplings get "Unnamed" labels
        self.params = {}
        self.ewald = None
        # per-site interactions of an inhomogeneous system (to_inhomogeneous)
        self.inhomog = None
        self.precision = int(precision)
        self.layout = 0  # device gather layout: 0 automatic (LDS tiles when they fit), 1 ghost-padded
        self.rng = np.random.default_rng(seed)
        self.noise_seed = int(self.rng.integers(0, 2**63))
        self.step_counter = 0
        ns = self.nsites
        self._dip = np.zeros((ns, 3))
        self._coh = np.zeros((ns, max(self.N, 1)), complex)
        self._handle = None
        self._host_dirty = True   # host state newer than the device copy
        self._dev_dirty = False   # device state newer than the host copy
        polarize_spins_(self, (0, 0, 1))

    # ---- sizes and state views
    @property
    def ncells(self):
        return int(np.prod(self.dims))

    @property
    def nsites(self):
        return self.ncells * self.na

    def _pull(self):
        if self._dev_dirty:
            h = self._handle
            d = np.empty((h.total, 3))
            h.call("snh_get_dipoles", _lib.dptr(d))
            self._dip[:] = d
            if self.N > 0:
                z = np.empty((h.total, self.N, 2))
                h.call("snh_get_coherents", _lib.dptr(z))
                self._coh[:] = z[..., 0] + 1j * z[..., 1]
            self._dev_dirty = False

    @property
    def dipoles(self):
        """View [c1, c2, c3, i, xyz]; reading may be followed by writes, so the
        next device call re-uploads the host state."""
        self._pull()
        self._host_dirty = True
        return _julia_view(self._dip, self.dims, self.na, 3)

    @property
    def coherents(self):
        self._pull()
        self._host_dirty = True
        return _julia_view(self._coh, self.dims, self.na, self.N)

    def dipoles_flat(self):
        self._pull()
        return self._dip.copy()

    def coherents_flat(self):
        self._pull()
        return self._coh.copy()

    def set_state_flat(self, dip=None, coh=None):
        self._pull()
        if coh is not None:
            self._coh[:] = np.asarray(coh).reshape(self.nsites, self.N)
            self._dip[:] = expected_spin_host(self._coh, self.N)
        elif dip is not None:
            self._dip[:] = np.asarray(dip).reshape(self.nsites, 3)
        self._host_dirty = True

    def site_kappas(self):
        """kappa of every site in Julia order (per site for an inhomogeneous
        system, where vacancies have kappa = 0)."""
        if self.inhomog is not None:
            return self.inhomog.kappa.copy()
        return np.repeat(self.kappas, self.ncells)

    def site_index(self, site):
        c1, c2, c3, i = site
        L1, L2, L3 = self.dims
        return c1 + L1 * (c2 + L2 * (c3 + L3 * i))

    def _invalidate(self):
        if self._handle is not None:
            self._pull()
            self._handle = None
            self._host_dirty = True

    # ---- coupling tables (the C-ABI spec)
    def spec(self, precision=None, dims_list=None):
        """Arrays describing the couplings, as consumed by snh_create."""
        if self.inhomog is not None:
            from .inhomogeneous import inhomogeneous_spec
            return inhomogeneous_spec(self, precision, dims_list)
        na, N = self.na, self.N
        stev = np.zeros((na, 28))
        kmax = np.zeros(na, np.int32)
        lam = np.zeros((na, max(N, 1), max(N, 1), 2))
        for i in range(na):
            if N == 0:
                stev[i] = self.onsite[i].packed()
                kmax[i] = self.onsite[i].kmax
            else:
                lam[i, :, :, 0] = self.onsite[i].real
                lam[i, :, :, 1] = self.onsite[i].imag
        bonds = sorted(self.pairs.items(), key=lambda kv: (kv[0].i, bond_parity(kv[0]), kv[0].j, kv[0].n))
        nb = len(bonds)
        bond = np.zeros((max(nb, 1), 6), np.int32)
        bilin = np.zeros((max(nb, 1), 9))
        biquad = np.zeros((max(nb, 1), 25))
        has_bq = np.zeros(max(nb, 1), np.int32)
        scalar = np.zeros(max(nb, 1))
        gen_start = np.zeros(max(nb, 1) + 1, np.int32)
        gen = []
        for k, (b, (sc, J, B, G)) in enumerate(bonds):
            bond[k] = (b.i, b.j, *b.n, int(bond_parity(b)))
            bilin[k] = np.asarray(J, float).reshape(9)
            if B is not None:
                biquad[k] = np.asarray(B, float).reshape(25)
                has_bq[k] = 1
            scalar[k] = sc
            gen_start[k] = len(gen)
            gen.extend(G or [])
        gen_start[nb:] = len(gen)
        gen_A = gen_B = None
        if gen:
            Nm = max(N, 1)
            gen_A = np.ascontiguousarray(np.stack([np.stack([np.asarray(A).real, np.asarray(A).imag], -1)
                                                   for A, _ in gen]).reshape(len(gen), Nm, Nm, 2))
            gen_B = np.ascontiguousarray(np.stack([np.stack([np.asarray(B).real, np.asarray(B).imag], -1)
                                                   for _, B in gen]).reshape(len(gen), Nm, Nm, 2))
        dims_list = [self.dims] if dims_list is None else dims_list
        ew = self.ewald
        return {
            "na": na, "N": N, "precision": int(precision or self.precision),
            "kappa": np.ascontiguousarray(self.kappas, float), "g": np.ascontiguousarray(self.gs.reshape(na, 9)),
            "field": np.ascontiguousarray(self.extfield), "stevens": stev, "kmax": kmax,
            "Lambda": np.ascontiguousarray(lam), "nbonds": nb, "bond": bond, "bilin": bilin, "biquad": biquad,
            "has_biquad": has_bq, "scalar": scalar, "ewald": int(ew is not None),
            "mu0_muB2": float(ew[0]) if ew else 0.0,
            "demag": np.asarray(ew[1] if ew else np.zeros((3, 3)), float).reshape(9),
            "latvecs": self.crystal.latvecs.reshape(9).copy(),
            "positions": np.ascontiguousarray(self.crystal.positions, float),
            "nreplicas": len(dims_list), "dims": np.ascontiguousarray(np.array(dims_list, np.int32).reshape(-1, 3)),
            "layout": int(getattr(self, "layout", 0)),
            "gen_start": gen_start if gen else None, "gen_A": gen_A, "gen_B": gen_B,
        }

    # ---- device copy
    def _device(self):
        if self._handle is None:
            self._handle = _lib.Handle(self.spec())
            self._host_dirty = True
        if self._host_dirty:
            h = self._handle
            if self.N == 0:
                h.call("snh_set_dipoles", _lib.dptr(np.ascontiguousarray(self._dip)))
            else:
                z = np.ascontiguousarray(np.stack([self._coh.real, self._coh.imag], -1))
                h.call("snh_set_coherents", _lib.dptr(z))
            self._host_dirty = False
        return self._handle

    def __repr__(self):
        mode = "[Dipole mode]" if self.N == 0 else f"[SU({self.N})]"
        return f"System({mode}, Supercell ({'x'.join(map(str, self.dims))})x{self.na})"


def clone_system(sys: System, seed=None) -> System:
    """Independent copy sharing no mutable state (src/System/System.jl:138-174).
    With `seed`, the clone gets its own generator (and noise key)."""
    c = System.__new__(System)
    c.__dict__.update({k: v for k, v in sys.__dict__.items() if not k.startswith("_")})
    c.onsite = [o.copy() if isinstance(o, np.ndarray) else o for o in sys.onsite]
    c.pairs = {b: [sc, np.array(J), None if B is None else np.array(B), None if G is None else list(G)]
               for b, (sc, J, B, G) in sys.pairs.items()}
    c.params = {k: p.copy() for k, p in sys.params.items()}
    c.extfield = sys.extfield.copy()
    c.kappas = sys.kappas.copy()
    c.inhomog = copy.deepcopy(sys.inhomog)
    sys._pull()
    c._dip = sys._dip.copy()
    c._coh = sys._coh.copy()
    c._handle = None
    c._host_dirty = True
    c._dev_dirty = False
    if seed is not None:
        c.rng = np.random.default_rng(seed)
        c.noise_seed = int(c.rng.integers(0, 2**63))
    c.step_counter = 0
    return c


# ---------------------------------------------------------------------------
# Host-side SU(N) helpers for setting states (src/Operators/Spin.jl:72-116)

def expected_spin_host(Z, N):
    Sx, Sy, Sz = spin_matrices((N - 1) / 2)
    Z = np.asarray(Z)
    return np.stack([np.einsum("si,ij,sj->s", Z.conj(), S, Z).real for S in (Sx, Sy, Sz)], -1)


def ket_from_dipole(d, N):
    Sx, Sy, Sz = spin_matrices((N - 1) / 2)
    w, v = np.linalg.eigh(d[0] * Sx + d[1] * Sy + d[2] * Sz)
    return v[:, np.argmax(w)]


def _normalize_ket(Z, kappa):
    return np.zeros_like(Z) if kappa == 0 else Z / np.sqrt(np.vdot(Z, Z).real / kappa)


def _normalize_dipole(s, kappa):
    return np.zeros(3) if kappa == 0 else kappa * s / np.linalg.norm(s)


# ---------------------------------------------------------------------------
# State setters (src/System/System.jl:425-570)

def set_dipole_(sys: System, direction, site):
    sys._pull()
    k = sys.site_index(site)
    kappa = sys.site_kappas()[k]
    if sys.N == 0:
        sys._dip[k] = _normalize_dipole(np.asarray(direction, float), kappa)
    else:
        Z = _normalize_ket(ket_from_dipole(np.asarray(direction, float), sys.N), kappa)
        sys._coh[k] = Z
        sys._dip[k] = expected_spin_host(Z[None], sys.N)[0]
    sys._host_dirty = True


def set_coherent_(sys: System, Z, site):
    if sys.N == 0:
        raise ValueError("Cannot set zero-length coherent state")
    sys._pull()
    k = sys.site_index(site)
    Z = _normalize_ket(np.asarray(Z, complex), sys.site_kappas()[k])
    sys._coh[k] = Z
    sys._dip[k] = expected_spin_host(Z[None], sys.N)[0]
    sys._host_dirty = True


def polarize_spins_(sys: System, direction):
    sys._pull()
    d = np.asarray(direction, float)
    kap = sys.site_kappas()
    if sys.N == 0:
        sys._dip[:] = kap[:, None] * (d / np.linalg.norm(d))
    else:
        Z = ket_from_dipole(d, sys.N)
        d = np.sqrt(np.vdot(Z, Z).real / np.where(kap > 0, kap, 1.0))[:, None]
        sys._coh[:] = np.where(kap[:, None] > 0, Z[None, :] / d, 0)
        sys._dip[:] = expected_spin_host(sys._coh, sys.N)
    sys._host_dirty = True


def randomize_spins_(sys: System):
    """Spins uniform on the sphere (dipoles) or normalised complex Gaussian
    kets (SU(N)), drawn from the system's own generator (System.jl:444-452)."""
    sys._pull()
    kap = sys.site_kappas()
    if sys.N == 0:
        v = sys.rng.standard_normal((sys.nsites, 3))
        sys._dip[:] = kap[:, None] * v / np.linalg.norm(v, axis=1, keepdims=True)
    else:
        z = (sys.rng.standard_normal((sys.nsites, sys.N)) + 1j * sys.rng.standard_normal((sys.nsites, sys.N))) / np.sqrt(2)
        d = np.sqrt((np.abs(z) ** 2).sum(1, keepdims=True) / np.where(kap > 0, kap, 1.0)[:, None])
        sys._coh[:] = np.where(kap[:, None] > 0, z / d, 0)
        sys._dip[:] = expected_spin_host(sys._coh, sys.N)
    sys._host_dirty = True


# ---------------------------------------------------------------------------
# Coupling setters

def set_field_(sys: System, B):
    """External field for every site (src/System/Interactions.jl:191-212)."""
    sys.extfield[:] = np.asarray(B, float)
    if sys.inhomog is not None:
        sys.inhomog.field[:] = np.asarray(B, float)
    sys._invalidate()


def _require_homogeneous(sys, alt):
    if sys.inhomog is not None:
        raise ValueError(f"Use `{alt}` for an inhomogeneous system.")


def _to_float_or_mat3(J):
    J = np.asarray(J, float)
    if J.ndim == 0:
        return float(J)
    J = J.reshape(3, 3)
    if np.allclose(J, J[0, 0] * np.eye(3), rtol=1e-12, atol=0):
        return float(J[0, 0])
    return J


class ModelParam:
    """A labeled coupling strength `val` and the unit couplings it scales
    (src/System/Types.jl:76-86).  pairs: Bond -> [scalar, J 3x3, B 5x5 or
    None, general (A, B) terms or None]; onsites: atom -> StevensExpansion
    (:dipole) or N x N matrix (:SUN)."""

    def __init__(self, label, val, pairs=None, onsites=None):
        self.label = label
        self.val = float(val)
        self.pairs = pairs or {}
        self.onsites = onsites or {}

    def copy(self):
        return ModelParam(self.label, self.val,
                          {b: [sc, np.array(J), None if B is None else np.array(B),
                               None if G is None else [(np.array(A), np.array(Bm)) for A, Bm in G]]
                           for b, (sc, J, B, G) in self.pairs.items()},
                          {i: (o.copy() if isinstance(o, np.ndarray) else o) for i, o in self.onsites.items()})


def _param_label(sys, param, key):
    """`param` is (label, val) like Julia's `:label => val`; None gives the
    coupling an "Unnamed" label tied to `key` (a bond or atom), so that a
    later unlabeled call on the same bond or atom replaces it
    (src/System/ModelParams.jl:9-48)."""
    if param is not None:
        label, val = param
        return str(label), float(val)
    for lab, p in sys.params.items():
        if lab.startswith("Unnamed") and getattr(p, "key", None) == key:
            return lab, 1.0
    n = sum(1 for lab in sys.params if lab.startswith("Unnamed"))
    return f"Unnamed{n + 1}", 1.0


def _replace_param(sys, label, val, pairs=None, onsites=None, key=None):
    p = ModelParam(label, val, pairs, onsites)
    p.key = key
    sys.params.pop(label, None)
    sys.params[label] = p
    repopulate_couplings_from_params_(sys)


def repopulate_couplings_from_params_(sys: System):
    """Couplings = sum over parameters of value x unit couplings
    (src/System/Interactions.jl:1-42)."""
    sys.pairs = {}
    sys.onsite = [StevensExpansion.zero() if sys.N == 0 else np.zeros((sys.N, sys.N), complex)
                  for _ in range(sys.na)]
    for p in sys.params.values():
        for b, (sc, J, B, G) in p.pairs.items():
            e = sys.pairs.setdefault(b, [0.0, np.zeros((3, 3)), None, None])
            e[0] += p.val * sc
            e[1] = e[1] + p.val * np.asarray(J)
            if B is not None:
                e[2] = p.val * np.asarray(B) if e[2] is None else e[2] + p.val * np.asarray(B)
            if G:  # TensorDecomposition: terms concatenate, scaled on the A side
                e[3] = (e[3] or []) + [(p.val * np.asarray(A), np.asarray(Bm)) for A, Bm in G]
        for i, o in p.onsites.items():
            sys.onsite[i] = sys.onsite[i] + (o.scaled(p.val) if sys.N == 0 else p.val * o)
    sys._invalidate()


def get_param(sys: System, label):
    return sys.params[str(label)].val


def get_params(sys: System, labels):
    return [get_param(sys, l) for l in labels]


def set_params_(sys: System, labels, vals):
    """Sets each labeled parameter and rebuilds the couplings
    (src/System/ModelParams.jl:130-150)."""
    if len(labels) != len(vals):
        raise ValueError("Mismatched lengths")
    for l, v in zip(labels, vals):
        if str(l) not in sys.params:
            raise KeyError(f"No parameter {l!r}")
        sys.params[str(l)].val = float(v)
    repopulate_couplings_from_params_(sys)


def set_param_(sys: System, label, val):
    set_params_(sys, [label], [val])


def param_spec(sys: System, label):
    """Coupling spec of one parameter's unit couplings on the system's bond
    list (the derivative of the couplings with respect to that parameter)."""
    p = sys.params[str(label)]
    unit = System.__new__(System)
    unit.__dict__.update({k: v for k, v in sys.__dict__.items() if not k.startswith("_")})
    unit.params = {p.label: ModelParam(p.label, 1.0, p.pairs, p.onsites)}
    unit.extfield = np.zeros_like(sys.extfield)
    unit.ewald = None
    unit.pairs = {b: [0.0, np.zeros((3, 3)), None, None] for b in sys.pairs}
    for b, (sc, J, B, G) in p.pairs.items():
        unit.pairs[b] = [sc, np.asarray(J, float), None if B is None else np.asarray(B, float), G]
    unit.onsite = [StevensExpansion.zero() if sys.N == 0 else np.zeros((sys.N, sys.N), complex)
                   for _ in range(sys.na)]
    for i, o in p.onsites.items():
        unit.onsite[i] = o
    return unit.spec()


def set_exchange_(sys: System, J, bond: Bond, biquad=0.0, propagate=None, param=None):
    """Exchange S_i.J S_j (+ biquad (S_i.S_j)^2) on `bond` and its equivalents
    (src/System/PairExchange.jl:379-385, 322-340, 185-260).  `param` =
    (label, val) scales the coupling by a labeled parameter (the reference's
    trailing `:label => val`).

    Scalar J propagates to every bond of the same distance shell (exact for
    Heisenberg couplings).  Matrix J applies to `bond` in every cell (pass
    `propagate=[bonds...]` with pre-rotated matrices for symmetry partners)."""
    _require_homogeneous(sys, "set_exchange_at_")
    if bond.i == bond.j and tuple(bond.n) == (0, 0, 0):
        raise ValueError("Bond must connect different atoms.")
    Jv = _to_float_or_mat3(J)
    scalar = 0.0
    biquad = float(biquad)
    if biquad != 0.0:
        # adapt_for_biquad (PairExchange.jl:322-340): (Si.Sj)^2 = Qi.g Qj - Si.Sj/2
        # + s1(s1+1)s2(s2+1)/3 for quantum spins; Qi.g Qj + s1^2 s2^2/3 in the
        # large-s limit (:dipole_uncorrected)
        s1, s2 = sys.spins[bond.i], sys.spins[bond.j]
        if sys.mode == "dipole_uncorrected":
            scalar += biquad * s1 ** 2 * s2 ** 2 / 3
        else:
            Jv = Jv - biquad / 2 if isinstance(Jv, float) else Jv - (biquad / 2) * np.eye(3)
            scalar += biquad * s1 * (s1 + 1) * s2 * (s2 + 1) / 3
    if isinstance(Jv, float):
        targets = sys.crystal.equivalent_bonds(bond) if propagate is None else list(propagate)
    else:
        targets = [bond] if propagate is None else list(propagate)
    pairs = {}
    for b in targets:
        Jm = Jv * np.eye(3) if isinstance(Jv, float) else np.asarray(Jv)
        bq = None
        if biquad != 0.0:
            f = biquad
            if sys.mode == "dipole":
                f *= rcs_factors(sys.spins[b.i])[2] * rcs_factors(sys.spins[b.j])[2]
            bq = f * np.diag(SCALAR_BIQUAD_METRIC)
        pairs[b] = [scalar, Jm, bq, None]
        pairs[b.reverse()] = [scalar, Jm.T.copy(), None if bq is None else bq.T.copy(), None]
    key = ("pair", min(bond, bond.reverse(), key=lambda x: (x.i, x.j, x.n)))
    label, val = _param_label(sys, param, key)
    _replace_param(sys, label, val, pairs=pairs, key=key)


def set_pair_biquad_(sys: System, B, bond: Bond, param=None):
    """Coupling Q_i . B Q_j of Stevens quadrupoles (5x5, q descending) on one
    bond in every cell; the reverse bond carries B' (PairExchange.jl:447-468).
    In :dipole mode B is renormalised by rcs_2(s_i) rcs_2(s_j) (not in the
    large-s :dipole_uncorrected mode)."""
    _require_homogeneous(sys, "set_pair_coupling_at_")
    B = np.asarray(B, float).reshape(5, 5)
    if sys.mode == "dipole":
        B = B * rcs_factors(sys.spins[bond.i])[2] * rcs_factors(sys.spins[bond.j])[2]
    pairs = {bond: [0.0, np.zeros((3, 3)), B.copy(), None],
             bond.reverse(): [0.0, np.zeros((3, 3)), B.T.copy(), None]}
    key = ("biquad", min(bond, bond.reverse(), key=lambda x: (x.i, x.j, x.n)))
    label, val = _param_label(sys, param, key)
    _replace_param(sys, label, val, pairs=pairs, key=key)


def to_product_space(Si, Sj):
    """Operators of two sites lifted to the pair's product space:
    A (x) I and I (x) B (to_product_space, src/Operators/TensorOperators.jl)."""
    Ni, Nj = Si[0].shape[0], Sj[0].shape[0]
    return ([np.kron(a, np.eye(Nj)) for a in Si], [np.kron(np.eye(Ni), b) for b in Sj])


def set_pair_coupling_(sys: System, op, bond: Bond, extract_parts=True, param=None):
    """General pair coupling on `bond` in every cell (set_pair_coupling!,
    src/System/PairExchange.jl:274-312): `op` is a Hermitian matrix on the
    pair's product space C^Ni (x) C^Nj, or a function of the lifted spin
    matrices (Si, Sj).  It is split into a scalar, bilinear and biquadratic
    part (extract_parts) and a general remainder sum_k A_k (x) B_k, whose
    bond energy is sum_k Re<Zi|A_k|Zi> Re<Zj|B_k|Zj>; the reverse bond
    carries the transposed parts and (B_k, A_k).  The general remainder needs
    :SUN mode.  No symmetry propagation: call once per bond (with the
    operator rotated as required)."""
    _require_homogeneous(sys, "set_pair_coupling_at_")
    if sys.mode == "dipole_uncorrected":
        raise ValueError("General coupling not currently supported for mode :dipole_uncorrected. "
                         "Use set_exchange! with option `biquad` for scalar biquadratic.")
    if bond.i == bond.j and tuple(bond.n) == (0, 0, 0):
        raise ValueError("Bond must connect different atoms.")
    si, sj = sys.spins[bond.i], sys.spins[bond.j]
    Ni, Nj = int(round(2 * si + 1)), int(round(2 * sj + 1))
    if callable(op):
        op = op(*to_product_space(spin_matrices(si), spin_matrices(sj)))
    op = np.asarray(op, complex)
    if op.shape != (Ni * Nj, Ni * Nj):
        raise ValueError("Invalid matrix size.")
    if not np.allclose(op, op.conj().T, rtol=0, atol=1e-12):
        raise ValueError("Operator is not Hermitian")
    if Ni == Nj == 2:
        raise ValueError("Use set_exchange_ instead. Quantum spin 1/2 allows only bilinear coupling.")
    scalar, J, B, G = decompose_pair_coupling(op, Ni, Nj, extract_parts=extract_parts)
    if G and sys.mode != "SUN":
        raise ValueError("Invalid pair coupling. In dipole mode, the most general allowed form is "
                         "(Si, Sj) -> Si'*J*Sj + [(Si'*K1*Si)*(Sj'*K2*Sj) + ...] where J is any 3x3 matrix, "
                         "while K1, K2 must be Hermitian and traceless.")
    if B is not None and sys.mode == "dipole":
        B = B * rcs_factors(si)[2] * rcs_factors(sj)[2]
    G = [(np.asarray(A, complex), np.asarray(Bm, complex)) for A, Bm in G] or None
    pairs = {bond: [scalar, J.copy(), None if B is None else B.copy(), G],
             bond.reverse(): [scalar, J.T.copy(), None if B is None else B.T.copy(),
                              None if G is None else [(Bm.copy(), A.copy()) for A, Bm in G]]}
    key = ("pair", min(bond, bond.reverse(), key=lambda x: (x.i, x.j, x.n)))
    label, val = _param_label(sys, param, key)
    _replace_param(sys, label, val, pairs=pairs, key=key)


def set_onsite_coupling_(sys: System, op, i: int, param=None):
    """Onsite anisotropy on atom i (src/System/OnsiteCoupling.jl:1-23, 159-205).
    `op` is an N x N matrix in the spin-s irrep, or a function of the spin
    matrices; `param` = (label, val) as in set_exchange_.  No symmetry
    propagation: call once per atom (with the operator rotated as required)."""
    _require_homogeneous(sys, "set_onsite_coupling_at_")
    unit = onsite_coupling(sys, op, sys.spins[i], sys.kappas[i])
    key = ("onsite", int(i))
    label, val = _param_label(sys, param, key)
    _replace_param(sys, label, val, onsites={int(i): unit}, key=key)


def onsite_coupling(sys: System, op, s, kappa):
    """The onsite coupling of an operator on a spin-s site (onsite_coupling,
    OnsiteCoupling.jl:1-44): the matrix itself in :SUN mode; in :dipole mode
    its Stevens coefficients renormalised by rcs_k(s); in
    :dipole_uncorrected mode a polynomial of the classical spin (a function
    receives spin_matrices(inf)) expanded on |S| = kappa, unrenormalised."""
    fn = callable(op) and not isinstance(op, SpinPolynomial)  # a SpinPolynomial evaluates when called
    if sys.mode == "dipole_uncorrected":
        if fn:
            op = op(spin_matrices(np.inf))
        if not isinstance(op, SpinPolynomial) and not np.ndim(op) == 0:
            raise ValueError("System with mode :dipole_uncorrected requires a symbolic operator.")
        return StevensExpansion.from_coefficients(polynomial_to_stevens_coefficients(op, kappa))
    if fn:
        op = op(spin_matrices(s))
    if isinstance(op, SpinPolynomial):
        raise ValueError("Symbolic operator only valid for system with mode :dipole_uncorrected.")
    A = np.asarray(op, complex)
    N = int(round(2 * s + 1))
    if A.shape != (N, N):
        raise ValueError("Invalid matrix size.")
    if not np.allclose(A, A.conj().T):
        raise ValueError("Operator is not Hermitian")
    if N == 2:
        if sys.mode == "dipole":
            raise ValueError("Consider mode :dipole_uncorrected. Onsite coupling is trivial for quantum spin 1/2.")
        raise ValueError("Onsite coupling is trivial for quantum spin 1/2.")
    if sys.mode == "SUN":
        return A.copy()
    c = matrix_to_stevens_coefficients((A + A.conj().T) / 2)
    lam = rcs_factors(s)
    return StevensExpansion.from_coefficients({k: lam[k] * c[k] for k in range(7)})


def set_stevens_expansion_(sys: System, stvexp: StevensExpansion, i: int, param=None):
    """Sets already-renormalised classical Stevens coefficients (:dipole)."""
    _require_homogeneous(sys, "set_onsite_coupling_at_")
    if sys.mode == "SUN":
        raise ValueError("Stevens expansions apply to :dipole mode")
    key = ("onsite", int(i))
    label, val = _param_label(sys, param, key)
    _replace_param(sys, label, val, onsites={int(i): stvexp}, key=key)


def enable_dipole_dipole_(sys: System, mu0_muB2=VACUUM_PERMEABILITY, demag=1 / 3):
    """Long-range dipole-dipole by Ewald summation (Interactions.jl:182-189)."""
    demag = np.asarray(demag, float)
    if demag.ndim == 0:
        demag = demag * np.eye(3)
    sys.ewald = (float(mu0_muB2), demag.reshape(3, 3).copy())
    sys._invalidate()


def set_spin_rescaling_(sys: System, kappas):
    """Per-sublattice |S| (dipole) or |Z|^2 (SU(N)) rescaling (System.jl:573-)."""
    sys._pull()
    sys.kappas = np.asarray(kappas, float).reshape(sys.na).copy()
    if sys.inhomog is not None:  # every non-vacant site of each sublattice
        kap = np.repeat(sys.kappas, sys.ncells)
        sys.inhomog.kappa = np.where(sys.inhomog.kappa == 0, 0.0, kap)
    sys._invalidate()
