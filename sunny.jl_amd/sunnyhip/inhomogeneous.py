"""Inhomogeneous systems: per-site interactions, vacancies, per-site fields
(src/System/Interactions.jl:59-100, 210-260; PairExchange.jl:390-500;
OnsiteCoupling.jl:206-230).

`to_inhomogeneous(sys)` returns a copy whose interactions are kept per site
(Sunny's Array{Interactions,4}); `set_exchange_at_`, `set_pair_coupling_at_`,
`set_onsite_coupling_at_`, `set_vacancy_at_` and `set_field_at_` then edit
single sites and bonds.

On the device an inhomogeneous system is the homogeneous system whose
chemical cell is the whole supercell: every site is its own sublattice (in
Julia's site order, so state arrays keep their layout), the lattice vectors
are the supercell's, and a bond (c, i) -> (c + n, j) becomes a bond between
the two sites' sublattices whose cell offset is the supercell image
floor((c + n) / L).  The kernels then read per-sublattice tables that are
per-site tables.  The Ewald term then couples every pair of sites directly
(one wavevector), so its size is bounded by the contraction's LDS.
"""
from __future__ import annotations

import copy

import numpy as np

from .crystal import Bond, bond_parity
from .operators import StevensExpansion, decompose_pair_coupling, rcs_factors, spin_matrices
from .system import SCALAR_BIQUAD_METRIC, System, clone_system, onsite_coupling, to_product_space


class SiteInteractions:
    """Per-site interactions of an inhomogeneous system: onsite[s] and
    pairs[s] (Bond -> [scalar, J, B or None, general terms or None]) with
    the bond written from site s; kappa, g, field per site."""

    def __init__(self, sys: System):
        nc, ns = sys.ncells, sys.nsites
        atom = np.arange(ns) // nc
        self.onsite = [copy.deepcopy(sys.onsite[i]) for i in atom]
        self.pairs = []
        for s in range(ns):
            i = atom[s]
            self.pairs.append({b: [e[0], np.array(e[1]), None if e[2] is None else np.array(e[2]),
                                   None if e[3] is None else list(e[3])]
                               for b, e in sys.pairs.items() if b.i == i})
        self.kappa = sys.kappas[atom].astype(float)
        self.g = sys.gs[atom].copy()
        self.field = sys.extfield[atom].copy()


def is_homogeneous(sys: System) -> bool:
    return getattr(sys, "inhomog", None) is None


def _require_inhomogeneous(sys):
    if is_homogeneous(sys):
        raise ValueError("Use `to_inhomogeneous` first.")


def to_inhomogeneous(sys: System) -> System:
    """A copy that allows per-site interactions (Interactions.jl:80-100);
    labelled parameters are dropped, as in the reference."""
    if not is_homogeneous(sys):
        raise ValueError("System is already inhomogeneous.")
    ret = clone_system(sys)
    ret.inhomog = SiteInteractions(sys)
    ret.params = {}
    ret._invalidate()
    return ret


def _site_tuple(sys, site):
    c1, c2, c3, i = (int(x) for x in site)
    L1, L2, L3 = sys.dims
    if not (0 <= c1 < L1 and 0 <= c2 < L2 and 0 <= c3 < L3 and 0 <= i < sys.na):
        raise ValueError(f"site {site} out of range")
    return c1, c2, c3, i


def _is_vacant(sys, k):
    return sys.inhomog.kappa[k] == 0


def sites_to_internal_bond(sys: System, site1, site2, offset=None) -> Bond:
    """The bond joining two sites (PairExchange.jl:390-428): with `offset`
    it must agree with the cells modulo dims; otherwise the shortest of the
    periodic wrappings, if it is shorter than the next by a factor 4."""
    c1, i1 = np.array(site1[:3]), site1[3]
    c2, i2 = np.array(site2[:3]), site2[3]
    dims = np.array(sys.dims)
    n0 = c2 - c1
    if offset is not None:
        offset = np.asarray(offset, int)
        if np.all(np.mod(offset - n0, dims) == 0):
            return Bond(i1, i2, tuple(int(x) for x in offset))
        raise ValueError(f"Cells {tuple(c1)} and {tuple(c2)} are not compatible with the offset {tuple(offset)} "
                         f"for a system with dimensions {tuple(dims)}.")
    cands = [Bond(i1, i2, tuple(int(x) for x in n0 + dims * np.array((a, b, c))))
             for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    dist = np.array([sys.crystal.distance(b) for b in cands])
    perm = np.argsort(dist, kind="stable")
    if 4 * dist[perm[0]] < dist[perm[1]] - 1e-12:
        return cands[perm[0]]
    raise ValueError(f"Ambiguous offset vector. Possibilities include {cands[perm[0]].n} and {cands[perm[1]].n}. "
                     "Try using a bigger system size, or pass an explicit offset.")


def _set_pair_at(sys, entry, entry_rev, site1, site2, offset):
    _require_inhomogeneous(sys)
    site1, site2 = _site_tuple(sys, site1), _site_tuple(sys, site2)
    k1, k2 = sys.site_index(site1), sys.site_index(site2)
    if _is_vacant(sys, k1) or _is_vacant(sys, k2):
        raise ValueError("Cannot couple vacant site")
    bond = sites_to_internal_bond(sys, site1, site2, offset)
    if bond.i == bond.j and tuple(bond.n) == (0, 0, 0):
        raise ValueError("Bond must connect different atoms.")
    zero = (entry[0] == 0 and not np.any(entry[1]) and entry[2] is None and not entry[3])
    for k, b, e in ((k1, bond, entry), (k2, bond.reverse(), entry_rev)):
        sys.inhomog.pairs[k].pop(b, None)  # replace_coupling_in_list!
        if not zero:
            sys.inhomog.pairs[k][b] = e
    sys._invalidate()


def set_exchange_at_(sys: System, J, site1, site2, biquad=0.0, offset=None):
    """Exchange on the single bond joining two sites, both directions
    (set_exchange_at!, PairExchange.jl:464-500 with adapt_for_biquad)."""
    Jm = np.asarray(J, float) * np.eye(3) if np.ndim(J) == 0 else np.asarray(J, float).reshape(3, 3)
    scalar, bq = 0.0, None
    biquad = float(biquad)
    s1, s2 = sys.spins[site1[3]], sys.spins[site2[3]]
    if biquad != 0.0:
        if sys.mode == "dipole_uncorrected":
            scalar += biquad * s1 ** 2 * s2 ** 2 / 3
        else:
            Jm = Jm - (biquad / 2) * np.eye(3)
            scalar += biquad * s1 * (s1 + 1) * s2 * (s2 + 1) / 3
        f = biquad
        if sys.mode == "dipole":
            f *= rcs_factors(s1)[2] * rcs_factors(s2)[2]
        bq = f * np.diag(SCALAR_BIQUAD_METRIC)
    _set_pair_at(sys, [scalar, Jm, bq, None], [scalar, Jm.T.copy(), None if bq is None else bq.T.copy(), None],
                 site1, site2, offset)


def set_pair_coupling_at_(sys: System, op, site1, site2, offset=None, extract_parts=True):
    """General pair operator on the single bond joining two sites
    (set_pair_coupling_at!, PairExchange.jl:502-530)."""
    if sys.mode == "dipole_uncorrected":
        raise ValueError("General coupling not currently supported for mode :dipole_uncorrected. "
                         "Use set_exchange! with option `biquad` for scalar biquadratic.")
    s1, s2 = sys.spins[site1[3]], sys.spins[site2[3]]
    N1, N2 = int(round(2 * s1 + 1)), int(round(2 * s2 + 1))
    if callable(op):
        op = op(*to_product_space(spin_matrices(s1), spin_matrices(s2)))
    op = np.asarray(op, complex)
    if op.shape != (N1 * N2, N1 * N2) or not np.allclose(op, op.conj().T, rtol=0, atol=1e-12):
        raise ValueError("Invalid or non-Hermitian pair operator")
    scalar, J, B, G = decompose_pair_coupling(op, N1, N2, extract_parts=extract_parts)
    if G and sys.mode != "SUN":
        raise ValueError("Invalid pair coupling. In dipole mode, the most general allowed form is "
                         "(Si, Sj) -> Si'*J*Sj + [(Si'*K1*Si)*(Sj'*K2*Sj) + ...]")
    if B is not None and sys.mode == "dipole":
        B = B * rcs_factors(s1)[2] * rcs_factors(s2)[2]
    G = [(np.asarray(A, complex), np.asarray(Bm, complex)) for A, Bm in G] or None
    _set_pair_at(sys, [scalar, J.copy(), B, G],
                 [scalar, J.T.copy(), None if B is None else B.T.copy(),
                  None if G is None else [(Bm.copy(), A.copy()) for A, Bm in G]], site1, site2, offset)


def set_onsite_coupling_at_(sys: System, op, site):
    """Onsite operator of one site (set_onsite_coupling_at!,
    OnsiteCoupling.jl:206-225; the renormalisation of onsite_coupling,
    :1-23)."""
    _require_inhomogeneous(sys)
    site = _site_tuple(sys, site)
    k = sys.site_index(site)
    if _is_vacant(sys, k):
        raise ValueError("Cannot couple vacant site")
    sys.inhomog.onsite[k] = onsite_coupling(sys, op, sys.spins[site[3]], sys.inhomog.kappa[k])
    sys._invalidate()


def set_vacancy_at_(sys: System, site):
    """Makes one site nonmagnetic (set_vacancy_at!, Interactions.jl:237-260):
    kappa = 0, zero state, no onsite term, and its bonds removed from both
    ends."""
    _require_inhomogeneous(sys)
    site = _site_tuple(sys, site)
    sys._pull()
    k = sys.site_index(site)
    H = sys.inhomog
    H.kappa[k] = 0.0
    sys._dip[k] = 0
    sys._coh[k] = 0
    H.onsite[k] = StevensExpansion.zero() if sys.N == 0 else np.zeros((sys.N, sys.N), complex)
    L = np.array(sys.dims)
    c = np.array(site[:3])
    for b in list(H.pairs[k]):
        cj = np.mod(c + np.array(b.n), L)
        k2 = sys.site_index((*cj, b.j))
        H.pairs[k2].pop(b.reverse(), None)
    H.pairs[k].clear()
    sys._invalidate()
    sys._host_dirty = True


def set_field_at_(sys: System, B, site):
    """External field (energy units) at one site (set_field_at!,
    Interactions.jl:227-229)."""
    _require_inhomogeneous(sys)
    site = _site_tuple(sys, site)
    sys.inhomog.field[sys.site_index(site)] = np.asarray(B, float).reshape(3)
    sys._invalidate()


def inhomogeneous_spec(sys: System, precision=None, dims_list=None) -> dict:
    """The snh_create spec of an inhomogeneous system: the supercell as the
    chemical cell, every site a sublattice (see the module docstring)."""
    H = sys.inhomog
    L = np.array(sys.dims)
    for d in (dims_list or [sys.dims]):
        if tuple(d) != tuple(sys.dims):
            raise ValueError("replicas of an inhomogeneous system share its dims")
    R = len(dims_list) if dims_list else 1
    nc, ns, N = sys.ncells, sys.nsites, sys.N

    def cell_of(s):
        c = s % nc
        return np.array([c % L[0], (c // L[0]) % L[1], c // (L[0] * L[1])])

    stev = np.zeros((ns, 28))
    kmax = np.zeros(ns, np.int32)
    lam = np.zeros((ns, max(N, 1), max(N, 1), 2))
    for s in range(ns):
        if N == 0:
            stev[s] = H.onsite[s].packed()
            kmax[s] = H.onsite[s].kmax
        else:
            lam[s, :, :, 0] = H.onsite[s].real
            lam[s, :, :, 1] = H.onsite[s].imag
    rows, bil, bqs, hbq, scs, gen_start, gen = [], [], [], [], [], [], []
    for s in range(ns):
        c = cell_of(s)
        for b, (sc, J, B, G) in sorted(H.pairs[s].items(), key=lambda kv: (bond_parity(kv[0]), kv[0].j, kv[0].n)):
            cj = c + np.array(b.n)
            t = sys.site_index((*np.mod(cj, L), b.j))
            image = np.floor_divide(cj, L)
            rows.append((s, t, *image, int(bond_parity(b))))
            bil.append(np.asarray(J, float).reshape(9))
            bqs.append(np.zeros(25) if B is None else np.asarray(B, float).reshape(25))
            hbq.append(0 if B is None else 1)
            scs.append(sc)
            gen_start.append(len(gen))
            gen.extend(G or [])
    nb = len(rows)
    gen_start.append(len(gen))
    gen_A = gen_B = None
    if gen:
        Nm = max(N, 1)
        gen_A = np.ascontiguousarray(np.stack([np.stack([np.asarray(A).real, np.asarray(A).imag], -1)
                                               for A, _ in gen]).reshape(len(gen), Nm, Nm, 2))
        gen_B = np.ascontiguousarray(np.stack([np.stack([np.asarray(B).real, np.asarray(B).imag], -1)
                                               for _, B in gen]).reshape(len(gen), Nm, Nm, 2))
    pos = np.array([(cell_of(s) + sys.crystal.positions[s // nc]) / L for s in range(ns)])
    ew = sys.ewald
    return {
        "na": ns, "N": N, "precision": int(precision or sys.precision),
        "kappa": np.ascontiguousarray(H.kappa, float), "g": np.ascontiguousarray(H.g.reshape(ns, 9)),
        "field": np.ascontiguousarray(H.field, float), "stevens": stev, "kmax": kmax,
        "Lambda": np.ascontiguousarray(lam), "nbonds": nb,
        "bond": np.ascontiguousarray(np.array(rows, np.int32).reshape(-1, 6)) if nb else np.zeros((1, 6), np.int32),
        "bilin": np.ascontiguousarray(np.array(bil).reshape(-1, 9)) if nb else np.zeros((1, 9)),
        "biquad": np.ascontiguousarray(np.array(bqs).reshape(-1, 25)) if nb else np.zeros((1, 25)),
        "has_biquad": np.array(hbq, np.int32) if nb else np.zeros(1, np.int32),
        "scalar": np.array(scs, float) if nb else np.zeros(1),
        "ewald": int(ew is not None), "mu0_muB2": float(ew[0]) if ew else 0.0,
        "demag": np.asarray(ew[1] if ew else np.zeros((3, 3)), float).reshape(9),
        "latvecs": (sys.crystal.latvecs * L[None, :]).reshape(9).copy(),
        "positions": np.ascontiguousarray(pos, float),
        "nreplicas": R, "dims": np.ones((R, 3), np.int32),
        "layout": int(getattr(sys, "layout", 0)),
        "gen_start": np.array(gen_start, np.int32) if gen else None, "gen_A": gen_A, "gen_B": gen_B,
    }
