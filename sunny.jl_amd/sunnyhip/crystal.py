"""Crystal cells and bonds for the lattices the spin-dynamics path runs on.

Sunny builds crystals with full space-group symmetry (src/Symmetry/, out of
scope here: it is host-side setup).  This module keeps only what the hot path
consumes: lattice vectors, fractional positions, and directed bonds
``Bond(i, j, n)`` (src/Symmetry/Bond.jl) with global displacements.
Symmetry propagation of couplings is replaced by distance shells, which is
exact for isotropic (Heisenberg) couplings because all bonds of one distance
shell that are related by symmetry carry the same scalar J.
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np


def lattice_vectors(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """Lattice vectors as columns, as src/Symmetry/LatticeUtils.jl:40-56."""
    sg, cg = np.sin(np.radians(gamma)), np.cos(np.radians(gamma))
    cb, ca = np.cos(np.radians(beta)), np.cos(np.radians(alpha))
    v1 = [a, 0.0, 0.0]
    v2 = [b * cg, b * sg, 0.0]
    v3x = c * cb
    v3y = c / sg * (ca - cb * cg)
    v3z = c / sg * np.sqrt(sg**2 - ca**2 - cb**2 + 2 * ca * cb * cg)
    return np.array([v1, v2, [v3x, v3y, v3z]], dtype=float).T.copy()


@dataclasses.dataclass(frozen=True)
class Bond:
    """Directed bond from atom i to atom j displaced by n cells (0-based atoms)."""

    i: int
    j: int
    n: tuple

    def reverse(self) -> "Bond":
        return Bond(self.j, self.i, tuple(-x for x in self.n))


def bond_parity(bond: Bond) -> bool:
    """`isculled` flag: true for the lexicographically positive direction
    (src/System/PairExchange.jl:11-15)."""
    delta = (bond.j - bond.i,) + tuple(bond.n)
    assert delta != (0, 0, 0, 0)
    return delta > (0, 0, 0, 0)


class Crystal:
    """Lattice vectors (columns) and fractional atom positions."""

    def __init__(self, latvecs, positions):
        self.latvecs = np.asarray(latvecs, dtype=float).reshape(3, 3)
        self.positions = np.asarray(positions, dtype=float).reshape(-1, 3)

    @property
    def natoms(self) -> int:
        return len(self.positions)

    def displacement(self, bond: Bond) -> np.ndarray:
        d = self.positions[bond.j] + np.asarray(bond.n) - self.positions[bond.i]
        return self.latvecs @ d

    def distance(self, bond: Bond) -> float:
        return float(np.linalg.norm(self.displacement(bond)))

    def bonds_within(self, rmax: float, nmax: int = 3):
        """Every directed bond (i, j, n) with 0 < |r| <= rmax."""
        out = []
        rng = range(-nmax, nmax + 1)
        for i in range(self.natoms):
            for j in range(self.natoms):
                for n in itertools.product(rng, rng, rng):
                    if i == j and n == (0, 0, 0):
                        continue
                    b = Bond(i, j, n)
                    if self.distance(b) <= rmax + 1e-9:
                        out.append(b)
        return out

    def equivalent_bonds(self, bond: Bond, tol: float = 1e-8):
        """Bonds of the same distance shell as `bond`, from every atom."""
        d = self.distance(bond)
        nmax = int(np.ceil(d / min(np.linalg.norm(self.latvecs, axis=0)))) + 1
        return [b for b in self.bonds_within(d + tol, nmax) if abs(self.distance(b) - d) < tol]


def cubic_crystal(a=1.0) -> Crystal:
    return Crystal(lattice_vectors(a, a, a, 90, 90, 90), [[0, 0, 0]])


def pyrochlore_crystal(a=1.0) -> Crystal:
    """Fd-3m (227), Wyckoff 16c in origin choice 2: the four tetrahedron
    vertices (0,0,0), (3/4,1/4,1/2), (1/4,1/2,3/4), (1/2,3/4,1/4) plus the
    F-centring translations; 16 sites in the conventional cubic cell."""
    base = np.array([[0, 0, 0], [3, 1, 2], [1, 2, 3], [2, 3, 1]], dtype=float) / 4
    fcc = np.array([[0, 0, 0], [0, 2, 2], [2, 0, 2], [2, 2, 0]], dtype=float) / 4
    pos = [np.mod(b + t, 1.0) for t in fcc for b in base]
    return Crystal(lattice_vectors(a, a, a, 90, 90, 90), pos)


def diamond_crystal(a=1.0) -> Crystal:
    """Same site list as Sunny.diamond_crystal (src/Symmetry/Crystal.jl)."""
    pos = np.array([[0, 0, 0], [2, 2, 0], [1, 1, 1], [3, 3, 1], [2, 0, 2], [0, 2, 2], [3, 1, 3], [1, 3, 3]]) / 4
    return Crystal(lattice_vectors(a, a, a, 90, 90, 90), pos)


def triangular_crystal(a=1.0, c=1.0) -> Crystal:
    return Crystal(lattice_vectors(a, a, c, 90, 90, 120), [[0, 0, 0]])


def pyrochlore_local_axes(cryst: Crystal) -> np.ndarray:
    """Local <111> axis of every pyrochlore site: the direction from a
    tetrahedron centre to the site (sign irrelevant for (S.n)^2)."""
    a = np.linalg.norm(cryst.latvecs[:, 0])
    dnn = a * np.sqrt(2) / 4
    axes = []
    for i in range(cryst.natoms):
        nbrs = [cryst.displacement(b) for b in cryst.bonds_within(dnn + 1e-6, 1) if b.i == i]
        assert len(nbrs) == 6, "pyrochlore site must have 6 nearest neighbours"
        d0 = nbrs[0]
        tet = [d0] + [d for d in nbrs[1:] if abs(np.linalg.norm(d - d0) - dnn) < 1e-6]
        assert len(tet) == 3
        n = -np.sum(tet, axis=0)
        axes.append(n / np.linalg.norm(n))
    return np.array(axes)
