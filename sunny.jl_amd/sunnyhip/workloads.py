"""Builders for the spin models the benchmarks and parity tests run.

Each returns a fully coupled `System` with random initial spins drawn from
`seed`.  Lattices and couplings follow the Sunny idioms they replace:
`System(crystal, [i => Moment(s, g)], mode; dims, seed)`, `set_exchange!`,
`set_onsite_coupling!`, `enable_dipole_dipole!`.
"""
from __future__ import annotations

import numpy as np

from .crystal import Bond, cubic_crystal, pyrochlore_crystal, pyrochlore_local_axes, triangular_crystal
from .system import (VACUUM_PERMEABILITY, Moment, System, enable_dipole_dipole_, randomize_spins_, set_exchange_,
                     set_onsite_coupling_)


def bond_at_distance(cryst, dist, i=0, tol=1e-6):
    """A directed bond from atom i of the given length (first found)."""
    for b in cryst.bonds_within(dist + tol, 2):
        if b.i == i and abs(cryst.distance(b) - dist) < tol:
            return b
    raise ValueError(f"no bond of length {dist}")


def cubic_model(L=16, J=1.0, D=-0.1, s=1.0, g=2.0, seed=0, precision=64, dims=None, params=False):
    """Simple cubic, a = 1 A, one site per cell: nearest-neighbour Heisenberg
    J and onsite D*Sz^2.  With params=True the two strengths are the labeled
    parameters "J" and "D" (set_exchange!(sys, 1.0, bond, :J => J))."""
    cryst = cubic_crystal(1.0)
    sys = System(cryst, [(0, Moment(s=s, g=g))], "dipole", dims=dims or (L, L, L), seed=seed, precision=precision)
    if params:
        set_exchange_(sys, 1.0, Bond(0, 0, (1, 0, 0)), param=("J", J))
        set_onsite_coupling_(sys, lambda S: S[2] @ S[2], 0, param=("D", D))
    else:
        set_exchange_(sys, J, Bond(0, 0, (1, 0, 0)))
        if D != 0:
            set_onsite_coupling_(sys, lambda S: D * S[2] @ S[2], 0)
    randomize_spins_(sys)
    return sys


def pyrochlore_model(L=32, J1=1.0, J2=0.1, A=-0.05, s=1.5, g=2.0, a=10.0, seed=0, precision=64, dims=None,
                     dipolar=False, mu0_muB2=VACUUM_PERMEABILITY, demag=1 / 3):
    """Pyrochlore (Fd-3m 16c, conventional cell a) with Heisenberg J1, J2 and
    easy-axis A*(S.n_i)^2 along each site's local <111> axis; optionally the
    Ewald dipole-dipole interaction."""
    cryst = pyrochlore_crystal(a)
    na = cryst.natoms
    sys = System(cryst, [(i, Moment(s=s, g=g)) for i in range(na)], "dipole", dims=dims or (L, L, L), seed=seed,
                 precision=precision)
    if J1 != 0:
        set_exchange_(sys, J1, bond_at_distance(cryst, a * np.sqrt(2) / 4))
    if J2 != 0:
        set_exchange_(sys, J2, bond_at_distance(cryst, a * np.sqrt(6) / 4))
    if A != 0:
        axes = pyrochlore_local_axes(cryst)
        for i in range(na):
            n = axes[i]
            set_onsite_coupling_(sys, lambda S, n=n: A * (n[0] * S[0] + n[1] * S[1] + n[2] * S[2]) @
                                 (n[0] * S[0] + n[1] * S[1] + n[2] * S[2]), i)
    if dipolar:
        enable_dipole_dipole_(sys, mu0_muB2, demag)
    randomize_spins_(sys)
    return sys


def triangular_sun_model(L1=32, L2=None, J1=-0.24, J2=0.05, J3=0.2, D=2.2, s=1.0, g=2.0, seed=0, precision=32):
    """Triangular lattice (a = 4 A, c = 6.75 A) in :SUN mode with J1, J2, J3
    selected by bond vector and easy-plane D*Sz^2."""
    cryst = triangular_crystal(4.0, 6.75)
    L2 = L1 if L2 is None else L2
    sys = System(cryst, [(0, Moment(s=s, g=g))], "SUN", dims=(L1, L2, 1), seed=seed, precision=precision)
    for J, n in ((J1, (1, 0, 0)), (J2, (1, 2, 0)), (J3, (2, 0, 0))):
        if J != 0:
            set_exchange_(sys, J, Bond(0, 0, n))
    if D != 0:
        set_onsite_coupling_(sys, lambda S: D * S[2] @ S[2], 0)
    randomize_spins_(sys)
    return sys


def sun3_tempering_batch(nreplicas=64, Lmin=32, Lmax=256, kT_min=0.1, kT_max=2.0, seed=0, precision=32):
    """Parallel-tempering set of SU(3) triangular replicas of uneven size:
    L_r drawn uniformly from {Lmin..Lmax} with the seed, L_r x L_r x 1 cells,
    temperatures geometric from kT_min to kT_max.  Returns (systems, kTs)."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(Lmin, Lmax + 1, size=nreplicas)
    systems = [triangular_sun_model(L1=int(L), seed=seed + 1 + r, precision=precision) for r, L in enumerate(Ls)]
    kTs = kT_min * (kT_max / kT_min) ** (np.arange(nreplicas) / max(1, nreplicas - 1))
    return systems, kTs
