"""Parallel tempering over one device replica batch.

Host mirror of Sunny's ParallelTempering (src/MonteCarlo/ParallelTempering.jl):
one `ReplicaBatch` holds every replica; rank k of the temperature schedule
samples replica `system_ids[k]`.  Each exchange interval advances all
replicas with one set of Langevin launches at their current temperatures,
then reads the R FP64 energies of the batch (one device reduction) and
attempts the exchanges of neighbouring ranks on the host.
"""
from __future__ import annotations

import math

import numpy as np

from .integrators import Langevin, ReplicaBatch, energy, step_
from .samplers import LocalSampler
from .system import clone_system


class ParallelTempering:
    """ParallelTempering(system, sampler, kT_sched) (ParallelTempering.jl:18-36).

    `sampler` is a Langevin integrator or a LocalSampler (its kT is replaced
    per rank).  With a LocalSampler each replica's energy is kept from its
    initial device energy plus the energy changes of the accepted trials, as
    the reference keeps it in the samplers' dE (ParallelTempering.jl:27-33,
    44-45, 63-66; tracking it per replica id instead of per rank makes the
    reference's swap of dE with the labels implicit), and every round makes
    exch_interval calls of nsweeps = 1.  With `systems` given, those
    replicas are used (they may differ in size: a device extension; Sunny
    clones one system).  Ranks and replica ids are 0-based here.  The
    acceptance draws come from one generator seeded by `seed` (Sunny draws
    from the rng of the first system of each pair)."""

    def __init__(self, system, sampler, kT_sched, *, systems=None, precision=None, seed=0):
        if not isinstance(sampler, (Langevin, LocalSampler)):
            raise TypeError("the device ensemble steps Langevin or LocalSampler samplers")
        if isinstance(sampler, LocalSampler):
            sampler = sampler.copy()
            sampler.nsweeps = 1.0
        self.kT_sched = np.asarray(kT_sched, float)
        self.n_replicas = len(self.kT_sched)
        self.sampler = sampler
        self.systems = list(systems) if systems is not None else \
            [clone_system(system, seed=int(system.noise_seed % (2**31)) + r) for r in range(self.n_replicas)]
        if len(self.systems) != self.n_replicas:
            raise ValueError("one system per temperature")
        prec = precision if precision is not None else self.systems[0].precision
        self.batch = ReplicaBatch(self.systems, precision=prec)
        self.system_ids = np.arange(self.n_replicas)
        self.n_accept = np.zeros(self.n_replicas, np.int64)
        self.n_exch = np.zeros(self.n_replicas, np.int64)
        self.rng = np.random.default_rng(seed)
        # LocalSampler: energy of every replica id, updated by accepted trials
        self.E = np.asarray(energy(self.batch), float).copy() if isinstance(sampler, LocalSampler) else None

    def replica_kTs(self):
        """Temperature of every replica (indexed by replica id)."""
        kT = np.empty(self.n_replicas)
        kT[self.system_ids] = self.kT_sched
        return kT

    def replica_exchange_(self, exch_start, energies=None):
        """replica_exchange!(PT, exch_start) (ParallelTempering.jl:39-73):
        ranks exch_start, exch_start + 2, ... try to swap with the next rank
        with probability min(1, exp(ln P)), ln P = (E1 - E2)(1/kT1 - 1/kT2).
        `energies` (by replica id) defaults to the device energies (with a
        LocalSampler: the tracked ones)."""
        if energies is None:
            energies = self.E if self.E is not None else energy(self.batch)
        E = np.asarray(energies, float)
        ids, kT = self.system_ids, self.kT_sched
        for rank in range(exch_start, self.n_replicas - 1, 2):
            r1, r2 = rank, rank + 1
            E1, E2 = E[ids[r1]], E[ids[r2]]
            dS1 = E1 / kT[r1] - E2 / kT[r1]
            dS2 = E2 / kT[r2] - E1 / kT[r2]
            ln_P = dS1 + dS2
            if ln_P >= 0 or self.rng.random() < math.exp(ln_P):
                ids[r1], ids[r2] = ids[r2], ids[r1]
                self.n_accept[r1] += 1
            self.n_exch[r1] += 1

    def step_ensemble_(self, nsteps, exch_interval):
        """step_ensemble!(PT, nsteps, exch_interval) (ParallelTempering.jl:79-95):
        cld(nsteps, exch_interval) rounds of exch_interval steps of every
        replica at its rank's temperature, each followed by an exchange
        attempt starting alternately at the second and the first rank."""
        n_exch = -(-nsteps // exch_interval)
        for i in range(1, n_exch + 1):
            out = step_(self.batch, self.sampler, nsteps=exch_interval, kTs=self.replica_kTs())
            if self.E is not None:
                self.E += out[:, 0]
            self.replica_exchange_(i % 2)

    def acceptance(self):
        """Exchange acceptance rate of every rank pair (rank, rank + 1)."""
        return np.divide(self.n_accept, np.maximum(self.n_exch, 1))
