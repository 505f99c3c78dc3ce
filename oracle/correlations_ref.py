"""Restatement of Sunny's sampled-correlation accumulation.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the library): numpy
versions of observable_values! / new_sample! / accum_sample!
(Sunny.jl src/SampledCorrelations/CorrelationSampling.jl:1-138) for :dipole
systems, with the buffers of SampledCorrelations (SampledCorrelations.jl:
194-268).  Arrays use Julia's index order, with numpy axes in the same
order: samplebuf[a, q1, q2, q3, i, t], data[c, i, j, q1, q2, q3, w].
Trajectories come from the caller (the C oracle's integrator), as a list of
Tsnap snapshots of the dipoles in Julia site order.
"""
from __future__ import annotations

import numpy as np


def fftfreq_int(n):
    """FFTW.fftfreq(n, n): 0, 1, ..., ceil(n/2) - 1, -floor(n/2), ..., -1."""
    return np.fft.fftfreq(n, 1.0 / n).round().astype(int)


def observable_values(S, dims, na, observables):
    """buf[a, q1, q2, q3, i] = obs[a, i] . S (CorrelationSampling.jl:1-8),
    S [sites][3] in Julia order (site = c1 + L1 (c2 + L2 (c3 + L3 i)))."""
    L1, L2, L3 = dims
    Sr = np.asarray(S, float).reshape(na, L3, L2, L1, 3).transpose(3, 2, 1, 0, 4)  # [c1, c2, c3, i, xyz]
    return np.einsum("aix,pqrix->apqri", np.asarray(observables, float), Sr)


class SampledCorrelationsRef:
    def __init__(self, dims, na, observables, corr_pairs, n_all_omega, window="cosine"):
        self.dims, self.na = tuple(dims), na
        self.obs = np.asarray(observables, float)
        self.pairs = list(corr_pairs)
        self.T = n_all_omega
        self.window = window
        L1, L2, L3 = dims
        self.data = np.zeros((len(self.pairs), na, na, L1, L2, L3, self.T), complex)
        self.nsamples = 0

    def new_sample(self, snapshots):
        """samplebuf of new_sample! (CorrelationSampling.jl:33-44): Tsnap
        snapshots, then zero padding."""
        Tsnap = self.T // 2 + 1
        assert len(snapshots) == Tsnap
        L1, L2, L3 = self.dims
        buf = np.zeros((len(self.obs), L1, L2, L3, self.na, self.T), complex)
        for n, S in enumerate(snapshots):
            buf[..., n] = observable_values(S, self.dims, self.na, self.obs)
        return buf

    def accum_sample(self, buf):
        """accum_sample! (CorrelationSampling.jl:46-138), without the error
        estimate M."""
        T = self.T
        Tsnap = T // 2 + 1
        ncell = int(np.prod(self.dims))
        # conj, fft over the cell axes with 1/sqrt(prod(dims)), conj
        buf = np.conj(np.fft.fftn(np.conj(buf), axes=(1, 2, 3))) / np.sqrt(ncell)
        buf = np.fft.fft(buf, axis=5)
        n_contrib = (Tsnap - np.abs(fftfreq_int(T))).astype(float)
        n_contrib[n_contrib == 0] = np.inf
        self.nsamples += 1
        count = self.nsamples
        window = np.cos(np.linspace(0, np.pi, T + 1)[:-1]) ** 2
        for c, (a, b) in enumerate(self.pairs):
            for i in range(self.na):
                for j in range(self.na):
                    corr = np.conj(buf[a, :, :, :, i, :]) * buf[b, :, :, :, j, :]
                    corr = np.fft.ifft(corr, axis=3) / n_contrib
                    if self.window == "cosine":
                        corr = corr * window
                    corr = np.fft.fft(corr, axis=3)
                    self.data[c, i, j] += (corr - self.data[c, i, j]) / count
