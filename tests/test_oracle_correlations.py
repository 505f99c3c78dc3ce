"""Pins of the sampled-correlation restatement (oracle/correlations_ref.py)
by properties the mathematics fixes: the lattice transform's sign and
normalisation, the frequency sum rule, and the resolution of the dt
adjustment."""
import math

import numpy as np

import sunnyhip as sh
from oracle.correlations_ref import SampledCorrelationsRef, fftfreq_int, observable_values
from sunnyhip.correlations import adjusted_dt_and_downsampling_factor


def test_fftfreq_order():
    assert list(fftfreq_int(6)) == [0, 1, 2, -3, -2, -1]
    assert list(fftfreq_int(5)) == [0, 1, 2, -2, -1]


def test_lattice_transform_sign_and_scale():
    """accum_sample!'s conj . fft . conj / sqrt(N) is A(q) = sum_r e^{+iq.r} A(r) / sqrt(N)
    (CorrelationSampling.jl:58-63), by a direct sum."""
    rng = np.random.default_rng(0)
    dims, na, T = (3, 2, 4), 2, 4
    ref = SampledCorrelationsRef(dims, na, np.eye(3)[:, None, :].repeat(na, 1), [(0, 1)], T, "rectangular")
    snaps = [rng.standard_normal((np.prod(dims) * na, 3)) for _ in range(T // 2 + 1)]
    buf = ref.new_sample(snaps)
    got = np.conj(np.fft.fftn(np.conj(buf), axes=(1, 2, 3))) / np.sqrt(np.prod(dims))
    L1, L2, L3 = dims
    want = np.zeros_like(buf)
    for m1 in range(L1):
        for m2 in range(L2):
            for m3 in range(L3):
                ph = np.exp(2j * np.pi * (m1 * np.arange(L1)[:, None, None] / L1 + m2 * np.arange(L2)[None, :, None] / L2
                                          + m3 * np.arange(L3)[None, None, :] / L3))
                want[:, m1, m2, m3] = np.einsum("xyz,axyzit->ait", ph, buf) / np.sqrt(np.prod(dims))
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_frequency_sum_rule():
    """(1/T) sum_w data(q, w) = mean over the recorded snapshots of
    conj(A_a,i(q, t)) A_b,j(q, t): Parseval, n_contrib(0) = Tsnap and
    window(0) = 1, for either window."""
    rng = np.random.default_rng(1)
    dims, na, T = (2, 3, 2), 2, 8
    obs = rng.standard_normal((3, na, 3))
    pairs = [(2, 2), (1, 2), (0, 2), (1, 1), (0, 1), (0, 0)]
    for window in ("cosine", "rectangular"):
        ref = SampledCorrelationsRef(dims, na, obs, pairs, T, window)
        snaps = [rng.standard_normal((np.prod(dims) * na, 3)) for _ in range(T // 2 + 1)]
        buf = ref.new_sample(snaps)
        ref.accum_sample(buf)
        Aq = np.conj(np.fft.fftn(np.conj(buf), axes=(1, 2, 3))) / np.sqrt(np.prod(dims))
        for c, (a, b) in enumerate(pairs):
            for i in range(na):
                for j in range(na):
                    lhs = ref.data[c, i, j].sum(-1) / T
                    rhs = (np.conj(Aq[a, ..., i, :]) * Aq[b, ..., j, :]).sum(-1) / (T // 2 + 1)
                    np.testing.assert_allclose(lhs, rhs, atol=1e-12)


def test_observables_are_projections():
    rng = np.random.default_rng(2)
    dims, na = (2, 2, 3), 3
    S = rng.standard_normal((np.prod(dims) * na, 3))
    obs = rng.standard_normal((2, na, 3))
    v = observable_values(S, dims, na, obs)
    # site (c1, c2, c3, i) in Julia order
    c1, c2, c3, i = 1, 0, 2, 2
    s = S[c1 + 2 * (c2 + 2 * (c3 + 3 * i))]
    np.testing.assert_allclose(v[:, c1, c2, c3, i], obs[:, i] @ s)


def test_adjusted_dt_resolves_omega_max():
    for dt, n_omega, wmax in ((0.05, 50, 5.0), (0.02, 6, 2.5), (0.1, 101, 3.0)):
        dt2, mp = adjusted_dt_and_downsampling_factor(dt, n_omega, wmax)
        assert dt2 <= dt + 1e-15 and mp == math.ceil(math.pi / (dt * wmax))
        # the Nyquist energy of the sampled trajectory is exactly omega_max
        assert abs(math.pi / (dt2 * mp) - wmax) <= 1e-12 * wmax
