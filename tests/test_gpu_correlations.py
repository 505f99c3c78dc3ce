"""GPU parity of the device SampledCorrelations with the restated
accum_sample! (oracle/correlations_ref.py) on trajectories of the FP64
oracle integrator."""
import numpy as np
import pytest

import sunnyhip as sh
from oracle import oracle as O
from oracle.correlations_ref import SampledCorrelationsRef

from helpers import announce, rel_err

pytestmark = pytest.mark.gpu


def model(name):
    if name == "cubic4":
        return sh.workloads.cubic_model(L=4, seed=5, precision=64)
    return sh.workloads.pyrochlore_model(L=2, J1=1.0, J2=0.1, A=-0.05, seed=6, precision=64)


@pytest.mark.parametrize("name,window", [("cubic4", "cosine"), ("pyro2", "cosine"), ("cubic4", "rectangular")])
def test_sampled_correlations_match_oracle(name, window):
    check_against_ref(model(name), name, window)


def test_sampled_correlations_library_fft(monkeypatch):
    """SNH_CORR_FFT=lib: the lattice transform by a batched hipFFT plan."""
    monkeypatch.setenv("SNH_CORR_FFT", "lib")
    s = model("pyro2")
    check_against_ref(s, "pyro2 (library FFT)", "cosine")
    assert s._device().query("corr_fft_lib") == 1


def test_sampled_correlations_long_axis():
    """An axis of 40 cells (beyond the register line transforms) takes the
    hipFFT plan."""
    s = sh.workloads.cubic_model(dims=(40, 2, 1), seed=9, precision=64)
    check_against_ref(s, "cubic 40x2x1", "cosine")
    assert s._device().query("corr_fft_lib") == 1


def check_against_ref(s, name, window):
    # n_all_omega = 10; omega_max = 20 keeps the trajectory short (measperiod 4,
    # 40 steps per two samples) so that rounding differences between two FP64
    # integrators stay far below the tolerance despite chaotic dynamics
    energies = np.linspace(0.0, 20.0, 6)
    sc = sh.SampledCorrelations(s, energies=energies, dt=0.05, window=window)
    announce(f"sampled correlations {name} {window}, T={sc.n_all_omega}, measperiod={sc.measperiod}")
    na = s.crystal.natoms
    ref = SampledCorrelationsRef(s.dims, na, sc.observables, sc.corr_pairs, sc.n_all_omega, window)
    o = O.Oracle(s.spec())
    S = s.dipoles_flat()
    Tsnap = sc.n_all_omega // 2 + 1
    for _ in range(2):
        snaps = [S]
        for _ in range(Tsnap - 1):
            for _ in range(sc.measperiod):
                S, _ = o.step_midpoint(S=S, dt=sc.dt, damping=0.0, tol=sc.tol)
            snaps.append(S)
        ref.accum_sample(ref.new_sample(snaps))
        sh.add_sample_(sc, s)
    assert sc.nsamples == 2
    assert rel_err(s.dipoles_flat(), S, scale=1.0) <= 1e-10
    data = sc.data
    scale = np.abs(ref.data).max()
    np.testing.assert_allclose(data, ref.data, rtol=0, atol=1e-9 * scale)
    # frequency sum rule on the device data: (1/T) sum_w S(q, w) is real and
    # non-negative on the diagonal pairs (a, a), i = j
    diag = [c for c, (a, b) in enumerate(sc.corr_pairs) if a == b]
    m = data[diag][:, np.arange(na), np.arange(na)].sum(-1) / sc.n_all_omega
    assert np.all(m.real >= -1e-12) and np.abs(m.imag).max() <= 1e-10 * scale


def test_sun_sampled_correlations_match_oracle():
    """:SUN trajectories: the dipole observables act on the expected spins
    (observable_values!, CorrelationSampling.jl:1-8, reads sys.dipoles)."""
    s = sh.workloads.triangular_sun_model(L1=4, L2=3, seed=7, precision=64)
    energies = np.linspace(0.0, 8.0, 6)
    sc = sh.SampledCorrelations(s, energies=energies, dt=0.05)
    announce(f"sampled correlations SU(3) triangular 4x3, T={sc.n_all_omega}, measperiod={sc.measperiod}")
    ref = SampledCorrelationsRef(s.dims, 1, sc.observables, sc.corr_pairs, sc.n_all_omega, "cosine")
    o = O.Oracle(s.spec())
    Z = s.coherents_flat()
    Tsnap = sc.n_all_omega // 2 + 1
    for _ in range(2):
        snaps = [o.expected_spin(Z)]
        for _ in range(Tsnap - 1):
            for _ in range(sc.measperiod):
                Z, _ = o.step_midpoint(Z=Z, dt=sc.dt, damping=0.0, tol=sc.tol)
            snaps.append(o.expected_spin(Z))
        ref.accum_sample(ref.new_sample(snaps))
        sh.add_sample_(sc, s)
    assert rel_err(s.coherents_flat(), Z, scale=1.0) <= 1e-10
    scale = np.abs(ref.data).max()
    np.testing.assert_allclose(sc.data, ref.data, rtol=0, atol=1e-9 * scale)
