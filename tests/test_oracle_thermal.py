"""Thermal checks that pin the oracle's Langevin noise term.

The stochastic Heun step (Integrators.jl:313-357, fill_noise! :288-298) must
sample exp(-E/kT).  Two known answers pin it:

* the energy distribution of a two-site chain, E = S1.S2, for :dipole and
  :SUN Heun: the check of test/test_samplers.jl:120-227 (same model: two
  sites, J = 1, |S| = 1; damping 0.1, kT 0.1, dt 0.02; 1000 equilibration
  steps, 2000 samples 500 steps apart; 10-bin distribution, atol 0.05);
* a paramagnet of |S| = kappa = 3/2 in a field h = g^T B:
  <S.h^> = -kappa L(kappa |h| / kT), L(x) = coth x - 1/x, which also pins how
  the noise and damping terms scale with kappa.

The reference writes its chain with to_inhomogeneous and set_exchange_at!
(one bond, offset (-1,0,0)); here the homogeneous J = 1/2 on Bond(0,0,(1,0,0))
over dims (2,1,1) gives the same Hamiltonian (both bond directions join the two
sites, E = 2 (1/2) S1.S2).  Each check is shown to fail with a plausible
mistake planted in the oracle (noise scaled by sqrt(dt lambda kT); corrector
without noise; damping with S/kappa).
"""
import math

import numpy as np

import sunnyhip as sh
from oracle import oracle as O

from helpers import langevin_function

CHAIN = dict(damping=0.1, kT=0.1, dt=0.02, n_equilib=1000, n_samples=2000, n_decorr=500, n_bins=10, atol=0.05)


def cubic_slice_area(alpha, k):
    """(k-1)-dimensional area of the slice sum(x) = alpha of the unit k-cube
    (test_samplers.jl:125-134)."""
    return sum((-1) ** i * math.comb(k, i) * (alpha - i) ** (k - 1) / math.factorial(k - 1)
               for i in range(int(math.floor(alpha)) + 1))


def chain_P(E, kT, n=2, J=1.0):
    """Energy density of an open n-site chain of unit classical spins
    (test_samplers.jl:136-140)."""
    E_min = -J * max(1.0, n - 1.0)
    return (2 * J) ** (n - 2) * cubic_slice_area((E - E_min) / (2 * J), n - 1) * math.exp(-E / kT) / \
        (2 * kT * math.sinh(J / kT)) ** (n - 1)


def empirical_distribution(data, numbins):
    """Histogram of `data` over its own range (test_samplers.jl:143-158)."""
    data = np.asarray(data, float)
    lo, hi = data.min(), data.max()
    delta = (hi - lo) / numbins
    boundaries = np.arange(numbins + 1) * delta + lo
    idx = np.minimum(np.floor((data - lo) / delta).astype(int), numbins - 1)
    counts = np.bincount(idx, minlength=numbins).astype(float)
    return counts / len(data), boundaries


def discretize_P(boundaries, kT, n=2, J=1.0, delta=0.001):
    """Riemann sums of chain_P over each bin (test_samplers.jl:162-171)."""
    out = np.zeros(len(boundaries) - 1)
    for b in range(len(out)):
        Es = np.arange(boundaries[b], boundaries[b + 1] + 1e-15, delta)
        out[b] = sum(chain_P(E, kT, n, J) * delta for E in Es)
    return out


def two_site_chain(mode, precision=64, seed=0):
    """Two sites joined by E = S1.S2 with |S| = 1: :dipole s = 1, or :SUN
    s = 1/2 with kappa = 2 (test_samplers.jl:174-186)."""
    cryst = sh.cubic_crystal(1.0)
    s = 1.0 if mode == "dipole" else 0.5
    sys = sh.System(cryst, [(0, sh.Moment(s=s, g=2))], mode, dims=(2, 1, 1), seed=seed, precision=precision)
    if mode == "SUN":
        sh.set_spin_rescaling_(sys, [2.0])
    b = sh.Bond(0, 0, (1, 0, 0))
    sh.set_exchange_(sys, 0.5, b, propagate=[b])  # the chain's axis only
    sh.randomize_spins_(sys)
    return sys


def chain_distribution_error(energies):
    Ps, boundaries = empirical_distribution(energies, CHAIN["n_bins"])
    return float(np.max(np.abs(Ps - discretize_P(boundaries, CHAIN["kT"]))))


def oracle_chain_energies(mode, seed=0):
    sys = two_site_chain(mode, seed=seed)
    o = O.Oracle(sys.spec())
    st = dict(S=sys.dipoles_flat()) if mode == "dipole" else dict(Z=sys.coherents_flat())
    key = O.replica_key(sys.noise_seed, 0)
    kw = dict(dt=CHAIN["dt"], damping=CHAIN["damping"], kT=CHAIN["kT"], key=key)
    state, _ = o.langevin_run(**st, **kw, step0=0, nsteps=CHAIN["n_equilib"], every=0)
    st = dict(S=state) if mode == "dipole" else dict(Z=state)
    n = CHAIN["n_samples"] * CHAIN["n_decorr"]
    _, E = o.langevin_run(**st, **kw, step0=CHAIN["n_equilib"], nsteps=n, every=CHAIN["n_decorr"])
    return E


def test_chain_density_normalised():
    """chain_P integrates to 1 over [-J, J] (a check of the restated formula)."""
    h = 1e-4
    Es = -1 + h * (np.arange(20000) + 0.5)  # midpoints: the density jumps to 0 at E = J
    for kT in (0.1, 0.5, 2.0):
        assert abs(h * sum(chain_P(E, kT) for E in Es) - 1) < 1e-6


def test_chain_energy_distribution():
    for mode in ("dipole", "SUN"):
        E = oracle_chain_energies(mode)
        assert np.all(np.abs(E) <= 1 + 1e-9)
        assert chain_distribution_error(E) <= CHAIN["atol"], mode


def test_chain_distribution_catches_planted_noise_mistakes():
    for m in (O.mutation.NOISE_SCALE, O.mutation.CORRECTOR_NOISE):
        with O.mutation(m):
            E = oracle_chain_energies("dipole")
        assert chain_distribution_error(E) > 2 * CHAIN["atol"], m


PARA = dict(kappa=1.5, g=2.0, kT=1.0, x=2.0, damping=0.5, dt=0.01, L=2000, n_equilib=500, nsteps=2000, every=10)


def paramagnet(precision=64, seed=1, L=None):
    """L independent sites of |S| = 3/2 (g = 2) in a field along z chosen so
    that kappa |h| / kT = 2."""
    p = PARA
    cryst = sh.cubic_crystal(1.0)
    sys = sh.System(cryst, [(0, sh.Moment(s=p["kappa"], g=p["g"]))], "dipole", dims=(L or p["L"], 1, 1),
                    seed=seed, precision=precision)
    b = p["x"] * p["kT"] / (p["kappa"] * p["g"])
    sh.set_field_(sys, [0.0, 0.0, b])
    sh.randomize_spins_(sys)
    return sys, b * p["g"]


def paramagnet_expected():
    """<S.h^> for E = h.S (the field term of Interactions.jl:401-432)."""
    return -PARA["kappa"] * langevin_function(PARA["x"])


def oracle_paramagnet_mean():
    p = PARA
    sys, h = paramagnet()
    o = O.Oracle(sys.spec())
    key = O.replica_key(sys.noise_seed, 0)
    kw = dict(dt=p["dt"], damping=p["damping"], kT=p["kT"], key=key)
    S, _ = o.langevin_run(S=sys.dipoles_flat(), **kw, step0=0, nsteps=p["n_equilib"], every=0)
    _, E = o.langevin_run(S=S, **kw, step0=p["n_equilib"], nsteps=p["nsteps"], every=p["every"])
    return float(np.mean(E)) / (h * sys.nsites)


def test_paramagnet_langevin_function():
    m = oracle_paramagnet_mean()
    assert abs(m - paramagnet_expected()) <= 0.01 * PARA["kappa"], (m, paramagnet_expected())


def test_paramagnet_catches_planted_kappa_mistake():
    """Damping with S/kappa heats a kappa != 1 moment to kappa kT: invisible
    to the |S| = 1 chain, caught here."""
    with O.mutation(O.mutation.DAMPING_KAPPA):
        m = oracle_paramagnet_mean()
    assert abs(m - paramagnet_expected()) > 0.05 * PARA["kappa"], m
