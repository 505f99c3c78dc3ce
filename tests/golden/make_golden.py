"""Writes the golden fixtures in this directory (FP64 oracle outputs).

Run from the repository root:  python tests/golden/make_golden.py
Each fixture holds the initial state, the raw noise of one Langevin step,
and the oracle's energy, gradient, Langevin step and midpoint step for a
small instance of one benchmark model family.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "sunny.jl_amd")]

import sunnyhip as sh  # noqa: E402
from oracle import oracle as O  # noqa: E402


def cases():
    yield "cubic_4", sh.workloads.cubic_model(L=4, seed=11), dict(dt=0.05, damping=0.1, kT=0.5)
    yield "pyrochlore_2", sh.workloads.pyrochlore_model(L=2, seed=12), dict(dt=0.025, damping=0.2, kT=0.3)
    yield "pyrochlore_dipolar_2", sh.workloads.pyrochlore_model(L=2, J1=0.05, J2=0, A=0, s=3.5, seed=13,
                                                                dipolar=True), dict(dt=0.01, damping=0.1, kT=0.2)
    yield "triangular_su3_8", sh.workloads.triangular_sun_model(L1=8, L2=8, seed=14, precision=64), \
        dict(dt=0.02, damping=0.1, kT=0.5)


def compute(s, par, noise):
    o = O.Oracle(s.spec())
    out = {}
    if s.N == 0:
        S = s.dipoles_flat()
        out["state"] = S
        out["energy"] = np.array(o.energy(S))
        out["grad"] = o.grad(S)
        out["langevin"] = o.step_langevin(S=S, noise=noise, **par)
        mp, it = o.step_midpoint(S=S, dt=par["dt"])
        out["midpoint"], out["midpoint_iters"] = mp, np.array(it)
    else:
        Z = s.coherents_flat()
        out["state"] = Z
        out["energy"] = np.array(o.energy(Z=Z))
        out["grad"] = o.grad_coherents(Z)
        out["langevin"] = o.step_langevin(Z=Z, noise=noise, **par)
        mp, it = o.step_midpoint(Z=Z, dt=par["dt"])
        out["midpoint"], out["midpoint_iters"] = mp, np.array(it)
    return out


def noise_for(s, name):
    rng = np.random.default_rng(sum(map(ord, name)))
    if s.N == 0:
        return rng.standard_normal((s.nsites, 3))
    return (rng.standard_normal((s.nsites, s.N)) + 1j * rng.standard_normal((s.nsites, s.N))) / np.sqrt(2)


if __name__ == "__main__":
    for name, s, par in cases():
        noise = noise_for(s, name)
        out = compute(s, par, noise)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), noise=noise,
                            dt=par["dt"], damping=par["damping"], kT=par["kT"], **out)
        print(name, s.nsites, float(out["energy"]), int(out["midpoint_iters"]))
