"""Golden fixtures (tests/golden/*.npz, written by make_golden.py): the host
builders regenerate the stored initial states, the oracle reproduces its
stored outputs, and the FP64 device path matches them."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

from helpers import announce, rel_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden as MG  # noqa: E402

CASES = {name: (s, par) for name, s, par in MG.cases()}


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz")))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    s, par = CASES[name]
    f = load(name)
    state = s.dipoles_flat() if s.N == 0 else s.coherents_flat()
    np.testing.assert_array_equal(state, f["state"])
    out = MG.compute(s, par, f["noise"])
    for key in ("energy", "grad", "langevin", "midpoint", "midpoint_iters"):
        np.testing.assert_allclose(out[key], f[key], rtol=0, atol=1e-12 * max(1.0, np.abs(f[key]).max()))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_matches_fixture(name):
    import sunnyhip as sh
    s, par = CASES[name]
    f = load(name)
    announce(f"golden {name} fp64")
    if s.N == 0:
        assert rel_err(sh.energy_grad_dipoles(s), f["grad"]) <= 1e-11
    else:
        assert rel_err(sh.energy_grad_coherents(s), f["grad"]) <= 1e-11
    assert abs(sh.energy(s) - float(f["energy"])) <= 1e-11 * max(1.0, abs(float(f["energy"])))
    integ = sh.Langevin(float(f["dt"]), damping=float(f["damping"]), kT=float(f["kT"]))
    sh.step_(s, integ, noise=f["noise"])
    got = s.dipoles_flat() if s.N == 0 else s.coherents_flat()
    assert rel_err(got, f["langevin"]) <= 1e-11
    # midpoint from the fixture's initial state
    if s.N == 0:
        s.set_state_flat(dip=f["state"])
    else:
        s.set_state_flat(coh=f["state"])
    it = sh.step_(s, sh.ImplicitMidpoint(float(f["dt"])))[0]
    got = s.dipoles_flat() if s.N == 0 else s.coherents_flat()
    assert rel_err(got, f["midpoint"]) <= 1e-11
    assert abs(int(it) - int(f["midpoint_iters"])) <= 1
