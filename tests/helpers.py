"""Shared helpers for the parity tests (oracle vs device)."""
import sys

import numpy as np


def announce(case):
    """Name the case before any GPU work so a failing run is attributable."""
    print(f"[case] {case}", flush=True)
    sys.stdout.flush()


def rel_err(a, b, scale=None):
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b)))


def energy_scale(sys):
    """Sum over bonds and sites of |J| kappa^2 + onsite magnitude: the natural
    scale for absolute energy errors of random configurations."""
    sp = sys.spec()
    kap = sp["kappa"]
    e = 0.0
    for b in range(sp["nbonds"]):
        i, j = sp["bond"][b][:2]
        e += np.abs(sp["bilin"][b]).max() * kap[i] * kap[j] * sys.ncells / 2
    return max(e, 1.0)
