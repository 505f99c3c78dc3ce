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


def site_coloring(spec, dims):
    """A proper colouring of single sites (period m = dims), greedy in site
    order: an independent colouring for the oracle's checkerboard sweeps."""
    L1, L2, L3 = dims
    na = spec["na"]
    bonds = np.asarray(spec["bond"]).reshape(-1, 6)
    color = -np.ones((na, L3, L2, L1), np.int32)
    for i in range(na):
        for c3 in range(L3):
            for c2 in range(L2):
                for c1 in range(L1):
                    used = set()
                    for b in bonds:
                        if b[0] != i:
                            continue
                        j = b[1]
                        cc = ((c1 + b[2]) % L1, (c2 + b[3]) % L2, (c3 + b[4]) % L3)
                        used.add(int(color[j, cc[2], cc[1], cc[0]]))
                    k = 0
                    while k in used:
                        k += 1
                    color[i, c3, c2, c1] = k
    return (L1, L2, L3), int(color.max()) + 1, color


def coloring_is_proper(spec, dims, m, color):
    """No bond joins two sites of one colour class colour, and m divides dims."""
    L1, L2, L3 = dims
    if any(L % mm for L, mm in zip(dims, m)):
        return False
    color = np.asarray(color).reshape(spec["na"], m[2], m[1], m[0])
    bonds = np.asarray(spec["bond"]).reshape(-1, 6)
    c1, c2, c3 = np.meshgrid(np.arange(L1), np.arange(L2), np.arange(L3), indexing="ij")
    for b in bonds:
        i, j = b[0], b[1]
        ci = color[i, c3 % m[2], c2 % m[1], c1 % m[0]]
        d1, d2, d3 = (c1 + b[2]) % L1, (c2 + b[3]) % L2, (c3 + b[4]) % L3
        cj = color[j, d3 % m[2], d2 % m[1], d1 % m[0]]
        if np.any(ci == cj):
            return False
    return True


def langevin_function(x):
    """coth(x) - 1/x: the mean of cos(theta) for a classical moment."""
    return 1 / np.tanh(x) - 1 / x
