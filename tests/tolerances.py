"""Tolerances of the parity tests that follow from the operators' own
conditioning (test infrastructure).

The exact projection's pressure increment solves L = D*G, whose wall-normal
block L_y (D1_P2C after D1_C2P with the wall faces reset) has a
one-dimensional null space (the gauge) and, on its range, the condition
number kappa = |lambda|_max / |lambda|_min,nonzero (1.9e4 at ny = 128 with
tanh beta 1.5).  Two backward-stable FP64 solves of the same system may
differ by kappa * eps relative, and the pressure is the running sum of one
increment per projection, so pressure comparisons (mean removed) are held to
4 * kappa * eps * (number of projections).

That bound misses the other source of pressure round-off: each increment
solves div(u*) / (alpha dt), and div(u*) is a difference of O(1) velocities
that is O(dt) small, so its relative rounding is ~eps |u| ||D|| / |div u*|
(~1e-12 at dt = 5e-4).  That floor is measured on the oracle itself: the
same steps from an initial velocity perturbed by one ulp per value
(ulp_perturbed); the pressure held to ten times what that moves, when that
is larger than the kappa bound."""
import numpy as np

EPS64 = 2.0 ** -52
EPS32 = 2.0 ** -23

_KAPPA = {}


def ly_kappa(dom):
    """Condition number on its range of the exact projection's wall-normal
    block L_y, assembled column by column from the oracle's line operators
    (D1_C2P of a unit pressure vector with Neumann walls, wall faces set to
    zero, then D1_P2C with Dirichlet walls): |lambda|_max over the smallest
    nonzero |lambda| (the smallest one is the gauge null mode)."""
    key = (tuple(dom.nc), round(float(dom.ymc[0]), 15))
    if key not in _KAPPA:
        ny = dom.nc[1]
        L = np.zeros((ny, ny))
        for k in range(ny):
            e = np.zeros(ny)
            e[k] = 1.0
            g = dom.op1d(1, 4, 4, [5, 5], e, ny + 1, np.zeros(4))
            g[0] = g[ny] = 0.0
            L[:, k] = dom.op1d(1, 5, 4, [4, 4], g, ny)
        ev = np.sort(np.abs(np.linalg.eigvals(L)))
        _KAPPA[key] = float(ev[-1] / ev[1])
    return _KAPPA[key]


def pressure_tol(dom, n_proj):
    """4 * kappa(L_y) * eps per projection, summed over the projections
    whose increments make up the pressure."""
    return 4.0 * ly_kappa(dom) * EPS64 * n_proj


def ulp_perturbed(fields, seed=12345):
    """qx, qy, qz each multiplied by (1 + e eps), e in {-1, 0, 1} per value;
    the pressure unchanged"""
    rng = np.random.default_rng(seed)
    out = [np.asfortranarray(f * (1.0 + EPS64 * rng.integers(-1, 2, f.shape))) for f in fields[:3]]
    return out + [fields[3]]


def rel_mean_removed(a, b):
    a, b = a - a.mean(), b - b.mean()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
