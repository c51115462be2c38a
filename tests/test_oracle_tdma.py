"""Pins the oracle's tridiagonal solvers (restating
src/basics_algorithms.f90:14-109) on the reference's own known-answer
systems and on dense brute force."""
import json
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _kat():
    with open(os.path.join(HERE, "golden", "tdma_kat.json")) as f:
        return json.load(f)


def test_noncyclic_known_answer(po):
    k = _kat()
    c, d = po.tdma_preprocess(k["a"], k["b"], k["c"])
    x = po.tdma_solve(False, k["r"], k["a"], k["b"], c, d)
    assert np.array_equal(np.round(x, 12), np.array(k["noncyclic_solution"], float))
    assert np.abs(x - np.array(k["noncyclic_solution"])).max() < 1e-13


def test_cyclic_known_answer(po):
    k = _kat()
    n = len(k["a"])
    a, b, c = (np.array(k[s]) for s in "abc")
    cp, dp = po.tdma_preprocess(a[: n - 1], b[: n - 1], c[: n - 1])
    cf = c.copy()
    cf[: n - 1] = cp
    df = np.zeros(n)
    df[: n - 1] = dp
    x = po.tdma_solve(True, k["r"], a, b, cf, df)
    ref = np.array([float(Fraction(p, q)) for p, q in k["cyclic_solution_rational"]])
    assert np.abs(x - ref).max() < 1e-13


def test_against_dense_solve(po):
    rng = np.random.default_rng(0)
    for n in (5, 17, 64):
        a = rng.uniform(0.1, 0.4, n)
        c = rng.uniform(0.1, 0.4, n)
        b = 1.0 + rng.uniform(0.0, 0.2, n)
        r = rng.standard_normal(n)
        cp, dp = po.tdma_preprocess(a, b, c)
        x = po.tdma_solve(False, r, a, b, cp, dp)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.abs(x - np.linalg.solve(M, r)).max() < 1e-13
        # cyclic: a[0] couples to x[n-1], c[n-1] to x[0]
        cp, dp = po.tdma_preprocess(a[: n - 1], b[: n - 1], c[: n - 1])
        cf, df = c.copy(), np.zeros(n)
        cf[: n - 1], df[: n - 1] = cp, dp
        x = po.tdma_solve(True, r, a, b, cf, df)
        Mc = M.copy()
        Mc[0, n - 1] = a[0]
        Mc[n - 1, 0] = c[n - 1]
        assert np.abs(x - np.linalg.solve(Mc, r)).max() < 1e-12
