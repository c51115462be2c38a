"""Pins the numpy oracle of the MLP surrogate: its hand-written backward
pass against central finite differences, and its BF16 rounding."""
import numpy as np


def test_backward_matches_finite_differences():
    import mlp_oracle as mo

    rng = np.random.default_rng(0)
    p = rng.normal(0.0, 0.3, mo.NP)
    x = rng.uniform(-1, 1, 37)
    t = rng.normal(0, 1, (37, mo.NO))
    _, g = mo.loss_grad(p, x, t)
    idx = np.concatenate([rng.choice(mo.NP, 60, replace=False), [0, mo.OB1, mo.OW2, mo.OB2, mo.OW3, mo.OB3]])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mo.loss_grad(pp, x, t)[0] - mo.loss_grad(pm, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    import mlp_oracle as mo

    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.14159265, 1e-3], dtype=np.float32)
    r = mo.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie rounds to even
    assert r[2] == 1.0 + 2**-6    # tie rounds to even (up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
    assert abs(r[5] - 1e-3) < 1e-3 * 2**-8
