"""The surrogate's property refresh on the matrix cores (BF16 activations,
FP32 accumulation, v_mfma_f32_16x16x16_bf16) against the vector-FMA refresh
on FP32 activations, over every cell of a heated-channel-sized batch: the
two differ only by the BF16 rounding of the 128 hidden activations of each
evaluation, so every output agrees to well within 1e-2 of its range (and
far closer in the mean); a wrong lane layout or block order would put the
outputs off by O(1)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_refresh_matrix_cores_vs_vector():
    from chapsim2_amd import _lib
    from chapsim2_amd.initial import HEATED_H_RANGE, heated_channel_table
    from chapsim2_amd.mlp import Surrogate
    from chapsim2_amd.props import PROPS_MLP, PropertyTable

    tab = heated_channel_table(4096, *HEATED_H_RANGE)
    pt = PropertyTable(tab, -1.0)
    sur = Surrogate(pt, nsamples=200_000, seed=5)
    sur.train(1500, batch=4096, lr=2e-3)
    pt.set_surrogate(sur)
    L = _lib.load()
    n = 1_000_003  # not a multiple of the 64 cells of a wave
    rng = np.random.default_rng(3)
    rhoh = np.interp(rng.uniform(0.8, 1.2, n), tab[:, 1], tab[:, 1] * tab[:, 3])
    res = {}
    for on in (1, 0):
        _lib.check(L.chs_mlp_set_matrix_cores(sur.h, on))
        res[on] = pt.from_rhoh(rhoh, mode=PROPS_MLP)
    _lib.check(L.chs_mlp_set_matrix_cores(sur.h, 1))
    for k in ("h", "d", "m", "k", "t"):
        a, b = res[1][k], res[0][k]
        span = np.abs(b).max()
        dmax, dmean = float(np.abs(a - b).max() / span), float(np.abs(a - b).mean() / span)
        print(f"{k}: matrix cores vs vector FMAs max {dmax:.2e} mean {dmean:.2e}", flush=True)
        assert dmax < 1e-2 and dmean < 2e-3, (k, dmax, dmean)
