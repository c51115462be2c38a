"""Batched channels (chs_batch_*): uneven instances advanced together, as
eager per-instance streams and as one captured graph per RK3 step, give the
same fields as each instance run on its own."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LX, LZ = 4 * np.pi, 4 * np.pi / 3


def _specs():
    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT

    base = dict(lxx=LX, lyb=-1.0, lyt=1.0, lzz=LZ, istret=2, dt=5e-4, idriven=1, projection=PROJ_EXACT)
    return [
        dict(base, nc=(32, 24, 16), rstret=0.075, ren=2800.0),
        dict(base, nc=(48, 32, 24), rstret=0.09, ren=6900.0, precision=PREC_MIXED),
        dict(base, nc=(40, 25, 20), rstret=0.06, ren=2800.0),
    ]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_instances(graph):
    from chapsim2_amd.domain import Batch, Domain
    from chapsim2_amd.initial import channel_fields

    specs = _specs()
    fields = [channel_fields(s["nc"], LX, -1.0, 1.0, LZ, 2, s["rstret"], noise=0.1, seed=40 + i)
              for i, s in enumerate(specs)]
    b = Batch(specs)
    singles = []
    try:
        for d, f in zip(b.domains, fields):
            d.set_state(*f)
            d.update_re(0)
            d.substep(0)
        b.run(0, 4, graph=graph)
        for s, f in zip(specs, fields):
            kw = {k: v for k, v in s.items() if k not in ("nc", "lxx", "lyb", "lyt", "lzz")}
            d = Domain(s["nc"], LX, -1.0, 1.0, LZ, **kw)
            singles.append(d)
            d.set_state(*f)
            d.update_re(0)
            d.substep(0)
            d.run(0, 4)
        for i, (db, ds) in enumerate(zip(b.domains, singles)):
            for name, x, y in zip(("qx", "qy", "qz", "pr"), db.get_state(), ds.get_state()):
                assert np.array_equal(x, y), (i, name, np.abs(x - y).max())
            assert np.all(db.mass_check() < 1e-12)
        # a new state for the mixed instance between runs: the next step (a
        # replay of the captured graph with graph=True) must round the new
        # state, not reuse the FP32 copies the last step left
        s1 = specs[1]
        f2 = channel_fields(s1["nc"], LX, -1.0, 1.0, LZ, 2, s1["rstret"], noise=0.1, seed=99)
        b.domains[1].set_state(*f2)
        b.run(4, 1, graph=graph)
        kw = {k: v for k, v in s1.items() if k not in ("nc", "lxx", "lyb", "lyt", "lzz")}
        d = Domain(s1["nc"], LX, -1.0, 1.0, LZ, **kw)
        singles.append(d)
        d.set_state(*f2)
        d.run(4, 1)
        for name, x, y in zip(("qx", "qy", "qz", "pr"), b.domains[1].get_state(), d.get_state()):
            assert np.array_equal(x, y), ("after set_state", name, np.abs(x - y).max())
    finally:
        for d in singles:
            d.close()
        b.close()
