"""chs_velocity_range_noalloc: the velocity range with its block partials in
the library's static per-device array, against chs_velocity_range (block
partials in the flow solver's own buffer) on the same states.  Max and min
are exact in any order, so the two must agree bit for bit: on a stepped
channel, then on a second, smaller channel (the static array still holds
the first call's partials and must be fully overwritten), and on a state
with a planted NaN and infinity (both return CHS_ENAN with the same range).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)


def _both(dm):
    a, b = np.zeros(6), np.zeros(6)
    ra = dm.lib.chs_velocity_range(dm.h, a.ctypes.data_as(C.POINTER(C.c_double)))
    rb = dm.lib.chs_velocity_range_noalloc(dm.h, b.ctypes.data_as(C.POINTER(C.c_double)))
    return ra, a, rb, b


def test_noalloc_velocity_range_matches_allocating_entry_point():
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import Domain
    from chapsim2_amd.initial import channel_fields

    for step, nc in enumerate(((64, 48, 32), (16, 12, 8))):
        print(f"no-alloc velocity range, call {step + 1}: channel {nc}", flush=True)
        fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=11 + step)
        dm = Domain(nc, LX, -1.0, 1.0, LZ, **CFG)
        try:
            dm.set_state(*fields)
            dm.update_re(0)
            dm.substep(0)
            dm.run(0, 1)
            ra, a, rb, b = _both(dm)
            assert ra == _lib.CHS_OK and rb == _lib.CHS_OK, (ra, rb)
            assert np.array_equal(a, b), (a, b)
            qx, qy, qz, _ = dm.get_state()
            assert np.array_equal(b, [qx.max(), qx.min(), qy.max(), qy.min(), qz.max(), qz.min()])
            if step == 1:
                qx, qy, qz, pr = dm.get_state()
                qy = np.array(qy, order="F")
                qy[3, 5, 7] = np.nan
                qz = np.array(qz, order="F")
                qz[1, 2, 3] = np.inf
                dm.set_state(qx, qy, qz, pr)
                ra, a, rb, b = _both(dm)
                assert ra == _lib.CHS_ENAN and rb == _lib.CHS_ENAN, (ra, rb)
                assert np.array_equal(a, b) and b[4] == np.inf, (a, b)
        finally:
            dm.close()
