"""``solve_fft_poisson`` of src/poisson_interface.f90:119-149 on the GPU
(FFT(x,z) + TDMA(y) path of src/poisson_fishpack.f90:354-490)."""
import numpy as np

from . import _lib


def solve_fft_poisson(rhs_xpencil, dm):
    """In-place solve of the (nc1,nc2,nc3) Fortran-ordered array."""
    if not (isinstance(rhs_xpencil, np.ndarray) and rhs_xpencil.dtype == np.float64
            and rhs_xpencil.flags.f_contiguous):
        raise TypeError("rhs must be a Fortran-contiguous float64 array")
    if rhs_xpencil.shape != dm.shapes["ccc"]:
        raise ValueError(f"rhs has shape {rhs_xpencil.shape}, expected {dm.shapes['ccc']}")
    _lib.check(dm.lib.chs_poisson(dm.h, rhs_xpencil.ctypes.data, _lib.MEM_HOST))
    return rhs_xpencil
