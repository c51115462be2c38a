"""The 18 compact operators of the reference's ``operations`` module
(src/basics_operations2.f90:270-312, bodies :3952-4790), with the same names
and argument meaning: ``Get_<dir>_<kind>_3D(fi3d, fo3d, dm, iacc, ibc, fbc2d)``.

``fi3d``/``fo3d`` are Fortran-ordered float64 arrays with the reference's
staggered extents; ``fo3d`` is filled in place (intent(out)).  ``fbc2d`` is
optional exactly as in the reference: when it is absent, interior, Neumann
(and, for centre-input operators, Dirichlet) ends fall back to one-sided
closures.  The computation runs on the GPU through ``chs_op3d``.
"""
import ctypes as C

import numpy as np

from . import _lib

_KINDS = {"midp_C2P": 0, "midp_P2C": 1, "1der_C2C": 2, "1der_P2P": 3, "1der_C2P": 4, "1der_P2C": 5}
_DIRS = {"x": 0, "y": 1, "z": 2}


def op3d(dm, d, kind, fi3d, fo3d, iacc, ibc, fbc2d=None):
    """Generic form: direction d (0,1,2), kind 0..5."""
    fi = np.asfortranarray(fi3d, dtype=np.float64)
    if not (isinstance(fo3d, np.ndarray) and fo3d.dtype == np.float64 and fo3d.flags.f_contiguous):
        raise TypeError("fo3d must be a Fortran-contiguous float64 array")
    fb = None if fbc2d is None else np.asfortranarray(fbc2d, dtype=np.float64)
    if fb is not None:
        want = list(fi.shape)
        want[d] = 4
        if list(fb.shape) != want:
            raise ValueError(f"fbc2d has shape {fb.shape}, expected {tuple(want)}")
    _lib.check(dm.lib.chs_op3d(dm.h, d, kind, int(iacc), _lib.ivec(ibc), fi.ctypes.data,
                               _lib.ivec(fi.shape), fo3d.ctypes.data, _lib.ivec(fo3d.shape),
                               None if fb is None else fb.ctypes.data, _lib.MEM_HOST))
    return fo3d


def _make(dname, kname):
    d, k = _DIRS[dname], _KINDS[kname]

    def f(fi3d, fo3d, dm, iacc, ibc, fbc2d=None):
        return op3d(dm, d, k, fi3d, fo3d, iacc, ibc, fbc2d)

    f.__name__ = f"Get_{dname}_{kname}_3D"
    f.__doc__ = (f"Get_{dname}_{kname}_3D of src/basics_operations2.f90 (direction {dname}, "
                 f"kind {kname}) on the GPU.")
    return f


for _d in _DIRS:
    for _k in _KINDS:
        globals()[f"Get_{_d}_{_k}_3D"] = _make(_d, _k)

__all__ = [f"Get_{d}_{k}_3D" for d in _DIRS for k in _KINDS] + ["op3d"]


def inspect_op1d(desc, d, kind, iacc, ibc, fi, n_out, fbc=None):
    """Host evaluation of the planned operator on one line (no GPU)."""
    L = _lib.load()
    fi = np.ascontiguousarray(fi, dtype=np.float64)
    fo = np.zeros(n_out)
    fb = None if fbc is None else np.ascontiguousarray(fbc, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    _lib.check(L.chs_inspect_op1d(C.byref(desc), d, kind, int(iacc), _lib.ivec(ibc), fi.ctypes.data_as(dp),
                                  len(fi), fo.ctypes.data_as(dp), n_out,
                                  None if fb is None else fb.ctypes.data_as(dp)))
    return fo


def inspect_coeff(fam, is_rhs, row, k, bc, acc):
    return _lib.load().chs_inspect_coeff(fam, is_rhs, row, k, bc, acc)
