"""Thermal properties from a tabulated equation of state: the host-side
counterpart of the reference's fluid-thermo-property table
(src/input_thermo.f90:121-693), backed by libchapsim_hip."""
import ctypes as C

import numpy as np

from . import _lib

PROPS_TABLE, PROPS_MLP = 0, 1
NAMES = ("h", "d", "m", "k", "t", "b", "cp")
NAMES_MLP = ("h", "d", "m", "k", "t")


class PropertyTable:
    """Non-dimensional property table built from dimensional rows of
    (P, H, T, D, M, K, CP, B) with reference temperature ref_t0."""

    def __init__(self, table, ref_t0):
        self.lib = _lib.load()
        tab = np.ascontiguousarray(table, dtype=np.float64)
        if tab.ndim != 2 or tab.shape[1] != 8:
            raise ValueError("table must have 8 columns (P, H, T, D, M, K, CP, B)")
        self.h = C.c_void_p()
        _lib.check(self.lib.chs_props_create(tab.ctypes.data_as(C.POINTER(C.c_double)), tab.shape[0],
                                             float(ref_t0), C.byref(self.h)))

    def info(self):
        lo, hi = C.c_double(), C.c_double()
        ref = (C.c_double * 7)()
        _lib.check(self.lib.chs_props_info(self.h, C.byref(lo), C.byref(hi), ref))
        return lo.value, hi.value, np.array(ref[:])

    def from_rhoh(self, rhoh, mode=PROPS_TABLE):
        """ftp_refresh_thermal_properties_from_DH on every element of rhoh;
        returns a dict of h, d, m, k, t, b, cp arrays.  Raises ChsError with
        code CHS_ERANGE when an element was outside [dhmin, dhmax]."""
        x = np.ascontiguousarray(rhoh, dtype=np.float64).ravel()
        names = NAMES if mode == PROPS_TABLE else NAMES_MLP
        outs = {n: np.zeros_like(x) for n in names}
        ptrs = [outs[n].ctypes.data if n in outs else None for n in NAMES]
        _lib.check(self.lib.chs_props_from_rhoh(self.h, x.ctypes.data, x.size, *ptrs, mode, _lib.MEM_HOST))
        return outs

    def set_surrogate(self, surrogate):
        """Attach a chapsim2_amd.mlp.Surrogate (None detaches it): mode
        PROPS_MLP then evaluates the properties from h with the network."""
        self._surrogate = surrogate
        _lib.check(self.lib.chs_props_set_surrogate(self.h, surrogate.h if surrogate is not None else None))

    def from_h(self, h):
        """ftp_refresh_thermal_properties_from_H on every element of h;
        returns a dict of d, m, k, t, b, cp arrays."""
        x = np.ascontiguousarray(h, dtype=np.float64).ravel()
        names = ("d", "m", "k", "t", "b", "cp")
        outs = {n: np.zeros_like(x) for n in names}
        _lib.check(self.lib.chs_props_from_h(self.h, x.ctypes.data, x.size, *[outs[n].ctypes.data for n in names],
                                             _lib.MEM_HOST))
        return outs

    def close(self):
        if self.h:
            self.lib.chs_props_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
