"""MLP surrogate of a property table (libchapsim_hip chs_mlp_*):
h -> (T, rho, mu, kappa, cp), trained on the GPU on samples of the table's
own interpolation; the table stays the fallback and the yardstick."""
import ctypes as C

import numpy as np

from . import _lib

NP = 4613
OUTPUTS = ("t", "d", "m", "k", "cp")


class Surrogate:
    def __init__(self, table, nsamples=1_000_000, seed=1):
        """table: a chapsim2_amd.props.PropertyTable."""
        self.lib = _lib.load()
        self.table = table
        self.h = C.c_void_p()
        _lib.check(self.lib.chs_mlp_create(table.h, int(nsamples), int(seed), C.byref(self.h)))

    def info(self):
        n = C.c_int()
        norm = (C.c_float * 12)()
        _lib.check(self.lib.chs_mlp_info(self.h, C.byref(n), norm))
        return n.value, np.array(norm[:], dtype=np.float32)

    def params(self):
        p32 = np.zeros(NP, dtype=np.float32)
        p16 = np.zeros(NP, dtype=np.uint16)
        _lib.check(self.lib.chs_mlp_get_params(self.h, p32.ctypes.data, p16.ctypes.data))
        return p32, p16

    def set_params(self, p32):
        p = np.ascontiguousarray(p32, dtype=np.float32)
        _lib.check(self.lib.chs_mlp_set_params(self.h, p.ctypes.data))

    def grad(self, x, t):
        """(loss, gradient) of the mean squared error on the given normalised
        samples x[B], t[B, 5] with the BF16 weights."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(NP, dtype=np.float32)
        loss = C.c_float()
        _lib.check(self.lib.chs_mlp_grad(self.h, x.ctypes.data, t.ctypes.data, x.size, g.ctypes.data,
                                         C.byref(loss)))
        return loss.value, g

    def train(self, steps, batch=4096, lr=1e-3):
        loss = C.c_float()
        _lib.check(self.lib.chs_mlp_train(self.h, int(steps), int(batch), C.c_float(lr), C.byref(loss)))
        return loss.value

    def __call__(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        outs = {n: np.zeros_like(h) for n in OUTPUTS}
        _lib.check(self.lib.chs_mlp_eval(self.h, h.ctypes.data, h.size, *[outs[n].ctypes.data for n in OUTPUTS],
                                         _lib.MEM_HOST))
        return outs

    def close(self):
        if self.h:
            self.lib.chs_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
