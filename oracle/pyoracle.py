"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  The HIP product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN, BC_INTRPL = range(7)
CD2, CD4, CP4, CP6 = 1, 2, 3, 4
MIDP_C2P, MIDP_P2C, D1_C2C, D1_P2P, D1_C2P, D1_P2C = range(6)
FAM_D1C2C, FAM_D1P2P, FAM_D1C2P, FAM_D1P2C, FAM_M1C2P, FAM_M1P2C = range(6)


class OrcDom(C.Structure):
    _fields_ = [
        ("nc", C.c_int * 3), ("np", C.c_int * 3), ("npgeo", C.c_int * 3),
        ("periodic", C.c_int * 3), ("stretching", C.c_int), ("istret", C.c_int),
        ("rstret", C.c_double),
        ("lxx", C.c_double), ("lyb", C.c_double), ("lyt", C.c_double), ("lzz", C.c_double),
        ("h", C.c_double * 3), ("h1r", C.c_double * 3), ("h2r", C.c_double * 3),
        ("yp", C.POINTER(C.c_double)), ("yc", C.POINTER(C.c_double)),
        ("ymp", C.POINTER(C.c_double)), ("ymc", C.POINTER(C.c_double)),
    ]


class OrcFlowCfg(C.Structure):
    _fields_ = [
        ("nc", C.c_int * 3),
        ("lxx", C.c_double), ("lyb", C.c_double), ("lyt", C.c_double), ("lzz", C.c_double),
        ("istret", C.c_int), ("rstret", C.c_double),
        ("wall_y", C.c_int), ("uwall", (C.c_double * 3) * 2),
        ("iacc", C.c_int), ("dt", C.c_double), ("ren", C.c_double), ("reninit", C.c_double),
        ("init_re_to", C.c_int), ("idriven", C.c_int), ("poisson", C.c_int),
        ("projection", C.c_int),
    ]


def _stale():
    if not os.path.exists(LIB_PATH):
        return True
    t = os.path.getmtime(LIB_PATH)
    srcs = [f for f in os.listdir(HERE) if f.endswith((".c", ".h")) or f == "Makefile"]
    return any(os.path.getmtime(os.path.join(HERE, f)) > t for f in srcs)


def build(force=False):
    """Build liboracle.so when it is missing or older than its sources (the
    sources are absent on a machine that only received the built library)."""
    have_src = os.path.exists(os.path.join(HERE, "orc_ops.c"))
    if force or (have_src and _stale()) or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_coeff.restype = C.c_double
        L.orc_coeff.argtypes = [C.c_int] * 6
        L.orc_tdma_preprocess.argtypes = [dp, dp, dp, dp, C.c_int]
        L.orc_tdma_solve.argtypes = [C.c_int, dp, dp, dp, dp, dp, C.c_int]
        L.orc_lhs.argtypes = [C.c_int] * 5 + [dp] * 4
        L.orc_dom_init.argtypes = [C.POINTER(OrcDom), ip, ip] + [C.c_double] * 4 + [C.c_int, C.c_double]
        L.orc_dom_free.argtypes = [C.POINTER(OrcDom)]
        L.orc_op1d.argtypes = [C.POINTER(OrcDom), C.c_int, C.c_int, C.c_int, ip, dp, C.c_int, dp, C.c_int, dp]
        L.orc_op3d.restype = C.c_int
        L.orc_op3d.argtypes = [C.POINTER(OrcDom), C.c_int, C.c_int, C.c_int, ip, dp, ip, dp, ip, dp]
        L.orc_rfftf.argtypes = [C.c_int, dp]
        L.orc_rfftb.argtypes = [C.c_int, dp]
        L.orc_flow_create.restype = C.c_void_p
        L.orc_flow_create.argtypes = [C.POINTER(OrcFlowCfg)]
        L.orc_flow_destroy.argtypes = [C.c_void_p]
        L.orc_flow_dom.restype = C.POINTER(OrcDom)
        L.orc_flow_dom.argtypes = [C.c_void_p]
        L.orc_flow_set_state.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_flow_get_state.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_flow_get_rhs0.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_flow_update_re.argtypes = [C.c_void_p, C.c_int]
        L.orc_solve_momentum_eq.argtypes = [C.c_void_p, C.c_int]
        L.orc_flow_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_project.argtypes = [C.c_void_p, C.c_int]
        L.orc_divergence.argtypes = [C.c_void_p, dp]
        L.orc_mass_check.argtypes = [C.c_void_p, dp]
        L.orc_cfl.argtypes = [C.c_void_p, dp]
        L.orc_bulk.argtypes = [C.c_void_p, dp]
        L.orc_momentum_rhs.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.orc_poisson_solve.argtypes = [C.c_void_p, dp]
        L.orc_vol_average_x.restype = C.c_double
        L.orc_vol_average_x.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_props_from_t.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_thermo_enable.restype = C.c_int
        L.orc_thermo_enable.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_double, C.c_double]
        L.orc_thermo_set_rhoh.restype = C.c_long
        L.orc_thermo_set_rhoh.argtypes = [C.c_void_p, dp]
        L.orc_thermo_get.argtypes = [C.c_void_p] + [dp] * 9
        L.orc_energy_rhs.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_solve_energy_eq.restype = C.c_long
        L.orc_solve_energy_eq.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return (C.c_int * len(a))(*a)


def coeff(fam, is_rhs, row, k, bc, acc):
    return lib().orc_coeff(fam, is_rhs, row, k, bc, acc)


def tdma_preprocess(a, b, c):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.array(c, dtype=np.float64)
    d = np.zeros_like(a)
    lib().orc_tdma_preprocess(_p(a), _p(b), _p(c), _p(d), len(a))
    return c, d


def tdma_solve(peri, x, a, b, c, d):
    x = np.array(x, dtype=np.float64)
    lib().orc_tdma_solve(int(peri), _p(x), _p(np.ascontiguousarray(a, float)),
                         _p(np.ascontiguousarray(b, float)), _p(np.ascontiguousarray(c, float)),
                         _p(np.ascontiguousarray(d, float)), len(x))
    return x


def lhs(fam, n, bc1, bc2, acc):
    a, b, c, d = (np.zeros(n) for _ in range(4))
    lib().orc_lhs(fam, n, bc1, bc2, acc, _p(a), _p(b), _p(c), _p(d))
    return a, b, c, d


class Dom:
    """Mesh of the oracle (src/geometry.f90:410-593)."""

    def __init__(self, nc, periodic, lxx=1.0, lyb=0.0, lyt=1.0, lzz=1.0, istret=0, rstret=0.0):
        self.s = OrcDom()
        lib().orc_dom_init(C.byref(self.s), _ip(nc), _ip([int(bool(p)) for p in periodic]),
                           lxx, lyb, lyt, lzz, istret, rstret)
        self._own = True

    @classmethod
    def from_ptr(cls, ptr, keep):
        self = cls.__new__(cls)
        self.s = ptr.contents
        self._own = False
        self._keep = keep
        return self

    def __del__(self):
        if getattr(self, "_own", False) and _lib is not None:
            lib().orc_dom_free(C.byref(self.s))
            self._own = False

    @property
    def nc(self):
        return list(self.s.nc)

    @property
    def np(self):
        return list(self.s.np)

    @property
    def h(self):
        return list(self.s.h)

    def _vec(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    @property
    def yp(self):
        return self._vec(self.s.yp, self.s.npgeo[1])

    @property
    def yc(self):
        return self._vec(self.s.yc, self.s.nc[1])

    @property
    def ymp(self):
        return self._vec(self.s.ymp, self.s.npgeo[1])

    @property
    def ymc(self):
        return self._vec(self.s.ymc, self.s.nc[1])

    def op1d(self, d, op, iacc, ibc, fi, nout, fbc=None):
        fi = np.ascontiguousarray(fi, dtype=np.float64)
        fo = np.zeros(nout)
        fb = None if fbc is None else np.ascontiguousarray(fbc, dtype=np.float64)
        lib().orc_op1d(C.byref(self.s), d, op, iacc, _ip(ibc), _p(fi), len(fi), _p(fo), nout,
                       None if fb is None else _p(fb))
        return fo

    def op3d(self, d, op, iacc, ibc, fi, out_shape, fbc=None):
        """fi and the result are Fortran-ordered arrays (x fastest)."""
        fi = np.asfortranarray(fi, dtype=np.float64)
        fo = np.zeros(out_shape, order="F")
        fb = None if fbc is None else np.asfortranarray(fbc, dtype=np.float64)
        rc = lib().orc_op3d(C.byref(self.s), d, op, iacc, _ip(ibc),
                            fi.ctypes.data_as(C.POINTER(C.c_double)), _ip(fi.shape),
                            fo.ctypes.data_as(C.POINTER(C.c_double)), _ip(out_shape),
                            None if fb is None else fb.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise ValueError("shape mismatch")
        return fo


def rfftf(r):
    r = np.array(r, dtype=np.float64)
    lib().orc_rfftf(len(r), _p(r))
    return r


def rfftb(r):
    r = np.array(r, dtype=np.float64)
    lib().orc_rfftb(len(r), _p(r))
    return r


class Flow:
    """Flow solver of the oracle; fields are Fortran-ordered numpy arrays."""

    def __init__(self, nc, lxx, lyb, lyt, lzz, istret=0, rstret=0.0, wall_y=True,
                 iacc=CP6, dt=1e-3, ren=1.0, reninit=None, init_re_to=0, idriven=0,
                 poisson=2, uwall=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), projection=0):
        cfg = OrcFlowCfg()
        for d in range(3):
            cfg.nc[d] = nc[d]
        cfg.lxx, cfg.lyb, cfg.lyt, cfg.lzz = lxx, lyb, lyt, lzz
        cfg.istret, cfg.rstret = istret, rstret
        cfg.wall_y = int(wall_y)
        for s in range(2):
            for v in range(3):
                cfg.uwall[s][v] = uwall[s][v]
        cfg.iacc, cfg.dt, cfg.ren = iacc, dt, ren
        cfg.reninit = ren if reninit is None else reninit
        cfg.init_re_to, cfg.idriven, cfg.poisson = init_re_to, idriven, poisson
        cfg.projection = int(projection)
        self.cfg = cfg
        self.h = lib().orc_flow_create(C.byref(cfg))
        self.dom = Dom.from_ptr(lib().orc_flow_dom(self.h), self)
        nx, ny, nz = nc
        npx, npy, npz = self.dom.np
        self.shapes = {"qx": (npx, ny, nz), "qy": (nx, npy, nz), "qz": (nx, ny, npz), "pr": (nx, ny, nz)}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().orc_flow_destroy(self.h)
            self.h = None

    def set_state(self, qx, qy, qz, pr):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (qx, qy, qz, pr)]
        for a, k in zip(arrs, ("qx", "qy", "qz", "pr")):
            assert a.shape == self.shapes[k], (k, a.shape, self.shapes[k])
        lib().orc_flow_set_state(self.h, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])

    def get_state(self):
        out = [np.zeros(self.shapes[k], order="F") for k in ("qx", "qy", "qz", "pr")]
        lib().orc_flow_get_state(self.h, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in out])
        return out

    def get_rhs0(self):
        out = [np.zeros(self.shapes[k], order="F") for k in ("qx", "qy", "qz")]
        lib().orc_flow_get_rhs0(self.h, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in out])
        return out

    def update_re(self, it):
        lib().orc_flow_update_re(self.h, it)

    def substep(self, isub):
        lib().orc_solve_momentum_eq(self.h, isub)

    def project(self, isub):
        lib().orc_project(self.h, isub)

    def step(self, it):
        lib().orc_flow_step(self.h, it)

    def momentum_rhs(self, isub):
        out = [np.zeros(self.shapes[k], order="F") for k in ("qx", "qy", "qz")]
        lib().orc_momentum_rhs(self.h, isub, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in out])
        return out

    def divergence(self):
        out = np.zeros(self.shapes["pr"], order="F")
        lib().orc_divergence(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def mass_check(self):
        out = np.zeros(3)
        lib().orc_mass_check(self.h, _p(out))
        return out

    def cfl(self):
        """(convective CFL, diffusion number), src/tools_solver.f90:220-375."""
        out = np.zeros(2)
        lib().orc_cfl(self.h, _p(out))
        return out

    def bulk(self):
        """(bulk mean kinetic energy, bulk qx), src/io_monitor.f90:207-229."""
        out = np.zeros(2)
        lib().orc_bulk(self.h, _p(out))
        return out

    def poisson(self, rhs):
        r = np.array(rhs, dtype=np.float64, order="F")
        lib().orc_poisson_solve(self.h, r.ctypes.data_as(C.POINTER(C.c_double)))
        return r

    # ---- variable-property path ----
    def thermo_enable(self, props, t_wall, prandtl, t_init):
        """Isothermal y walls at t_wall (non-dimensional), reference Prandtl
        number, uniform initial state at t_init; g from the current q."""
        self._props = props  # the table must outlive the solver's use of it
        tw = np.ascontiguousarray(t_wall, dtype=np.float64)
        rc = lib().orc_thermo_enable(self.h, C.addressof(props.s), _p(tw), float(prandtl), float(t_init))
        if rc != 0:
            raise ValueError("variable properties need stationary y walls")
        nx, ny, nz = self.dom.nc
        self.shapes["ccc"] = (nx, ny, nz)

    def thermo_set_rhoh(self, rhoh):
        r = np.asfortranarray(rhoh, dtype=np.float64)
        assert r.shape == self.shapes["pr"]
        return int(lib().orc_thermo_set_rhoh(self.h, r.ctypes.data_as(C.POINTER(C.c_double))))

    def thermo_get(self):
        """dict of rhoh, h, t, k, d, m (cell centres) and gx, gy, gz"""
        names = ("rhoh", "h", "t", "k", "d", "m")
        out = {k: np.zeros(self.shapes["pr"], order="F") for k in names}
        for k, sh in (("gx", "qx"), ("gy", "qy"), ("gz", "qz")):
            out[k] = np.zeros(self.shapes[sh], order="F")
        lib().orc_thermo_get(self.h, *[out[k].ctypes.data_as(C.POINTER(C.c_double))
                                       for k in names + ("gx", "gy", "gz")])
        return out

    def energy_rhs(self, isub):
        out = np.zeros(self.shapes["pr"], order="F")
        lib().orc_energy_rhs(self.h, isub, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def energy_substep(self, isub):
        return int(lib().orc_solve_energy_eq(self.h, isub))


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class OrcProps(C.Structure):
    _MAX = 16384
    _fields_ = [("n", C.c_int)] + [(nm, C.c_double * 16384) for nm in ("t", "d", "m", "k", "b", "cp", "h", "rhoh")] + [
        ("ref", C.c_double * 7), ("dhmin", C.c_double), ("dhmax", C.c_double), ("dimensional", C.c_int)]


class Props:
    """Thermal-property table of the oracle (src/input_thermo.f90:121-693)."""

    def __init__(self, raw, ref_t0):
        raw = np.ascontiguousarray(raw, dtype=np.float64)
        self.s = OrcProps()
        rc = lib().orc_props_build(_p(raw), raw.shape[0], C.c_double(ref_t0), C.byref(self.s))
        if rc != 0:
            raise ValueError("bad table")

    @property
    def dhmin(self):
        return self.s.dhmin

    @property
    def dhmax(self):
        return self.s.dhmax

    def column(self, name):
        return np.array(getattr(self.s, name)[: self.s.n])

    def from_h(self, h):
        out = (C.c_double * 6)()
        lib().orc_props_from_h(C.byref(self.s), C.c_double(h), out)
        return np.array(out[:])

    def from_t(self, t):
        """(h, d, m, k, t, b, cp) from the temperature (table branch of
        ftp_refresh_thermal_properties_from_T_undim)"""
        out = (C.c_double * 7)()
        lib().orc_props_from_t(C.byref(self.s), C.c_double(t), out)
        return np.array(out[:])

    def from_dh(self, rhoh):
        """(h, d, m, k, t, b, cp), clamped flag"""
        rhoh = np.atleast_1d(np.asarray(rhoh, dtype=np.float64))
        res = np.zeros((len(rhoh), 7))
        flags = np.zeros(len(rhoh), dtype=bool)
        out = (C.c_double * 7)()
        for i, v in enumerate(rhoh):
            flags[i] = lib().orc_props_from_dh(C.byref(self.s), C.c_double(v), out) != 0
            res[i] = out[:]
        return res, flags
