"""Domain / solver handle: the host-side counterpart of the reference's
``t_domain`` + ``t_flow`` set-up (src/geometry.f90:410-593,
src/bc_general.f90:137-757, src/input_general.f90:842-888), backed by a
``chs_handle`` of libchapsim_hip.
"""
import ctypes as C

import numpy as np

from . import _lib

# boundary kinds (src/modules.f90:223-234) and accuracies (:248-251)
IBC_INTERIOR, IBC_PERIODIC, IBC_SYMMETRIC, IBC_ASYMMETRIC, IBC_DIRICHLET, IBC_NEUMANN, IBC_INTRPL = range(7)
IACCU_CD2, IACCU_CD4, IACCU_CP4, IACCU_CP6 = 1, 2, 3, 4
FFT_FISHPACK_2DFFT, FFT_2DECOMP_3DFFT = 2, 3
PREC_FP64, PREC_MIXED = 0, 1
PROJ_REFERENCE, PROJ_EXACT = 0, 1


def make_desc(nc, lxx, lyb, lyt, lzz, periodic_y=False, istret=0, rstret=0.0, iacc=IACCU_CP6,
              dt=1e-3, ren=1.0, reninit=None, init_re_to=0, idriven=0, poisson=FFT_FISHPACK_2DFFT,
              uwall=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), precision=PREC_FP64, projection=0):
    d = _lib.ChsDesc()
    for k in range(3):
        d.nc[k] = int(nc[k])
    d.periodic_y = int(bool(periodic_y))
    d.lxx, d.lyb, d.lyt, d.lzz = float(lxx), float(lyb), float(lyt), float(lzz)
    d.istret, d.rstret = int(istret), float(rstret)
    d.iacc = int(iacc)
    d.dt = float(dt)
    d.ren = float(ren)
    d.reninit = float(ren if reninit is None else reninit)
    d.init_re_to = int(init_re_to)
    d.idriven = int(idriven)
    d.poisson = int(poisson)
    for s in range(2):
        for v in range(3):
            d.uwall[s][v] = float(uwall[s][v])
    d.precision = int(precision)
    d.projection = int(projection)
    return d


class Domain:
    """One channel (y walls) or periodic box on the GPU."""

    def __init__(self, nc, lxx, lyb, lyt, lzz, **kw):
        self.lib = _lib.load()
        self.desc = make_desc(nc, lxx, lyb, lyt, lzz, **{k: v for k, v in kw.items() if k != "_handle"})
        self.nc = tuple(int(n) for n in nc)
        py = self.nc[1] if kw.get("periodic_y", False) else self.nc[1] + 1
        self.np = (self.nc[0], py, self.nc[2])
        nx, ny, nz = self.nc
        self.shapes = {
            "qx": (nx, ny, nz), "qy": (nx, py, nz), "qz": (nx, ny, nz), "pr": (nx, ny, nz),
            "pcc": (nx, ny, nz), "cpc": (nx, py, nz), "ccp": (nx, ny, nz), "ccc": (nx, ny, nz),
            "ppc": (nx, py, nz), "pcp": (nx, ny, nz), "cpp": (nx, py, nz),
        }
        handle = kw.get("_handle")
        if handle is not None:
            self.h = C.c_void_p(handle)  # owned by a Batch
            self.owned = False
            return
        h = C.c_void_p()
        _lib.check(self.lib.chs_create(C.byref(self.desc), C.byref(h)))
        self.h = h
        self.owned = True

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "owned", True):
                self.lib.chs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ---------------------------------------------------------
    def set_state(self, qx, qy, qz, pr):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (qx, qy, qz, pr)]
        for a, k in zip(arrs, ("qx", "qy", "qz", "pr")):
            if a.shape != self.shapes[k]:
                raise ValueError(f"{k} has shape {a.shape}, expected {self.shapes[k]}")
        _lib.check(self.lib.chs_set_state(self.h, *[a.ctypes.data for a in arrs], _lib.MEM_HOST))

    def get_state(self):
        out = [np.zeros(self.shapes[k], order="F") for k in ("qx", "qy", "qz", "pr")]
        _lib.check(self.lib.chs_get_state(self.h, *[a.ctypes.data for a in out], _lib.MEM_HOST))
        return out

    def update_re(self, it):
        _lib.check(self.lib.chs_update_re(self.h, int(it)))

    def substep(self, isub):
        _lib.check(self.lib.chs_substep(self.h, int(isub)))

    def run(self, iter0, nsteps):
        _lib.check(self.lib.chs_run(self.h, int(iter0), int(nsteps)))

    def sync(self):
        _lib.check(self.lib.chs_sync(self.h))

    def momentum_rhs(self, isub):
        out = [np.zeros(self.shapes[k], order="F") for k in ("qx", "qy", "qz")]
        _lib.check(self.lib.chs_momentum_rhs(self.h, int(isub), *[a.ctypes.data for a in out], _lib.MEM_HOST))
        return out

    def mass_check(self):
        out = np.zeros(3)
        _lib.check(self.lib.chs_mass_check(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def cfl(self):
        """(convective CFL number, diffusion number) of the current state,
        Check_cfl_convection / Check_cfl_diffusion
        (src/tools_solver.f90:220-375)."""
        out = np.zeros(2)
        _lib.check(self.lib.chs_cfl(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def bulk(self):
        """(bulk mean kinetic energy, bulk qx) of the current state, the
        values of write_monitor_bulk (src/io_monitor.f90:207-229)."""
        out = np.zeros(2)
        _lib.check(self.lib.chs_bulk(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def velocity_range(self, check=True):
        """[max qx, min qx, max qy, min qy, max qz, min qz] of the current
        state, the per-step check of Find_max_min_3d
        (src/tools_general.f90:1071-1113).  The library reports NaN or
        infinite velocity values as CHS_ENAN: with ``check`` that raises
        ChsError; without it the return code comes back as the second item."""
        out = np.zeros(6)
        rc = self.lib.chs_velocity_range(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        if check:
            _lib.check(rc)
            return out
        return out, rc

    # ---- variable properties (heated channel) ------------------------------
    def thermo_enable(self, table, t_wall, prandtl, t_init, mode=0):
        """Isothermal y walls at t_wall (bottom, top; non-dimensional), the
        reference Prandtl number, a uniform initial state at t_init, the
        property table (chapsim2_amd.props.PropertyTable, kept alive here)
        and its mode (0 table, 1 surrogate attached to the table)."""
        d = _lib.ChsThermoDesc()
        d.props = table.h.value if hasattr(table.h, "value") else table.h
        d.mode = int(mode)
        d.t_wall[0], d.t_wall[1] = float(t_wall[0]), float(t_wall[1])
        d.prandtl, d.t_init = float(prandtl), float(t_init)
        self._table = table
        _lib.check(self.lib.chs_thermo_enable(self.h, C.byref(d)))

    def thermo_set_mode(self, mode):
        _lib.check(self.lib.chs_thermo_set_mode(self.h, int(mode)))

    def thermo_set_rhoh(self, rhoh):
        r = np.asfortranarray(rhoh, dtype=np.float64)
        assert r.shape == self.shapes["ccc"], r.shape
        _lib.check(self.lib.chs_thermo_set_rhoh(self.h, r.ctypes.data, _lib.MEM_HOST))

    def thermo_get(self):
        """dict of rhoh, h, t, k, d, m (cell centres) and gx, gy, gz"""
        names = ("rhoh", "h", "t", "k", "d", "m")
        out = {k: np.zeros(self.shapes["ccc"], order="F") for k in names}
        for k, sh in (("gx", "qx"), ("gy", "qy"), ("gz", "qz")):
            out[k] = np.zeros(self.shapes[sh], order="F")
        _lib.check(self.lib.chs_thermo_get(self.h, *[out[k].ctypes.data for k in names + ("gx", "gy", "gz")],
                                           _lib.MEM_HOST))
        return out

    def energy_rhs(self, isub):
        out = np.zeros(self.shapes["ccc"], order="F")
        _lib.check(self.lib.chs_energy_rhs(self.h, int(isub), out.ctypes.data, _lib.MEM_HOST))
        return out

    def divergence(self):
        out = np.zeros(self.shapes["ccc"], order="F")
        _lib.check(self.lib.chs_divergence(self.h, out.ctypes.data, _lib.MEM_HOST))
        return out


class Batch:
    """Independent channels advanced together (chs_batch_*): each spec is a
    dict of Domain arguments (nc, lxx, lyb, lyt, lzz and keywords).  Every
    RK3 step of all instances runs as one captured graph when graph=True."""

    def __init__(self, specs):
        self.lib = _lib.load()
        n = len(specs)
        descs = (_lib.ChsDesc * n)()
        for i, sp in enumerate(specs):
            kw = {k: v for k, v in sp.items() if k not in ("nc", "lxx", "lyb", "lyt", "lzz")}
            descs[i] = make_desc(sp["nc"], sp["lxx"], sp["lyb"], sp["lyt"], sp["lzz"], **kw)
        h = C.c_void_p()
        _lib.check(self.lib.chs_batch_create(descs, n, C.byref(h)))
        self.h = h
        self.domains = []
        for i, sp in enumerate(specs):
            kw = {k: v for k, v in sp.items() if k not in ("nc", "lxx", "lyb", "lyt", "lzz")}
            hi = self.lib.chs_batch_instance(self.h, i)
            self.domains.append(Domain(sp["nc"], sp["lxx"], sp["lyb"], sp["lyt"], sp["lzz"], _handle=hi, **kw))

    def run(self, iter0, nsteps, graph=True):
        _lib.check(self.lib.chs_batch_run(self.h, int(iter0), int(nsteps), int(bool(graph))))

    def close(self):
        if getattr(self, "h", None):
            for d in self.domains:
                d.close()
            self.lib.chs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
