"""ctypes binding of libchapsim_hip.so (the HIP product library).

The library is built in-tree by ``chapsim2_amd.build``; loading fails loudly
when it is missing.  There is no CPU fallback for any compute entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libchapsim_hip.so")

CHS_OK, CHS_EINVAL, CHS_ENOMEM, CHS_EHIP, CHS_ENAN, CHS_EUNSUP, CHS_ERANGE = 0, -1, -2, -3, -4, -5, -6
MEM_HOST, MEM_DEVICE = 0, 1

# every entry point of include/chapsim_hip.h and include/chapsim_hip_inspect.h
EXPORTS = [
    "chs_abi_version", "chs_last_error", "chs_create", "chs_destroy", "chs_sync", "chs_op3d",
    "chs_poisson", "chs_set_state", "chs_get_state", "chs_update_re", "chs_substep", "chs_run",
    "chs_momentum_rhs", "chs_mass_check", "chs_cfl", "chs_bulk", "chs_velocity_range", "chs_velocity_range_noalloc",
    "chs_divergence",
    "chs_batch_create", "chs_batch_destroy", "chs_batch_instance", "chs_batch_run",
    "chs_props_create", "chs_props_destroy", "chs_props_info", "chs_props_from_rhoh", "chs_props_from_h",
    "chs_mlp_create", "chs_mlp_destroy", "chs_mlp_info", "chs_mlp_get_params", "chs_mlp_set_params",
    "chs_mlp_grad", "chs_mlp_train", "chs_mlp_eval",
    "chs_inspect_coeff", "chs_inspect_op1d", "chs_inspect_projection", "chs_profile_begin", "chs_profile_end", "chs_tune_report",
    "chs_op3d_f32", "chs_poisson_f32", "chs_tune_force", "chs_tune_set",
    "chs_props_set_surrogate", "chs_thermo_enable", "chs_thermo_set_rhoh", "chs_thermo_set_mode", "chs_thermo_get",
    "chs_energy_rhs", "chs_mlp_set_matrix_cores",
]


class ChsThermoDesc(C.Structure):
    _fields_ = [("props", C.c_void_p), ("mode", C.c_int), ("t_wall", C.c_double * 2), ("prandtl", C.c_double),
                ("t_init", C.c_double)]


class ChsDesc(C.Structure):
    _fields_ = [
        ("nc", C.c_int * 3),
        ("periodic_y", C.c_int),
        ("lxx", C.c_double), ("lyb", C.c_double), ("lyt", C.c_double), ("lzz", C.c_double),
        ("istret", C.c_int),
        ("rstret", C.c_double),
        ("iacc", C.c_int),
        ("dt", C.c_double),
        ("ren", C.c_double), ("reninit", C.c_double),
        ("init_re_to", C.c_int),
        ("idriven", C.c_int),
        ("poisson", C.c_int),
        ("uwall", (C.c_double * 3) * 2),
        ("precision", C.c_int),
        ("projection", C.c_int),
    ]


class ChsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"chs error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m chapsim2_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    L.chs_abi_version.restype = C.c_int
    L.chs_last_error.restype = C.c_char_p
    L.chs_create.argtypes = [C.POINTER(ChsDesc), C.POINTER(vp)]
    L.chs_destroy.argtypes = [vp]
    L.chs_sync.argtypes = [vp]
    L.chs_op3d.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, vp, ip, vp, ip, vp, C.c_int]
    L.chs_poisson.argtypes = [vp, vp, C.c_int]
    L.chs_op3d_f32.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, vp, ip, vp, ip, vp]
    L.chs_poisson_f32.argtypes = [vp, vp]
    L.chs_set_state.argtypes = [vp, vp, vp, vp, vp, C.c_int]
    L.chs_get_state.argtypes = [vp, vp, vp, vp, vp, C.c_int]
    L.chs_update_re.argtypes = [vp, C.c_int]
    L.chs_substep.argtypes = [vp, C.c_int]
    L.chs_run.argtypes = [vp, C.c_int, C.c_int]
    L.chs_momentum_rhs.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int]
    L.chs_mass_check.argtypes = [vp, dp]
    L.chs_cfl.argtypes = [vp, dp]
    L.chs_bulk.argtypes = [vp, dp]
    L.chs_velocity_range.argtypes = [vp, dp]
    L.chs_velocity_range_noalloc.argtypes = [vp, dp]
    L.chs_props_set_surrogate.argtypes = [vp, vp]
    L.chs_thermo_enable.argtypes = [vp, C.POINTER(ChsThermoDesc)]
    L.chs_thermo_set_rhoh.argtypes = [vp, vp, C.c_int]
    L.chs_thermo_set_mode.argtypes = [vp, C.c_int]
    L.chs_thermo_get.argtypes = [vp] + [vp] * 9 + [C.c_int]
    L.chs_energy_rhs.argtypes = [vp, C.c_int, vp, C.c_int]
    L.chs_mlp_set_matrix_cores.argtypes = [vp, C.c_int]
    L.chs_divergence.argtypes = [vp, vp, C.c_int]
    L.chs_inspect_coeff.restype = C.c_double
    L.chs_inspect_coeff.argtypes = [C.c_int] * 6
    L.chs_batch_create.argtypes = [C.POINTER(ChsDesc), C.c_int, C.POINTER(vp)]
    L.chs_batch_destroy.argtypes = [vp]
    L.chs_batch_instance.argtypes = [vp, C.c_int]
    L.chs_batch_instance.restype = vp
    L.chs_batch_run.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.chs_props_create.argtypes = [dp, C.c_int, C.c_double, C.POINTER(vp)]
    L.chs_props_destroy.argtypes = [vp]
    L.chs_props_info.argtypes = [vp, dp, dp, dp]
    L.chs_props_from_rhoh.argtypes = [vp, vp, C.c_long] + [vp] * 7 + [C.c_int, C.c_int]
    L.chs_props_from_h.argtypes = [vp, vp, C.c_long] + [vp] * 6 + [C.c_int]
    L.chs_mlp_create.argtypes = [vp, C.c_int, C.c_uint, C.POINTER(vp)]
    L.chs_mlp_destroy.argtypes = [vp]
    L.chs_mlp_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_float)]
    L.chs_mlp_get_params.argtypes = [vp, vp, vp]
    L.chs_mlp_set_params.argtypes = [vp, vp]
    L.chs_mlp_grad.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_float)]
    L.chs_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]
    L.chs_mlp_eval.argtypes = [vp, vp, C.c_long] + [vp] * 5 + [C.c_int]
    L.chs_inspect_op1d.argtypes = [C.POINTER(ChsDesc), C.c_int, C.c_int, C.c_int, ip, dp, C.c_int, dp,
                                   C.c_int, dp]
    L.chs_inspect_projection.argtypes = [C.POINTER(ChsDesc), dp, dp, dp, dp, dp]
    L.chs_profile_begin.argtypes = [vp, C.c_int]
    L.chs_profile_end.argtypes = [vp, dp, C.POINTER(C.c_longlong), dp]
    L.chs_tune_report.argtypes = [C.c_char_p, C.c_int]
    L.chs_tune_force.argtypes = [C.c_char_p, C.c_int]
    L.chs_tune_set.argtypes = [C.c_char_p, C.c_int]
    for name in EXPORTS:
        if name != "chs_batch_instance":
            getattr(L, name).restype = getattr(L, name).restype or C.c_int
    L.chs_last_error.restype = C.c_char_p
    L.chs_inspect_coeff.restype = C.c_double
    _lib = L
    return L


def check(rc):
    if rc != CHS_OK:
        msg = load().chs_last_error()
        raise ChsError(rc, msg.decode() if msg else "")
    return rc


def ivec(v):
    return (C.c_int * len(v))(*[int(x) for x in v])
