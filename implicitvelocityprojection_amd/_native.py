"""ctypes binding of libivp.so (the C ABI declared in include/ivp.h).

The library must have been built (``python -m implicitvelocityprojection_amd.build``); there is
no fallback: a missing or unloadable library raises ImportError.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libivp.so")

IVP_OK = 0
IVP_NOT_CONVERGED = 1
IVP_NOT_PD = 2
IVP_NEIGHBOR_OVERFLOW = 3
IVP_ERR_ARG = -1
IVP_ERR_HIP = -2
IVP_ERR_CAPACITY = -3
IVP_ERR_STATE = -4
IVP_ERR_GRID = -5
IVP_ERR_INTERNAL = -6

FIELD_X, FIELD_V, FIELD_DENSITY, FIELD_K, FIELD_DIVERGENCE = 0, 1, 2, 3, 4
OPT_GRADIENT_CACHE = 1
ROWS_IDS, ROWS_F32, ROWS_F16 = 0, 1, 2     # values of OPT_GRADIENT_CACHE
OPT_SCENE_CG_CHUNK = 2
OPT_SCENE_GRAPH = 3
OPT_SCENE_LANES = 4
OPT_BUILD_LANES = 5
OPT_W18_WARM_START = 6
OPT_W18_PRECONDITIONER = 7
OPT_TILE_WINDOWS = 8
MIN_SCENE_DYNAMIC = 256

# kernel classes of ivp_profile_read / ivp_profile_enable masks
KCLASSES = ["sort", "build", "ip_pack", "ip_rhs", "cg_J", "cg_JT", "cg_update", "cg_p", "visc", "w18_Av",
            "w18_update", "w18_p", "w18_pre", "cg_refresh", "st_setup", "st_Av", "prj_head", "prj_tail", "glue",
            "window", "scan"]


class Config(ctypes.Structure):
    _fields_ = [("support_radius", ctypes.c_float), ("domain_start", ctypes.c_float * 3),
                ("domain_size", ctypes.c_float * 3), ("cache_size", ctypes.c_int), ("material_fluid", ctypes.c_int)]


class IPParams(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_float), ("tol_opt", ctypes.c_double), ("max_iteration_opt", ctypes.c_int),
                ("use_fixed_iters", ctypes.c_int), ("tol_pcg", ctypes.c_double), ("max_iteration_pcg", ctypes.c_int)]


class IPStats(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_int), ("density_errors", ctypes.POINTER(ctypes.c_double)),
                ("pcg_iters", ctypes.POINTER(ctypes.c_int)), ("n_density_errors", ctypes.c_int),
                ("n_pcg", ctypes.c_int), ("outer_iters", ctypes.c_int), ("final_density_error", ctypes.c_double),
                ("pAp_bad", ctypes.c_double), ("pcg_iter_bad", ctypes.c_int), ("pairs_traversed", ctypes.c_longlong),
                ("rows_traversed", ctypes.c_longlong), ("cg_launch_chunks", ctypes.c_int),
                ("solve_ms", ctypes.c_double), ("df_outer_iters", ctypes.c_int), ("df_ms", ctypes.c_double),
                ("iter_costs", ctypes.POINTER(ctypes.c_double)), ("n_iter_costs", ctypes.c_int)]


class PRJParams(ctypes.Structure):
    _fields_ = [("tol_opt", ctypes.c_double), ("max_iteration_opt", ctypes.c_int), ("use_fixed_iters", ctypes.c_int),
                ("omega", ctypes.c_float), ("warm_start", ctypes.c_int), ("enable_df", ctypes.c_int),
                ("df_num_threshold", ctypes.c_int), ("df_tol_opt", ctypes.c_double)]


class CGStats(ctypes.Structure):
    _fields_ = [("cg_iters", ctypes.c_int), ("residual_norm2", ctypes.c_double)]


class StepParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("gravity", ctypes.c_float * 3), ("viscosity_method", ctypes.c_int),
                ("viscosity_coeff", ctypes.c_float), ("ip", IPParams), ("surface_tension_method", ctypes.c_int),
                ("surface_tension_coeff", ctypes.c_float), ("pressure_method", ctypes.c_int), ("prj", PRJParams)]


class Scene(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int), ("domain_start", ctypes.c_float * 3),
                ("domain_size", ctypes.c_float * 3)]


class SceneStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("outer_iters", ctypes.c_int), ("cg_iters", ctypes.c_int),
                ("final_density_error", ctypes.c_double)]


class KernelTiming(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double)]


# every symbol include/ivp.h declares, with its ctypes signature
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
SIGNATURES = {
    "ivp_version": (ctypes.c_char_p, []),
    "ivp_status_string": (ctypes.c_char_p, [_I]),
    "ivp_create": (_I, [ctypes.POINTER(Config), _I, _P, ctypes.POINTER(_P)]),
    "ivp_destroy": (_I, [_P]),
    "ivp_set_particles": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ivp_set_option": (_I, [_P, _I, _I]),
    "ivp_get_option": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "ivp_write": (_I, [_P, _I, _P]),
    "ivp_read": (_I, [_P, _I, _P]),
    "ivp_neighbors_build": (_I, [_P]),
    "ivp_neighbors_export": (_I, [_P, _I, _P, _P, _P]),
    "ivp_ip_solve": (_I, [_P, _F, ctypes.POINTER(IPParams), ctypes.POINTER(IPStats)]),
    "ivp_prj_solve": (_I, [_P, _F, ctypes.POINTER(PRJParams), ctypes.POINTER(IPStats)]),
    "ivp_prj_solve_df": (_I, [_P, _F, ctypes.POINTER(PRJParams), ctypes.POINTER(IPStats)]),
    "ivp_viscosity_explicit": (_I, [_P, _F, _F]),
    "ivp_viscosity_weiler18": (_I, [_P, _F, _F, ctypes.POINTER(CGStats)]),
    "ivp_apply_gravity": (_I, [_P, ctypes.POINTER(_F), _F]),
    "ivp_advect_boundary": (_I, [_P, _F]),
    "ivp_step": (_I, [_P, ctypes.POINTER(StepParams), ctypes.POINTER(IPStats), ctypes.POINTER(CGStats),
                      ctypes.POINTER(CGStats)]),
    "ivp_surface_tension_jeske23": (_I, [_P, _F, _F, ctypes.POINTER(CGStats)]),
    "ivp_profile_enable": (_I, [_P, _I]),
    "ivp_profile_read": (_I, [_P, ctypes.POINTER(KernelTiming), _I]),
    "ivp_profile_reset": (_I, [_P]),
    "ivp_reset": (_I, [_P]),
    "ivp_counts": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong)]),
    "ivp_tile_windows": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong)]),
    "ivp_set_scenes": (_I, [_P, _I, ctypes.POINTER(Scene)]),
    "ivp_step_scenes": (_I, [_P, ctypes.POINTER(StepParams), ctypes.POINTER(SceneStats)]),
    "ivp_max_speed2": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "ivp_rigid_transform": (_I, [_P, _P, _P, _I, _I, ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F),
                                 _P]),
    "ivp_voxelize": (_I, [_P, _I, _P, _I, ctypes.c_double, _I, _P, ctypes.c_longlong,
                          ctypes.POINTER(ctypes.c_longlong), _P]),
    "ivp_voxelize_noalloc": (_I, [_P, _I, _P, _I, ctypes.c_double, _I, _P, ctypes.c_longlong,
                                  ctypes.POINTER(ctypes.c_longlong), _P]),
    "ivp_scene_history": (_I, [_P, _I, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I), ctypes.POINTER(_I),
                               ctypes.POINTER(_I)]),
}

_lib = None


def load():
    """Load libivp.so and bind every entry point.  Raises ImportError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m implicitvelocityprojection_amd.build`")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise ImportError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class IVPError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = load().ivp_status_string(status).decode()
        super().__init__(f"{where}: {msg} (status {status})" if where else f"{msg} (status {status})")


def check(status, where=""):
    if status < 0:
        raise IVPError(status, where)
    return status


def check_ok(status, where=""):
    """Raise IVPError on any status but IVP_OK (entry points without a solver status to report)."""
    if status != IVP_OK:
        raise IVPError(status, where)
    return status
