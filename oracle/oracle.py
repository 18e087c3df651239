"""Python front end of the CPU oracle (test infrastructure only).

The oracle is the checker for the HIP product path: only tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py import it.  It wraps ``oracle/_build/libivp_oracle.so`` (built
from oracle/ivp_oracle.c by oracle/Makefile) with numpy arrays in the reference's layout and
composes the reference's step (scripts/framework.py:423-544, velocity-based branch) from the
restated kernels.
"""

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libivp_oracle.so")
_OMP_PATH = os.path.join(_HERE, "_build", "libivp_oracle_omp.so")
_lib = None
_serial = None

ORC_OK, ORC_NOT_CONVERGED, ORC_NOT_PD = 0, 1, 2
VISC_INPLACE, VISC_JACOBI = 0, 1


class OrcGrid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("inv_w", ctypes.c_float), ("dims", ctypes.c_int * 3)]


class OrcIPParams(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_float), ("tol_opt", ctypes.c_double), ("max_iteration_opt", ctypes.c_int),
                ("use_fixed_iters", ctypes.c_int), ("tol_pcg", ctypes.c_double),
                ("max_iteration_pcg", ctypes.c_int)]


class OrcIPStats(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_int), ("density_errors", ctypes.POINTER(ctypes.c_double)),
                ("pcg_iters", ctypes.POINTER(ctypes.c_int)), ("n_density_errors", ctypes.c_int),
                ("n_pcg", ctypes.c_int), ("outer_iters", ctypes.c_int),
                ("final_density_error", ctypes.c_double), ("final_err_loop", ctypes.c_double),
                ("pAp_bad", ctypes.c_double), ("pcg_iter_bad", ctypes.c_int)]


class OrcPRJParams(ctypes.Structure):
    _fields_ = [("tol_opt", ctypes.c_double), ("max_iteration_opt", ctypes.c_int),
                ("use_fixed_iters", ctypes.c_int), ("omega", ctypes.c_float), ("warm_start", ctypes.c_int),
                ("enable_df", ctypes.c_int), ("df_num_threshold", ctypes.c_int), ("df_tol_opt", ctypes.c_double)]


class OrcPRJStats(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_int), ("density_errors", ctypes.POINTER(ctypes.c_double)),
                ("n_density_errors", ctypes.c_int), ("outer_iters", ctypes.c_int),
                ("final_density_error", ctypes.c_double), ("final_err_loop", ctypes.c_double),
                ("df_outer_iters", ctypes.c_int), ("df_final_err", ctypes.c_double)]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH) or not os.path.exists(_OMP_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s"], cwd=_HERE, check=True)


def _open(path):
    if not os.path.exists(path):
        build()
    h = ctypes.CDLL(path)
    h.orc_cubic_kernel.restype = ctypes.c_float
    h.orc_cubic_kernel.argtypes = [ctypes.c_float, ctypes.c_float]
    return h


def lib():
    global _lib, _serial
    if _lib is None:
        _serial = _lib = _open(_LIB_PATH)
    return _lib


def use_threads(parallel):
    """Select the serial checker (False, the default) or the OpenMP build of the same source
    (True: row loops on every core, atomics in the scatter, per-thread partial sums), which
    only the all-cores CPU baseline uses.  Returns the thread count of the selected build."""
    global _lib, _serial
    lib()
    _lib = _open(_OMP_PATH) if parallel else _serial
    return int(_lib.orc_threads())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def cubic_kernel(r, h):
    return float(lib().orc_cubic_kernel(ctypes.c_float(r), ctypes.c_float(h)))


def spiky_grad(r, h):
    r = _f32(r)
    out = np.zeros(3, np.float32)
    lib().orc_spiky_grad(_p(r), ctypes.c_float(h), _p(out))
    return out


def grid_init(domain_start, domain_size, h):
    g = OrcGrid()
    ds = _f32(domain_start)
    dz = _f32(domain_size)
    lib().orc_grid_init(_p(ds), _p(dz), ctypes.c_float(h), ctypes.byref(g))
    return g


def cell_keys(grid, x):
    x = _f32(x)
    key = np.zeros(x.shape[0], np.int32)
    lib().orc_cell_keys(ctypes.byref(grid), ctypes.c_int(x.shape[0]), _p(x), _p(key))
    return key


def build_neighbors(grid, x, is_dynamic, h, cache_size):
    """Reference-layout neighbour cache: (count [n], ids [n*C], gradW [n*C, 3], max_total)."""
    x = _f32(x)
    dyn = _i32(is_dynamic)
    n = x.shape[0]
    count = np.zeros(n, np.int32)
    ids = np.zeros(n * cache_size, np.int32)
    gradW = np.zeros((n * cache_size, 3), np.float32)
    mx = ctypes.c_int(0)
    rc = lib().orc_build_neighbors(ctypes.byref(grid), ctypes.c_int(n), _p(x), _p(dyn), ctypes.c_float(h),
                                   ctypes.c_int(cache_size), _p(count), _p(ids), _p(gradW), ctypes.byref(mx))
    if rc != 0:
        raise MemoryError("oracle neighbour build failed")
    return count, ids, gradW, int(mx.value)


def brute_neighbors(x, h, cap):
    x = _f32(x)
    n = x.shape[0]
    count = np.zeros(n, np.int32)
    ids = np.zeros(n * cap, np.int32)
    lib().orc_brute_neighbors(ctypes.c_int(n), _p(x), ctypes.c_float(h), ctypes.c_int(cap), _p(count), _p(ids))
    return count, ids


def density(x, m, is_dynamic, h, C, count, ids, out=None):
    x, m, dyn, count, ids = _f32(x), _f32(m), _i32(is_dynamic), _i32(count), _i32(ids)
    out = np.zeros(x.shape[0], np.float32) if out is None else out
    lib().orc_density(ctypes.c_int(x.shape[0]), _p(x), _p(m), _p(dyn), ctypes.c_float(h), ctypes.c_int(C),
                      _p(count), _p(ids), _p(out))
    return out


def aii(m, m_inv, is_dynamic, C, count, ids, gradW):
    m, m_inv, dyn = _f32(m), _f32(m_inv), _i32(is_dynamic)
    count, ids, gradW = _i32(count), _i32(ids), _f32(gradW)
    out = np.zeros(m.shape[0], np.float32)
    lib().orc_aii(ctypes.c_int(m.shape[0]), _p(m), _p(m_inv), _p(dyn), ctypes.c_int(C), _p(count), _p(ids),
                  _p(gradW), _p(out))
    return out


def J(xv, m, is_dynamic, C, count, ids, gradW):
    xv, m, dyn = _f32(xv), _f32(m), _i32(is_dynamic)
    out = np.zeros(m.shape[0], np.float32)
    lib().orc_J(ctypes.c_int(m.shape[0]), _p(xv), _p(m), _p(dyn), ctypes.c_int(C), _p(_i32(count)),
                _p(_i32(ids)), _p(_f32(gradW)), _p(out))
    return out


def JT(y, m, is_dynamic, C, count, ids, gradW):
    y, m, dyn = _f32(y), _f32(m), _i32(is_dynamic)
    out = np.zeros((m.shape[0], 3), np.float32)
    lib().orc_JT(ctypes.c_int(m.shape[0]), _p(y), _p(m), _p(dyn), ctypes.c_int(C), _p(_i32(count)),
                 _p(_i32(ids)), _p(_f32(gradW)), _p(out))
    return out


def ip_apply_A(m, is_dynamic, C, count, ids, gradW, k, dfdt, omega, p):
    m, dyn, k, dfdt, p = _f32(m), _i32(is_dynamic), _f32(k), _f32(dfdt), _f32(p)
    out = np.zeros((m.shape[0], 3), np.float32)
    lib().orc_ip_apply_A(ctypes.c_int(m.shape[0]), _p(m), _p(dyn), ctypes.c_int(C), _p(_i32(count)),
                         _p(_i32(ids)), _p(_f32(gradW)), _p(k), _p(dfdt), ctypes.c_float(omega), _p(p), _p(out))
    return out


class IPResult:
    def __init__(self, status, stats, density_errors, pcg_iters):
        self.status = status
        self.outer_iters = stats.outer_iters
        self.final_density_error = stats.final_density_error
        self.density_errors = density_errors
        self.pcg_iters = pcg_iters
        self.pAp_bad = stats.pAp_bad
        self.pcg_iter_bad = stats.pcg_iter_bad
        self.last_err = stats.final_err_loop


def ip_solve(v, m, dens, density0, is_dynamic, C, count, ids, gradW, k, dt, omega=10.0, tol_opt=1e-3,
             max_iteration_opt=1000, use_fixed_iters=False, tol_pcg=1e-6, max_iteration_pcg=1000):
    """IP.solve (ip.py:331-404) on copies; returns (v_new, divergence, IPResult)."""
    v = _f32(v).copy()
    n = v.shape[0]
    div = np.zeros(n, np.float32)
    prm = OrcIPParams(omega, tol_opt, max_iteration_opt, int(bool(use_fixed_iters)), tol_pcg, max_iteration_pcg)
    cap = max_iteration_opt + 2
    errs = np.zeros(cap, np.float64)
    its = np.zeros(cap, np.int32)
    st = OrcIPStats()
    st.capacity = cap
    st.density_errors = errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    st.pcg_iters = its.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    status = lib().orc_ip_solve(ctypes.c_int(n), _p(v), _p(_f32(m)), _p(_f32(dens)), _p(_f32(density0)),
                                _p(_i32(is_dynamic)), ctypes.c_int(C), _p(_i32(count)), _p(_i32(ids)),
                                _p(_f32(gradW)), _p(_f32(k)), ctypes.c_float(dt), ctypes.byref(prm), _p(div),
                                ctypes.byref(st))
    res = IPResult(status, st, list(errs[:min(st.n_density_errors, cap)]), list(its[:min(st.n_pcg, cap)]))
    return v, div, res


class PRJResult:
    def __init__(self, status, stats, density_errors):
        self.status = status
        self.outer_iters = stats.outer_iters
        self.final_density_error = stats.final_density_error
        self.density_errors = density_errors
        self.last_err = stats.final_err_loop
        self.df_outer_iters = stats.df_outer_iters
        self.df_last_err = stats.df_final_err


def prj_solve(v, m, dens, density0, is_dynamic, C, count, ids, gradW, k, dt, tol_opt=1e-3, max_iteration_opt=1000,
              use_fixed_iters=False, omega=0.5, warm_start=False, enable_df=False, df_num_threshold=25,
              df_tol_opt=1e-2, lam_prev=None, lam_df_prev=None):
    """PRJ.solve (prj.py:292-340, + solve_df when enable_df) on copies.  lam_prev / lam_df_prev
    are the previous solve's multipliers (p_previous, dt2_lam_df_previous), used with warm_start.
    Returns (v_new, divergence, PRJResult, lam_prev_new, lam_df_prev_new)."""
    v = _f32(v).copy()
    n = v.shape[0]
    div = np.zeros(n, np.float32)
    lp = np.zeros(n, np.float32) if lam_prev is None else _f32(lam_prev).copy()
    ldp = np.zeros(n, np.float32) if lam_df_prev is None else _f32(lam_df_prev).copy()
    prm = OrcPRJParams(tol_opt, max_iteration_opt, int(bool(use_fixed_iters)), omega, int(bool(warm_start)),
                       int(bool(enable_df)), df_num_threshold, df_tol_opt)
    cap = max_iteration_opt + 2
    errs = np.zeros(cap, np.float64)
    st = OrcPRJStats()
    st.capacity = cap
    st.density_errors = errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    status = lib().orc_prj_solve(ctypes.c_int(n), _p(v), _p(_f32(m)), _p(_f32(dens)), _p(_f32(density0)),
                                 _p(_i32(is_dynamic)), ctypes.c_int(C), _p(_i32(count)), _p(_i32(ids)),
                                 _p(_f32(gradW)), _p(_f32(k)), ctypes.c_float(dt), ctypes.byref(prm), _p(lp),
                                 _p(ldp), _p(div), ctypes.byref(st))
    res = PRJResult(status, st, list(errs[:min(st.n_density_errors, cap)]))
    return v, div, res, lp, ldp


def viscosity(x, v, m, dens, density0, material, is_dynamic, h, C, count, ids, gradW, coeff, dt,
              mode=VISC_JACOBI, material_fluid=1):
    v = _f32(v).copy()
    lib().orc_viscosity(ctypes.c_int(v.shape[0]), _p(_f32(x)), _p(v), _p(_f32(m)), _p(_f32(dens)),
                        _p(_f32(density0)), _p(_i32(material)), _p(_i32(is_dynamic)), ctypes.c_float(h),
                        ctypes.c_int(C), _p(_i32(count)), _p(_i32(ids)), _p(_f32(gradW)), ctypes.c_float(coeff),
                        ctypes.c_float(dt), ctypes.c_int(material_fluid), ctypes.c_int(mode))
    return v


def inverse33(a):
    a = _f32(a).reshape(9)
    out = np.zeros(9, np.float32)
    lib().orc_inverse33(_p(a), _p(out))
    return out.reshape(3, 3)


def w18_apply(x, m, dens, density0, material, is_dynamic, h, C, count, ids, gradW, coeff, dt, vin=None,
              material_fluid=1):
    """(A v_in, block-Jacobi inverses) of the Weiler18 operator (weiler18.py:21-172)."""
    n = _f32(x).shape[0]
    Av = np.zeros((n, 3), np.float32)
    P = np.zeros((n, 9), np.float32)
    vin_a = _f32(vin) if vin is not None else None
    lib().orc_w18_apply_A(ctypes.c_int(n), _p(_f32(x)), _p(_f32(m)), _p(_f32(dens)), _p(_f32(density0)),
                          _p(_i32(material)), _p(_i32(is_dynamic)), ctypes.c_float(h), ctypes.c_int(C),
                          _p(_i32(count)), _p(_i32(ids)), _p(_f32(gradW)), ctypes.c_float(coeff),
                          ctypes.c_float(dt), ctypes.c_int(material_fluid),
                          _p(vin_a) if vin_a is not None else None, _p(Av) if vin_a is not None else None,
                          _p(P))
    return Av, P.reshape(n, 3, 3)


def weiler18(x, v, m, dens, density0, material, is_dynamic, h, C, count, ids, gradW, coeff, dt,
             material_fluid=1, warm_start=False, use_preconditioner=True, dv=None):
    """Weiler18.solve (weiler18.py:371-401).  dv is the solver's correction buffer (self.dv): the
    initial guess with warm_start (zeros when None), replaced by the new correction; returns
    (v, CG iterations) and, when dv is given, fills it in place."""
    v = _f32(v).copy()
    n = v.shape[0]
    buf = np.zeros((n, 3), np.float32) if dv is None else dv
    assert buf.dtype == np.float32 and buf.flags.c_contiguous and buf.shape == (n, 3)
    it = lib().orc_weiler18_ex(ctypes.c_int(n), _p(_f32(x)), _p(v), _p(_f32(m)), _p(_f32(dens)),
                               _p(_f32(density0)), _p(_i32(material)), _p(_i32(is_dynamic)), ctypes.c_float(h),
                               ctypes.c_int(C), _p(_i32(count)), _p(_i32(ids)), _p(_f32(gradW)),
                               ctypes.c_float(coeff), ctypes.c_float(dt), ctypes.c_int(material_fluid),
                               ctypes.c_int(int(bool(warm_start))), ctypes.c_int(int(bool(use_preconditioner))),
                               _p(buf))
    return v, int(it)


def wst(r, d, h_st):
    lib().orc_wst.restype = ctypes.c_float
    return float(lib().orc_wst(ctypes.c_float(r), ctypes.c_float(d), ctypes.c_float(h_st)))


def grad_wst(x_ij, r, d, h_st):
    out = np.zeros(3, np.float32)
    lib().orc_grad_wst(_p(_f32(x_ij)), ctypes.c_float(r), ctypes.c_float(d), ctypes.c_float(h_st), _p(out))
    return out


def jeske23_setup(x, v, m, dens, material, is_dynamic, C, count, ids, gradW, sigma, dt, d, h_st,
                  material_fluid=1):
    """compute_divergence / compute_g_tilde / compute_rhs (jeske23.py:73-242): (div, g~[n*C], b)."""
    n = x.shape[0]
    div = np.zeros(n, np.float32)
    gt = np.zeros(n * C, np.float32)
    b = np.zeros((n, 3), np.float32)
    lib().orc_jeske23_setup(ctypes.c_int(n), _p(_f32(x)), _p(_f32(v)), _p(_f32(m)), _p(_f32(dens)),
                            _p(_i32(material)), _p(_i32(is_dynamic)), ctypes.c_int(C), _p(_i32(count)),
                            _p(_i32(ids)), _p(_f32(gradW)), ctypes.c_float(sigma), ctypes.c_float(dt),
                            ctypes.c_float(d), ctypes.c_float(h_st), ctypes.c_int(material_fluid),
                            _p(div), _p(gt), _p(b))
    return div, gt, b


def jeske23_apply_M(m, material, is_dynamic, C, count, ids, g_tilde, sigma, dt, vin, material_fluid=1):
    """(M v, block-diagonal preconditioner inverse) of jeske23.py:245-350."""
    n = vin.shape[0]
    Mv = np.zeros((n, 3), np.float32)
    P = np.zeros((n, 9), np.float32)
    lib().orc_jeske23_apply_M(ctypes.c_int(n), _p(_f32(m)), _p(_i32(material)), _p(_i32(is_dynamic)),
                              ctypes.c_int(C), _p(_i32(count)), _p(_i32(ids)), _p(_f32(g_tilde)),
                              ctypes.c_float(sigma), ctypes.c_float(dt), ctypes.c_int(material_fluid),
                              _p(_f32(vin)), _p(Mv), _p(P))
    return Mv, P


def jeske23(x, v, m, dens, material, is_dynamic, C, count, ids, gradW, sigma, dt, d, h_st, material_fluid=1):
    """Jeske23.solve (jeske23.py:590-655); returns (v, CG iterations)."""
    v = _f32(v).copy()
    it = lib().orc_jeske23(ctypes.c_int(v.shape[0]), _p(_f32(x)), _p(v), _p(_f32(m)), _p(_f32(dens)),
                           _p(_i32(material)), _p(_i32(is_dynamic)), ctypes.c_int(C), _p(_i32(count)),
                           _p(_i32(ids)), _p(_f32(gradW)), ctypes.c_float(sigma), ctypes.c_float(dt),
                           ctypes.c_float(d), ctypes.c_float(h_st), ctypes.c_int(material_fluid))
    return v, int(it)


def gravity(v, is_dynamic, g, dt):
    v = _f32(v).copy()
    lib().orc_gravity(ctypes.c_int(v.shape[0]), _p(v), _p(_i32(is_dynamic)), _p(_f32(g)), ctypes.c_float(dt))
    return v


def advect(x, v, is_dynamic, dt):
    x, v = _f32(x).copy(), _f32(v).copy()
    lib().orc_advect(ctypes.c_int(x.shape[0]), _p(x), None, _p(v), _p(_i32(is_dynamic)), ctypes.c_float(dt))
    return x, v


def enforce_boundary(x, v, is_dynamic, domain_size, padding):
    x, v = _f32(x).copy(), _f32(v).copy()
    lib().orc_enforce_boundary(ctypes.c_int(x.shape[0]), _p(x), _p(v), _p(_i32(is_dynamic)),
                               _p(_f32(domain_size)), ctypes.c_float(padding))
    return x, v


# ----------------------------------------------------------------------------------------------
# Whole-step composition (Framework, scripts/framework.py)
# ----------------------------------------------------------------------------------------------

class OracleSim:
    """The reference's velocity-based IP step on numpy state, built from the restated kernels.

    State arrays are in the reference's particle order.  ``visc_mode`` selects how the in-place
    explicit viscosity kernel (viscosity.py:9-108) is executed: VISC_JACOBI (every row reads the
    pre-step velocities) or VISC_INPLACE (serial index order)."""

    class Diverged(RuntimeError):
        pass

    def __init__(self, config, particles, visc_mode=VISC_JACOBI):
        from implicitvelocityprojection_amd.scenes import particle_masses
        self.cfg = config
        self.h = float(np.float32(config.support_radius))
        self.C = int(config.get_cfg("neighborCacheSize", 60))
        self.x = particles.x.copy()
        self.v = particles.v.copy()
        self.density0 = particles.density0.copy()
        self.density = particles.density0.copy()
        self.is_dynamic = particles.is_dynamic.copy()
        self.material = particles.material.copy()
        self.m, self.m_inv = particle_masses(config, particles)
        self.k = np.zeros(particles.n, np.float32)
        self.divergence = np.zeros(particles.n, np.float32)
        self.grid = grid_init(config.domain_start, config.domain_size, self.h)
        self.dt = float(config.dt)
        self.g = np.array(config.gravity, np.float32)
        self.visc_mode = visc_mode
        self.ip_kw = dict(
            omega=float(config.get_cfg("omega", 10.0)),
            tol_opt=float(config.get_cfg("densityErrorTol", 1e-3)),
            max_iteration_opt=int(config.get_cfg("maxOuterIters", 1000)),
            use_fixed_iters=bool(config.get_cfg("fixedOuterIters", False)),
            tol_pcg=float(config.get_cfg("tol_pcg", 1e-6)),
        )
        self.pressure_method = int(config.get_cfg("pressureMethod", 1))
        self.prj_kw = dict(
            tol_opt=self.ip_kw["tol_opt"], max_iteration_opt=self.ip_kw["max_iteration_opt"],
            use_fixed_iters=self.ip_kw["use_fixed_iters"],
        )
        self.lam_prev = np.zeros(particles.n, np.float32)
        self.lam_df_prev = np.zeros(particles.n, np.float32)
        self.time = 0.0
        self.spawns = []
        self.particle_max_num = particles.n
        self.object_id = particles.object_id.copy()
        self.x0 = particles.x.copy()
        # CFL time step (framework.py:150-152, 427-440)
        self.dt_default = self.dt
        self.use_cfl = bool(config.get_cfg("useCFL", False))
        self.cfl_factor = float(config.get_cfg("cflFactor", 0.4))
        self.cfl_min_dt = float(config.get_cfg("cflMinDt", 1e-4))
        # keyframe tracks: (object id, t -> (R (3, 3), T (3,), pivot (3,)) float32)
        self.tracks = []
        self.last_ip = None
        self.last_prj = None
        self.last_visc_iters = 0
        self.last_st_iters = 0
        self.max_total = 0

    def rebuild(self):
        """Framework._rebuild_neighbors_and_density (framework.py:282-290)."""
        self.count, self.ids, self.gradW, self.max_total = build_neighbors(
            self.grid, self.x, self.is_dynamic, self.h, self.C)
        density(self.x, self.m, self.is_dynamic, self.h, self.C, self.count, self.ids, out=self.density)
        self.k = aii(self.m, self.m_inv, self.is_dynamic, self.C, self.count, self.ids, self.gradW)

    def non_pressure(self):
        """compute_non_pressure_forces (framework.py:400-421): gravity, surface tension, viscosity.
        Surface tension at coefficient 0 is an exact no-op and is skipped; the explicit (Akinci)
        model of surfaceTensionMethod 0 is not restated."""
        self.v = gravity(self.v, self.is_dynamic, self.g, self.dt)
        st_coeff = float(self.cfg.surface_tension_coeff)
        if st_coeff != 0.0:
            if int(self.cfg.surface_tension_method) != 1:
                raise NotImplementedError("explicit surface tension (surfaceTensionMethod 0) is not restated")
            d = 2.0 * float(self.cfg.particle_radius)
            self.v, self.last_st_iters = jeske23(self.x, self.v, self.m, self.density, self.material,
                                                 self.is_dynamic, self.C, self.count, self.ids, self.gradW,
                                                 st_coeff, self.dt, d, 2.0 * d)
        method = int(self.cfg.viscosity_method)
        coeff = float(self.cfg.viscosity_coeff)
        if method == 1:
            self.v, self.last_visc_iters = weiler18(self.x, self.v, self.m, self.density, self.density0,
                                                    self.material, self.is_dynamic, self.h, self.C, self.count,
                                                    self.ids, self.gradW, coeff, self.dt)
        elif coeff != 0.0:
            self.v = viscosity(self.x, self.v, self.m, self.density, self.density0, self.material,
                               self.is_dynamic, self.h, self.C, self.count, self.ids, self.gradW, coeff,
                               self.dt, mode=self.visc_mode)

    def pressure(self):
        if self.pressure_method == 0:
            return self.pressure_prj()
        if self.pressure_method != 1:
            raise NotImplementedError(f"pressureMethod {self.pressure_method} is not restated")
        v, div, res = ip_solve(self.v, self.m, self.density, self.density0, self.is_dynamic, self.C,
                               self.count, self.ids, self.gradW, self.k, self.dt, **self.ip_kw)
        self.last_ip = res
        if res.status == ORC_NOT_PD:
            raise self.Diverged(f"[IP] non-positive definite matrix at PCG iter {res.pcg_iter_bad} "
                                f"(pAp={res.pAp_bad:.2e})")
        if res.status == ORC_NOT_CONVERGED:
            raise self.Diverged(f"[IP] does not converge! density err={res.last_err:.2e} "
                                f"(tol={self.ip_kw['tol_opt']:.2e}) after {res.outer_iters} iters")
        self.v, self.divergence = v, div

    def pressure_prj(self):
        """PRJ.solve (prj.py:292-340) and, with prj_kw enable_df, solve_df (framework.py:483-487)."""
        v, div, res, self.lam_prev, self.lam_df_prev = prj_solve(
            self.v, self.m, self.density, self.density0, self.is_dynamic, self.C, self.count, self.ids,
            self.gradW, self.k, self.dt, lam_prev=self.lam_prev, lam_df_prev=self.lam_df_prev, **self.prj_kw)
        self.last_prj = res
        if res.status == ORC_NOT_CONVERGED:
            raise self.Diverged(f"[PRJ] does not converge! density err={res.last_err:.2e} "
                                f"(tol={self.prj_kw['tol_opt']:.2e}) after {res.outer_iters} iters")
        self.v, self.divergence = v, div

    def add_spawns(self, entries, particle_max_num):
        """Scheduled bodies (scenes.SpawnEntry list) and the particle capacity
        (scheduled_spawn.py:68-117)."""
        self.spawns = [[e, False] for e in sorted(entries, key=lambda e: e.time)]
        self.particle_max_num = int(particle_max_num)

    def _fire_spawns(self):
        """ScheduledSpawnSystem.step(time): append every due body (cut to the capacity left),
        masses by add_particle's convention (simulation_data.py:40-57)."""
        for rec in self.spawns:
            e, fired = rec
            if fired or self.time < e.time:
                continue
            rec[1] = True
            remain = self.particle_max_num - self.x.shape[0]
            if remain <= 0:
                continue
            n = min(e.positions.shape[0], remain)
            rho = np.full(n, e.density, np.float32)
            m = (np.float32(0.8 * self.cfg.particle_diameter ** 3) * rho).astype(np.float32)
            m_inv = np.zeros(n, np.float32)
            if e.is_dynamic == 1:
                m_inv[:] = np.float32(1.0) / m
            else:
                m = m * np.float32(10.0)
            cat = lambda a, b: np.concatenate([a, b])
            self.x = cat(self.x, e.positions[:n].astype(np.float32))
            self.v = cat(self.v, np.tile(e.velocity[None, :].astype(np.float32), (n, 1)))
            self.m, self.m_inv = cat(self.m, m), cat(self.m_inv, m_inv)
            self.density0 = cat(self.density0, rho)
            self.density = cat(self.density, rho)
            self.is_dynamic = cat(self.is_dynamic, np.full(n, e.is_dynamic, np.int32))
            self.material = cat(self.material, np.full(n, e.material, np.int32))
            self.k = cat(self.k, np.zeros(n, np.float32))
            self.divergence = cat(self.divergence, np.zeros(n, np.float32))
            self.lam_prev = cat(self.lam_prev, np.zeros(n, np.float32))
            self.lam_df_prev = cat(self.lam_df_prev, np.zeros(n, np.float32))
            self.object_id = cat(self.object_id, np.full(n, e.object_id, np.int32))
            self.x0 = cat(self.x0, e.positions[:n].astype(np.float32))

    def cfl_dt(self):
        """framework.py:427-440: dt = min(dt_default, cflFactor h / v_max), floored at cflMinDt;
        v_max^2 the float32 max of (vx vx + vy vy) + vz vz over dynamic particles."""
        dyn = self.is_dynamic == 1
        v = self.v[dyn]
        vsq = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
        v_max_sq = float(vsq.max()) if vsq.size else 0.0
        if math.isfinite(v_max_sq) and v_max_sq > 1e-16:
            dt = min(self.dt_default, self.cfl_factor * self.cfg.support_radius / v_max_sq ** 0.5)
        else:
            dt = self.dt_default
        return max(dt, self.cfl_min_dt)

    def animate(self):
        """apply_transform_kernel (animation.py:7-34) for every track at the current time:
        x = (pivot + T) + R (x0 - pivot), float32, products summed left to right."""
        for oid, pose in self.tracks:
            sel = self.object_id == oid
            if not sel.any():
                continue
            R, T, pivot = pose(self.time)
            R, T, pivot = np.float32(R), np.float32(T), np.float32(pivot)
            d = self.x0[sel] - pivot
            r = np.stack([(R[i, 0] * d[:, 0] + R[i, 1] * d[:, 1]) + R[i, 2] * d[:, 2] for i in range(3)], axis=1)
            self.x[sel] = (pivot + T) + r

    def step(self):
        """Framework.forward (framework.py:423-544), velocity-based branch: CFL time step, keyframe
        poses, forces, pressure, advection, boundary, spawns, rebuild."""
        if self.use_cfl:
            self.dt = self.cfl_dt()
        if self.tracks:
            self.animate()
        self.non_pressure()
        self.pressure()
        self.x, self.v = advect(self.x, self.v, self.is_dynamic, self.dt)
        self.x, self.v = enforce_boundary(self.x, self.v, self.is_dynamic,
                                          np.array(self.cfg.domain_size, np.float32), self.h)
        if self.spawns:
            self._fire_spawns()
        self.rebuild()
        self.time += self.dt
