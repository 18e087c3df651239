"""Thin object wrapper over the libivp C ABI, taking torch device tensors as buffers.

torch is used only for device memory and the stream: every array handed to the library is a
contiguous CUDA (ROCm) tensor whose ``data_ptr()`` goes through ctypes unchanged.
"""

import ctypes

import numpy as np
import torch

from . import _native as N


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("libivp buffers must be device tensors")
    if not t.is_contiguous():
        raise ValueError("libivp buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class IPParams:
    """IP solver knobs (pressure_base.py:314-317, ip.py:228-229)."""

    def __init__(self, omega=10.0, tol_opt=1e-3, max_iteration_opt=1000, use_fixed_iters=False, tol_pcg=1e-6,
                 max_iteration_pcg=1000):
        self.omega = float(omega)
        self.tol_opt = float(tol_opt)
        self.max_iteration_opt = int(max_iteration_opt)
        self.use_fixed_iters = bool(use_fixed_iters)
        self.tol_pcg = float(tol_pcg)
        self.max_iteration_pcg = int(max_iteration_pcg)

    def c(self):
        return N.IPParams(self.omega, self.tol_opt, self.max_iteration_opt, int(self.use_fixed_iters), self.tol_pcg,
                          self.max_iteration_pcg)


class PRJParams:
    """PRJ solver knobs (pressure_base.py:314-317, prj.py:150-161)."""

    def __init__(self, tol_opt=1e-3, max_iteration_opt=1000, use_fixed_iters=False, omega=0.5, warm_start=False,
                 enable_df=False, df_num_threshold=25, df_tol_opt=1e-2):
        self.tol_opt = float(tol_opt)
        self.max_iteration_opt = int(max_iteration_opt)
        self.use_fixed_iters = bool(use_fixed_iters)
        self.omega = float(omega)
        self.warm_start = bool(warm_start)
        self.enable_df = bool(enable_df)
        self.df_num_threshold = int(df_num_threshold)
        self.df_tol_opt = float(df_tol_opt)

    def c(self):
        return N.PRJParams(self.tol_opt, self.max_iteration_opt, int(self.use_fixed_iters), self.omega,
                           int(self.warm_start), int(self.enable_df), self.df_num_threshold, self.df_tol_opt)


class IPResult:
    """Statistics of one pressure solve (IP, or PRJ: no CG iterations, df_ fields filled)."""

    def __init__(self, status, st, errs, its):
        self.status = status
        self.density_errors = [float(e) for e in errs[:min(st.n_density_errors, len(errs))]]
        self.pcg_iters = [int(i) for i in its[:min(st.n_pcg, len(its))]]
        self.outer_iters = st.outer_iters
        self.final_density_error = st.final_density_error
        self.pAp_bad = st.pAp_bad
        self.pcg_iter_bad = st.pcg_iter_bad
        self.pairs_traversed = st.pairs_traversed
        self.rows_traversed = st.rows_traversed
        self.cg_launch_chunks = st.cg_launch_chunks
        self.solve_ms = st.solve_ms
        self.df_outer_iters = st.df_outer_iters
        self.df_ms = st.df_ms
        costs = getattr(st, "_costs", None)
        self.iter_costs = [] if costs is None else [float(c) for c in costs[:min(st.n_iter_costs, len(costs))]]

    @property
    def last_err(self):
        return self.density_errors[-1] if self.density_errors else 0.0


class Context:
    """One libivp context: a particle store plus neighbour rows and solver buffers."""

    def __init__(self, support_radius, domain_start, domain_size, max_particles, cache_size=60, material_fluid=1,
                 stream=None):
        self.lib = N.load()
        cfg = N.Config()
        cfg.support_radius = float(np.float32(support_radius))
        cfg.domain_start[:] = [float(v) for v in np.asarray(domain_start, np.float32)]
        cfg.domain_size[:] = [float(v) for v in np.asarray(domain_size, np.float32)]
        cfg.cache_size = int(cache_size)
        cfg.material_fluid = int(material_fluid)
        self.cache_size = int(cache_size)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.check(self.lib.ivp_create(ctypes.byref(cfg), int(max_particles), ctypes.c_void_p(self.stream.cuda_stream),
                                    ctypes.byref(h)), "ivp_create")
        self.h = h
        self.max_particles = int(max_particles)
        self.loaded = False          # a particle set has been loaded (ivp_set_particles)

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.ivp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counts(self):
        n, nd = ctypes.c_int(), ctypes.c_int()
        N.check(self.lib.ivp_counts(self.h, ctypes.byref(n), ctypes.byref(nd), None))
        return n.value, nd.value

    def pair_count(self):
        """Dynamic-dynamic neighbour pairs stored by the last build."""
        p = ctypes.c_longlong()
        N.check(self.lib.ivp_counts(self.h, None, None, ctypes.byref(p)))
        return p.value

    def window_stats(self):
        """(tiles, tiles with a window, neighbour ids staged per pass) of the last build."""
        t, w, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        N.check(self.lib.ivp_tile_windows(self.h, ctypes.byref(t), ctypes.byref(w), ctypes.byref(k)),
                "ivp_tile_windows")
        return t.value, w.value, k.value

    @property
    def gradient_cache(self):
        """Neighbour-row format: ROWS_IDS (0, gradW evaluated per pass), ROWS_F32 (1, float32
        gradW cached per pair) or ROWS_F16 (2, float16 gradW cached for the IP CG operator)."""
        v = ctypes.c_int()
        N.check(self.lib.ivp_get_option(self.h, N.OPT_GRADIENT_CACHE, ctypes.byref(v)))
        return v.value

    @gradient_cache.setter
    def gradient_cache(self, fmt):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_GRADIENT_CACHE, int(fmt)), "ivp_set_option")

    # -- particle store ----------------------------------------------------------------------
    def set_particles(self, x, v, m, m_inv, density0, is_dynamic, material):
        n = int(x.shape[0])
        N.check(self.lib.ivp_set_particles(self.h, n, _ptr(x), _ptr(v), _ptr(m), _ptr(m_inv), _ptr(density0),
                                           _ptr(is_dynamic), _ptr(material)), "ivp_set_particles")
        self.loaded = True

    def reset(self):
        """Density back to density0; k, divergence and PRJ's warm-start multipliers zeroed;
        neighbour rows dropped."""
        N.check(self.lib.ivp_reset(self.h), "ivp_reset")

    def write(self, field, t):
        N.check(self.lib.ivp_write(self.h, field, _ptr(t)), "ivp_write")

    def read(self, field, out):
        N.check(self.lib.ivp_read(self.h, field, _ptr(out)), "ivp_read")
        return out

    # -- neighbours ---------------------------------------------------------------------------
    def neighbors_build(self):
        """Returns the status (IVP_OK or IVP_NEIGHBOR_OVERFLOW)."""
        return N.check(self.lib.ivp_neighbors_build(self.h), "ivp_neighbors_build")

    def neighbors_export(self, count, ids, gradW, cache_size=None):
        C = self.cache_size if cache_size is None else int(cache_size)
        N.check(self.lib.ivp_neighbors_export(self.h, C, _ptr(count), _ptr(ids), _ptr(gradW)), "ivp_neighbors_export")

    # -- solvers -------------------------------------------------------------------------------
    def _stats(self, cap):
        errs = np.zeros(cap, np.float64)
        its = np.zeros(cap, np.int32)
        costs = np.zeros(cap, np.float64)
        st = N.IPStats()
        st.capacity = cap
        st.density_errors = errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        st.pcg_iters = its.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        st.iter_costs = costs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        st._costs = costs          # keeps the array alive with the struct
        return st, errs, its

    def ip_solve(self, dt, params: IPParams):
        st, errs, its = self._stats(params.max_iteration_opt + 2)
        prm = params.c()
        status = N.check(self.lib.ivp_ip_solve(self.h, ctypes.c_float(dt), ctypes.byref(prm), ctypes.byref(st)),
                         "ivp_ip_solve")
        return IPResult(status, st, errs, its)

    def prj_solve(self, dt, params: PRJParams):
        st, errs, its = self._stats(params.max_iteration_opt + 2)
        prm = params.c()
        status = N.check(self.lib.ivp_prj_solve(self.h, ctypes.c_float(dt), ctypes.byref(prm), ctypes.byref(st)),
                         "ivp_prj_solve")
        return IPResult(status, st, errs, its)

    def prj_solve_df(self, dt, params: PRJParams):
        st, errs, its = self._stats(2)
        prm = params.c()
        status = N.check(self.lib.ivp_prj_solve_df(self.h, ctypes.c_float(dt), ctypes.byref(prm),
                                                   ctypes.byref(st)), "ivp_prj_solve_df")
        return IPResult(status, st, errs, its)

    def viscosity_explicit(self, coeff, dt):
        N.check_ok(self.lib.ivp_viscosity_explicit(self.h, ctypes.c_float(coeff), ctypes.c_float(dt)),
                   "ivp_viscosity_explicit")

    def viscosity_weiler18(self, coeff, dt):
        """Weiler18.solve; returns the CG iteration count (raises IVPError on any other status,
        e.g. rows that the last rebuild truncated)."""
        cs = N.CGStats()
        N.check_ok(self.lib.ivp_viscosity_weiler18(self.h, ctypes.c_float(coeff), ctypes.c_float(dt),
                                                   ctypes.byref(cs)), "ivp_viscosity_weiler18")
        return cs.cg_iters

    def max_speed2(self):
        """max over dynamic particles of |v|^2 (float32 sum of squares) of the stored velocities."""
        out = ctypes.c_double()
        N.check(self.lib.ivp_max_speed2(self.h, ctypes.byref(out)), "ivp_max_speed2")
        return out.value

    def apply_gravity(self, g, dt):
        ga = (ctypes.c_float * 3)(*[float(v) for v in g])
        N.check(self.lib.ivp_apply_gravity(self.h, ga, ctypes.c_float(dt)), "ivp_apply_gravity")

    def advect_boundary(self, dt):
        N.check(self.lib.ivp_advect_boundary(self.h, ctypes.c_float(dt)), "ivp_advect_boundary")

    def step(self, dt, gravity, viscosity_method, viscosity_coeff, params, surface_tension_method=0,
             surface_tension_coeff=0.0):
        """One step with the pressure solver of ``params`` (IPParams: IP, PRJParams: PRJ);
        returns (pressure result, viscosity CG iterations).  The surface-tension CG iterations of
        the step are left in ``last_st_iters``."""
        sp = N.StepParams()
        sp.dt = float(dt)
        sp.gravity[:] = [float(v) for v in gravity]
        sp.viscosity_method = int(viscosity_method)
        sp.viscosity_coeff = float(viscosity_coeff)
        if isinstance(params, PRJParams):
            sp.pressure_method = 0
            sp.prj = params.c()
        else:
            sp.pressure_method = 1
            sp.ip = params.c()
        sp.surface_tension_method = int(surface_tension_method)
        sp.surface_tension_coeff = float(surface_tension_coeff)
        st, errs, its = self._stats(params.max_iteration_opt + 2)
        cs = N.CGStats()
        cst = N.CGStats()
        status = N.check(self.lib.ivp_step(self.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(cs),
                                           ctypes.byref(cst)), "ivp_step")
        self.last_st_iters = cst.cg_iters
        return IPResult(status, st, errs, its), cs.cg_iters

    def surface_tension_jeske23(self, sigma, dt):
        """Jeske23.solve; returns the CG iteration count."""
        cs = N.CGStats()
        N.check_ok(self.lib.ivp_surface_tension_jeske23(self.h, ctypes.c_float(sigma), ctypes.c_float(dt),
                                                        ctypes.byref(cs)), "ivp_surface_tension_jeske23")
        return cs.cg_iters

    # -- independent scenes (one context, many scenes) ----------------------------------------
    def set_scenes(self, sizes, domain_starts, domain_sizes):
        """Declare the next set_particles call's particles as consecutive scenes: sizes[s]
        particles with domain (domain_starts[s], domain_sizes[s]) each (see include/ivp.h)."""
        S = len(sizes)
        arr = (N.Scene * S)()
        for s in range(S):
            arr[s].n_particles = int(sizes[s])
            arr[s].domain_start[:] = [float(v) for v in np.asarray(domain_starts[s], np.float32)]
            arr[s].domain_size[:] = [float(v) for v in np.asarray(domain_sizes[s], np.float32)]
        N.check(self.lib.ivp_set_scenes(self.h, S, arr), "ivp_set_scenes")
        self.n_scenes = S

    def step_scenes(self, dt, gravity, params: IPParams):
        """One step of every scene (viscosity off).  Returns (status, list of per-scene
        (status, outer_iters, cg_iters, final_density_error))."""
        sp = N.StepParams()
        sp.dt = float(dt)
        sp.gravity[:] = [float(v) for v in gravity]
        sp.viscosity_method = 0
        sp.viscosity_coeff = 0.0
        sp.pressure_method = 1
        sp.ip = params.c()
        out = (N.SceneStats * self.n_scenes)()
        status = N.check(self.lib.ivp_step_scenes(self.h, ctypes.byref(sp), out), "ivp_step_scenes")
        return status, [(o.status, o.outer_iters, o.cg_iters, o.final_density_error) for o in out]

    def scene_history(self, scene, capacity=1026):
        """(density errors, CG iterations per update) of one scene in the last step_scenes."""
        errs = np.zeros(capacity, np.float64)
        its = np.zeros(capacity, np.int32)
        ne, npcg = ctypes.c_int(), ctypes.c_int()
        N.check(self.lib.ivp_scene_history(self.h, int(scene), capacity,
                                           errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           its.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(ne),
                                           ctypes.byref(npcg)), "ivp_scene_history")
        return [float(e) for e in errs[:min(ne.value, capacity)]], [int(i) for i in its[:min(npcg.value, capacity)]]

    def _opt(self, opt):
        v = ctypes.c_int()
        N.check(self.lib.ivp_get_option(self.h, opt, ctypes.byref(v)))
        return v.value

    @property
    def build_lanes(self):
        """The neighbour build's candidate-scan form: 0 by the dynamic count, 1, 2, 4, 8 or 16 lanes
        per row, or 64 for the scan by cells (the rows are identical in every form)."""
        return self._opt(N.OPT_BUILD_LANES)

    @build_lanes.setter
    def build_lanes(self, k):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_BUILD_LANES, int(k)), "ivp_set_option")

    @property
    def scene_cg_chunk(self):
        """CG iterations per replay of the scene solve's launch sequence."""
        return self._opt(N.OPT_SCENE_CG_CHUNK)

    @scene_cg_chunk.setter
    def scene_cg_chunk(self, k):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_SCENE_CG_CHUNK, int(k)), "ivp_set_option")

    @property
    def scene_lanes(self):
        """Lanes per neighbour row in the scene solve's pair passes (1, 2 or 4)."""
        return self._opt(N.OPT_SCENE_LANES)

    @scene_lanes.setter
    def scene_lanes(self, k):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_SCENE_LANES, int(k)), "ivp_set_option")

    @property
    def scene_graph(self):
        """True: the scene solve replays a captured hipGraph; False: it enqueues the launches."""
        return bool(self._opt(N.OPT_SCENE_GRAPH))

    @scene_graph.setter
    def scene_graph(self, on):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_SCENE_GRAPH, int(bool(on))), "ivp_set_option")

    @property
    def w18_warm_start(self):
        """Weiler18.warm_start: start the viscosity PCG from the previous solve's correction."""
        return bool(self._opt(N.OPT_W18_WARM_START))

    @w18_warm_start.setter
    def w18_warm_start(self, on):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_W18_WARM_START, int(bool(on))), "ivp_set_option")

    @property
    def w18_preconditioner(self):
        """Weiler18.use_preconditioner: block-Jacobi PCG (True) or plain CG (False)."""
        return bool(self._opt(N.OPT_W18_PRECONDITIONER))

    @w18_preconditioner.setter
    def w18_preconditioner(self, on):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_W18_PRECONDITIONER, int(bool(on))), "ivp_set_option")

    @property
    def tile_windows(self):
        """Pair passes load each 64-row tile's distinct neighbours into on-chip memory once
        (True, default) instead of gathering every pair's neighbour; a change applies from the
        next neighbour build on.  Results are bitwise the same."""
        return bool(self._opt(N.OPT_TILE_WINDOWS))

    @tile_windows.setter
    def tile_windows(self, on):
        N.check(self.lib.ivp_set_option(self.h, N.OPT_TILE_WINDOWS, int(bool(on))), "ivp_set_option")

    # -- instrumentation ----------------------------------------------------------------------
    def profile_enable(self, mask=-1):
        N.check(self.lib.ivp_profile_enable(self.h, int(mask)))

    def profile_reset(self):
        N.check(self.lib.ivp_profile_reset(self.h))

    def profile_read(self):
        arr = (N.KernelTiming * len(N.KCLASSES))()
        n = N.check(self.lib.ivp_profile_read(self.h, arr, len(N.KCLASSES)))
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms, bytes=arr[i].bytes)
                for i in range(n)}


def kclass_mask(*names):
    m = 0
    for nm in names:
        m |= 1 << N.KCLASSES.index(nm)
    return m
