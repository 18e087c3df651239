"""Pressure solvers, mirroring the reference's ``PressureBase``, ``IP`` and ``PRJ``
(scripts/sim/forces/pressure/pressure_base.py:280-506, ip.py:222-404, prj.py:143-340).

Same constructor ``(particle_system, neighbour_search, config)``, the same mutable solver knobs
(``omega``, ``tol_opt``, ``max_iteration_opt``, ``use_fixed_iters``, ``tol_pcg``,
``max_iteration_pcg``; PRJ: ``omega`` = 0.5, ``warm_start``, ``enable_df``,
``df_num_threshold``, ``df_tol_opt``), the same ``stats_*`` attributes and the same exceptions;
``solve(dt)`` runs the whole outer loop in libivp.
"""

import torch

from . import _native as N
from .context import IPParams, PRJParams


class PressureBase:
    is_position_based = False

    def __init__(self, particle_system, neighbour_search, config):
        self.ps = particle_system
        self.ns = neighbour_search
        M = self.ps.particle_max_num
        dev = self.ps.device
        self.k = torch.zeros(M, dtype=torch.float32, device=dev)
        self.max_iteration_opt = int(config.get_cfg("maxOuterIters", 1000))
        self.use_fixed_iters = bool(config.get_cfg("fixedOuterIters", False))
        self.tol_opt = float(config.get_cfg("densityErrorTol", 1e-3))
        self.omega = float(config.get_cfg("omega", 10.0))
        self.stats_density_errors = []
        self.stats_iter_costs = []
        self.stats_pcg_iters = []
        self.stats_pos_residuals = []
        self.stats_outer_iters = 0
        self.stats_final_density_error = 0.0
        self.stats_final_pos_residual = 0.0
        self.stats_total_cost_ms = 0.0

    def compute_density(self):
        """pressure_base.py:341-352 -- density is evaluated by the neighbour build."""
        self.ps.sync_from_library(density=True)

    def compute_Aii(self):
        """pressure_base.py:354-365 -- k = 1/(Aii + 1e-3) is evaluated by the neighbour build."""
        n = self.ps.particle_num
        self.ps.ctx.read(N.FIELD_K, self.k[:n])

    def reset_stats(self):
        self.stats_density_errors = []
        self.stats_iter_costs = []
        self.stats_pcg_iters = []
        self.stats_pos_residuals = []

    def check_converged(self, err, n_updates, pos_res=0.0):
        """pressure_base.py:487-495"""
        if (not self.use_fixed_iters) and err >= self.tol_opt and n_updates >= self.max_iteration_opt:
            raise RuntimeError(f"[{type(self).__name__}] does not converge! density err={err:.2e} "
                               f"(tol={self.tol_opt:.2e}) after {n_updates} iters")

    def solve(self, dt):
        pass


class IP(PressureBase):
    """Velocity-based density solver: implicit projection with a CG inner solve (ip.py:222)."""

    def __init__(self, particle_system, neighbour_search, config):
        super().__init__(particle_system, neighbour_search, config)
        self.tol_pcg = float(config.get_cfg("tol_pcg", 1e-6))
        self.max_iteration_pcg = 1000
        self.D_inv = self.k
        self.last_result = None

    def params(self):
        return IPParams(omega=self.omega, tol_opt=self.tol_opt, max_iteration_opt=self.max_iteration_opt,
                        use_fixed_iters=self.use_fixed_iters, tol_pcg=self.tol_pcg,
                        max_iteration_pcg=self.max_iteration_pcg)

    def record(self, res):
        """Copy one solve's statistics into the reference's attributes and raise its errors."""
        self.last_result = res
        self.reset_stats()
        self.stats_density_errors = list(res.density_errors)
        self.stats_iter_costs = list(res.iter_costs)
        self.stats_pcg_iters = list(res.pcg_iters)
        self.stats_outer_iters = res.outer_iters
        self.stats_total_cost_ms = res.solve_ms
        if res.status == N.IVP_NOT_PD:
            raise RuntimeError(f"[IP] non-positive definite matrix at PCG iter {res.pcg_iter_bad} "
                               f"(pAp={res.pAp_bad:.2e})")
        if res.status == N.IVP_NEIGHBOR_OVERFLOW:
            raise RuntimeError("[IP] neighbour rows overflow neighborCacheSize")
        self.check_converged(res.last_err, res.outer_iters)
        self.stats_final_density_error = res.final_density_error

    def solve(self, dt):
        """ip.py:331-404: reads ps.v, writes ps.v and ps.divergence."""
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        res = ps.ctx.ip_solve(dt, self.params())
        self.record(res)
        ps.sync_from_library(v=True, divergence=True)


class PRJ(PressureBase):
    """Projected Jacobi on the constant-density constraint (IISPH-style), with the optional
    divergence-free (DFSPH-style) second pass (prj.py:143-340)."""

    def __init__(self, particle_system, neighbour_search, config):
        super().__init__(particle_system, neighbour_search, config)
        self.outer_iter = 0
        self.warm_start = False
        self.omega = 0.5
        self.enable_df = False
        self.df_num_threshold = 25
        self.df_tol_opt = 1e-2
        self.stats_df_outer_iters = 0
        self.stats_df_total_cost_ms = 0.0
        self.D_inv = self.k
        self.last_result = None

    def params(self):
        return PRJParams(tol_opt=self.tol_opt, max_iteration_opt=self.max_iteration_opt,
                         use_fixed_iters=self.use_fixed_iters, omega=self.omega, warm_start=self.warm_start,
                         enable_df=self.enable_df, df_num_threshold=self.df_num_threshold,
                         df_tol_opt=self.df_tol_opt)

    def record(self, res):
        """Copy one solve's statistics into the reference's attributes and raise its errors
        (the divergence-free pass's, when it ran, into stats_df_*)."""
        self.last_result = res
        self.reset_stats()
        self.stats_density_errors = list(res.density_errors)
        self.stats_iter_costs = list(res.iter_costs)
        self.outer_iter = res.outer_iters
        self.stats_outer_iters = res.outer_iters
        self.stats_total_cost_ms = res.solve_ms
        if res.status == N.IVP_NEIGHBOR_OVERFLOW:
            raise RuntimeError("[PRJ] neighbour rows overflow neighborCacheSize")
        self.check_converged(res.last_err, res.outer_iters)
        self.stats_final_density_error = res.final_density_error
        if self.enable_df:
            self.stats_df_outer_iters = res.df_outer_iters
            self.stats_df_total_cost_ms = res.df_ms

    def solve(self, dt):
        """prj.py:292-340: reads ps.v, writes ps.v and ps.divergence."""
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        res = ps.ctx.prj_solve(dt, self.params())
        self.record(res)
        ps.sync_from_library(v=True, divergence=True)

    def solve_df(self, dt):
        """prj.py:213-290: the divergence-free pass on ps.v."""
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        res = ps.ctx.prj_solve_df(dt, self.params())
        self.stats_df_outer_iters = res.df_outer_iters
        self.stats_df_total_cost_ms = res.df_ms
        ps.sync_from_library(v=True, divergence=True)
