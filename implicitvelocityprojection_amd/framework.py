"""Step orchestrator, mirroring the reference's ``Framework`` (scripts/framework.py:143-544) for
the velocity-based solvers (IP, and PRJ with its optional divergence-free pass).

``forward()`` runs one reference step -- gravity, viscosity, pressure solve, advection, domain clamp,
neighbour rebuild, density, Aii -- as a single library call (ivp_step) on the library's particle
store, then copies x, v, density and divergence back into the SimulationData arrays (skip that
with ``sync_arrays = False`` when only the final state is read).  ``ps.v`` is read at the start of
every step, so velocity edits between steps take effect as in the reference; position edits need
a rebuild (``initialize()``), as the reference's cached neighbour rows do.
"""

import math
import time

from . import _native as N
from .neighbour_search import NeighbourSearch
from .pressure import IP, PRJ
from .viscosity import Viscosity, Weiler18

_UNSUPPORTED_PRESSURE = {3: "SISPH", 5: "PBF"}
_PRESSURE_CLASSES = {0: PRJ, 1: IP}       # framework.py:168-172


class Framework:
    def __init__(self, simulation_data, neighbour_search: NeighbourSearch, config):
        self.ps = simulation_data
        self.ns = neighbour_search
        self.cfg = config
        self.dt = config.dt
        self.dt_default = config.dt
        # CFL time step (framework.py:150-152, 427-440)
        self.use_cfl = bool(config.get_cfg("useCFL", False))
        self.cfl_factor = float(config.get_cfg("cflFactor", 0.4))
        self.cfl_min_dt = float(config.get_cfg("cflMinDt", 1e-4))
        self.time = 0.0
        self.initial_particle_num = simulation_data.particle_num
        self.stats_pressure_ms = 0.0
        self.stats_elasticity_ms = 0.0
        self.g = tuple(float(v) for v in config.gravity)
        pressure_method = int(config.get_cfg("pressureMethod", 1))
        if pressure_method in _UNSUPPORTED_PRESSURE:
            raise NotImplementedError(f"pressureMethod {pressure_method} ({_UNSUPPORTED_PRESSURE[pressure_method]}) "
                                      "is not part of this build; use pressureMethod 1 (IP) or 0 (PRJ)")
        if pressure_method not in _PRESSURE_CLASSES:
            raise ValueError(f"Unknown pressureMethod {pressure_method} (valid: [0, 1, 3, 5]).")
        self.pressure = _PRESSURE_CLASSES[pressure_method](simulation_data, neighbour_search, config)
        self.viscosity = (Weiler18 if config.viscosity_method == 1 else Viscosity)(simulation_data, neighbour_search,
                                                                                   config)
        self.viscosity_coeff = config.viscosity_coeff
        # framework.py:181-187: surfaceTensionMethod 1 is Jeske23 (implicit); the explicit model
        # (method 0) is provided only at coefficient 0, where it does nothing
        self.surface_tension_method = int(config.surface_tension_method)
        self.surface_tension_coeff = config.surface_tension_coeff
        if self.surface_tension_coeff != 0.0 and self.surface_tension_method != 1:
            raise NotImplementedError("explicit surface tension (surfaceTensionMethod 0) is not part of this build; "
                                      "use surfaceTensionMethod 1 (Jeske23) or surfaceTensionCoeff 0")
        self.last_st_iters = 0
        self.pressure_enabled = bool(config.get_cfg("pressureEnabled", True))
        if not self.pressure_enabled:
            raise NotImplementedError("pressureEnabled false is not supported")
        self.sync_arrays = True
        self.last_visc_iters = 0
        self.scheduled_spawn_system = None        # set by the caller (main.py:73)
        self.animation = None                     # set by the caller (main.py:76-80)

    def initialize(self):
        """framework.py:282-290, 323-330: neighbour rebuild, density and Aii at the current x."""
        ps = self.ps
        self.ns.build(ps.x, ps.particle_num)
        self.ns.build_cache(ps.x, ps.particle_num, ps.cache_size)
        self.pressure.compute_density()
        self.pressure.compute_Aii()

    def reset(self):
        """framework.py:227-280 for the state this build keeps: x = x0, v = v0, density0, t = 0."""
        ps = self.ps
        ps.particle_num = self.initial_particle_num
        n = ps.particle_num
        ps.x[:n] = ps.x0[:n]
        ps.x_old[:n] = ps.x0[:n]
        ps.v[:n] = ps.v0[:n]
        ps.density.copy_(ps.density0)
        self.pressure.k.zero_()
        if self.scheduled_spawn_system is not None:
            self.scheduled_spawn_system.reset()
        # framework.py:265-270: the solvers' persistent multipliers (PRJ's warm-start lambdas) are
        # zeroed with k
        if ps.ctx is not None and ps.ctx.loaded:
            ps.ctx.reset()
        self.time = 0.0
        ps._dirty = True
        self.initialize()

    def forward(self):
        """One step (framework.py:423-544, velocity-based branch)."""
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        if self.use_cfl:
            # dt = cflFactor h / v_max, at most the scene's dt and at least cflMinDt
            v_max_sq = ps.ctx.max_speed2()
            if math.isfinite(v_max_sq) and v_max_sq > 1e-16:
                self.dt = min(self.dt_default, self.cfl_factor * ps.support_radius / v_max_sq ** 0.5)
            else:
                self.dt = self.dt_default
            self.dt = max(self.dt, self.cfl_min_dt)
        if isinstance(self.viscosity, Weiler18):
            self.viscosity.apply_options(ps.ctx)
        t0 = time.perf_counter()
        res, vit = ps.ctx.step(self.dt, self.g, self.cfg.viscosity_method, self.viscosity_coeff,
                               self.pressure.params(), self.surface_tension_method, self.surface_tension_coeff)
        self.stats_pressure_ms = res.solve_ms
        self.last_visc_iters = vit
        self.last_st_iters = ps.ctx.last_st_iters
        self.pressure.record(res)
        ps._cache_stale = True
        spawning = self.scheduled_spawn_system is not None
        if self.sync_arrays or spawning or self.animation is not None:
            n = ps.particle_num
            ps.x_old[:n] = ps.x[:n]
            ps.sync_from_library(x=True, v=True, density=True, divergence=True)
            ps.ctx.read(N.FIELD_K, self.pressure.k[:n])
        # Animated bodies take their pose at this step's start time (framework.py:442-444).  The
        # reference moves them before the physics but solves with the neighbour cache, density and
        # Aii of the previous build; animated bodies are static rigid here, so moving them after
        # the library step, before the rebuild that ends the step, gives the same step.
        moved = False
        if self.animation is not None:
            moved = self.animation.apply(self.time, self.dt)
            self.animation.clear_kinematic()
        # bodies due at this step's start time join before the rebuild (framework.py:528-540): the
        # library step has rebuilt for the old set, so a step that spawns or moves bodies rebuilds
        # once more, with the set reloaded
        if spawning and self.scheduled_spawn_system.step(self.time) > 0:
            moved = True
        if moved:
            ps._dirty = True
            self.initialize()
        self.time += self.dt
        self.stats_step_ms = (time.perf_counter() - t0) * 1000.0
