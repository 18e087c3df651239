"""Particle data, mirroring the reference's ``SimulationData`` (scripts/sim/simulation_data.py).

Arrays are device tensors sized ``particle_max_num`` in the caller's particle order, with the
reference's names and meaning.  The arrays the IP hot path works on are mirrored in a libivp
context (its particle store keeps dynamic particles sorted by cell); ``sync_to_library`` and
``sync_from_library`` move data between the two, and the drop-in classes call them around every
library operation.

The neighbour cache arrays ``neighbor_count`` / ``neighbor_ids`` / ``neighbor_gradW`` keep the
reference layout ([i * cache_size + k], simulation_data.py:206-209) but are produced on demand:
the solvers read the library's own tiled rows, and the reference-layout cache is exported when
one of these attributes is read after a rebuild.
"""

import numpy as np
import torch

from . import _native as N
from .config import SimConfig
from .context import Context

_OWNERS = {}


def owner_of(positions):
    """The SimulationData whose position array is ``positions``."""
    ps = _OWNERS.get(positions.data_ptr())
    if ps is None:
        raise ValueError("positions must be a SimulationData.x array")
    return ps


class SimulationData:
    material_rigid = 0
    material_fluid = 1

    def __init__(self, config: SimConfig, device=None):
        self.cfg = config
        self.dim = 3
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.domain_size = np.array(config.domain_size, dtype=np.float32)
        self.particle_radius = config.particle_radius
        self.particle_diameter = config.particle_diameter
        self.support_radius = config.support_radius
        self.padding = self.support_radius
        self.m_V0_init = 0.8 * self.particle_diameter ** self.dim
        self.num_objects = 1
        self.num_rigid_bodies = 0
        self.rigid_particle_num = 0
        self.particle_num = 0
        self.particle_max_num = 0
        self.num_deformable = 0
        self.cache_size = int(config.get_cfg("neighborCacheSize", 60))
        self.allocated = False
        self.ctx = None
        self._dirty = True          # particle set changed since the library last loaded it
        self._cache_stale = True    # reference-layout cache not exported since the last rebuild
        self._cache = None

    # -- allocation (simulation_data.py:165-212) ------------------------------------------------
    def allocate(self, particle_max_num: int, num_objects: int = 1, num_rigid_bodies: int = 0):
        self.particle_max_num = int(particle_max_num)
        self.num_objects = int(max(num_objects, 1))
        self.num_rigid_bodies = int(num_rigid_bodies)
        z = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=self.device)
        M = self.particle_max_num
        self.object_id = z(M, dt=torch.int32)
        self.x = z(M, 3)
        self.x_old = z(M, 3)
        self.x0 = z(M, 3)
        self.v = z(M, 3)
        self.v0 = z(M, 3)
        self.m = z(M)
        self.m_inv = z(M)
        self.density = z(M)
        self.density0 = z(M)
        self.divergence = z(M)
        self.material = z(M, dt=torch.int32)
        self.color = z(M, 4)
        self.is_dynamic = z(M, dt=torch.int32)
        self.is_kinematic = z(M, dt=torch.int32)
        _OWNERS[self.x.data_ptr()] = self
        cfg = self.cfg
        self.ctx = Context(np.float32(self.support_radius), cfg.domain_start, cfg.domain_size, max(M, 1),
                           cache_size=self.cache_size, material_fluid=self.material_fluid)
        # extension key (no counterpart in the reference): neighbour-row format, 0 ids only
        # (default), 1 cached float32 gradW, 2 the float16 precision form (include/ivp.h)
        self.ctx.gradient_cache = int(cfg.get_cfg("gradientCache", 0))
        self.allocated = True

    # -- particles (simulation_data.py:214-356) --------------------------------------------------
    def add_particles(self, object_id, new_particles_num, new_particles_positions, new_particles_velocity,
                      new_particle_density, new_particles_material, new_particles_is_dynamic, new_particles_color,
                      m_V0=None):
        if not self.allocated:
            raise RuntimeError("Must call allocate() before adding particles")
        n = int(new_particles_num)
        if self.particle_num + n > self.particle_max_num:
            print(f"Warning: Cannot add {n} particles. Only {self.particle_max_num - self.particle_num} slots available.")
            return
        s, e = self.particle_num, self.particle_num + n
        dev = self.device
        t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
        pos = t(np.asarray(new_particles_positions, np.float32).reshape(n, 3))
        vel = t(np.asarray(new_particles_velocity, np.float32).reshape(n, 3))
        dens = np.asarray(new_particle_density, np.float32).reshape(n)
        dyn = np.asarray(new_particles_is_dynamic, np.int32).reshape(n)
        v0 = np.float32(self.m_V0_init if m_V0 is None else float(m_V0))
        # add_particle (simulation_data.py:40-57): m = V0 rho; static: m *= 10, m_inv = 0
        m = (v0 * dens).astype(np.float32)
        m_inv = np.zeros(n, np.float32)
        m_inv[dyn == 1] = np.float32(1.0) / m[dyn == 1]
        m[dyn != 1] = m[dyn != 1] * np.float32(10.0)
        color = np.asarray(new_particles_color, np.float32).reshape(n, -1)
        if color.shape[1] == 3:
            color = np.hstack([color, np.full((n, 1), 255.0, np.float32)])
        self.object_id[s:e] = int(object_id)
        self.x[s:e] = pos
        self.x_old[s:e] = pos
        self.x0[s:e] = pos
        self.v[s:e] = vel
        self.density0[s:e] = t(dens)
        self.density[s:e] = t(dens)
        self.m[s:e] = t(m)
        self.m_inv[s:e] = t(m_inv)
        self.material[s:e] = t(np.asarray(new_particles_material, np.int32).reshape(n), torch.int32)
        self.is_dynamic[s:e] = t(dyn, torch.int32)
        self.color[s:e] = t(color)
        self.particle_num = e
        self._dirty = True

    def add_cube(self, object_id, lower_corner, cube_size, material, is_dynamic, color=(0, 0, 0), density=None,
                 velocity=None):
        from .scenes import cube_lattice
        pos = cube_lattice(lower_corner, cube_size, self.particle_diameter)
        n = pos.shape[0]
        vel = np.zeros_like(pos) if velocity is None else np.tile(np.asarray(velocity, np.float32), (n, 1))
        self.add_particles(object_id, n, pos, vel, np.full(n, 1000.0 if density is None else density, np.float32),
                           np.full(n, material, np.int32), np.full(n, is_dynamic, np.int32),
                           np.tile(np.asarray(color, np.float32), (n, 1)))

    # -- library synchronisation ------------------------------------------------------------------
    def sync_to_library(self, positions=True, velocities=False):
        """Load the particle set (first call or after a change), else copy x (dynamic) / v in."""
        n = self.particle_num
        if self._dirty:
            self.ctx.set_particles(self.x[:n], self.v[:n], self.m[:n], self.m_inv[:n], self.density0[:n],
                                   self.is_dynamic[:n], self.material[:n])
            self._dirty = False
            return
        if positions:
            self.ctx.write(N.FIELD_X, self.x[:n])
        if velocities:
            self.ctx.write(N.FIELD_V, self.v[:n])

    def sync_from_library(self, x=False, v=False, density=False, divergence=False):
        n = self.particle_num
        if x:
            self.ctx.read(N.FIELD_X, self.x[:n])
        if v:
            self.ctx.read(N.FIELD_V, self.v[:n])
        if density:
            self.ctx.read(N.FIELD_DENSITY, self.density[:n])
        if divergence:
            self.ctx.read(N.FIELD_DIVERGENCE, self.divergence[:n])

    def _export_cache(self):
        n, C, M = self.particle_num, self.cache_size, self.particle_max_num
        if self._cache is None:
            self._cache = (torch.zeros(M, dtype=torch.int32, device=self.device),
                           torch.zeros(M * C, dtype=torch.int32, device=self.device),
                           torch.zeros(M * C, 3, dtype=torch.float32, device=self.device))
        count, ids, gradW = self._cache
        if self._cache_stale and n > 0:
            self.ctx.neighbors_export(count, ids, gradW.view(-1), cache_size=C)
            self._cache_stale = False
        return self._cache

    @property
    def neighbor_count(self):
        return self._export_cache()[0]

    @property
    def neighbor_ids(self):
        return self._export_cache()[1]

    @property
    def neighbor_gradW(self):
        return self._export_cache()[2]
