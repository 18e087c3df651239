"""Scheduled body spawns, mirroring the reference's ``ScheduledSpawnSystem``
(scripts/sim/utils/scheduled_spawn.py:35-117): pre-voxelised bodies added to the particle set at
their times.  Growing ``particle_num`` makes the next library sync load the grown set
(``ivp_set_particles``, up to the capacity given at allocation); ``Framework.forward`` then
rebuilds neighbours, density and Aii in the same step, as the reference does.
"""

import numpy as np


class ScheduledSpawn:
    """scheduled_spawn.py:14-32 for the fluid and rigid materials."""

    def __init__(self, time, object_id, positions, velocity, density, color, material="fluid", is_dynamic=True):
        self.time = float(time)
        self.object_id = int(object_id)
        self.positions = np.asarray(positions).astype(np.float32, copy=False)
        self.velocity = np.array(velocity, dtype=np.float32)
        self.density = float(density)
        self.color = np.array(color, dtype=np.int32)
        self.material_code = 0 if material == "rigid" else 1
        self.is_dynamic = 1 if is_dynamic else 0
        self.fired = False


class ScheduledSpawnSystem:
    def __init__(self, ps):
        self.ps = ps
        self.spawns = []
        self._sorted = True
        self.rigid_capacity = 0
        self.rigid_spawned = 0

    def add(self, spawn: ScheduledSpawn):
        self.spawns.append(spawn)
        if spawn.material_code == 0:
            self.rigid_capacity += spawn.positions.shape[0]
        self._sorted = False

    def reset(self):
        self.ps.rigid_particle_num -= self.rigid_spawned
        self.rigid_spawned = 0
        for s in self.spawns:
            s.fired = False

    def step(self, t):
        """Fire the spawns whose time has come, in time order (stable).  Returns the particles
        added; a body that does not fit in the remaining capacity is cut to it."""
        if not self._sorted:
            self.spawns.sort(key=lambda s: s.time)
            self._sorted = True
        added = 0
        for s in self.spawns:
            if s.fired or t < s.time:
                continue
            remain = int(self.ps.particle_max_num) - int(self.ps.particle_num)
            if remain <= 0:
                s.fired = True
                continue
            n = min(s.positions.shape[0], remain)
            self.ps.add_particles(
                s.object_id, n, s.positions[:n], np.tile(s.velocity[None, :], (n, 1)),
                np.full(n, s.density, dtype=np.float32), np.full(n, s.material_code, dtype=np.int32),
                np.full(n, s.is_dynamic, dtype=np.int32), np.tile(s.color[None, :3], (n, 1)))
            s.fired = True
            added += n
            if s.material_code == 0:
                self.ps.rigid_particle_num += n
                self.rigid_spawned += n
        return added
