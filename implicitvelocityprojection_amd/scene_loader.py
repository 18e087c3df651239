"""Scene loading, mirroring the reference's ``SceneLoader`` (scripts/io/scene_loader.py:40-644):
``prepare_scene`` builds every block's particles (meshes voxelised on the GPU by ``ivp_voxelize``)
and the scheduled spawns, ``populate_scene`` allocates and adds them in the reference's order,
``setup_scheduled_spawn_system`` returns the spawn system for ``Framework``.  See scenes.py for
the block semantics.
"""

import os

import numpy as np

from . import mesh as _mesh
from .config import SimConfig
from .scenes import load_particles, spawn_entries
from .spawn import ScheduledSpawn, ScheduledSpawnSystem


class SceneLoader:
    def __init__(self, config: SimConfig, scene_file_path: str = "", model_root=None):
        """Relative ``geometryFile`` paths are opened as given (relative to the working directory,
        as the reference does) unless `model_root` names the directory to resolve them against."""
        self.cfg = config
        self.scene_file_path = scene_file_path
        self.model_root = model_root
        self.particle_diameter = config.particle_diameter
        self._voxels = {}

    def get_scene_name(self, scene_file_path: str) -> str:
        import pathlib
        return pathlib.Path(scene_file_path).stem if scene_file_path else "scene"

    def _path(self, p):
        return p if (self.model_root is None or os.path.isabs(p)) else os.path.join(self.model_root, p)

    def load_transformed_mesh(self, obj: dict):
        """scene_loader.py:532-545: the OBJ file scaled, rotated about its vertex mean, translated."""
        return _mesh.place(_mesh.load_obj(self._path(obj["geometryFile"])), obj)

    def load_mesh(self, obj: dict, a: float = 1.0) -> np.ndarray:
        """scene_loader.py:547-580: voxel centres (float64) of the placed mesh at pitch a * d."""
        m = self.load_transformed_mesh(obj)
        pts = _mesh.voxelize(m, a * self.particle_diameter, max_iter=15).cpu().numpy()
        print(f"Mesh object {obj['objectId']} loaded. Particle count: {pts.shape[0]}")
        return pts

    def prepare_scene(self):
        ps = load_particles(self.cfg, mesh_points=self.load_mesh)
        spawns = spawn_entries(self.cfg, self.load_mesh)
        spawn_capacity = sum(int(e.positions.shape[0]) for e in spawns)
        object_ids = sorted(set(int(o) for o in np.unique(ps.object_id))) if ps.n else []
        object_ids = sorted(set(object_ids) | {e.object_id for e in spawns})
        return {
            "particles": ps,
            "spawn_defs": spawns,
            "object_map": self._object_map(ps, spawns),
            "particle_max_num": ps.n + spawn_capacity,
            "fluid_particle_num": int((ps.material == 1).sum()),
            "rigid_particle_num": int((ps.material == 0).sum()),
            "solid_particle_num": 0,
            "num_objects": (max(object_ids) + 1) if object_ids else 1,
            "num_rigid_bodies": len(self.cfg.get_rigid_bodies()),
            "object_ids": object_ids,
        }

    def _object_map(self, ps, spawns):
        """Scene object per id with its particles' rest positions (the "voxelizedPoints" /
        "generatedPositions" the reference keeps, scene_loader.py:86-186), read by the animation
        tracks' named pivots.  A spawn id maps to the last body spawned under it."""
        cfg = self.cfg
        out = {}
        blocks = (cfg.get_fluid_blocks(), cfg.get_fluid_bodies(), cfg.get_rigid_bodies(), cfg.get_generated_bodies())
        for group in blocks:
            for obj in group:
                oid = int(obj["objectId"])
                entry = dict(obj)
                entry["voxelizedPoints"] = ps.x[ps.object_id == oid]
                out[oid] = entry
        spawn_objs = {}
        for sp in cfg.get_scheduled_spawns():
            spawn_objs[int(sp["objectId"])] = sp
        for e in spawns:
            base = max(k for k in spawn_objs if k <= e.object_id)
            entry = dict(spawn_objs[base])
            entry["voxelizedPoints"] = e.positions
            entry["objectId"] = e.object_id
            out[e.object_id] = entry
        return out

    def populate_scene(self, ps, scene_data: dict):
        ps.allocate(particle_max_num=scene_data["particle_max_num"], num_objects=scene_data["num_objects"],
                    num_rigid_bodies=scene_data["num_rigid_bodies"])
        ps.rigid_particle_num = scene_data["rigid_particle_num"]
        p = scene_data["particles"]
        # one add_particles call per block, in insertion order
        cuts = [0] + [i for i in range(1, p.n) if p.object_id[i] != p.object_id[i - 1]] + [p.n]
        for s, e in zip(cuts[:-1], cuts[1:]):
            if e > s:
                ps.add_particles(int(p.object_id[s]), e - s, p.x[s:e], p.v[s:e], p.density0[s:e], p.material[s:e],
                                 p.is_dynamic[s:e], np.zeros((e - s, 3), np.float32))
        ps.v0.copy_(ps.v)

    def setup_scheduled_spawn_system(self, ps, scene_data):
        """scene_loader.py:582-616: None without spawns."""
        defs = scene_data.get("spawn_defs", [])
        if not defs:
            return None
        system = ScheduledSpawnSystem(ps)
        for e in defs:
            system.add(ScheduledSpawn(e.time, e.object_id, e.positions, e.velocity, e.density, e.color,
                                      material="rigid" if e.material == 0 else "fluid", is_dynamic=bool(e.is_dynamic)))
        return system
