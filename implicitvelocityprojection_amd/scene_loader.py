"""Scene loading, mirroring the reference's ``SceneLoader.prepare_scene`` / ``populate_scene``
(scripts/io/scene_loader.py:70-429) for the procedural blocks (FluidBlocks, GeneratedBodies,
boundaryLayers); see scenes.py for the semantics and the unsupported mesh blocks.
"""

import numpy as np

from .config import SimConfig
from .scenes import load_particles


class SceneLoader:
    def __init__(self, config: SimConfig, scene_file_path: str = ""):
        self.cfg = config
        self.scene_file_path = scene_file_path

    def prepare_scene(self):
        ps = load_particles(self.cfg)
        object_ids = sorted(set(int(o) for o in np.unique(ps.object_id))) if ps.n else []
        return {
            "particles": ps,
            "particle_max_num": ps.n,
            "fluid_particle_num": int((ps.material == 1).sum()),
            "rigid_particle_num": int((ps.material == 0).sum()),
            "solid_particle_num": 0,
            "num_objects": (max(object_ids) + 1) if object_ids else 1,
            "num_rigid_bodies": 0,
            "object_ids": object_ids,
        }

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
