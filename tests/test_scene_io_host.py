"""Host logic of scene loading with meshes: block order and mesh semantics of load_particles,
ScheduledSpawns expansion (scene_loader.py:143-186) and the spawn system's firing rules
(scheduled_spawn.py:68-117).  Voxel centres come from the oracle's voxeliser here."""

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.mesh import load_obj, place
from implicitvelocityprojection_amd.scenes import grid_offsets, load_particles, spawn_entries
from implicitvelocityprojection_amd.spawn import ScheduledSpawn, ScheduledSpawnSystem
from oracle import voxelize as OV
from tests import meshes as M


def _scene(tmp_path):
    (tmp_path / "ball.obj").write_text(M.to_obj(*M.icosphere(0.12, 2)))
    (tmp_path / "cup.obj").write_text(M.to_obj(*M.open_container((-0.3, 0.0, -0.3), (0.3, 0.4, 0.3), 0.06)))
    ball, cup = str(tmp_path / "ball.obj"), str(tmp_path / "cup.obj")
    return {
        "Configuration": {"domainStart": [0, 0, 0], "domainEnd": [1.2, 1.2, 1.2], "particleRadius": 0.01,
                          "timeStepSize": 0.001, "boundaryLayers": 1},
        "FluidBlocks": [{"objectId": 0, "start": [0.1, 0.8, 0.1], "end": [0.2, 0.9, 0.2]}],
        "FluidBodies": [{"objectId": 3, "geometryFile": ball, "translation": [0.6, 0.5, 0.6],
                         "velocity": [0.0, -0.5, 0.0], "angularVelocity": [0.0, 2.0, 0.0]}],
        "RigidBodies": [{"objectId": 1, "geometryFile": cup, "translation": [0.6, 0.1, 0.6], "isDynamic": False,
                         "velocity": [1.0, 0.0, 0.0]}],
        "ScheduledSpawns": [{"objectId": 10, "geometryFile": ball, "translation": [0.35, 0.95, 0.6], "spawnNum": 3,
                             "startTime": 0.002, "interval": 0.003, "velocity": [0, -1, 0],
                             "waveOffset": [0.25, 0.0, 0.0], "colorCycle": [[1, 2, 3], [4, 5, 6]]},
                            {"objectId": 20, "geometryFile": ball, "translation": [0.3, 0.65, 0.3],
                             "grid": {"counts": [2, 1, 2], "spacing": [0.3, 0.0, 0.3]}, "spawnNum": 1,
                             "startTime": 0.001, "interval": 0.004, "material": "rigid", "isDynamic": False}],
    }


def _mesh_points(cfg):
    def f(obj):
        m = place(load_obj(obj["geometryFile"]), obj)
        return OV.voxelize(m.vertices, m.faces, 1.0 * cfg.particle_diameter)
    return f


def test_mesh_blocks_order_and_semantics(tmp_path):
    cfg = SimConfig.from_dict(_scene(tmp_path))
    mp = _mesh_points(cfg)
    ps = load_particles(cfg, mesh_points=mp)
    # insertion order: fluid blocks, fluid bodies, rigid bodies, generated bodies, boundary walls
    ids = [int(o) for o in ps.object_id[np.r_[0, np.nonzero(np.diff(ps.object_id))[0] + 1]]]
    assert ids == [0, 3, 1, 24]          # walls take max(object ids incl. spawn ids 20..23) + 1
    ball = mp(cfg.get_fluid_bodies()[0])
    sel = ps.object_id == 3
    np.testing.assert_array_equal(ps.x[sel], ball.astype(np.float32))
    assert np.all(ps.material[sel] == 1) and np.all(ps.is_dynamic[sel] == 1)
    # v = v_lin + omega x (r - r_cm)
    r = ball - ball.mean(axis=0)
    want = np.cross(np.array([0, 2, 0], np.float32), r).astype(np.float32) + np.array([0, -0.5, 0], np.float32)
    np.testing.assert_array_equal(ps.v[sel], want)
    cup = ps.object_id == 1
    assert np.all(ps.material[cup] == 0) and np.all(ps.is_dynamic[cup] == 0) and np.all(ps.v[cup] == 0.0)
    with pytest.raises(NotImplementedError):
        load_particles(cfg)                       # mesh blocks need a voxeliser


def test_spawn_expansion(tmp_path):
    cfg = SimConfig.from_dict(_scene(tmp_path))
    mp = _mesh_points(cfg)
    e = spawn_entries(cfg, mp)
    ball0 = mp(cfg.get_scheduled_spawns()[0])
    a = [x for x in e if 10 <= x.object_id < 20]
    assert [x.time for x in a] == [0.002, 0.002 + 0.003, 0.002 + 2 * 0.003]
    assert [x.object_id for x in a] == [10, 11, 12]
    assert [list(x.color) for x in a] == [[1, 2, 3], [4, 5, 6], [1, 2, 3]]
    for k, x in enumerate(a):
        shift = k * np.array([0.25, 0.0, 0.0], np.float32)
        np.testing.assert_array_equal(x.positions, ((ball0 + shift) if k else ball0).astype(np.float32))
        assert x.material == 1 and x.is_dynamic == 1
    g = [x for x in e if x.object_id >= 20]
    # a 2 x 1 x 2 grid: one id per slot (reused by every wave), copies centred on the translation
    assert len(g) == 4 and sorted({x.object_id for x in g}) == [20, 21, 22, 23]
    off = grid_offsets(cfg.get_scheduled_spawns()[1])
    np.testing.assert_array_equal(off, np.array([[-0.15, 0, -0.15], [-0.15, 0, 0.15], [0.15, 0, -0.15],
                                                 [0.15, 0, 0.15]], np.float32))
    assert all(x.material == 0 and x.is_dynamic == 0 for x in g)


class _PS:
    def __init__(self, n, cap):
        self.particle_num, self.particle_max_num, self.rigid_particle_num = n, cap, 0
        self.calls = []

    def add_particles(self, object_id, n, pos, vel, dens, mat, dyn, color, m_V0=None):
        self.calls.append((object_id, n))
        self.particle_num += n


def test_spawn_system_fires_in_time_order_within_capacity():
    ps = _PS(100, 100 + 10 + 7)
    sys_ = ScheduledSpawnSystem(ps)
    pts = lambda n: np.zeros((n, 3))
    sys_.add(ScheduledSpawn(0.5, 2, pts(10), [0, 0, 0], 1000.0, [0, 0, 0]))
    sys_.add(ScheduledSpawn(0.2, 1, pts(5), [0, 0, 0], 1000.0, [0, 0, 0], material="rigid", is_dynamic=False))
    sys_.add(ScheduledSpawn(0.5, 3, pts(4), [0, 0, 0], 1000.0, [0, 0, 0]))
    assert sys_.step(0.1) == 0
    assert sys_.step(0.2) == 5 and ps.calls == [(1, 5)] and ps.rigid_particle_num == 5
    assert sys_.step(0.6) == 12 and ps.calls[1:] == [(2, 10), (3, 2)]   # the last one cut to capacity
    assert sys_.step(1.0) == 0
    sys_.reset()
    assert ps.rigid_particle_num == 0 and not any(s.fired for s in sys_.spawns)
