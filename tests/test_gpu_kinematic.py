"""CFL time step and keyframe-animated rigid bodies on the GPU: ivp_rigid_transform and
ivp_max_speed2 against numpy float32, and Framework.forward with useCFL and an animated static
paddle against the oracle's step (teacher-forced per step)."""

import ctypes

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from oracle import oracle as O
from tests import meshes as M
from tests.helpers import rel_inf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_rigid_transform_and_max_speed_match_numpy():
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.animation import rotation_matrix
    from implicitvelocityprojection_amd.context import Context
    lib = N.load()
    rng = np.random.default_rng(1)
    n = 5000
    x0 = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    oid = rng.integers(0, 3, size=n).astype(np.int32)
    R = rotation_matrix(np.array([0.3, 0.8, -0.52]) / np.linalg.norm([0.3, 0.8, -0.52]), 0.77)
    pivot = np.array([0.1, -0.2, 0.3], np.float32)
    T = np.array([0.5, 0.25, -1.0], np.float32)
    dx0, dx = torch.from_numpy(x0).cuda(), torch.from_numpy(x0 * 3).cuda()
    doid = torch.from_numpy(oid).cuda()
    f3 = lambda a: (ctypes.c_float * 3)(*[float(v) for v in a])
    rc = lib.ivp_rigid_transform(ctypes.c_void_p(dx0.data_ptr()), ctypes.c_void_p(dx.data_ptr()),
                                 ctypes.c_void_p(doid.data_ptr()), n, 2, f3(pivot),
                                 (ctypes.c_float * 9)(*[float(v) for v in R.reshape(9)]), f3(T), None)
    assert rc == 0
    got = dx.cpu().numpy()
    want = x0 * 3
    d = x0[oid == 2] - pivot
    r = np.stack([(R[i, 0] * d[:, 0] + R[i, 1] * d[:, 1]) + R[i, 2] * d[:, 2] for i in range(3)], axis=1)
    want[oid == 2] = (pivot + T) + r
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # max |v|^2 over dynamic particles
    ctx = Context(0.04, (0, 0, 0), (2, 2, 2), n)
    dyn = (rng.uniform(size=n) < 0.7).astype(np.int32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    m = np.full(n, 0.0064, np.float32)
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dt)
    ctx.set_particles(t(x0 + 1.0), t(v), t(m), t(np.where(dyn == 1, 1 / m, 0).astype(np.float32)),
                      t(np.full(n, 1000.0, np.float32)), t(dyn, torch.int32), t(np.ones(n, np.int32), torch.int32))
    vd = v[dyn == 1]
    assert ctx.max_speed2() == float(((vd[:, 0] * vd[:, 0] + vd[:, 1] * vd[:, 1]) + vd[:, 2] * vd[:, 2]).max())
    ctx.close()


def _scene(tmp_path):
    (tmp_path / "paddle.obj").write_text(M.to_obj(*M.box((-0.15, -0.02, -0.15), (0.15, 0.02, 0.15))))
    return {
        "Configuration": {"domainStart": [0, 0, 0], "domainEnd": [1.0, 1.0, 1.0], "particleRadius": 0.01,
                          "timeStepSize": 0.001, "useCFL": True, "cflFactor": 0.4, "cflMinDt": 2e-4,
                          "viscosityMethod": 0, "viscosityCoeff": 0.0, "surfaceTensionCoeff": 0.0},
        "FluidBlocks": [{"objectId": 0, "start": [0.3, 0.1, 0.3], "end": [0.7, 0.3, 0.7],
                         "velocity": [20.0, 0.0, 0.0]}],
        "RigidBodies": [{"objectId": 1, "geometryFile": str(tmp_path / "paddle.obj"),
                         "translation": [0.5, 0.36, 0.5], "isDynamic": False,
                         "Animation": {"tracks": [{
                             "channels": {"translation": [{"t": 0.0, "value": [0, 0, 0]},
                                                          {"t": 0.01, "value": [0.02, -0.06, 0.0]}],
                                          "rotation_angle_deg": [{"t": 0.0, "value": 0.0},
                                                                 {"t": 0.01, "value": 25.0}]},
                             "pivot": "object_restCOM", "rotation_axis": "y"}]}}],
        "GeneratedBodies": [{"objectId": 2, "type": "cube", "translation": [0.5, 0.5, 0.5], "size": [0.96, 0.96, 0.96],
                             "layers": 2}],
    }


def test_framework_cfl_and_animated_paddle_match_oracle(tmp_path):
    from implicitvelocityprojection_amd.animation import AnimationEngine
    from implicitvelocityprojection_amd.framework import Framework
    from implicitvelocityprojection_amd.neighbour_search import NeighbourSearch
    from implicitvelocityprojection_amd.scene_loader import SceneLoader
    from implicitvelocityprojection_amd.simulation_data import SimulationData
    cfg = SimConfig.from_dict(_scene(tmp_path))
    loader = SceneLoader(cfg)
    data = loader.prepare_scene()
    ps = SimulationData(cfg)
    loader.populate_scene(ps, data)
    ns = NeighbourSearch(cfg.domain_size, cfg.support_radius)
    fw = Framework(ps, ns, cfg)
    anim = AnimationEngine(ps)
    anim.build(data)
    assert anim.has_tracks()
    fw.animation = anim
    fw.initialize()
    sim = O.OracleSim(cfg, data["particles"])
    sim.tracks = [(tr.object_id, lambda t, tr=tr: (*tr.pose(t), tr.pivot)) for tr in anim.tracks]
    sim.rebuild()
    n = ps.particle_num
    paddle = data["particles"].object_id == 1
    dts = []
    for step in range(12):
        ps.x[:n] = torch.from_numpy(sim.x).cuda()
        ps.v[:n] = torch.from_numpy(sim.v).cuda()
        ps._dirty = True                  # the paddle's (static) positions changed: reload
        fw.initialize()
        fw.forward()
        sim.step()
        dts.append(fw.dt)
        assert fw.dt == sim.dt, (step, fw.dt, sim.dt)
        dyn = sim.is_dynamic == 1
        x_g = ps.x[:n].cpu().numpy()
        v_g = ps.v[:n].cpu().numpy()
        assert rel_inf(v_g[dyn], sim.v[dyn]) <= 1e-5, step
        assert rel_inf(x_g[dyn], sim.x[dyn]) <= 1e-5, step
        assert np.array_equal(x_g[paddle], sim.x[paddle]), step
    print("dt per step:", dts, flush=True)
    assert min(dts) < cfg.dt                 # the CFL bound was active
    # the paddle moved down and turned
    assert sim.x[paddle][:, 1].mean() < data["particles"].x[paddle][:, 1].mean() - 0.005
