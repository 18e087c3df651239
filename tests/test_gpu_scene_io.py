"""Scene loading with meshes on the GPU: SceneLoader's particles and scheduled spawns (meshes
voxelised by ivp_voxelize) equal the oracle's, and Framework.forward with a ScheduledSpawnSystem
follows the oracle's step with the same spawns (teacher-forced per step)."""

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.scenes import load_particles, spawn_entries
from oracle import oracle as O
from tests.helpers import rel_inf
from tests.test_scene_io_host import _mesh_points, _scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _cfg(tmp_path):
    sc = _scene(tmp_path)
    sc["Configuration"].update(surfaceTensionCoeff=0.0, viscosityCoeff=0.01, viscosityMethod=0)
    return sc, SimConfig.from_dict(sc)


def test_scene_loader_particles_and_spawns_match_oracle(tmp_path):
    from implicitvelocityprojection_amd.scene_loader import SceneLoader
    sc, cfg = _cfg(tmp_path)
    loader = SceneLoader(cfg)
    data = loader.prepare_scene()
    ps_o = load_particles(cfg, mesh_points=_mesh_points(cfg))
    ps_g = data["particles"]
    print(f"scene io: {ps_g.n} particles, {len(data['spawn_defs'])} spawn entries", flush=True)
    assert ps_g.n == ps_o.n
    for f in ("x", "v", "density0", "material", "is_dynamic", "object_id"):
        assert np.array_equal(getattr(ps_g, f), getattr(ps_o, f)), f
    e_o = spawn_entries(cfg, _mesh_points(cfg))
    assert len(data["spawn_defs"]) == len(e_o)
    for a, b in zip(data["spawn_defs"], e_o):
        assert (a.time, a.object_id, a.material, a.is_dynamic) == (b.time, b.object_id, b.material, b.is_dynamic)
        assert np.array_equal(a.positions, b.positions)
    assert data["particle_max_num"] == ps_o.n + sum(e.positions.shape[0] for e in e_o)


def test_framework_spawns_match_oracle(tmp_path):
    from implicitvelocityprojection_amd.framework import Framework
    from implicitvelocityprojection_amd.neighbour_search import NeighbourSearch
    from implicitvelocityprojection_amd.scene_loader import SceneLoader
    from implicitvelocityprojection_amd.simulation_data import SimulationData
    sc, cfg = _cfg(tmp_path)
    loader = SceneLoader(cfg)
    data = loader.prepare_scene()
    ps = SimulationData(cfg)
    loader.populate_scene(ps, data)
    ns = NeighbourSearch(cfg.domain_size, cfg.support_radius)
    fw = Framework(ps, ns, cfg)
    fw.scheduled_spawn_system = loader.setup_scheduled_spawn_system(ps, data)
    fw.initialize()
    sim = O.OracleSim(cfg, data["particles"])
    sim.add_spawns(spawn_entries(cfg, _mesh_points(cfg)), data["particle_max_num"])
    sim.rebuild()
    counts = []
    for step in range(10):
        n = ps.particle_num
        assert n == sim.x.shape[0]
        ps.x[:n] = torch.from_numpy(sim.x).cuda()
        ps.v[:n] = torch.from_numpy(sim.v).cuda()
        fw.initialize()
        fw.forward()
        sim.step()
        n = ps.particle_num
        counts.append(n)
        assert n == sim.x.shape[0], (step, n, sim.x.shape[0])
        dyn = sim.is_dynamic == 1
        x_g = ps.x[:n].cpu().numpy()
        v_g = ps.v[:n].cpu().numpy()
        assert rel_inf(v_g[dyn], sim.v[dyn]) <= 1e-5, step
        assert rel_inf(x_g[dyn], sim.x[dyn]) <= 1e-5, step
        assert np.array_equal(x_g[~dyn], sim.x[~dyn])
        assert abs(fw.time - sim.time) < 1e-12
    print("particle counts per step:", counts, flush=True)
    assert counts[0] < counts[-1] == data["particle_max_num"]
    # reset drops the spawned bodies and re-arms the spawns
    fw.reset()
    assert ps.particle_num == data["particles"].n and fw.time == 0.0
    assert not any(s.fired for s in fw.scheduled_spawn_system.spawns)
