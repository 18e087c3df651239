"""The reference-shaped Python interface (SimConfig -> SceneLoader -> SimulationData ->
NeighbourSearch -> Framework / IP) driven like the reference's main.py, checked against the CPU
oracle on the same scene."""

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.scenes import synthetic_scene
from oracle import oracle as O
from tests.helpers import rel_inf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _setup(scene):
    from implicitvelocityprojection_amd.framework import Framework
    from implicitvelocityprojection_amd.neighbour_search import NeighbourSearch
    from implicitvelocityprojection_amd.scene_loader import SceneLoader
    from implicitvelocityprojection_amd.simulation_data import SimulationData
    cfg = SimConfig.from_dict(scene)
    loader = SceneLoader(cfg)
    sd = loader.prepare_scene()
    ps = SimulationData(cfg)
    loader.populate_scene(ps, sd)
    ns = NeighbourSearch(cfg.domain_size, cfg.support_radius)
    fw = Framework(ps, ns, cfg)
    fw.initialize()
    return cfg, ps, ns, fw, sd["particles"]


def test_framework_forward_matches_oracle_and_exposes_reference_attributes():
    cfg, ps, ns, fw, particles = _setup(synthetic_scene(0))
    sim = O.OracleSim(cfg, particles)
    sim.rebuild()
    n = ps.particle_num
    np.testing.assert_array_equal(ps.density[:n].cpu().numpy().view(np.uint32), sim.density.view(np.uint32))
    np.testing.assert_array_equal(fw.pressure.k[:n].cpu().numpy().view(np.uint32), sim.k.view(np.uint32))
    for _ in range(5):
        # teacher forcing: start both from the oracle's state
        ps.x[:n] = torch.from_numpy(sim.x).cuda()
        ps.v[:n] = torch.from_numpy(sim.v).cuda()
        fw.initialize()
        fw.forward()
        sim.step()
        dyn = sim.is_dynamic == 1
        assert rel_inf(ps.v[:n].cpu().numpy()[dyn], sim.v[dyn]) <= 1e-5
        assert fw.pressure.stats_outer_iters == sim.last_ip.outer_iters
        assert fw.pressure.stats_final_density_error <= fw.pressure.tol_opt
        assert len(fw.pressure.stats_density_errors) == len(sim.last_ip.density_errors)
        # stats_iter_costs (ip.py:369, 392): one cost per outer-loop evaluation, within the solve
        costs = fw.pressure.stats_iter_costs
        assert len(costs) == len(fw.pressure.stats_density_errors)
        assert all(c > 0.0 for c in costs) and sum(costs) <= fw.pressure.stats_total_cost_ms
    # reference-layout cache on demand
    cnt = ps.neighbor_count[:n].cpu().numpy()
    np.testing.assert_array_equal(cnt, sim.count)


def test_framework_precision_form_from_scene_key():
    """"gradientCache": 2 in the scene's Configuration selects the float16 precision form; the
    step then meets the float16 bars against the oracle (velocities within 5e-4)."""
    scene = synthetic_scene(0)
    scene["Configuration"]["gradientCache"] = 2
    cfg, ps, ns, fw, particles = _setup(scene)
    assert ps.ctx.gradient_cache == 2
    sim = O.OracleSim(cfg, particles)
    sim.rebuild()
    n = ps.particle_num
    for _ in range(3):
        ps.x[:n] = torch.from_numpy(sim.x).cuda()     # teacher forcing, as above
        ps.v[:n] = torch.from_numpy(sim.v).cuda()
        fw.initialize()
        fw.forward()
        sim.step()
        dyn = sim.is_dynamic == 1
        assert rel_inf(ps.v[:n].cpu().numpy()[dyn], sim.v[dyn]) <= 5e-4
        assert fw.pressure.stats_final_density_error <= fw.pressure.tol_opt


def test_ip_solve_dropin_and_exceptions():
    from implicitvelocityprojection_amd.pressure import IP
    cfg, ps, ns, fw, _ = _setup(synthetic_scene(0))
    ip = fw.pressure
    assert isinstance(ip, IP) and ip.omega == 10.0 and ip.max_iteration_pcg == 1000
    ip.solve(cfg.dt)
    assert ip.stats_outer_iters >= 1 and len(ip.stats_pcg_iters) == ip.stats_outer_iters
    # unreachable tolerance with one update allowed: the reference's RuntimeError
    ip.tol_opt, ip.max_iteration_opt = 1e-12, 1
    v_before = ps.v.clone()
    with pytest.raises(RuntimeError, match="does not converge"):
        ip.solve(cfg.dt)
    assert torch.equal(ps.v, v_before)
    ip.tol_opt, ip.max_iteration_opt = 1e-3, 100
    # an indefinite operator: "non-positive definite matrix"
    ip.omega = -1e6
    with pytest.raises(RuntimeError, match="non-positive definite"):
        ip.solve(cfg.dt)


def test_framework_prj_matches_oracle():
    """pressureMethod 0 selects PRJ (framework.py:168-172) with the reference's knobs; forward()
    matches the oracle's PRJ step, and the divergence-free pass runs when enable_df is set."""
    from implicitvelocityprojection_amd.pressure import PRJ
    sc = synthetic_scene(0)
    sc["Configuration"]["pressureMethod"] = 0
    cfg, ps, ns, fw, particles = _setup(sc)
    assert isinstance(fw.pressure, PRJ)
    assert fw.pressure.omega == 0.5 and fw.pressure.enable_df is False and fw.pressure.warm_start is False
    assert fw.pressure.df_num_threshold == 25 and fw.pressure.df_tol_opt == 1e-2
    sim = O.OracleSim(cfg, particles)
    sim.rebuild()
    assert sim.pressure_method == 0
    n = ps.particle_num
    dyn = sim.is_dynamic == 1
    for step in range(4):
        if step == 3:
            fw.pressure.enable_df = True
            sim.prj_kw["enable_df"] = True
        ps.x[:n] = torch.from_numpy(sim.x).cuda()
        ps.v[:n] = torch.from_numpy(sim.v).cuda()
        fw.initialize()
        fw.forward()
        sim.step()
        assert rel_inf(ps.v[:n].cpu().numpy()[dyn], sim.v[dyn]) <= 1e-5
        assert fw.pressure.stats_outer_iters == sim.last_prj.outer_iters
        assert len(fw.pressure.stats_density_errors) == len(sim.last_prj.density_errors)
        assert fw.pressure.stats_final_density_error <= fw.pressure.tol_opt
    assert fw.pressure.stats_df_outer_iters == sim.last_prj.df_outer_iters >= 1


def test_unsupported_solvers_are_explicit():
    from implicitvelocityprojection_amd.framework import Framework
    sc = synthetic_scene(0, scale=(2, 2, 2))
    sc["Configuration"]["pressureMethod"] = 3
    with pytest.raises(NotImplementedError):
        _setup(sc)
    sc["Configuration"]["pressureMethod"] = 7
    with pytest.raises(ValueError):
        _setup(sc)
    sc["Configuration"]["pressureMethod"] = 1
    sc["Configuration"]["surfaceTensionMethod"] = 0
    sc["Configuration"]["surfaceTensionCoeff"] = 0.1
    with pytest.raises(NotImplementedError):
        _setup(sc)
    assert Framework is not None
