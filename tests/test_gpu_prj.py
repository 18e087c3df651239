"""GPU parity of the PRJ solver (prj.py: projected Jacobi, optional divergence-free pass)
against the CPU oracle on the same teacher-forced inputs, in every neighbour-row format.

Bars: pass counts equal; density error per pass within 1e-4 relative (+1e-9 absolute);
velocities within 1e-5 relative (max-norm); divergence within 1e-4; warm-start multipliers
follow the particles through a re-sort.
"""

import numpy as np
import pytest

from tests.helpers import gpu_context, oracle_at, rel_inf
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

V_TOL = 1e-5
CACHE_ALL = pytest.mark.parametrize("cache", [0, 1, 2], ids=["ids_only", "gradW_f32", "gradW_f16"])


def _np(t):
    return t.detach().cpu().numpy()


def _oracle(sim, v_in, **kw):
    return O.prj_solve(v_in, sim.m, sim.density, sim.density0, sim.is_dynamic, sim.C, sim.count, sim.ids,
                       sim.gradW, sim.k, sim.dt, **kw)


def _errors_close(a, b):
    assert len(a) == len(b), (a, b)
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-4 * abs(y) + 1e-9, (a, b)


@CACHE_ALL
@pytest.mark.parametrize("step,tol", [(0, 1e-3), (5, 1e-3), (20, 1e-5)])
def test_prj_solve_matches_oracle(step, tol, cache):
    from implicitvelocityprojection_amd.context import PRJParams
    sim = oracle_at(0, step)
    v_in = O.gravity(sim.v, sim.is_dynamic, sim.g, sim.dt)
    v_o, div_o, res_o, lam_o, _ = _oracle(sim, v_in, tol_opt=tol)
    assert res_o.status == O.ORC_OK
    print(f"prj step {step} tol {tol} cache {cache}: oracle passes {res_o.outer_iters}", flush=True)
    ctx = gpu_context(sim, torch, cache)
    assert ctx.neighbors_build() == 0
    n = sim.x.shape[0]
    ctx.write(1, torch.from_numpy(v_in).cuda())
    res_g = ctx.prj_solve(sim.dt, PRJParams(tol_opt=tol))
    v_g = _np(ctx.read(1, torch.zeros((n, 3), device="cuda")))
    div_g = _np(ctx.read(4, torch.zeros(n, device="cuda")))
    dyn = sim.is_dynamic == 1
    assert res_g.status == 0
    assert res_g.outer_iters == res_o.outer_iters
    _errors_close(res_g.density_errors, res_o.density_errors)
    assert res_g.final_density_error < tol
    assert res_g.pcg_iters == []
    # stats_iter_costs (prj.py:189): one cost per pass
    assert len(res_g.iter_costs) == len(res_g.density_errors) and all(c > 0.0 for c in res_g.iter_costs)
    assert rel_inf(v_g[dyn], v_o[dyn]) <= V_TOL, rel_inf(v_g[dyn], v_o[dyn])
    assert np.array_equal(v_g[~dyn], v_in[~dyn])
    assert rel_inf(div_g, div_o) <= 1e-4
    ctx.close()


@CACHE_ALL
def test_prj_divergence_free_pass_matches_oracle(cache):
    from implicitvelocityprojection_amd.context import PRJParams
    sim = oracle_at(0, 5)
    v_in = O.gravity(sim.v, sim.is_dynamic, sim.g, sim.dt)
    kw = dict(enable_df=True, df_num_threshold=25, df_tol_opt=1e-3)
    v_o, div_o, res_o, _, ldf_o = _oracle(sim, v_in, **kw)
    assert res_o.status == O.ORC_OK and res_o.df_outer_iters >= 1
    ctx = gpu_context(sim, torch, cache)
    assert ctx.neighbors_build() == 0
    n = sim.x.shape[0]
    ctx.write(1, torch.from_numpy(v_in).cuda())
    # PRJ.solve then PRJ.solve_df, as Framework.forward calls them (framework.py:481-487)
    res_g = ctx.prj_solve(sim.dt, PRJParams(**kw))
    assert res_g.status == 0 and res_g.outer_iters == res_o.outer_iters and res_g.df_outer_iters == 0
    res_df = ctx.prj_solve_df(sim.dt, PRJParams(**kw))
    assert res_df.status == 0 and res_df.df_outer_iters == res_o.df_outer_iters
    v_g = _np(ctx.read(1, torch.zeros((n, 3), device="cuda")))
    div_g = _np(ctx.read(4, torch.zeros(n, device="cuda")))
    dyn = sim.is_dynamic == 1
    assert rel_inf(v_g[dyn], v_o[dyn]) <= V_TOL, rel_inf(v_g[dyn], v_o[dyn])
    assert rel_inf(div_g, div_o) <= 1e-4
    ctx.close()


def test_prj_warm_start_follows_particles_through_resort():
    """Two solves with warm_start: the second starts from half the first's multipliers, which
    the library keeps per particle (user order) across the neighbour rebuild that re-sorts its
    store.  Oracle: the same two solves with lam_prev handed over in user order."""
    from implicitvelocityprojection_amd.context import PRJParams
    sim_a = oracle_at(0, 3)
    sim_b = oracle_at(0, 4)
    va = O.gravity(sim_a.v, sim_a.is_dynamic, sim_a.g, sim_a.dt)
    vb = O.gravity(sim_b.v, sim_b.is_dynamic, sim_b.g, sim_b.dt)
    kw = dict(tol_opt=1e-5, warm_start=True)
    _, _, ra, lam_a, _ = _oracle(sim_a, va, **kw)
    vb_o, _, rb, _, _ = _oracle(sim_b, vb, lam_prev=lam_a, **kw)
    _, _, rb_cold, _, _ = _oracle(sim_b, vb, tol_opt=1e-5)
    assert rb.density_errors[0] != rb_cold.density_errors[0]
    ctx = gpu_context(sim_a, torch, 0)
    n = sim_a.x.shape[0]
    assert ctx.neighbors_build() == 0
    ctx.write(1, torch.from_numpy(va).cuda())
    ga = ctx.prj_solve(sim_a.dt, PRJParams(**kw))
    assert ga.outer_iters == ra.outer_iters
    dyn = sim_b.is_dynamic == 1
    xb = torch.from_numpy(sim_b.x).cuda()
    ctx.write(0, xb)
    assert ctx.neighbors_build() == 0
    ctx.write(1, torch.from_numpy(vb).cuda())
    gb = ctx.prj_solve(sim_b.dt, PRJParams(**kw))
    v_g = _np(ctx.read(1, torch.zeros((n, 3), device="cuda")))
    assert gb.outer_iters == rb.outer_iters
    _errors_close(gb.density_errors, rb.density_errors)
    assert rel_inf(v_g[dyn], vb_o[dyn]) <= V_TOL
    # ivp_reset zeroes the kept multipliers: a warm start is then a cold start
    ctx.reset()
    ctx.write(0, xb)
    assert ctx.neighbors_build() == 0
    ctx.write(1, torch.from_numpy(vb).cuda())
    gc = ctx.prj_solve(sim_b.dt, PRJParams(**kw))
    _errors_close(gc.density_errors, rb_cold.density_errors)
    ctx.close()


def test_prj_stop_rules_match_oracle():
    from implicitvelocityprojection_amd.context import PRJParams
    from implicitvelocityprojection_amd import _native as N
    sim = oracle_at(0, 0)
    v_in = O.gravity(sim.v, sim.is_dynamic, sim.g, sim.dt)
    ctx = gpu_context(sim, torch, 0)
    assert ctx.neighbors_build() == 0
    n = sim.x.shape[0]
    vt = torch.from_numpy(v_in).cuda()
    # no update allowed: IVP_NOT_CONVERGED, v untouched
    ctx.write(1, vt)
    res = ctx.prj_solve(sim.dt, PRJParams(max_iteration_opt=0))
    _, _, ro, _, _ = _oracle(sim, v_in, max_iteration_opt=0)
    assert ro.status == O.ORC_NOT_CONVERGED and res.status == N.IVP_NOT_CONVERGED
    assert res.outer_iters == 0 and len(res.density_errors) == 1
    assert np.array_equal(_np(ctx.read(1, torch.zeros((n, 3), device="cuda"))), v_in)
    # fixed iteration count: exactly that many updates
    ctx.write(1, vt)
    res = ctx.prj_solve(sim.dt, PRJParams(max_iteration_opt=7, use_fixed_iters=True))
    v_o, _, ro, _, _ = _oracle(sim, v_in, max_iteration_opt=7, use_fixed_iters=True)
    assert res.status == 0 and res.outer_iters == 7 == ro.outer_iters
    _errors_close(res.density_errors, ro.density_errors)
    dyn = sim.is_dynamic == 1
    assert rel_inf(_np(ctx.read(1, torch.zeros((n, 3), device="cuda")))[dyn], v_o[dyn]) <= V_TOL
    with pytest.raises(N.IVPError):
        ctx.prj_solve(sim.dt, PRJParams(max_iteration_opt=-1))
    ctx.close()


@pytest.mark.parametrize("df", [False, True], ids=["prj", "prj_df"])
def test_prj_full_step_matches_oracle(df):
    from implicitvelocityprojection_amd.context import PRJParams
    sim = oracle_at(0, 8)
    sim.pressure_method = 0
    sim.prj_kw.update(enable_df=df)
    ctx = gpu_context(sim, torch, 0)
    assert ctx.neighbors_build() == 0
    kw = dict(sim.prj_kw)
    res_g, _ = ctx.step(sim.dt, sim.g, sim.cfg.viscosity_method, sim.cfg.viscosity_coeff, PRJParams(**kw))
    sim.step()
    n = sim.x.shape[0]
    x_g = _np(ctx.read(0, torch.zeros((n, 3), device="cuda")))
    v_g = _np(ctx.read(1, torch.zeros((n, 3), device="cuda")))
    dyn = sim.is_dynamic == 1
    assert res_g.outer_iters == sim.last_prj.outer_iters
    assert res_g.df_outer_iters == (sim.last_prj.df_outer_iters if df else 0)
    assert rel_inf(v_g[dyn], sim.v[dyn]) <= V_TOL
    assert rel_inf(x_g[dyn], sim.x[dyn]) <= V_TOL
    assert np.array_equal(x_g[~dyn], sim.x[~dyn])
    ctx.close()
