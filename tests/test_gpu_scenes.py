"""Independent scenes in one context (ivp_set_scenes / ivp_step_scenes) against the same scenes
run one per context, and against the oracle.

A scene in the batch has its own grid with the single-scene grid's origin, dims and key order,
so its neighbour rows, density and k are bit-identical to a context holding only that scene.
The solve's reductions group rows into blocks differently (a block can span two scenes), so the
FP64 sums, and from them alpha and beta, can differ in the last bits: bars for the solve are the
single-scene parity bars (velocities and positions within 1e-5 relative, CG iterations within
+-1, same outer iterations, final density error at or below densityErrorTol).  The hipGraph
replay and the launch-by-launch form of the batched solve are bitwise equal over several steps.
"""

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.scenes import load_particles, particle_masses, synthetic_batch
from oracle import oracle as O
from tests.helpers import ip_kwargs, rel_inf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

V_TOL = 1e-5
STEPS = 3


def _scenes():
    return synthetic_batch(n_scenes=5, seed=4, count_range=(2048, 6000))


def _single(scene):
    from implicitvelocityprojection_amd.context import Context
    cfg = SimConfig.from_dict(scene)
    ps = load_particles(cfg)
    m, m_inv = particle_masses(cfg, ps)
    dev = torch.device("cuda")
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    ctx = Context(float(np.float32(cfg.support_radius)), cfg.domain_start, cfg.domain_size, ps.n,
                  cache_size=int(cfg.get_cfg("neighborCacheSize", 60)))
    ctx.set_particles(t(ps.x), t(ps.v), t(m), t(m_inv), t(ps.density0), t(ps.is_dynamic, torch.int32),
                      t(ps.material, torch.int32))
    assert ctx.neighbors_build() == 0
    return cfg, ps, ctx


def _np(t):
    return t.detach().cpu().numpy()


def test_scene_build_bit_exact_against_single_contexts():
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()
    batch = SceneBatch(scenes)
    dens_b = batch.read(2, (batch.n,))
    k_b = batch.read(3, (batch.n,))
    for s, scene in enumerate(scenes):
        _, ps, ctx = _single(scene)
        sl = batch.scene_slice(s)
        dens = _np(ctx.read(2, torch.zeros(ps.n, device="cuda")))
        k = _np(ctx.read(3, torch.zeros(ps.n, device="cuda")))
        assert np.array_equal(dens_b[sl].view(np.uint32), dens.view(np.uint32)), s
        assert np.array_equal(k_b[sl].view(np.uint32), k.view(np.uint32)), s
        ctx.close()
    batch.close()


@pytest.mark.parametrize("lanes", [1, 2, 64], ids=["one_lane", "two_lanes", "by_cells"])
def test_scene_build_forms_give_identical_rows(lanes):
    """Every candidate-scan form builds the batch's rows (each scene in its own grid) exactly as
    the default does: counts, ids, density and k bitwise."""
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    batch = SceneBatch(_scenes())
    C = 60
    out = {}
    for form in (0, lanes):
        assert N.check(batch.ctx.lib.ivp_set_option(batch.ctx.h, N.OPT_BUILD_LANES, form)) == 0
        assert batch.ctx.neighbors_build() == 0
        cnt = torch.zeros(batch.n, dtype=torch.int32, device="cuda")
        ids = torch.zeros(batch.n * C, dtype=torch.int32, device="cuda")
        g = torch.zeros(batch.n * C * 3, dtype=torch.float32, device="cuda")
        batch.ctx.neighbors_export(cnt, ids, g)
        out[form] = (_np(cnt), _np(ids).reshape(batch.n, C), batch.read(2, (batch.n,)), batch.read(3, (batch.n,)))
    c0, i0, d0, k0 = out[0]
    c1, i1, d1, k1 = out[lanes]
    assert np.array_equal(c0, c1)
    live = np.arange(C)[None, :] < c0[:, None]
    assert np.array_equal(i0[live], i1[live])
    assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
    assert np.array_equal(k0.view(np.uint32), k1.view(np.uint32))
    batch.close()


def test_scene_neighbour_export_matches_single_contexts():
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()[:3]
    batch = SceneBatch(scenes)
    C = 60
    cnt_b = torch.zeros(batch.n, dtype=torch.int32, device="cuda")
    ids_b = torch.zeros(batch.n * C, dtype=torch.int32, device="cuda")
    g_b = torch.zeros(batch.n * C * 3, dtype=torch.float32, device="cuda")
    batch.ctx.neighbors_export(cnt_b, ids_b, g_b)
    cnt_b, ids_b = _np(cnt_b), _np(ids_b).reshape(batch.n, C)
    for s, scene in enumerate(scenes):
        _, ps, ctx = _single(scene)
        cnt = torch.zeros(ps.n, dtype=torch.int32, device="cuda")
        ids = torch.zeros(ps.n * C, dtype=torch.int32, device="cuda")
        g = torch.zeros(ps.n * C * 3, dtype=torch.float32, device="cuda")
        ctx.neighbors_export(cnt, ids, g)
        cnt, ids = _np(cnt), _np(ids).reshape(ps.n, C)
        sl = batch.scene_slice(s)
        off = int(batch.offsets[s])
        assert np.array_equal(cnt_b[sl], cnt)
        live = np.arange(C)[None, :] < cnt[:, None]
        assert np.array_equal(ids_b[sl][live] - off, ids[live])
        ctx.close()
    batch.close()


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_scene_steps_match_single_contexts(lanes):
    from implicitvelocityprojection_amd.context import IPParams
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()
    batch = SceneBatch(scenes, chunk=2, lanes=lanes)
    hist_b = []
    for _ in range(STEPS):
        status, stats = batch.step()
        assert status == 0
        hist_b.append([batch.ctx.scene_history(s) for s in range(len(scenes))] + [stats])
    x_b = batch.read(0, (batch.n, 3))
    v_b = batch.read(1, (batch.n, 3))
    for s, scene in enumerate(scenes):
        cfg, ps, ctx = _single(scene)
        kw = ip_kwargs(cfg)
        dyn = ps.is_dynamic == 1
        for step in range(STEPS):
            res, _ = ctx.step(cfg.dt, cfg.gravity, 0, 0.0, IPParams(**kw))
            errs_b, pcg_b = hist_b[step][s]
            st_b = hist_b[step][-1][s]
            assert st_b[0] == 0
            assert st_b[1] == res.outer_iters, (s, step)
            assert len(pcg_b) == len(res.pcg_iters)
            for a, b in zip(pcg_b, res.pcg_iters):
                assert abs(a - b) <= 1, (s, step, pcg_b, res.pcg_iters)
            assert st_b[3] <= kw["tol_opt"]
            assert len(errs_b) == len(res.density_errors)
        x = _np(ctx.read(0, torch.zeros((ps.n, 3), device="cuda")))
        v = _np(ctx.read(1, torch.zeros((ps.n, 3), device="cuda")))
        sl = batch.scene_slice(s)
        assert rel_inf(x_b[sl][dyn], x[dyn]) <= V_TOL, s
        assert rel_inf(v_b[sl][dyn], v[dyn]) <= V_TOL, s
        assert np.array_equal(x_b[sl][~dyn], x[~dyn])
        ctx.close()
    batch.close()


def test_scene_step_matches_oracle():
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()[:2]
    batch = SceneBatch(scenes)
    status, stats = batch.step()
    assert status == 0
    x_b = batch.read(0, (batch.n, 3))
    v_b = batch.read(1, (batch.n, 3))
    for s, scene in enumerate(scenes):
        cfg = SimConfig.from_dict(scene)
        sim = O.OracleSim(cfg, load_particles(cfg))
        sim.rebuild()
        sim.step()
        dyn = sim.is_dynamic == 1
        sl = batch.scene_slice(s)
        assert stats[s][1] == sim.last_ip.outer_iters
        assert rel_inf(v_b[sl][dyn], sim.v[dyn]) <= V_TOL
        assert rel_inf(x_b[sl][dyn], sim.x[dyn]) <= V_TOL
    batch.close()


@pytest.mark.parametrize("lanes", [1, 4])
def test_scene_graph_replays_equal_direct_launches(lanes):
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()[:3]
    outs = []
    for graph in (True, False):
        batch = SceneBatch(scenes, graph=graph, chunk=2, lanes=lanes)
        for _ in range(STEPS):
            status, _ = batch.step()
            assert status == 0
            torch.cuda.synchronize()
        outs.append((batch.read(0, (batch.n, 3)), batch.read(1, (batch.n, 3))))
        batch.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


def test_scene_entry_points_of_one_scene_are_refused():
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.context import IPParams
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    batch = SceneBatch(_scenes()[:2])
    with pytest.raises(N.IVPError):
        batch.ctx.ip_solve(1e-3, IPParams())
    with pytest.raises(N.IVPError):
        batch.ctx.step(1e-3, [0.0, -9.81, 0.0], 0, 0.0, IPParams())
    with pytest.raises(N.IVPError):
        batch.ctx.gradient_cache = 1
    batch.close()


def test_scene_reload_recaptures_graphs():
    """Loading the particle set again frees and reallocates the neighbour rows; the replay
    graphs captured before must not be reused (they hold the old rows' addresses)."""
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = _scenes()[:2]
    batch = SceneBatch(scenes, chunk=2)
    for _ in range(2):
        assert batch.step()[0] == 0
    ref = SceneBatch(scenes, chunk=2)
    t = batch._t
    batch.ctx.set_particles(t(batch.x0), t(batch.v0), t(batch.m), t(batch.m_inv), t(batch.density0),
                            t(batch.is_dynamic, torch.int32), t(batch.material, torch.int32))
    assert batch.ctx.neighbors_build() == 0
    for _ in range(2):
        assert batch.step()[0] == 0
        assert ref.step()[0] == 0
    a = batch.read(1, (batch.n, 3))
    b = ref.read(1, (ref.n, 3))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    batch.close()
    ref.close()


def test_scene_batch_256_full_size():
    """The batched workload at its full size: 256 independent columns of 2 048-32 768 fluid
    particles (seed 4, 4.58M fluid particles), autotuned chunk and lanes.  Three steps, every
    scene converged (status 0, final density error at or below densityErrorTol), static particles
    untouched, all values finite; four sampled scenes (the smallest, the largest, two more) equal
    their one-scene contexts at the single-scene bars after the same three steps."""
    from implicitvelocityprojection_amd.context import IPParams
    from implicitvelocityprojection_amd.scene_batch import SceneBatch
    scenes = synthetic_batch(n_scenes=256, seed=4)
    batch = SceneBatch(scenes)
    batch.autotune(chunks=(2, 4), lanes=(1, 4), steps=1)
    x0 = batch.read(0, (batch.n, 3))
    for _ in range(STEPS):
        status, stats = batch.step()
        assert status == 0
        assert len(stats) == 256
        assert all(st[0] == 0 for st in stats)
        assert max(st[3] for st in stats) <= 1e-3
    x_b = batch.read(0, (batch.n, 3))
    v_b = batch.read(1, (batch.n, 3))
    assert np.all(np.isfinite(x_b)) and np.all(np.isfinite(v_b))
    stat = batch.is_dynamic == 0
    assert np.array_equal(x_b[stat].view(np.uint32), x0[stat].view(np.uint32))
    assert np.all(v_b[stat] == 0.0)
    sizes = [int(np.count_nonzero(load_particles(SimConfig.from_dict(sc)).is_dynamic == 1)) for sc in scenes]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 200})
    for s in picks:
        cfg, ps, ctx = _single(scenes[s])
        kw = ip_kwargs(cfg)
        for _ in range(STEPS):
            res, _ = ctx.step(cfg.dt, cfg.gravity, 0, 0.0, IPParams(**kw))
            assert res.status == 0
        x = _np(ctx.read(0, torch.zeros((ps.n, 3), device="cuda")))
        v = _np(ctx.read(1, torch.zeros((ps.n, 3), device="cuda")))
        dyn = ps.is_dynamic == 1
        sl = batch.scene_slice(s)
        assert rel_inf(x_b[sl][dyn], x[dyn]) <= V_TOL, s
        assert rel_inf(v_b[sl][dyn], v[dyn]) <= V_TOL, s
        ctx.close()
    batch.close()
