"""Tile windows: the pair passes read each 64-row tile's distinct neighbours from on-chip memory
instead of gathering them from the particle arrays.  They read the same values in the same order,
so every result must be bitwise the one of the same context with windows off: whole steps
(gravity, explicit or Weiler18 viscosity, Jeske23 surface tension, the IP or PRJ solve,
advection, the rebuild) on the 8k and 102k scenes, in the ids-only and float16 row formats, with
every tile windowed and with a window cap that leaves part of the tiles without one."""

import numpy as np
import pytest

from tests.helpers import gpu_context, ip_kwargs, oracle_at

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

STEPS = 3


def _np(t):
    return t.detach().cpu().numpy()


def _run(sim, cache, windows, solver):
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.context import IPParams, PRJParams
    ctx = gpu_context(sim, torch, cache)
    N.check(ctx.lib.ivp_set_option(ctx.h, N.OPT_TILE_WINDOWS, int(windows)), "ivp_set_option")
    assert ctx.neighbors_build() == 0
    tiles, windowed, staged = ctx.window_stats()
    n_dyn = int(np.count_nonzero(sim.is_dynamic == 1))
    if windows == 0:
        assert (tiles, windowed, staged) == (0, 0, 0)
    else:
        assert tiles == (n_dyn + 63) // 64
        # every tile fits the full cap; the 300 cap leaves a part of them without a window
        assert windowed == tiles if windows == 1 else 0 < windowed < tiles, (tiles, windowed)
        assert n_dyn <= staged <= 12 * n_dyn
    cfg = sim.cfg
    kw = ip_kwargs(cfg)
    visc = (1, 0.05) if solver in ("ip_w18", "st") else (0, 0.01)
    st = (1, 0.1) if solver == "st" else (0, 0.0)
    params = PRJParams(tol_opt=kw["tol_opt"], enable_df=True) if solver == "prj" else IPParams(**kw)
    log = []
    for _ in range(STEPS):
        res, vit = ctx.step(sim.dt, sim.g, visc[0], visc[1], params, st[0], st[1])
        log.append((res.status, res.outer_iters, tuple(res.pcg_iters), res.final_density_error,
                    res.df_outer_iters, vit, ctx.last_st_iters))
    n = sim.x.shape[0]
    out = {f: _np(ctx.read(f, torch.zeros((n, 3) if f < 2 else (n,), device="cuda"))) for f in range(5)}
    ctx.close()
    return log, out


@pytest.mark.parametrize("cfg_id,step", [(0, 10), (1, 1)], ids=["8k", "102k"])
@pytest.mark.parametrize("cache", [0, 2], ids=["ids_only", "gradW_f16"])
@pytest.mark.parametrize("cap", [1, 300], ids=["all_tiles", "cap300"])
@pytest.mark.parametrize("solver", ["ip_w18", "ip_explicit", "prj", "st"])
def test_windows_bitwise(cfg_id, step, cache, cap, solver):
    if cfg_id == 1 and (cache, cap) != (0, 1):
        pytest.skip("the 102k scene runs the default format with every tile windowed")
    sim = oracle_at(cfg_id, step)
    log_on, on = _run(sim, cache, cap, solver)
    log_off, off = _run(sim, cache, 0, solver)
    assert all(e[0] == 0 for e in log_on), log_on
    assert log_on == log_off
    for f in range(5):
        assert np.array_equal(on[f].view(np.uint32), off[f].view(np.uint32)), f
