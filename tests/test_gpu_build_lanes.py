"""The neighbour build's candidate scan with 1, 2, 4, 8 or 16 lanes per row, or by cells (64;
IVP_OPT_BUILD_LANES) fills the rows identically: neighbour caches, density and k bitwise equal
to the oracle's in every setting, on the 8k and 102k columns and on a ragged count."""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import gpu_context, oracle_at, scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("cfg_id,step", [(0, 6), (1, 0)])
@pytest.mark.parametrize("cache", [0, 1, 2], ids=["ids_only", "gradW_f32", "gradW_f16"])
def test_build_lanes_give_identical_rows(cfg_id, step, cache):
    from implicitvelocityprojection_amd import _native as N
    sim = oracle_at(cfg_id, step)
    n, C = sim.x.shape[0], sim.C
    ctx = gpu_context(sim, torch, cache)
    mask = np.zeros(n * C, bool)
    for i in range(n):
        mask[i * C:i * C + sim.count[i]] = True
    for lanes in (1, 2, 4, 8, 16, 64, 0):
        print(f"build lanes {lanes}: config {cfg_id} step {step} cache {cache}", flush=True)
        assert N.check(ctx.lib.ivp_set_option(ctx.h, N.OPT_BUILD_LANES, lanes)) == 0
        assert ctx.build_lanes == lanes
        assert ctx.neighbors_build() == 0
        count = torch.zeros(n, dtype=torch.int32, device="cuda")
        ids = torch.zeros(n * C, dtype=torch.int32, device="cuda")
        gradW = torch.zeros(n * C * 3, dtype=torch.float32, device="cuda")
        ctx.neighbors_export(count, ids, gradW)
        dens = _np(ctx.read(2, torch.zeros(n, device="cuda")))
        k = _np(ctx.read(3, torch.zeros(n, device="cuda")))
        assert np.array_equal(_np(count), sim.count)
        assert np.array_equal(_np(ids)[mask], sim.ids[mask])
        assert np.array_equal(dens.view(np.uint32), sim.density.view(np.uint32))
        assert np.array_equal(k.view(np.uint32), sim.k.view(np.uint32))
        assert ctx.pair_count() > 0
    with pytest.raises(N.IVPError):
        N.check(ctx.lib.ivp_set_option(ctx.h, N.OPT_BUILD_LANES, 3), "ivp_set_option")
    ctx.close()


def test_build_lanes_overflow_reported():
    """A row with more neighbours than the capacity is reported in every setting."""
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.context import Context
    cfg, ps = scene(0)
    sim = O.OracleSim(cfg, ps)
    n = sim.x.shape[0]
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dt)
    ctx = Context(sim.h, cfg.domain_start, cfg.domain_size, n, cache_size=20)
    ctx.set_particles(t(sim.x), t(sim.v), t(sim.m), t(sim.m_inv), t(sim.density0), t(sim.is_dynamic, torch.int32),
                      t(sim.material, torch.int32))
    for lanes in (1, 2, 8, 64):
        N.check(ctx.lib.ivp_set_option(ctx.h, N.OPT_BUILD_LANES, lanes))
        assert ctx.neighbors_build() == N.IVP_NEIGHBOR_OVERFLOW
    ctx.close()
