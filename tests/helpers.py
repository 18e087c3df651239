"""Shared test helpers: synthetic scenes, oracle states and GPU contexts."""

import functools

import numpy as np

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.scenes import load_particles, particle_masses, synthetic_scene
from oracle import oracle as O


def scene(cfg_id, scale=None):
    cfg = SimConfig.from_dict(synthetic_scene(cfg_id, scale=scale))
    return cfg, load_particles(cfg)


def ip_kwargs(cfg):
    return dict(omega=float(cfg.get_cfg("omega", 10.0)), tol_opt=float(cfg.get_cfg("densityErrorTol", 1e-3)),
                max_iteration_opt=int(cfg.get_cfg("maxOuterIters", 1000)),
                use_fixed_iters=bool(cfg.get_cfg("fixedOuterIters", False)),
                tol_pcg=float(cfg.get_cfg("tol_pcg", 1e-6)))


@functools.lru_cache(maxsize=8)
def oracle_states(cfg_id, steps, scale=None):
    """Oracle trajectory: list of (x, v) at the start of each step 0..steps (after rebuild)."""
    cfg, ps = scene(cfg_id, scale)
    sim = O.OracleSim(cfg, ps)
    sim.rebuild()
    out = [(sim.x.copy(), sim.v.copy())]
    for _ in range(steps):
        sim.step()
        out.append((sim.x.copy(), sim.v.copy()))
    return out


def oracle_at(cfg_id, step, scale=None):
    """An OracleSim positioned at the start of `step` (rebuilt at that state)."""
    cfg, ps = scene(cfg_id, scale)
    sim = O.OracleSim(cfg, ps)
    x, v = oracle_states(cfg_id, step, scale)[step]
    sim.x, sim.v = x.copy(), v.copy()
    sim.rebuild()
    return sim


def gpu_context(sim, torch, gradient_cache=0):
    """A libivp Context loaded with the oracle simulation's particle state.  gradient_cache
    selects the neighbour-row format: 0 ids only, 1 cached float32 gradW, 2 cached float16 gradW."""
    from implicitvelocityprojection_amd.context import Context
    cfg = sim.cfg
    dev = torch.device("cuda")
    ctx = Context(sim.h, cfg.domain_start, cfg.domain_size, max_particles=sim.x.shape[0], cache_size=sim.C)
    ctx.gradient_cache = gradient_cache
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    ctx.set_particles(t(sim.x), t(sim.v), t(sim.m), t(sim.m_inv), t(sim.density0), t(sim.is_dynamic, torch.int32),
                      t(sim.material, torch.int32))
    return ctx


def rel_inf(a, b):
    """max |a - b| / max |b| (0 when both vanish)."""
    den = float(np.max(np.abs(b))) if b.size else 0.0
    num = float(np.max(np.abs(a - b))) if a.size else 0.0
    if den == 0.0:
        return num
    return num / den
