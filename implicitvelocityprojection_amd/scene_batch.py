"""Many independent scenes stepped in one libivp context (ivp_set_scenes / ivp_step_scenes).

No counterpart in the reference, which runs one ``Framework`` per scene: this is the host side
of the batched form.  Every scene keeps its own coordinates, cell grid, IP outer loop, CG and
boundary clamp; each pass of a step runs once over all scenes.  The scenes must share the
support radius, neighbour cache size, time step, gravity and IP parameters, and run without
viscosity.
"""

import time

import numpy as np
import torch

from .config import SimConfig
from .context import Context, IPParams
from .scenes import load_particles, particle_masses


def _ip_params(cfg):
    return IPParams(omega=float(cfg.get_cfg("omega", 10.0)), tol_opt=float(cfg.get_cfg("densityErrorTol", 1e-3)),
                    max_iteration_opt=int(cfg.get_cfg("maxOuterIters", 1000)),
                    use_fixed_iters=bool(cfg.get_cfg("fixedOuterIters", False)),
                    tol_pcg=float(cfg.get_cfg("tol_pcg", 1e-6)))


class SceneBatch:
    """Scene JSON dicts -> one context holding all of them, scene-major in user order."""

    def __init__(self, scene_dicts, device="cuda", graph=True, chunk=4, lanes=1):
        self.cfgs = [SimConfig.from_dict(d) for d in scene_dicts]
        c0 = self.cfgs[0]
        for c in self.cfgs[1:]:
            same = (np.float32(c.support_radius) == np.float32(c0.support_radius) and c.dt == c0.dt
                    and list(c.gravity) == list(c0.gravity)
                    and int(c.get_cfg("neighborCacheSize", 60)) == int(c0.get_cfg("neighborCacheSize", 60)))
            if not same:
                raise ValueError("scenes of one batch must share support radius, dt, gravity and cache size")
            if c.viscosity_coeff != 0.0:
                raise ValueError("the batched step runs without viscosity")
        if c0.viscosity_coeff != 0.0:
            raise ValueError("the batched step runs without viscosity")
        self.particles = [load_particles(c) for c in self.cfgs]
        self.sizes = [p.n for p in self.particles]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        masses = [particle_masses(c, p) for c, p in zip(self.cfgs, self.particles)]
        cat = lambda xs: np.ascontiguousarray(np.concatenate(xs))
        self.x0 = cat([p.x for p in self.particles])
        self.v0 = cat([p.v for p in self.particles])
        self.m = cat([m for m, _ in masses])
        self.m_inv = cat([mi for _, mi in masses])
        self.density0 = cat([p.density0 for p in self.particles])
        self.is_dynamic = cat([p.is_dynamic for p in self.particles])
        self.material = cat([p.material for p in self.particles])
        self.n_dynamic = int(np.count_nonzero(self.is_dynamic == 1))
        self.dt = float(c0.dt)
        self.gravity = [float(g) for g in c0.gravity]
        self.params = _ip_params(c0)
        self.device = torch.device(device)
        h = float(np.float32(c0.support_radius))
        self.ctx = Context(h, c0.domain_start, c0.domain_size, self.n, cache_size=int(c0.get_cfg("neighborCacheSize", 60)))
        self.ctx.set_scenes(self.sizes, [c.domain_start for c in self.cfgs], [c.domain_size for c in self.cfgs])
        self.ctx.scene_graph = graph
        self.ctx.scene_cg_chunk = chunk
        self.ctx.scene_lanes = lanes
        t = self._t
        self.ctx.set_particles(t(self.x0), t(self.v0), t(self.m), t(self.m_inv), t(self.density0),
                               t(self.is_dynamic, torch.int32), t(self.material, torch.int32))
        if self.ctx.neighbors_build() != 0:
            raise RuntimeError("neighbour overflow at load")

    def _t(self, a, dt=torch.float32):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device, dt)

    def step(self):
        """One step of every scene; returns (status, per-scene stats tuples)."""
        return self.ctx.step_scenes(self.dt, self.gravity, self.params)

    def read(self, field, shape):
        return self.ctx.read(field, torch.zeros(shape, device=self.device)).cpu().numpy()

    def scene_slice(self, s):
        return slice(int(self.offsets[s]), int(self.offsets[s + 1]))

    def snapshot(self):
        return self.read(0, (self.n, 3)), self.read(1, (self.n, 3))

    def restore(self, snap):
        x, v = snap
        self.ctx.write(0, self._t(x))
        self.ctx.write(1, self._t(v))
        if self.ctx.neighbors_build() != 0:
            raise RuntimeError("neighbour overflow at restore")

    def autotune(self, chunks=(1, 2, 4, 8), lanes=(1, 2, 4), steps=2, build_forms=(2, 4, 64), builds=3):
        """Choose the neighbour build's candidate-scan form (rows identical in every form), then the
        CG iterations per replay and the lanes per neighbour row of the pair passes: time `builds`
        builds per form and `steps` steps per combination from the same saved state, keep the
        fastest, restore the state.  Returns {"build lanes L": ms per build, "chunk K, lanes G": ms
        per step}."""
        snap = self.snapshot()
        times = {}
        build_ms = {}
        for f in build_forms:
            self.ctx.build_lanes = f
            self.restore(snap)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(builds):
                if self.ctx.neighbors_build() != 0:
                    raise RuntimeError("neighbour overflow in the build autotune")
            torch.cuda.synchronize()
            build_ms[f] = (time.perf_counter() - t0) * 1e3 / builds
        if build_ms:
            self.ctx.build_lanes = min(build_ms, key=build_ms.get)
        for g in lanes:
            for k in chunks:
                self.restore(snap)
                self.ctx.scene_cg_chunk = k
                self.ctx.scene_lanes = g
                self.step()                      # captures the replay graphs for this setting
                self.restore(snap)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    self.step()
                torch.cuda.synchronize()
                times[(k, g)] = (time.perf_counter() - t0) * 1e3 / steps
        best = min(times, key=times.get)
        self.ctx.scene_cg_chunk, self.ctx.scene_lanes = best
        self.restore(snap)
        out = {f"build lanes {f}": ms for f, ms in build_ms.items()}
        out.update({f"chunk {k}, lanes {g}": ms for (k, g), ms in times.items()})
        return out

    def close(self):
        self.ctx.close()
