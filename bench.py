"""Benchmark: the reference's velocity-based IP step on the 1M-particle dam break, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--workload NAME]

`--workload` also accepts "column_102k" (40 x 64 x 40 fluid column, explicit viscosity 0.01,
dt = 1 ms) and "column_8k" (16 x 32 x 16 column, no viscosity, dt = 1 ms), the parity-test
scenes, for profiling; the default and the bench line's workload is "dambreak_1m".
`--pressure prj` times the steps with the PRJ solver (pressureMethod 0) instead of IP; the
"solvers" object of the line compares IP, PRJ and PRJ with its divergence-free pass on a few
steps from the same state either way.

Workload ("dambreak_1m"): a 100^3 fluid block (r = 0.005 m, jitter +-0.01 d, seed 2) in the corner
of a 4.0 x 1.5 x 1.0 m static two-layer particle box, implicit (Weiler 2018) viscosity 0.01,
dt = 0.5 ms, densityErrorTol 1e-3, maxOuterIters 100, omega 10, tol_pcg 1e-6.  One step is the
reference's Framework.forward: gravity, Weiler18 viscosity, IP solve, advection, domain clamp,
neighbour rebuild, density and Aii -- all of it on the GPU, inputs resident in HBM.

Prints one JSON line.  `value` is dynamic-particle-steps per second over the K timed steps.
`roofline` reports the kernel with the largest share of the timed region: its algorithmic bytes
per launch (the byte model the library attaches to each launch: every stored neighbour-pair
record once -- a 2 B tile-window index with the windows' lists once, or a 4 B id without them,
plus 12 B of gradW when gradients are cached -- and every per-particle array once) over its
average launch time, measured with HIP events that the kernel dispatch
itself stamps (hipExtLaunchKernelGGL) on the library's stream.  `ip_matvec_roofline` gives the
same figures for the IP CG matvec's two pair passes (cg_J, cg_JT) on a few steps after the timed
region.  `cpu_baseline` is the CPU oracle (a serial float32 C restatement of the reference
kernels) timed on 10 steps of the same workload, starting from the GPU state at the start of the
timed region, on one core; `all_cores` times the same steps with the oracle's OpenMP build on
every core the process may use, and the CPU model and core counts are reported with both.
`precision` compares `--precision-steps` consecutive steps (5 by default) of each GPU form, and
of the float32 oracle, with the float64 restatement of each step.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particle-steps/s + CG iters/s at 1M particles, ≤0.1% density err; % HBM peak"
HBM_PEAK_GBS = 8000.0


def copy_calibration(torch, mib):
    """A copy kernel of known size: each dispatch reads and writes exactly `mib` MiB."""
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    print(json.dumps({"copy_calibration_bytes": n * 4, "dispatches": 5}))


WORKLOADS = {"column_8k": 0, "column_102k": 1, "dambreak_1m": 2}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(kernel, workload="dambreak_1m"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    measurement (scaled by the copy-kernel calibration), or None when none was recorded."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("traffic_bytes")


def precision_study(ctxs, cfg, ps, m, m_inv, prm, steps, torch, dev, n, timed_steps=20):
    """Per-step error of each GPU form against the float64 restatement of the same step
    (oracle/fp64.py).  ctxs maps a form name to a context; every step starts all of them from
    the first context's state (neighbour sets for the restatement from the float32 oracle's
    build of that state).  Then each form times `timed_steps` steps from one common state."""
    from oracle import fp64
    from oracle import oracle as O
    grid = O.grid_init(cfg.domain_start, cfg.domain_size, float(np.float32(cfg.support_radius)))
    h = float(np.float32(cfg.support_radius))
    C = int(cfg.get_cfg("neighborCacheSize", 60))
    dyn = ps.is_dynamic == 1
    ip_kw = dict(omega=prm.omega, tol_opt=prm.tol_opt, max_iteration_opt=prm.max_iteration_opt,
                 use_fixed_iters=prm.use_fixed_iters, tol_pcg=prm.tol_pcg)
    names = list(ctxs)
    first = ctxs[names[0]]
    rd = lambda c, f: c.read(f, torch.zeros((n, 3), device=dev))

    def load_state(x, v):
        for nm in names[1:]:
            c = ctxs[nm]
            c.write(0, x)
            c.write(1, v)
            if c.neighbors_build() != 0:
                raise RuntimeError("neighbour overflow")

    rows = {nm: [] for nm in names}
    rows_oracle = []
    t0 = time.perf_counter()
    for _ in range(steps):
        xd, vd = rd(first, 0), rd(first, 1)
        load_state(xd, vd)
        x, v = xd.cpu().numpy(), vd.cpu().numpy()
        count, ids, _, _ = O.build_neighbors(grid, x, ps.is_dynamic, h, C)
        _, x64, v64, st = fp64.step(x, v, m, m_inv, ps.density0, ps.is_dynamic, ps.material, h, count, ids, C,
                                    cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff,
                                    cfg.domain_size, ip_kw)
        # the float32 oracle's own step from the same state: the float32 algorithm's distance to
        # float64, next to the GPU forms'
        sim = O.OracleSim(cfg, ps)
        sim.x, sim.v = x.astype(np.float32), v.astype(np.float32)
        sim.rebuild()
        sim.step()
        vo, xo = sim.v.astype(np.float64), sim.x.astype(np.float64)
        rows_oracle.append({"velocity_rel_l2": float(np.linalg.norm(vo[dyn] - v64[dyn]) / np.linalg.norm(v64[dyn])),
                            "velocity_rel_max": float(np.abs(vo[dyn] - v64[dyn]).max() / np.abs(v64[dyn]).max()),
                            "position_max_abs": float(np.abs(xo[dyn] - x64[dyn]).max()),
                            "density_error_oracle": sim.last_ip.final_density_error,
                            "density_error_fp64": st["final_density_error"],
                            "outer_iters_oracle": sim.last_ip.outer_iters, "outer_iters_fp64": st["outer_iters"]})
        for nm in names:
            c = ctxs[nm]
            res, _ = c.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, prm)
            xg = rd(c, 0).cpu().numpy().astype(np.float64)
            vg = rd(c, 1).cpu().numpy().astype(np.float64)
            rows[nm].append({"velocity_rel_l2": float(np.linalg.norm(vg[dyn] - v64[dyn]) / np.linalg.norm(v64[dyn])),
                             "velocity_rel_max": float(np.abs(vg[dyn] - v64[dyn]).max() / np.abs(v64[dyn]).max()),
                             "position_max_abs": float(np.abs(xg[dyn] - x64[dyn]).max()),
                             "density_error_gpu": res.final_density_error,
                             "density_error_fp64": st["final_density_error"],
                             "outer_iters_gpu": res.outer_iters, "outer_iters_fp64": st["outer_iters"]})
    seconds = time.perf_counter() - t0
    # speed of each form over the same steps from one state
    timing = {"steps": timed_steps}
    xd, vd = rd(first, 0), rd(first, 1)
    for nm in names:
        c = ctxs[nm]
        # three untimed steps first (the GPU idled through the CPU restatement above), then the
        # timed steps from the common state
        for warm in (True, False):
            c.write(0, xd)
            c.write(1, vd)
            if c.neighbors_build() != 0:
                raise RuntimeError("neighbour overflow")
            if warm:
                for _ in range(3):
                    c.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, prm)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        errs = 0.0
        for _ in range(timed_steps):
            res, _ = c.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, prm)
            errs = max(errs, res.final_density_error)
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - t1
        timing[nm] = {"ms_per_step": dt_ * 1e3 / timed_steps,
                      "particle_steps_per_s": int(np.count_nonzero(dyn)) * timed_steps / dt_,
                      "max_final_density_error": errs}
    gw = {0: "evaluated in float32", 1: "float32 cache", 2: "float16 cache in the CG operator, float32 "
                                                             "residual replacement every 8 CG iterations"}
    forms = {nm: {"gradW": gw[ctxs[nm].gradient_cache], "steps": rows[nm]} for nm in names}
    forms["oracle_f32"] = {"gradW": "float32 cache (CPU oracle, the reference's arithmetic)", "steps": rows_oracle}
    return {"reference": "float64 restatement of the step (CPU), from the GPU state", "dtype_gpu": "f32",
            "forms": forms, "timing": timing, "seconds": seconds}


def solver_study(src, cfg, ps, m, m_inv, prm_ip, steps, torch, dev, n, gradient_cache):
    """The reference's solver choice side by side (framework.py:168-172): from the state of
    context `src`, `steps` steps of the workload with IP (pressureMethod 1), with PRJ
    (pressureMethod 0, the reference's knobs: omega 0.5, cold start) and with PRJ plus its
    divergence-free pass (enable_df), each from the same state in a fresh context, with the
    solve's own pair passes timed by dispatch events."""
    from implicitvelocityprojection_amd.context import Context, PRJParams, kclass_mask
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    x = src.read(0, torch.zeros((n, 3), device=dev))
    v = src.read(1, torch.zeros((n, 3), device=dev))
    nd = int(np.count_nonzero(ps.is_dynamic == 1))
    prm_prj = PRJParams(tol_opt=prm_ip.tol_opt, max_iteration_opt=prm_ip.max_iteration_opt)
    prm_df = PRJParams(tol_opt=prm_ip.tol_opt, max_iteration_opt=prm_ip.max_iteration_opt, enable_df=True)
    forms = {"IP": (prm_ip, ("ip_pack", "ip_rhs", "cg_J", "cg_JT")), "PRJ": (prm_prj, ("prj_head", "prj_tail")),
             "PRJ+DF": (prm_df, ("prj_head", "prj_tail"))}
    out = {"steps": steps, "densityErrorTol": prm_ip.tol_opt, "maxOuterIters": prm_ip.max_iteration_opt}
    for name, (prm, kernels) in forms.items():
        c = Context(float(np.float32(cfg.support_radius)), cfg.domain_start, cfg.domain_size, ps.n,
                    cache_size=int(cfg.get_cfg("neighborCacheSize", 60)))
        c.gradient_cache = gradient_cache
        c.set_particles(x, v, t(m), t(m_inv), t(ps.density0), t(ps.is_dynamic, torch.int32),
                        t(ps.material, torch.int32))
        if c.neighbors_build() != 0:
            raise RuntimeError("neighbour overflow")
        c.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, prm)     # warm-up
        c.write(0, x)
        c.write(1, v)
        if c.neighbors_build() != 0:
            raise RuntimeError("neighbour overflow")
        c.profile_reset()
        c.profile_enable(kclass_mask(*kernels))
        outer, df_outer, errs, solve_ms, status = [], [], [], 0.0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            res, _ = c.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, prm)
            status = status or res.status
            outer.append(res.outer_iters)
            df_outer.append(res.df_outer_iters)
            errs.append(res.final_density_error)
            solve_ms += res.solve_ms + res.df_ms
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = c.profile_read()
        c.close()
        rec = {"status": status, "ms_per_step": wall * 1e3 / steps, "particle_steps_per_s": nd * steps / wall,
               "pressure_solve_ms_per_step": solve_ms / steps, "outer_iters": outer,
               "max_final_density_error": max(errs), "kernels": {}}
        if name == "PRJ+DF":
            rec["df_outer_iters"] = df_outer
        for k in kernels:
            p = prof[k]
            if p["launches"]:
                us = p["total_ms"] * 1e3 / p["launches"]
                rec["kernels"][k] = {"launches": p["launches"], "avg_launch_us": us,
                                     "achieved_GBs": p["bytes"] / p["launches"] / (us * 1e-6) / 1e9}
        out[name] = rec
    return out


def _box_mesh(lo, hi):
    v = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float64)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    return v, np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))], np.int64)


def scene_io_study(torch, repeats=5, cpu=True):
    """Mesh voxelisation (SceneLoader.load_mesh) on the GPU: a thick open-top container the size of
    the shipped dam break's (3.5 x 9.0 x 8.0 m, walls 0.24 m: five closed boxes, 60 triangles)
    at pitch 0.06 (particle radius 0.03), timed over `repeats` calls; the CPU baseline is the
    numpy restatement of the reference's trimesh voxeliser on the same mesh."""
    from implicitvelocityprojection_amd.mesh import Mesh, voxelize
    lo, hi, w = np.array([1.25, -0.12, 0.0]), np.array([4.75, 8.88, 8.0]), 0.24
    parts = [_box_mesh(lo, (hi[0], lo[1] + w, hi[2])), _box_mesh(lo, (lo[0] + w, hi[1], hi[2])),
             _box_mesh((hi[0] - w, lo[1], lo[2]), hi), _box_mesh(lo, (hi[0], hi[1], lo[2] + w)),
             _box_mesh((lo[0], lo[1], hi[2] - w), hi)]
    v = np.concatenate([p[0] for p in parts])
    f = np.concatenate([p[1] + 8 * k for k, p in enumerate(parts)])
    mesh, pitch = Mesh(v, f), 0.06
    pts = voxelize(mesh, pitch)          # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        pts = voxelize(mesh, pitch)
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) * 1e3 / repeats
    n = int(pts.shape[0])
    out = {"workload": "thick open container 3.5 x 9.0 x 8.0 m, walls 0.24 m, pitch 0.06 (60 triangles)",
           "points": n, "gpu_ms_per_mesh": gpu_ms, "gpu_points_per_s": n / (gpu_ms * 1e-3)}
    if cpu:
        from oracle import voxelize as OV
        t0 = time.perf_counter()
        ref = OV.voxelize(v, f, pitch)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds_per_mesh": cpu_s, "points_per_s": ref.shape[0] / cpu_s, "cores": 1,
                               "kind": "port", "sample": "the same mesh once, numpy restatement of trimesh's "
                                                         "subdivide voxeliser + scipy hole fill"}
        out["bitwise_equal_to_oracle"] = bool(np.array_equal(pts.cpu().numpy(), ref))
    return out


def _progress(msg):
    """A progress line on stderr (the result line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batched_study(n_scenes, steps, torch):
    """The batched form: n_scenes independent columns (synthetic_batch, scene-0 physics) stepped
    in one context, against the same scenes stepped one by one, each in its own context.
    Inputs resident in HBM in both; the CG iterations per replay are autotuned first."""
    from implicitvelocityprojection_amd.context import Context
    from implicitvelocityprojection_amd.scene_batch import SceneBatch, _ip_params
    from implicitvelocityprojection_amd.config import SimConfig
    from implicitvelocityprojection_amd.scenes import load_particles, particle_masses, synthetic_batch
    scenes = synthetic_batch(n_scenes=n_scenes)
    _progress(f"batched: {n_scenes} scenes loaded, autotuning")
    batch = SceneBatch(scenes)
    tune = batch.autotune(chunks=(1, 2, 4, 8), lanes=(1, 2, 4), steps=2)
    _progress("batched: autotuned, timing")
    chosen = {"cg_chunk": batch.ctx.scene_cg_chunk, "lanes_per_row": batch.ctx.scene_lanes,
              "build_lanes": batch.ctx.build_lanes}
    for _ in range(2):
        batch.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cg = 0
    max_err = 0.0
    outer = 0
    replays = 0
    for _ in range(steps):
        status, stats = batch.step()
        if status != 0:
            raise RuntimeError(f"scene step status {status}")
        cg += sum(s[2] for s in stats)
        outer = max(outer, max(s[1] for s in stats))
        max_err = max(max_err, max(s[3] for s in stats))
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    _progress(f"batched: {steps} steps timed; the scenes one by one")
    nd = batch.n_dynamic
    n_all = batch.n
    batch.close()
    # the same scenes one by one
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    ctxs = []
    for sc in scenes:
        cfg = SimConfig.from_dict(sc)
        ps = load_particles(cfg)
        m, mi = particle_masses(cfg, ps)
        ctx = Context(float(np.float32(cfg.support_radius)), cfg.domain_start, cfg.domain_size, ps.n,
                      cache_size=int(cfg.get_cfg("neighborCacheSize", 60)))
        ctx.set_particles(t(ps.x), t(ps.v), t(m), t(mi), t(ps.density0), t(ps.is_dynamic, torch.int32),
                          t(ps.material, torch.int32))
        if ctx.neighbors_build() != 0:
            raise RuntimeError("neighbour overflow at load")
        ctxs.append((cfg, ctx))
        if len(ctxs) % 64 == 0:
            _progress(f"one by one: {len(ctxs)} contexts loaded")
    prm = _ip_params(ctxs[0][0])
    for cfg, ctx in ctxs:
        ctx.step(cfg.dt, cfg.gravity, 0, 0.0, prm)
    one_steps = min(steps, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for cfg, ctx in ctxs:
        for _ in range(one_steps):
            res, _ = ctx.step(cfg.dt, cfg.gravity, 0, 0.0, prm)
            max_err = max(max_err, res.final_density_error)
    torch.cuda.synchronize()
    wall_o = time.perf_counter() - t0
    for _, ctx in ctxs:
        ctx.close()
    batched = nd * steps / wall_b
    one = nd * one_steps / wall_o
    return {"workload": f"{n_scenes} independent columns, 2048-32768 fluid particles each (seed 4), "
                        "r = 0.01, dt = 1 ms, no viscosity",
            "fluid_particles": nd, "particles": n_all, "steps": steps,
            "value": batched, "unit": "particle-steps/s", "ms_per_step": wall_b * 1e3 / steps,
            "cg_iters_per_s": cg / wall_b, "max_outer_iters": outer, "max_final_density_error": max_err,
            "autotune_ms_per_step": tune, "autotuned": chosen,
            "one_by_one": {"value": one, "unit": "particle-steps/s", "steps": one_steps,
                           "ms_per_step_of_all_scenes": wall_o * 1e3 / one_steps},
            "speedup_vs_one_by_one": batched / one}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--gradient-cache", type=int, default=0, choices=[0, 1, 2],
                    help="neighbour-row format: 0 ids only, gradW from positions (4 B per pair); 1 float32 gradW "
                         "cached (16 B per pair); 2 float16 gradW cached for the IP CG operator (12 B per pair) "
                         "with a float32 true-residual refresh every 8 CG iterations")
    ap.add_argument("--tile-windows", type=int, default=1, choices=[0, 1],
                    help="1 (default): the pair passes stage each 64-row tile's distinct neighbours in LDS; "
                         "0: every pair gathers its neighbour from memory (same results)")
    ap.add_argument("--precision-steps", type=int, default=5,
                    help="consecutive steps (from the state after the timed region) whose GPU result, in float32 and "
                         "in the float16 precision form, is compared with the float64 restatement of each step "
                         "(about 10 s of CPU per step at 1M); then 20 steps of each form are timed from one state; "
                         "0 skips both")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="dambreak_1m")
    ap.add_argument("--surface-tension", type=float, default=0.0, metavar="SIGMA",
                    help="add Jeske23 implicit surface tension (surfaceTensionMethod 1) with this coefficient "
                         "to the workload; 0 (default) leaves it off, as in the workload's scene")
    ap.add_argument("--solver-steps", type=int, default=5,
                    help="steps of the IP / PRJ / PRJ+divergence-free comparison from the state after the timed "
                         "region; 0 skips it")
    ap.add_argument("--pressure", choices=["ip", "prj"], default="ip",
                    help="pressure solver of the timed steps: IP (the workload's, default) or PRJ (pressureMethod 0)")
    ap.add_argument("--scene-io", action="store_true",
                    help="also time the GPU mesh voxeliser of scene loading against the CPU restatement "
                         "(about 30 s of CPU work)")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="independent scenes of the batched-form comparison; 0 skips it")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--event-scope", choices=["dominant", "all"], default="dominant",
                    help="kernel classes timed with dispatch events during the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device-to-device copies of MIB MiB (the reference kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes), then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the IP hot path here is a single-device build "
              "(the multi-GPU decomposition is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        copy_calibration(torch, args.copy_calibration)
        return
    from implicitvelocityprojection_amd.config import SimConfig
    from implicitvelocityprojection_amd.context import Context, IPParams, PRJParams, kclass_mask
    from implicitvelocityprojection_amd.scenes import load_particles, particle_masses, synthetic_scene

    scene = synthetic_scene(WORKLOADS[args.workload])
    if args.surface_tension != 0.0:
        scene["Configuration"]["surfaceTensionMethod"] = 1
        scene["Configuration"]["surfaceTensionCoeff"] = args.surface_tension
    cfg = SimConfig.from_dict(scene)
    ps = load_particles(cfg)
    m, m_inv = particle_masses(cfg, ps)
    h = float(np.float32(cfg.support_radius))
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    ctx = Context(h, cfg.domain_start, cfg.domain_size, ps.n, cache_size=int(cfg.get_cfg("neighborCacheSize", 60)))
    ctx.gradient_cache = args.gradient_cache
    ctx.tile_windows = bool(args.tile_windows)
    ctx.set_particles(t(ps.x), t(ps.v), t(m), t(m_inv), t(ps.density0), t(ps.is_dynamic, torch.int32),
                      t(ps.material, torch.int32))
    if ctx.neighbors_build() != 0:
        raise RuntimeError("neighbour overflow at load")
    n, nd = ctx.counts()
    ns = n - nd
    prm = IPParams(omega=cfg.get_cfg("omega", 10.0), tol_opt=cfg.get_cfg("densityErrorTol", 1e-3),
                   max_iteration_opt=cfg.get_cfg("maxOuterIters", 100), tol_pcg=cfg.get_cfg("tol_pcg", 1e-6))

    step_prm = prm if args.pressure == "ip" else PRJParams(tol_opt=prm.tol_opt, max_iteration_opt=prm.max_iteration_opt)
    st_method, st_coeff = int(cfg.surface_tension_method), float(cfg.surface_tension_coeff)
    if st_method != 1:
        st_coeff = 0.0     # the synthetic scenes set coefficient 0: the explicit model is a no-op

    def step():
        res, vit = ctx.step(cfg.dt, cfg.gravity, cfg.viscosity_method, cfg.viscosity_coeff, step_prm, st_method,
                            st_coeff)
        if res.status != 0:
            raise RuntimeError(f"{args.pressure.upper()} solve status {res.status}")
        return res, vit

    # warm-up, with every kernel class timed, to find the dominant kernel
    ctx.profile_enable(-1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    prof = ctx.profile_read()
    dominant = max(prof, key=lambda k: prof[k]["total_ms"])

    # state at the start of the timed region (for the CPU baseline)
    x0 = ctx.read(0, torch.zeros((n, 3), device=dev))
    v0 = ctx.read(1, torch.zeros((n, 3), device=dev))

    # timed region: events on the dominant kernel (or on every class)
    ctx.profile_reset()
    ctx.profile_enable(kclass_mask(dominant) if args.event_scope == "dominant" else -1)
    pairs_sum = 0
    cg_iters = 0
    solve_ms = 0.0
    visc_iters = 0
    st_iters = 0
    outer = 0
    max_err = 0.0
    cg_hist = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, vit = step()
        cg_step = sum(res.pcg_iters)
        cg_iters += cg_step
        cg_hist[cg_step] = cg_hist.get(cg_step, 0) + 1
        solve_ms += res.solve_ms
        visc_iters += vit
        st_iters += ctx.last_st_iters
        outer += res.outer_iters
        max_err = max(max_err, res.final_density_error)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile_enable(0)
    pd = prof[dominant]
    avg_ms = pd["total_ms"] / max(pd["launches"], 1)
    bytes_per_launch = pd["bytes"] / max(pd["launches"], 1)
    pairs_end = ctx.pair_count()
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    # kernel time of one IP CG iteration (J, J^T with p.Ap, x/r update with r.r, p update), from
    # dispatch events on a few steps after the timed region (launches past convergence excluded)
    cg_us_per_iter = None
    matvec_roofline = {}
    if args.pressure == "ip":
        cg_classes = ("cg_J", "cg_JT", "cg_update", "cg_p")
        ctx.profile_reset()
        ctx.profile_enable(kclass_mask(*cg_classes))
        it = 0
        for _ in range(5):
            res, _vit = step()
            it += sum(res.pcg_iters)
        torch.cuda.synchronize()
        pc = ctx.profile_read()
        ctx.profile_enable(0)
        if it > 0:
            cg_us_per_iter = {k: pc[k]["total_ms"] * 1e3 / it for k in cg_classes if k in pc}
            cg_us_per_iter["total"] = sum(cg_us_per_iter.values())
            # the IP matvec's two pair passes against the HBM roofline, on their own algorithmic
            # bytes (the byte model attached to each launch)
            for k in ("cg_J", "cg_JT"):
                q = pc[k]
                if q["launches"]:
                    us = q["total_ms"] * 1e3 / q["launches"]
                    b = q["bytes"] / q["launches"]
                    ach = b / (us * 1e-6) / 1e9
                    matvec_roofline[k] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": ach / HBM_PEAK_GBS, "avg_launch_us": us,
                                          "launches": q["launches"], "algorithmic_bytes_per_launch": b,
                                          "traffic": measured_traffic(k, args.workload)}

    precision = None
    if args.precision_steps > 0 and st_coeff == 0.0 and args.pressure == "ip":
        # the float64 restatement has the IP solve and no surface tension
        forms = {"f32": ctx}
        if args.gradient_cache != 2:
            # the precision form: float16 gradW rows in the IP CG operator
            c16 = Context(h, cfg.domain_start, cfg.domain_size, ps.n,
                          cache_size=int(cfg.get_cfg("neighborCacheSize", 60)))
            c16.gradient_cache = 2
            c16.set_particles(t(ps.x), t(ps.v), t(m), t(m_inv), t(ps.density0), t(ps.is_dynamic, torch.int32),
                              t(ps.material, torch.int32))
            forms["f16"] = c16
        precision = precision_study(forms, cfg, ps, m, m_inv, prm, args.precision_steps, torch, dev, n)
        if "f16" in forms:
            forms["f16"].close()

    solvers = None
    if args.solver_steps > 0 and st_coeff == 0.0:
        solvers = solver_study(ctx, cfg, ps, m, m_inv, prm, args.solver_steps, torch, dev, n, args.gradient_cache)

    scene_io = scene_io_study(torch, cpu=not args.no_cpu_baseline) if args.scene_io else None

    batched = None
    if args.batch_scenes > 0:
        batched = batched_study(args.batch_scenes, args.batch_steps, torch)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        def cpu_run(parallel):
            threads = O.use_threads(parallel)
            sim = O.OracleSim(cfg, ps)
            if args.pressure == "prj":
                sim.pressure_method = 0
            sim.x = x0.cpu().numpy().astype(np.float32)
            sim.v = v0.cpu().numpy().astype(np.float32)
            sim.rebuild()
            tc = time.perf_counter()
            for _ in range(args.cpu_steps):
                sim.step()
            return threads, time.perf_counter() - tc

        _, tcpu = cpu_run(False)
        threads, tomp = cpu_run(True)
        O.use_threads(False)
        cpu = {"value": nd * args.cpu_steps / tcpu, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of {args.workload} from the GPU state at the start of the timed "
                         f"region, serial float32 C oracle ({tcpu:.1f} s)",
               "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
               "all_cores": {"value": nd * args.cpu_steps / tomp, "cores": threads, "seconds": tomp,
                             "sample": f"the same {args.cpu_steps} steps, the oracle's OpenMP build (row loops on "
                                       f"{threads} threads, atomics in the J^T scatter)"}}

    out = {
        "metric": METRIC,
        "value": nd * args.steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "fluid_particles": nd, "static_particles": ns,
                   "gradient_cache": ["ids", "f32", "f16"][args.gradient_cache],
                   "dt": cfg.dt, "viscosity": (["explicit", "Weiler18"][cfg.viscosity_method] + f" {cfg.viscosity_coeff:g}"
                                               if cfg.viscosity_coeff != 0.0 else "off"),
                   "surface_tension": f"Jeske23 {st_coeff:g}" if st_coeff != 0.0 else "off",
                   "pressure": {"ip": "IP", "prj": "PRJ"}[args.pressure],
                   "densityErrorTol": prm.tol_opt,
                   "maxOuterIters": prm.max_iteration_opt, "neighbour_pairs": pairs_end,
                   "tile_windows": dict(zip(("tiles", "windowed", "staged_ids"), ctx.window_stats()))},
        "cg_iters_per_s": cg_iters / (solve_ms * 1e-3) if solve_ms > 0 else 0.0,
        "cg_iters": cg_iters,
        "cg_iters_per_step_hist": {str(k): cg_hist[k] for k in sorted(cg_hist)},
        "cg_kernel_us_per_iter": cg_us_per_iter,
        "ip_matvec_roofline": matvec_roofline or None,
        "ip_outer_iters": outer,
        "weiler18_cg_iters": visc_iters,
        "surface_tension_cg_iters": st_iters,
        "ip_solve_ms_per_step": solve_ms / args.steps,
        "max_final_density_error": max_err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant, args.workload), "kernel": dominant,
                     "avg_launch_us": avg_ms * 1e3, "launches": pd["launches"], "event_scope": args.event_scope,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "precision": precision,
        "solvers": solvers,
        "scene_io": scene_io,
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
