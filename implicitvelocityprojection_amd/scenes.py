"""Scene construction: the reference's ``SceneLoader`` (scripts/io/scene_loader.py) block by
block, plus the synthetic benchmark scenes.

Supported scene JSON blocks (reference semantics):
  * ``FluidBlocks``      scene_loader.py:90-102, 319-336 and the lattice of
                         ``SimulationData.add_cube`` (simulation_data.py:290-356):
                         ``np.arange(start, end, d)`` per axis, ``indexing='ij'``.
  * ``GeneratedBodies``  ``"type": "cube"`` hollow shells, scene_loader.py:132-140, 402-419,
                         447-492 (every face is a full grid, so edge/corner particles repeat).
  * ``boundaryLayers``   domain walls, scene_loader.py:188-194, 421-426, 494-545.

Extension keys (not in the reference), used by the synthetic scenes:
  * FluidBlocks ``"jitter"`` (fraction of the particle diameter d) and ``"jitterSeed"``: every
    coordinate gets ``numpy.random.default_rng(seed).uniform(-jitter*d, +jitter*d)`` added in
    float64 before the float32 store.

Mesh blocks (voxelised meshes, scene_loader.py:104-186, 359-400, 547-580):
  * ``FluidBodies``      fluid particles at the voxel centres, optional ``angularVelocity``
                         (v = v_lin + omega x (r - r_cm), r_cm the mean of the centres).
  * ``RigidBodies``      rigid particles, dynamic only with ``"isDynamic": true`` (velocity 0
                         otherwise).
  * ``ScheduledSpawns``  expanded into timed entries: ``spawnNum`` waves ``interval`` apart from
                         ``startTime``, an optional ``grid`` of copies centred on the
                         translation (one object id per grid slot), ``waveOffset`` per wave,
                         ``colorCycle``; material ``fluid`` or ``rigid``.
The voxel centres come from a callable ``mesh_points(obj) -> float64 (N, 3)`` (the GPU
voxeliser in ``SceneLoader``; the oracle's restatement in tests).  ``SolidBodies`` and
deformable spawn materials belong to the elasticity solver, which is not part of this build,
and raise ``NotImplementedError``.
"""

from dataclasses import dataclass
from functools import reduce

import numpy as np

from .config import SimConfig

MATERIAL_RIGID = 0
MATERIAL_FLUID = 1


@dataclass
class ParticleSet:
    """Particles in the reference's insertion order (scene_loader.py:217-429)."""
    x: np.ndarray           # (n, 3) float32
    v: np.ndarray           # (n, 3) float32
    density0: np.ndarray    # (n,) float32
    material: np.ndarray    # (n,) int32
    is_dynamic: np.ndarray  # (n,) int32
    object_id: np.ndarray   # (n,) int32

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def cube_lattice(lower_corner, cube_size, diam):
    """Lattice of ``SimulationData.add_cube`` (simulation_data.py:316-328)."""
    axes = [np.arange(lower_corner[i], lower_corner[i] + cube_size[i], diam) for i in range(3)]
    n = reduce(lambda a, b: a * b, [len(a) for a in axes])
    pos = np.array(np.meshgrid(*axes, sparse=False, indexing="ij"), dtype=np.float32)
    return pos.reshape(-1, n).transpose().copy()


def _shell_layer(lo, hi, diam):
    """One layer of a box's particle shell between corners lo and hi: the faces y = lo[1],
    y = hi[1], x = lo[0], x = hi[0], z = lo[2], z = hi[2] in that order, each a lattice at spacing
    diam over the two other axes (the lower-numbered axis outer), edge and corner particles
    repeated on every face that holds them -- the order and the np.arange bounds that the
    reference's generators use, so the particle arrays are identical (scene_loader.py:447-545)."""
    axes = [np.arange(lo[a], hi[a] + diam * 0.5, diam) for a in range(3)]
    faces = []
    for normal in (1, 0, 2):
        u, w = (a for a in range(3) if a != normal)
        gu, gw = np.meshgrid(axes[u], axes[w], indexing="ij")
        for level in (lo[normal], hi[normal]):
            face = np.empty((gu.size, 3), np.float64)
            face[:, u], face[:, w], face[:, normal] = gu.ravel(), gw.ravel(), level
            faces.append(face)
    return faces


def cube_shell(translation, size, layers, diam):
    """``SceneLoader._generate_body`` for ``type == "cube"``: `layers` shells around the box of
    the given centre and size, each one diam further out."""
    centre = np.array(translation, dtype=np.float32)
    half = np.array(size, dtype=np.float32) / 2.0
    lo, hi = centre - half, centre + half
    faces = [f for k in range(int(layers)) for f in _shell_layer(lo - k * diam, hi + k * diam, diam)]
    return np.concatenate(faces, axis=0).astype(np.float32)


def boundary_walls(domain_start, domain_size, num_layers, diam):
    """``SceneLoader._generate_boundary_walls``: `num_layers` shells around the domain, each one
    diam further out."""
    lo = np.array(domain_start, dtype=np.float32)
    hi = lo + np.array(domain_size, dtype=np.float32)
    faces = [f for k in range(int(num_layers)) for f in _shell_layer(lo - k * diam, hi + k * diam, diam)]
    return np.concatenate(faces, axis=0).astype(np.float32)


DEFORMABLE_MATERIALS = ("solid", "deformable", "elastic")    # scheduled_spawn.py:11


def _need_mesh_points(config, mesh_points):
    if config.get_solid_bodies():
        raise NotImplementedError("SolidBodies belong to the elasticity solver, which is not part of this build")
    for sp in config.get_scheduled_spawns():
        if str(sp.get("material", "fluid")).lower() in DEFORMABLE_MATERIALS:
            raise NotImplementedError("deformable ScheduledSpawns belong to the elasticity solver, which is not "
                                      "part of this build")
    if mesh_points is None and (config.get_fluid_bodies() or config.get_rigid_bodies()
                                or config.get_scheduled_spawns()):
        raise NotImplementedError("mesh blocks need a voxeliser: load the scene through SceneLoader")


def _spawn_object_ids(config):
    """Object ids the ScheduledSpawns claim (scene_loader.py:158-169)."""
    ids = []
    for sp in config.get_scheduled_spawns():
        n_slots = len(grid_offsets(sp))
        base = int(sp["objectId"])
        ids += [base + i for i in range(n_slots if n_slots > 1 else int(sp.get("spawnNum", 1)))]
    return ids


def load_particles(config: SimConfig, mesh_points=None) -> ParticleSet:
    """Particles of a scene, in the reference's insertion order: fluid blocks, fluid bodies, rigid
    bodies, generated bodies, boundary walls (scene_loader.py:319-426).  `mesh_points(obj)`
    returns the float64 voxel centres of a mesh block."""
    _need_mesh_points(config, mesh_points)
    diam = config.particle_diameter
    xs, vs, rho0, mat, dyn, oid = [], [], [], [], [], []

    def push(pos, vel, density, material, is_dynamic, obj):
        n = pos.shape[0]
        xs.append(pos.astype(np.float32))
        vs.append(np.broadcast_to(np.asarray(vel, dtype=np.float32), (n, 3)).copy())
        rho0.append(np.full(n, density, dtype=np.float32))
        mat.append(np.full(n, material, dtype=np.int32))
        dyn.append(np.full(n, is_dynamic, dtype=np.int32))
        oid.append(np.full(n, obj, dtype=np.int32))

    object_ids = []
    for fluid in config.get_fluid_blocks():
        object_ids.append(int(fluid["objectId"]))
        t = np.array(fluid.get("translation", [0, 0, 0]), dtype=np.float32)
        s = np.array(fluid.get("scale", [1, 1, 1]), dtype=np.float32)
        start = np.array(fluid["start"], dtype=np.float32) + t
        size = (np.array(fluid["end"], dtype=np.float32) - np.array(fluid["start"], dtype=np.float32)) * s
        end = start + size
        pos = cube_lattice(tuple(start), tuple(end - start), diam)
        jitter = float(fluid.get("jitter", 0.0))
        if jitter > 0.0:
            rng = np.random.default_rng(int(fluid.get("jitterSeed", 0)))
            amp = jitter * diam
            pos = (pos.astype(np.float64) + rng.uniform(-amp, amp, size=pos.shape)).astype(np.float32)
        push(pos, fluid.get("velocity", [0.0, 0.0, 0.0]), float(fluid.get("density", 1000.0)),
             MATERIAL_FLUID, 1, int(fluid["objectId"]))
    for fb in config.get_fluid_bodies():
        object_ids.append(int(fb["objectId"]))
        pts = np.asarray(mesh_points(fb), dtype=np.float64)
        vel = np.array(fb.get("velocity", [0.0, 0.0, 0.0]), dtype=np.float32)
        omega = np.array(fb.get("angularVelocity", [0.0, 0.0, 0.0]), dtype=np.float32)
        if np.linalg.norm(omega) > 0.0:
            vel = np.cross(omega, pts - pts.mean(axis=0)).astype(np.float32) + vel
        push(pts, vel, float(fb.get("density", 1000.0)), MATERIAL_FLUID, 1, int(fb["objectId"]))
    for rb in config.get_rigid_bodies():
        object_ids.append(int(rb["objectId"]))
        pts = np.asarray(mesh_points(rb), dtype=np.float64)
        rdyn = int(bool(rb.get("isDynamic", False)))
        vel = np.array(rb.get("velocity", [0.0, 0.0, 0.0]), dtype=np.float32) if rdyn else np.zeros(3, np.float32)
        push(pts, vel, float(rb.get("density", 1000.0)), MATERIAL_RIGID, rdyn, int(rb["objectId"]))
    for gb in config.get_generated_bodies():
        object_ids.append(int(gb["objectId"]))
        if gb.get("type", "cube") != "cube":
            raise ValueError(f"Unknown generated body type: {gb.get('type')}")
        pos = cube_shell(gb.get("translation", [0, 0, 0]), gb.get("size", [1, 1, 1]), int(gb.get("layers", 2)), diam)
        push(pos, [0.0, 0.0, 0.0], float(gb.get("density", 1000.0)), MATERIAL_RIGID,
             int(bool(gb.get("isDynamic", False))), int(gb["objectId"]))
    layers = int(config.get_cfg("boundaryLayers", 0))
    if layers > 0:
        pos = boundary_walls(config.domain_start, config.domain_size, layers, diam)
        object_ids += _spawn_object_ids(config)
        num_objects = (max(object_ids) + 1) if object_ids else 1
        push(pos, [0.0, 0.0, 0.0], float(config.density0), MATERIAL_RIGID, 0, num_objects)
    if not xs:
        z = np.zeros((0, 3), np.float32)
        return ParticleSet(z, z.copy(), np.zeros(0, np.float32), np.zeros(0, np.int32),
                           np.zeros(0, np.int32), np.zeros(0, np.int32))
    return ParticleSet(np.concatenate(xs), np.concatenate(vs), np.concatenate(rho0),
                       np.concatenate(mat), np.concatenate(dyn), np.concatenate(oid))


def grid_offsets(sp):
    """Offsets of a ScheduledSpawn's grid copies, centred on its translation
    (scene_loader.py:431-445): ``"grid": {"counts": [nx, ny, nz], "spacing": [dx, dy, dz]}``."""
    grid = sp.get("grid")
    if not grid:
        return np.zeros((1, 3), dtype=np.float32)
    counts = [max(1, int(c)) for c in grid.get("counts", [1, 1, 1])]
    spacing = [float(s) for s in grid.get("spacing", [0.0, 0.0, 0.0])]
    axes = [(np.arange(n) - 0.5 * (n - 1)) * s for n, s in zip(counts, spacing)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1).astype(np.float32)


@dataclass
class SpawnEntry:
    """One timed body of a ScheduledSpawns block (scheduled_spawn.py:14-32)."""
    time: float
    object_id: int
    positions: np.ndarray      # (n, 3) float32
    velocity: np.ndarray       # (3,) float32
    density: float
    material: int              # 0 rigid, 1 fluid
    is_dynamic: int
    color: np.ndarray


def spawn_entries(config: SimConfig, mesh_points):
    """ScheduledSpawns expanded as prepare_scene does (scene_loader.py:143-186): spawnNum waves,
    grid copies, waveOffset shifts (added in float64 to the float64 voxel centres), colours
    cycling per wave and slot."""
    _need_mesh_points(config, mesh_points)
    out = []
    for sp in config.get_scheduled_spawns():
        material = str(sp.get("material", "fluid")).lower()
        vox = np.asarray(mesh_points(sp), dtype=np.float64)
        num = int(sp.get("spawnNum", 1))
        start = float(sp.get("startTime", sp.get("time", 0.0)))
        interval = float(sp.get("interval", 1.0))
        base = int(sp["objectId"])
        slots = [vox + off for off in grid_offsets(sp)]
        has_grid = len(slots) > 1
        colors = sp.get("colorCycle") or [sp.get("color", [10, 100, 200])]
        wave = np.array(sp.get("waveOffset", [0.0, 0.0, 0.0]), dtype=np.float32)
        for k in range(num):
            shift = k * wave
            for g, pos in enumerate(slots):
                pts = (pos + shift) if k else pos
                out.append(SpawnEntry(time=start + k * interval, object_id=base + (g if has_grid else k),
                                      positions=pts.astype(np.float32),
                                      velocity=np.array(sp.get("velocity", [0.0, 0.0, 0.0]), dtype=np.float32),
                                      density=float(sp.get("density", 1000.0)),
                                      material=0 if material == "rigid" else 1,
                                      is_dynamic=1 if bool(sp.get("isDynamic", True)) else 0,
                                      color=np.array(colors[(k + g) % len(colors)], dtype=np.int32)))
    return out


# ----------------------------------------------------------------------------------------------
# Synthetic benchmark scenes
# ----------------------------------------------------------------------------------------------

# (fluid counts nx, ny, nz; box L; particle radius; margin; fluid placement; dt; steps;
#  viscosity method, coefficient; seed)
_SYNTHETIC = {
    0: dict(n=(16, 32, 16), L=(0.64, 1.0, 0.64), r=0.01, o=0.1, place="centred", dt=1e-3, steps=100,
            visc=(0, 0.0), seed=0, name="column_8k"),
    1: dict(n=(40, 64, 40), L=(1.6, 3.0, 1.6), r=0.01, o=0.1, place="centred", dt=1e-3, steps=200,
            visc=(0, 0.01), seed=1, name="column_102k"),
    2: dict(n=(100, 100, 100), L=(4.0, 1.5, 1.0), r=0.005, o=0.05, place="corner", dt=5e-4, steps=100,
            visc=(1, 0.01), seed=2, name="dambreak_1m"),
}


def synthetic_scene(cfg_id: int, scale=None, box=None, seed=None) -> dict:
    """Scene JSON dict of a synthetic benchmark scene.

    A fluid block of nx*ny*nz particles sits at voxel centres (k + 1/2) d with a uniform jitter of
    +-0.01 d, resting on the floor of a static two-layer particle box whose inner faces are at 0
    and L (shifted by the margin o inside the domain [0, L + 2o], so the domain clamp never
    engages).  ``scale`` optionally replaces the fluid counts, ``box`` the box size L and ``seed``
    the jitter seed (reduced test sizes, the scenes of ``synthetic_batch``).
      0: 16x32x16 = 8192 particles, box 0.64 x 1.0 x 0.64 m, r = 0.01, dt = 1 ms, no viscosity.
      1: 40x64x40 = 102400 particles, box 1.6 x 3.0 x 1.6 m, r = 0.01, dt = 1 ms,
         explicit viscosity 0.01.
      2: 100^3 = 1M particles in the corner [0,1]^3 of a 4.0 x 1.5 x 1.0 m box, r = 0.005,
         dt = 0.5 ms, implicit (Weiler 2018) viscosity 0.01.
    y is the gravity axis.
    """
    s = _SYNTHETIC[cfg_id]
    n = tuple(scale) if scale is not None else s["n"]
    r = s["r"]
    d = 2.0 * r
    L = tuple(box) if box is not None else s["L"]
    o = s["o"]
    if s["place"] == "centred":
        off = ((L[0] - n[0] * d) / 2.0, 0.0, (L[2] - n[2] * d) / 2.0)
    else:
        off = (0.0, 0.0, 0.0)
    start = [o + off[a] + 0.5 * d for a in range(3)]
    # end sits half a spacing past the last centre so np.arange yields exactly n points
    end = [start[a] + (n[a] - 0.5) * d for a in range(3)]
    visc_method, visc_coeff = s["visc"]
    return {
        "Configuration": {
            "domainStart": [0.0, 0.0, 0.0],
            "domainEnd": [L[0] + 2 * o, L[1] + 2 * o, L[2] + 2 * o],
            "particleRadius": r,
            "density0": 1000.0,
            "gravitation": [0.0, -9.81, 0.0],
            "timeStepSize": s["dt"],
            "viscosityMethod": visc_method,
            "viscosityCoeff": visc_coeff,
            "surfaceTensionMethod": 0,
            "surfaceTensionCoeff": 0.0,
            "pressureMethod": 1,
            "densityErrorTol": 1e-3,
            "maxOuterIters": 100,
            "omega": 10.0,
            "tol_pcg": 1e-6,
            "neighborCacheSize": 60,
            "steps": s["steps"],
            "name": s["name"],
        },
        "FluidBlocks": [{
            "objectId": 0,
            "start": start,
            "end": end,
            "velocity": [0.0, 0.0, 0.0],
            "density": 1000.0,
            "jitter": 0.01,
            "jitterSeed": s["seed"] if seed is None else int(seed),
        }],
        "GeneratedBodies": [{
            "objectId": 1,
            "type": "cube",
            "translation": [o + L[0] / 2.0, o + L[1] / 2.0, o + L[2] / 2.0],
            "size": list(L),
            "layers": 2,
            "density": 1000.0,
            "isDynamic": False,
        }],
    }


def synthetic_batch(n_scenes=256, seed=4, count_range=(2048, 32768), aspect_range=(0.5, 4.0)):
    """Scene JSON dicts of the batched workload: independent fluid columns with the physics of
    synthetic scene 0 (r = 0.01, dt = 1 ms, no viscosity).  With rng = default_rng(seed), the
    particle counts n_s = rng.integers(lo, hi + 1, n_scenes) and then the aspect ratios
    a_s = height / width = rng.uniform(a_lo, a_hi, n_scenes) are drawn; scene s has
    nx = nz = round((n_s / a_s)^(1/3)) and ny = round(n_s / nx^2) fluid particles per axis, a box
    (2 w, 1.5625 hgt, 2 w) with w = nx d and hgt = ny d (scene 0's proportions), and jitter seed
    1000 + s."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(count_range[0], count_range[1] + 1, size=n_scenes)
    aspects = rng.uniform(aspect_range[0], aspect_range[1], size=n_scenes)
    d = 2.0 * _SYNTHETIC[0]["r"]
    out = []
    for s in range(n_scenes):
        nx = max(1, int(round((float(counts[s]) / float(aspects[s])) ** (1.0 / 3.0))))
        ny = max(1, int(round(float(counts[s]) / (nx * nx))))
        box = (2.0 * nx * d, 1.5625 * ny * d, 2.0 * nx * d)
        scene = synthetic_scene(0, scale=(nx, ny, nx), box=box, seed=1000 + s)
        scene["Configuration"]["name"] = f"batch_scene_{s}"
        out.append(scene)
    return out


def particle_masses(config: SimConfig, ps: ParticleSet):
    """Mass convention of ``add_particle`` (simulation_data.py:40-57, 137):
    m = 0.8 d^3 rho0; static particles m *= 10 and m_inv = 0; dynamic m_inv = 1 / m."""
    v0 = np.float32(0.8 * config.particle_diameter ** 3)
    m = (v0 * ps.density0).astype(np.float32)
    dyn = ps.is_dynamic == 1
    m_inv = np.zeros_like(m)
    m_inv[dyn] = np.float32(1.0) / m[dyn]
    m[~dyn] = m[~dyn] * np.float32(10.0)
    return m, m_inv
