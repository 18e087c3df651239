"""Host-side logic: scene construction (scene_loader.py / simulation_data.py semantics), scene
JSON parsing (config_builder.py) and the mass convention (simulation_data.py:40-57, 137)."""

import json

import numpy as np
import pytest

from implicitvelocityprojection_amd.config import SimConfig
from implicitvelocityprojection_amd.scenes import (boundary_walls, cube_lattice, cube_shell, load_particles,
                                                   particle_masses, synthetic_scene)


def test_synthetic_scene_sizes():
    counts = {0: (8192, 18780), 1: (102400, 126620), 2: (1000000, 470460)}
    for cid, (nf, ns) in counts.items():
        if cid == 2:
            sc = synthetic_scene(2)
            assert sc["FluidBlocks"][0]["start"][0] == pytest.approx(0.05 + 0.005)
            continue
        ps = load_particles(SimConfig.from_dict(synthetic_scene(cid)))
        assert int((ps.is_dynamic == 1).sum()) == nf and int((ps.is_dynamic == 0).sum()) == ns


def test_fluid_lattice_at_voxel_centres_with_bounded_jitter():
    cfg = SimConfig.from_dict(synthetic_scene(0))
    ps = load_particles(cfg)
    x = ps.x[ps.is_dynamic == 1].astype(np.float64)
    d = 0.02
    ideal = cube_lattice((0.1 + 0.16 + 0.01, 0.1 + 0.01, 0.1 + 0.16 + 0.01), (0.32 - 0.01, 0.64 - 0.01, 0.32 - 0.01), d)
    assert ideal.shape == x.shape
    assert np.abs(x - ideal).max() <= 0.01 * d * (1 + 1e-3)
    assert np.abs(x - ideal).max() > 0.009 * d
    assert np.all(ps.v == 0.0)


def test_cube_shell_repeats_edges_like_reference():
    pts = cube_shell((0.5, 0.5, 0.5), (0.2, 0.2, 0.2), 1, 0.1)
    # 3 points per axis, 6 full 3x3 faces: 54 points, 26 distinct (edges and corners repeat)
    assert pts.shape == (54, 3)
    assert len({tuple(p) for p in np.round(pts, 6)}) == 26
    pts2 = cube_shell((0.5, 0.5, 0.5), (0.2, 0.2, 0.2), 2, 0.1)
    assert pts2.shape == (54 + 6 * 25, 3)


def test_boundary_walls_layers():
    pts = boundary_walls((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2, 0.5)
    assert pts.shape[0] == 6 * 9 + 6 * 25
    assert pts.min() == pytest.approx(-0.5) and pts.max() == pytest.approx(1.5)


def test_config_parsing_and_get_cfg(tmp_path):
    scene = {"Configuration": {"domainStart": [0.0, 0.0, 0.0], "domainEnd": [2.0, 3.0, 4.0], "particleRadius": 0.02,
                               "timeStepSize": 0.002, "density0": 1000, "gravitation": [0.0, -9.81, 0.0],
                               "viscosityMethod": 1, "viscosityCoeff": 0.1, "pressureMethod": 1,
                               "densityErrorTol": 1e-3, "omega": 10, "tol_pcg": 1e-6}}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(scene))
    cfg = SimConfig.from_json(str(p))
    assert cfg.domain_size == (2.0, 3.0, 4.0)
    assert cfg.support_radius == pytest.approx(0.08) and cfg.particle_diameter == pytest.approx(0.04)
    assert cfg.viscosity_method == 1 and cfg.dt == 0.002
    assert cfg.get_cfg("maxOuterIters", 1000) == 1000 and cfg.get_cfg("omega") == 10


def test_mass_convention():
    cfg = SimConfig.from_dict(synthetic_scene(0))
    ps = load_particles(cfg)
    m, minv = particle_masses(cfg, ps)
    dyn = ps.is_dynamic == 1
    m0 = np.float32(np.float32(0.8 * 0.02 ** 3) * np.float32(1000.0))
    assert np.all(m[dyn] == m0) and np.all(m[~dyn] == m0 * np.float32(10.0))
    assert np.all(minv[dyn] == np.float32(1.0) / m0) and np.all(minv[~dyn] == 0.0)


def test_mesh_blocks_are_rejected():
    cfg = SimConfig.from_dict({"Configuration": {}, "RigidBodies": [{"objectId": 1, "geometryFile": "x.obj"}]})
    with pytest.raises(NotImplementedError):
        load_particles(cfg)


@pytest.mark.parametrize("h", [0.02, 0.08, 0.1, 0.0625, 1.0])
def test_neighbour_predicate_band(h):
    """The build decides sqrtf(d2) < h without the square root outside [h2_lo, h2_hi), the
    bounds computed as in grid_build (h^2 (1 -/+ 2^-20), stepped one float outward).  Check
    every float32 d2 within h^2 (1 -/+ 2^-17) against the correctly rounded square root."""
    h = np.float32(h)
    hh = float(h) * float(h)
    lo = np.nextafter(np.float32(hh * (1.0 - 2.0 ** -20)), np.float32(0.0))
    hi = np.nextafter(np.float32(hh * (1.0 + 2.0 ** -20)), np.float32(1.0))
    a = np.float32(hh * (1.0 - 2.0 ** -17)).view(np.uint32)
    b = np.float32(hh * (1.0 + 2.0 ** -17)).view(np.uint32)
    d2 = np.arange(a, b + 1, dtype=np.uint32).view(np.float32)
    exact = np.sqrt(d2) < h
    assert np.all(exact[d2 < lo])
    assert not np.any(exact[d2 >= hi])
    assert np.count_nonzero((d2 >= lo) & (d2 < hi)) < 64   # the band needing the root is a few ulps


def test_synthetic_batch_scenes():
    """The batched workload's scene list: seeded counts and aspect ratios, scene-0 proportions,
    at least 256 fluid particles per scene (the library's per-scene minimum)."""
    from implicitvelocityprojection_amd.config import SimConfig
    from implicitvelocityprojection_amd.scenes import synthetic_batch
    a = synthetic_batch(n_scenes=8)
    b = synthetic_batch(n_scenes=8)
    assert a == b
    rng = np.random.default_rng(4)
    counts = rng.integers(2048, 32769, size=8)
    aspects = rng.uniform(0.5, 4.0, size=8)
    for s, scene in enumerate(a):
        cfg = SimConfig.from_dict(scene)
        fb = scene["FluidBlocks"][0]
        nx = int(round((counts[s] / aspects[s]) ** (1.0 / 3.0)))
        ny = int(round(counts[s] / nx ** 2))
        d = 0.02
        n_axis = [int(round((fb["end"][q] - fb["start"][q]) / d + 0.5)) for q in range(3)]
        assert n_axis == [nx, ny, nx]
        assert nx * ny * nx >= 256
        L = scene["GeneratedBodies"][0]["size"]
        assert np.allclose(L, [2 * nx * d, 1.5625 * ny * d, 2 * nx * d])
        assert fb["jitterSeed"] == 1000 + s
        assert cfg.viscosity_coeff == 0.0
