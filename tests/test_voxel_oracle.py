"""Pins of the voxelisation oracle (oracle/voxelize.py, the restated trimesh voxeliser of
scene_loader.py:547-580) and of the host-side OBJ reader and placement (mesh.py)."""

import math

import numpy as np
import pytest

from implicitvelocityprojection_amd.mesh import parse_obj, place, rotation_matrix
from oracle import voxelize as OV
from tests import meshes as M

P = 0.06


def _idx(points, pitch=P):
    return np.round(points / pitch).astype(np.int64)


def test_box_fills_its_rounded_index_box():
    """An axis-aligned box: the surface samples round onto the faces of the index box
    [round(lo/p), round(hi/p)], and the fill makes it solid."""
    lo, hi = np.array([0.313, -0.071, 1.004]), np.array([0.911, 0.4402, 1.3777])
    v, f = M.box(lo, hi)
    pts = OV.voxelize(v, f, P)
    a, b = np.round(lo / P).astype(int), np.round(hi / P).astype(int)
    assert pts.shape[0] == int(np.prod(b - a + 1))
    idx = _idx(pts)
    grid = np.stack(np.meshgrid(*[np.arange(a[k], b[k] + 1) for k in range(3)], indexing="ij"), -1).reshape(-1, 3)
    np.testing.assert_array_equal(idx, grid)          # also the x-major order
    np.testing.assert_allclose(pts, grid * P, rtol=0, atol=1e-12)


def test_subdivision_leaves_respect_the_edge_limit_and_keep_the_area():
    v, f = M.icosphere(0.7, 1)
    tri = v[f]
    samples, leaves, levels = OV.subdivide_samples(tri, P / 2, 15)
    assert samples.shape[0] == 3 * leaves
    # each level quarters a triangle: equal-sized faces all stop at the same level here
    assert leaves == tri.shape[0] * 4 ** levels
    s = samples.reshape(-1, 3, 3)
    e = np.linalg.norm(np.diff(s[:, [0, 1, 2, 0]], axis=1), axis=2)
    assert e.max() <= P / 2
    area = lambda t: 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1).sum()
    assert area(s) == pytest.approx(area(tri), rel=1e-12)
    # the original corners are among the samples
    orig = {tuple(x) for x in v}
    assert orig <= {tuple(x) for x in samples}
    with pytest.raises(OV.MaxIterExceeded):
        OV.subdivide_samples(tri, P / 2, levels - 1)


def test_sphere_volume_and_extent():
    R, c = 0.61, np.array([0.2, 0.35, -0.15])
    v, f = M.icosphere(R, 4, c)
    pts = OV.voxelize(v, f, P)
    # surface voxels reach at most half a voxel diagonal beyond the surface, the interior is solid
    vol = pts.shape[0] * P ** 3
    ball = lambda r: 4.0 / 3.0 * math.pi * r ** 3
    assert ball(R - P * math.sqrt(3) / 2) < vol < ball(R + P * math.sqrt(3) / 2)
    d = np.linalg.norm(pts - c, axis=1)
    assert d.max() <= R + P * math.sqrt(3) / 2 + 1e-9
    # every lattice centre well inside the sphere is present
    a = np.floor((c - R) / P).astype(int)
    b = np.ceil((c + R) / P).astype(int)
    g = np.stack(np.meshgrid(*[np.arange(a[k], b[k] + 1) for k in range(3)], indexing="ij"), -1).reshape(-1, 3)
    inner = g[np.linalg.norm(g * P - c, axis=1) < R - P * math.sqrt(3)]
    have = {tuple(x) for x in _idx(pts)}
    assert all(tuple(x) in have for x in inner)


def test_open_container_stays_hollow_and_torus_keeps_its_hole():
    v, f = M.open_container((0.0, 0.0, 0.0), (1.2, 0.9, 1.2), 0.13)
    idx = {tuple(x) for x in _idx(OV.voxelize(v, f, P))}
    assert (10, 8, 10) not in idx and (10, 14, 10) not in idx        # inside the cavity
    assert (10, 0, 10) in idx and (0, 8, 10) in idx                   # floor and wall
    v, f = M.torus(0.5, 0.15, 48, 16)
    idx = {tuple(x) for x in _idx(OV.voxelize(v, f, P))}
    assert (0, 0, 0) not in idx and (8, 0, 0) in idx


def test_enclosed_cavity_is_filled():
    """binary_fill_holes semantics: a cavity with no 6-connected path to the border is filled
    (a sphere inside a sphere becomes one solid ball)."""
    outer = M.icosphere(0.6, 4)
    inner = M.icosphere(0.35, 3, inward=True)
    v, f = M.merge(outer, inner)
    a = OV.voxelize(v, f, P)
    b = OV.voxelize(*outer, P)
    np.testing.assert_array_equal(a, b)


def _bfs_fill(occ):
    """Reference flood fill: empty voxels 6-connected to the border stay empty."""
    from collections import deque
    out = np.zeros_like(occ)
    q = deque()
    for idx in np.argwhere(~occ):
        if any(idx[k] in (0, occ.shape[k] - 1) for k in range(3)):
            out[tuple(idx)] = True
            q.append(tuple(idx))
    while q:
        x = q.popleft()
        for k in range(3):
            for s in (-1, 1):
                y = list(x)
                y[k] += s
                y = tuple(y)
                if all(0 <= y[j] < occ.shape[j] for j in range(3)) and not occ[y] and not out[y]:
                    out[y] = True
                    q.append(y)
    return ~out


def test_fill_is_six_connected_flood_from_the_border():
    rng = np.random.default_rng(3)
    for _ in range(20):
        occ = rng.uniform(size=(7, 6, 8)) < 0.45
        np.testing.assert_array_equal(OV.fill_holes(occ), _bfs_fill(occ))
    # a cavity that touches the outside only through an edge-diagonal gap is enclosed
    occ = np.ones((5, 5, 5), bool)
    occ[2, 2, 2] = False
    occ[1, 1, 2] = False          # diagonal neighbour of the cavity, not 6-adjacent to it
    occ[0, 1, 2] = False          # ... and connected to the border
    assert OV.fill_holes(occ)[2, 2, 2] and not OV.fill_holes(occ)[1, 1, 2]


def test_translation_by_whole_voxels_shifts_the_indices():
    v, f = M.icosphere(0.4, 3, (0.013, 0.021, -0.007))
    a = _idx(OV.voxelize(v, f, P))
    b = _idx(OV.voxelize(v + np.array([3, -2, 5]) * P, f, P))
    np.testing.assert_array_equal(b, a + np.array([3, -2, 5]))


OBJ = """# comment
o Cube
mtllib x.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 5 5 5
v 1 0 0
vn 0 0 1
vt 0.5 0.5
usemtl m
s 1
f 1/1/1 2/1/1 3/1/1 4/1/1
f -6//1 6//1 -3//1
"""


def test_obj_reader():
    m = parse_obj(OBJ)
    # quad fanned from its first corner; unreferenced vertex 5 dropped; duplicate of vertex 2 merged
    np.testing.assert_array_equal(m.vertices, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(m.faces, [[0, 1, 2], [0, 2, 3], [0, 1, 3]])
    with pytest.raises(ValueError):
        parse_obj("v 0 0 0\nf 1 2 3\n")


def test_placement_scale_rotate_translate():
    v, f = M.box((-1.0, 0.0, -2.0), (1.0, 1.0, 2.0))
    from implicitvelocityprojection_amd.mesh import Mesh
    m = place(Mesh(v, f), {"scale": [2.0, 3.0, 0.5], "rotationAxis": [0, 1, 0], "rotationAngle": 90,
                           "translation": [10.0, 0.0, 0.0]})
    # scaled box spans x in [-2, 2], y in [0, 3], z in [-1, 1]; its vertex mean is (0, 1.5, 0);
    # +90 degrees about y maps (x, z) -> (z, -x)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    np.testing.assert_allclose(lo, [9.0, 0.0, -2.0], atol=1e-12)
    np.testing.assert_allclose(hi, [11.0, 3.0, 2.0], atol=1e-12)
    R = rotation_matrix(math.pi / 2, [0, 1, 0], [0.0, 0.0, 0.0])[:3, :3]
    np.testing.assert_allclose(R @ [1.0, 0.0, 0.0], [0.0, 0.0, -1.0], atol=1e-15)
    # angles within 1e-6 rad of zero skip the rotation
    m2 = place(Mesh(v, f), {"rotationAngle": 1e-5})
    np.testing.assert_array_equal(m2.vertices, v)
