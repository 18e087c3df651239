"""Mesh voxelisation of the reference's scene loader, restated in numpy (TEST INFRASTRUCTURE ONLY:
only tests/ and bench.py's cpu_baseline leg load this module).

The reference (scripts/io/scene_loader.py:547-580, ``load_mesh``) calls trimesh:
``mesh.voxelized(pitch=a * particle_diameter, max_iter=15).fill()`` and takes ``.points``.
trimesh is not installed here.  This restates that published algorithm (trimesh's
``voxel.creation.voxelize_subdivide``, ``remesh.subdivide_to_size`` and ``voxel.morphology.fill``
with its default ``holes`` method, i.e. ``scipy.ndimage.binary_fill_holes``):

1. split every triangle with an edge longer than pitch / 2 into four at its edge midpoints
   ((a + b) / 2 in float64), level by level, at most ``max_iter`` times; the corners of the
   triangles that end below the limit are the surface samples;
2. a voxel is occupied when a sample rounds to it: ``round(v / pitch)`` (half to even);
3. the occupied set, in the box from its smallest to its largest index, is filled: every empty
   voxel not 6-connected to the box's border becomes occupied;
4. the points are ``(index - min) * pitch + min * pitch`` in float64, in x-major order.

Edge lengths are ``sqrt((dx^2 + dy^2) + dz^2)`` as numpy's sum over the last axis evaluates it.
Parity against trimesh itself is unpinned (no trimesh here); the tests pin this restatement by
closed forms (boxes), geometry (spheres, a hollow container) and subdivision invariants.
"""

import numpy as np


class MaxIterExceeded(ValueError):
    pass


def subdivide_samples(tri, max_edge, max_iter):
    """Corners of the final triangles of the recursive midpoint subdivision of `tri`
    (T, 3, 3) float64.  Returns (samples (S, 3), leaf triangle count, levels used)."""
    out = []
    leaves = 0
    level = 0
    for level in range(max_iter + 1):
        if tri.shape[0] == 0:
            break
        d = np.diff(tri[:, [0, 1, 2, 0]], axis=1)
        length = (d ** 2).sum(axis=2) ** 0.5
        long_ = (length > max_edge).any(axis=1)
        ok = tri[~long_]
        out.append(ok.reshape(-1, 3))
        leaves += ok.shape[0]
        if not long_.any():
            tri = tri[:0]
            break
        t = tri[long_]
        a, b, c = t[:, 0], t[:, 1], t[:, 2]
        ab, bc, ca = (a + b) / 2.0, (b + c) / 2.0, (c + a) / 2.0
        tri = np.concatenate([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                              np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)])
    if tri.shape[0]:
        raise MaxIterExceeded("max_iter exceeded")
    samples = np.concatenate(out) if out else np.zeros((0, 3))
    return samples, leaves, level


def fill_holes(occ):
    """scipy.ndimage.binary_fill_holes with the default (6-connected) structure."""
    from scipy import ndimage
    return ndimage.binary_fill_holes(occ)


def voxelize(vertices, faces, pitch, max_iter=15, edge_factor=2.0):
    """float64 (N, 3) voxel centres, x-major order (trimesh VoxelGrid.points after fill())."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if f.shape[0] == 0:
        return np.zeros((0, 3))
    samples, _, _ = subdivide_samples(v[f], pitch / edge_factor, max_iter)
    hit = np.round(samples / pitch).astype(np.int64)
    occupied = np.unique(hit, axis=0)
    lo = occupied.min(axis=0)
    shape = occupied.max(axis=0) - lo + 1
    dense = np.zeros(tuple(shape), dtype=bool)
    rel = occupied - lo
    dense[rel[:, 0], rel[:, 1], rel[:, 2]] = True
    filled = fill_holes(dense)
    idx = np.column_stack(np.nonzero(filled))
    origin = lo * pitch
    return idx.astype(np.float64) * pitch + origin
