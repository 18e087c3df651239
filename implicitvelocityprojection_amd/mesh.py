"""Triangle meshes for scene loading: Wavefront OBJ reading and the placement transform of the
reference's ``SceneLoader.load_transformed_mesh`` (scripts/io/scene_loader.py:532-545), in
float64 on the host.  Voxelisation of the placed mesh runs in libivp (``voxelize``).

The reference loads meshes with trimesh (absent here); this reader keeps what its voxelisation
reads: the vertex positions referenced by faces (identical positions merged, as trimesh's
default processing merges them) and the triangles, polygons split into a fan from their first
corner.
"""

import ctypes

import numpy as np
import torch

from . import _native as N


class Mesh:
    """vertices: float64 (V, 3); faces: int64 (F, 3) indices into vertices."""

    def __init__(self, vertices, faces):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.int64).reshape(-1, 3)


def _index(tok, nv):
    """1-based OBJ vertex index (negative: relative to the end) of a face token 'v/vt/vn'."""
    i = int(tok.split("/")[0])
    if i < 0:
        i = nv + i
    else:
        i -= 1
    if not 0 <= i < nv:
        raise ValueError(f"OBJ face index {tok!r} out of range (vertices so far: {nv})")
    return i


def parse_obj(text):
    """Mesh from the text of a Wavefront OBJ file (v and f records; other records ignored)."""
    verts, faces = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
        elif parts[0] == "f":
            idx = [_index(t, len(verts)) for t in parts[1:]]
            if len(idx) < 3:
                raise ValueError(f"OBJ face with {len(idx)} corners")
            for k in range(1, len(idx) - 1):
                faces.append([idx[0], idx[k], idx[k + 1]])
    v = np.array(verts, dtype=np.float64).reshape(-1, 3)
    f = np.array(faces, dtype=np.int64).reshape(-1, 3)
    # keep referenced vertices, identical positions merged (first occurrence order)
    used = np.unique(f) if f.size else np.zeros(0, np.int64)
    uniq, first, inv = np.unique(v[used], axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    remap = np.full(v.shape[0], -1, np.int64)
    remap[used] = rank[inv.reshape(-1)]
    return Mesh(uniq[order], remap[f])


def load_obj(path):
    with open(path) as fh:
        return parse_obj(fh.read())


def rotation_matrix(angle, direction, point):
    """4x4 rotation by `angle` (radians) about the axis `direction` through `point`
    (Rodrigues' formula: R = cos I + (1 - cos) d d^T + sin [d]x; translation p - R p)."""
    d = np.asarray(direction[:3], dtype=np.float64)
    d = d / np.sqrt(np.dot(d, d))
    s, c = np.sin(angle), np.cos(angle)
    R = np.diag([c, c, c])
    R += np.outer(d, d) * (1.0 - c)
    ds = d * s
    R += np.array([[0.0, -ds[2], ds[1]], [ds[2], 0.0, -ds[0]], [-ds[1], ds[0], 0.0]])
    M = np.identity(4)
    M[:3, :3] = R
    p = np.asarray(point[:3], dtype=np.float64)
    M[:3, 3] = p - np.dot(R, p)
    return M


def _apply(M, v):
    """Homogeneous transform of (V, 3) points, as one matrix product."""
    h = np.column_stack((v, np.ones(len(v))))
    return np.ascontiguousarray(np.dot(M, h.T).T[:, :3])


def place(mesh, obj):
    """load_transformed_mesh (scene_loader.py:532-545): scale, then rotate about the vertex mean
    (when |angle| > 1e-6), then translate; float64."""
    v = mesh.vertices.copy()
    S = np.identity(4)
    S[:3, :3] *= np.asarray(obj.get("scale", [1, 1, 1]), dtype=np.float64)
    v = _apply(S, v)
    angle = obj.get("rotationAngle", 0) / 360 * 2 * np.pi
    if abs(angle) > 1e-6:
        v = _apply(rotation_matrix(angle, obj.get("rotationAxis", [0, 1, 0]), v.mean(axis=0)), v)
    v = v + np.array(obj.get("translation", [0, 0, 0]), dtype=np.float64)
    return Mesh(v, mesh.faces)


def voxelize(mesh, pitch, max_iter=15, device="cuda"):
    """Voxel centres of the solid bounded by `mesh` (libivp ivp_voxelize: trimesh's subdivide
    voxeliser and hole fill, scene_loader.py:565-572), float64 (N, 3) on `device`, in the
    voxel grid's x-major order."""
    dev = torch.device(device)
    lib = N.load()
    verts = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
    faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    count = ctypes.c_longlong(0)
    # the points fit in the occupied index box of the rounded vertices: with a modest box, one
    # call with that capacity; otherwise a counting call first
    cap = 0
    if faces.shape[0] and faces.min() >= 0 and faces.max() < verts.shape[0]:
        used = verts[np.unique(faces)]
        box = np.rint(used.max(axis=0) / pitch) - np.rint(used.min(axis=0) / pitch) + 1
        nvox = float(np.prod(box))
        if nvox <= float(1 << 26):
            cap = int(nvox)
    out = torch.zeros(cap, 3, dtype=torch.float64, device=dev)
    for _ in range(2):
        rc = lib.ivp_voxelize(verts.ctypes.data_as(ctypes.c_void_p), verts.shape[0],
                              faces.ctypes.data_as(ctypes.c_void_p), faces.shape[0], ctypes.c_double(pitch),
                              int(max_iter), ctypes.c_void_p(out.data_ptr()), ctypes.c_longlong(cap),
                              ctypes.byref(count), stream)
        if rc == N.IVP_ERR_CAPACITY and count.value > cap:
            cap = count.value
            out = torch.zeros(cap, 3, dtype=torch.float64, device=dev)
            continue
        N.check(rc, "ivp_voxelize")
        return out[:count.value]
    raise RuntimeError("ivp_voxelize: point count changed between calls")
