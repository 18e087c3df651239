"""GPU mesh voxelisation (ivp_voxelize) against the numpy restatement of the reference's
trimesh voxeliser (oracle/voxelize.py): the same points, bit for bit, in the same order."""

import math

import numpy as np
import pytest

from oracle import voxelize as OV
from tests import meshes as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

P = 0.06


def _cases():
    yield "box", M.box((0.313, -0.071, 1.004), (0.911, 0.4402, 1.3777)), P
    v, f = M.icosphere(0.61, 4, (0.2, 0.35, -0.15))
    yield "sphere", (v, f), P
    yield "sphere_rotated", (M.rotate(v, 0.7, (1.0, 2.0, 0.5), (0.1, 0.0, 0.3)), f), P
    yield "container", M.open_container((0.0, 0.0, 0.0), (1.2, 0.9, 1.2), 0.13), P
    yield "torus", M.torus(0.5, 0.15, 48, 16, (1.0, 2.0, 3.0)), P
    yield "hollow_ball", M.merge(M.icosphere(0.6, 4), M.icosphere(0.35, 3, inward=True)), P
    yield "big_box_fine", M.box((-1.3, 0.2, -0.7), (1.9, 1.0, 0.5)), 0.02


@pytest.mark.parametrize("name,mesh,pitch", list(_cases()), ids=[c[0] for c in _cases()])
def test_voxelize_matches_oracle_bitwise(name, mesh, pitch):
    from implicitvelocityprojection_amd.mesh import Mesh, voxelize
    v, f = mesh
    print(f"voxelize {name}: {v.shape[0]} vertices, {f.shape[0]} faces, pitch {pitch}", flush=True)
    want = OV.voxelize(v, f, pitch)
    got = voxelize(Mesh(v, f), pitch).cpu().numpy()
    assert got.shape == want.shape and got.dtype == np.float64
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_voxelize_errors_and_capacity():
    import ctypes
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.mesh import Mesh, voxelize
    v, f = M.box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    # more levels needed than allowed: trimesh's "max_iter exceeded"
    with pytest.raises(N.IVPError):
        voxelize(Mesh(v, f), P, max_iter=2)
    # a face index out of range is refused on the host
    with pytest.raises(N.IVPError):
        voxelize(Mesh(v, np.array([[0, 1, 99]])), P)
    # no faces: no points
    assert voxelize(Mesh(v, np.zeros((0, 3), np.int64)), P).shape == (0, 3)
    # too small an output buffer: the count comes back, nothing is written
    lib = N.load()
    verts = np.ascontiguousarray(v)
    faces = np.ascontiguousarray(f, dtype=np.int32)
    out = torch.full((5, 3), -7.0, dtype=torch.float64, device="cuda")
    n = ctypes.c_longlong(0)
    rc = lib.ivp_voxelize(verts.ctypes.data_as(ctypes.c_void_p), 8, faces.ctypes.data_as(ctypes.c_void_p), 12,
                          ctypes.c_double(P), 15, ctypes.c_void_p(out.data_ptr()), 5, ctypes.byref(n), None)
    assert rc == N.IVP_ERR_CAPACITY and n.value == 18 ** 3
    assert torch.all(out == -7.0)


def test_voxelize_level_limit_matches_oracle():
    """The subdivision depth limit: a mesh that needs exactly L levels voxelises with
    max_iter = L (bitwise as the oracle) and is refused with L - 1, as the oracle refuses it."""
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.mesh import Mesh, voxelize
    for v, f in (M.icosphere(0.7, 1), M.box((0.0, 0.0, 0.0), (0.02, 0.025, 0.01)),
                 M.box((0.0, 0.0, 0.0), (0.01, 0.01, 0.01))):
        _, _, L = OV.subdivide_samples(v[f], P / 2, 15)
        want = OV.voxelize(v, f, P, max_iter=L)
        got = voxelize(Mesh(v, f), P, max_iter=L).cpu().numpy()
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), L
        if L > 0:
            with pytest.raises(OV.MaxIterExceeded):
                OV.voxelize(v, f, P, max_iter=L - 1)
            with pytest.raises(N.IVPError):
                voxelize(Mesh(v, f), P, max_iter=L - 1)
        else:
            assert got.shape[0] >= 1
