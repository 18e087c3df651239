"""ivp_voxelize_noalloc (the voxeliser over a fixed device array, no allocation) against
ivp_voxelize on the same meshes: the same point count and the same points, bit for bit, on two
calls in a row (the second mesh smaller than the first); meshes whose voxel grid, or one of
whose subdivision levels, exceeds the array are refused with IVP_ERR_CAPACITY, and the next call
works again."""

import ctypes

import numpy as np
import pytest

from tests import meshes as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

P = 0.06


def _call(fn, v, f, pitch, cap, out):
    verts = np.ascontiguousarray(v, dtype=np.float64)
    faces = np.ascontiguousarray(f, dtype=np.int32)
    n = ctypes.c_longlong(-1)
    ptr = ctypes.c_void_p(out.data_ptr()) if out is not None else None
    rc = fn(verts.ctypes.data_as(ctypes.c_void_p), verts.shape[0], faces.ctypes.data_as(ctypes.c_void_p),
            faces.shape[0], ctypes.c_double(pitch), 15, ptr, cap, ctypes.byref(n),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, n.value


def _both(lib, v, f, pitch):
    from implicitvelocityprojection_amd import _native as N
    rc, n = _call(lib.ivp_voxelize, v, f, pitch, 0, None)
    assert rc == N.IVP_ERR_CAPACITY and n > 0
    want = torch.full((n, 3), -1.0, dtype=torch.float64, device="cuda")
    got = torch.full((n, 3), -2.0, dtype=torch.float64, device="cuda")
    assert _call(lib.ivp_voxelize, v, f, pitch, n, want) == (0, n)
    assert _call(lib.ivp_voxelize_noalloc, v, f, pitch, n, got) == (0, n)
    torch.cuda.synchronize()
    w, g = want.cpu().numpy(), got.cpu().numpy()
    assert np.array_equal(w.view(np.uint64), g.view(np.uint64))
    return n


def test_voxelize_noalloc_matches_voxelize():
    from implicitvelocityprojection_amd import _native as N
    lib = N.load()
    v, f = M.icosphere(0.61, 4, (0.2, 0.35, -0.15))
    print("noalloc: sphere, then a smaller box, then a grid too large, then the box again", flush=True)
    n1 = _both(lib, v, f, P)
    vb, fb = M.box((0.313, -0.071, 1.004), (0.911, 0.4402, 1.3777))
    n2 = _both(lib, vb, fb, P)
    assert 0 < n2 < n1
    out = torch.zeros((8, 3), dtype=torch.float64, device="cuda")
    # 501^3 voxels: a grid larger than the 64 MiB array
    vl, fl = M.box((0.0, 0.0, 0.0), (5.0, 5.0, 5.0))
    assert _call(lib.ivp_voxelize_noalloc, vl, fl, 0.01, 8, out)[0] == N.IVP_ERR_CAPACITY
    # 401^3 voxels fit, but a subdivision level outgrows the rest of the array
    vl, fl = M.box((0.0, 0.0, 0.0), (4.0, 4.0, 4.0))
    assert _call(lib.ivp_voxelize_noalloc, vl, fl, 0.01, 8, out)[0] == N.IVP_ERR_CAPACITY
    assert _both(lib, vb, fb, P) == n2
