"""The neighbour build's scan by cells (IVP_OPT_BUILD_LANES 64) on inputs that reach its rare
paths: a packed cluster whose cells see more than 256 candidates (several candidate chunks per
segment) above a static floor, a lattice of spacing h/2 (pairs within float32 rounding of h,
decided by the correctly rounded root) with no static particles and a ragged fluid count, and
sparse fluid particles inside a static block (rows with static neighbours only).  Rows, density
and k must equal the one-lane build's (k_build: scan_ids, the reference's visiting order)
bitwise, and every row's neighbour count (dynamic and static) a brute-force float32 search over
all pairs."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

H = np.float32(0.04)


def _cluster():
    """24^3 fluid particles at spacing h/2.6 (about 74 neighbours, ~470 candidates per cell) on a
    two-layer static floor."""
    d = np.float32(H / np.float32(2.6))
    g = np.arange(24, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    fl = np.stack([X, Y, Z], -1).reshape(-1, 3) * d + np.float32(0.3)
    f = np.arange(-12, 36, dtype=np.float32)
    A, B = np.meshgrid(f, f, indexing="ij")
    fl0 = np.stack([A.ravel(), np.zeros(A.size, np.float32), B.ravel()], -1) * np.float32(H / 2)
    floor = np.concatenate([fl0 + np.float32([0.3, 0.3 - H / 2, 0.3]), fl0 + np.float32([0.3, 0.3 - H, 0.3])])
    x = np.concatenate([fl, floor]).astype(np.float32)
    dyn = np.concatenate([np.ones(len(fl), np.int32), np.zeros(len(floor), np.int32)])
    return x, dyn, 120


def _lattice():
    """4000 fluid particles on a lattice of spacing h/2 (a ragged count: not a multiple of 16 or
    64), no static particles."""
    d = np.float32(H / 2)
    g = np.arange(16, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = (np.stack([X, Y, Z], -1).reshape(-1, 3) * d + np.float32(0.25))[:4000].astype(np.float32)
    return x, np.ones(len(x), np.int32), 60


def _sparse_in_static():
    """A static block at spacing h/2 with 600 fluid particles spread through it at spacing 1.5 h
    (each row's neighbours are static only)."""
    d = np.float32(H / 2)
    g = np.arange(30, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    st = np.stack([X, Y, Z], -1).reshape(-1, 3) * d + np.float32(0.2)
    q = np.arange(10, dtype=np.float32) * np.float32(1.5 * H) + np.float32(0.2 + H / 4)
    A, B, C = np.meshgrid(q[:6], q[:10], q[:10], indexing="ij")
    fl = np.stack([A.ravel(), B.ravel(), C.ravel()], -1).astype(np.float32)
    x = np.concatenate([fl, st]).astype(np.float32)
    dyn = np.concatenate([np.ones(len(fl), np.int32), np.zeros(len(st), np.int32)])
    return x, dyn, 60


def _brute_counts(x, dyn, h):
    """Neighbours (dynamic and static) of every dynamic particle: j != i,
    sqrtf((dx*dx + dy*dy) + dz*dz) < h."""
    rows = np.flatnonzero(dyn == 1)
    out = np.zeros(len(rows), np.int64)
    for s in range(0, len(rows), 256):
        ri = rows[s:s + 256]
        r = x[ri, None, :] - x[None, :, :]
        d2 = (r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1]) + r[..., 2] * r[..., 2]
        hit = np.sqrt(d2) < h
        hit[np.arange(len(ri)), ri] = False
        out[s:s + 256] = hit.sum(1)
    return out


@pytest.mark.parametrize("case", [_cluster, _lattice, _sparse_in_static], ids=["cluster", "lattice", "sparse"])
def test_cell_scan_matches_one_lane_build(case):
    from implicitvelocityprojection_amd import _native as N
    from implicitvelocityprojection_amd.context import Context
    x, dyn, C = case()
    n = len(x)
    print(f"cell scan case {case.__name__}: n {n}, dynamic {int(dyn.sum())}, cache {C}", flush=True)
    dev = torch.device("cuda")
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    m = np.full(n, np.float32(0.8 * 0.02 ** 3 * 1000.0), np.float32)
    m_inv = np.where(dyn == 1, np.float32(1.0) / m, np.float32(0.0)).astype(np.float32)
    ctx = Context(float(H), (0.0, 0.0, 0.0), (1.5, 1.5, 1.5), n, cache_size=C)
    ctx.set_particles(t(x), t(np.zeros_like(x)), t(m), t(m_inv), t(np.full(n, 1000.0, np.float32)), t(dyn, torch.int32),
                      t(dyn, torch.int32))
    out = {}
    for lanes in (1, 64):
        assert N.check(ctx.lib.ivp_set_option(ctx.h, N.OPT_BUILD_LANES, lanes)) == 0
        assert ctx.neighbors_build() == 0
        count = torch.zeros(n, dtype=torch.int32, device=dev)
        ids = torch.zeros(n * C, dtype=torch.int32, device=dev)
        gradW = torch.zeros(n * C * 3, dtype=torch.float32, device=dev)
        ctx.neighbors_export(count, ids, gradW)
        dens = ctx.read(2, torch.zeros(n, device=dev)).cpu().numpy()
        k = ctx.read(3, torch.zeros(n, device=dev)).cpu().numpy()
        out[lanes] = (count.cpu().numpy(), ids.cpu().numpy(), dens, k)
    c1, i1, d1, k1 = out[1]
    c64, i64, d64, k64 = out[64]
    assert np.array_equal(c1, c64)
    mask = np.zeros(n * C, bool)
    for i in range(n):
        mask[i * C:i * C + c1[i]] = True
    assert np.array_equal(i1[mask], i64[mask])
    assert np.array_equal(d1.view(np.uint32), d64.view(np.uint32))
    assert np.array_equal(k1.view(np.uint32), k64.view(np.uint32))
    # counts in user order: the dynamic particles' rows against brute force
    bc = _brute_counts(x, dyn, H)
    assert np.array_equal(c64[dyn == 1].astype(np.int64), bc)
    assert bc.max() <= C
    if case is _cluster:
        assert bc.mean() > 60   # candidates well past 256 per cell
    if case is _sparse_in_static:   # static neighbours only
        nf = int(dyn.sum())   # the fluid particles come first in user order
        assert bc.min() > 0 and (i64[:nf * C][mask[:nf * C]] >= nf).all()
    ctx.close()
