// ivp_voxel.hip -- mesh voxelisation for scene loading (scene_loader.py:547-580: the reference's
// trimesh call mesh.voxelized(pitch, max_iter).fill().points).
//
// 1. Subdivision, level-synchronous: one thread per triangle of the level.  A triangle whose
//    three edges are at most pitch / 2 (sqrt((dx^2 + dy^2) + dz^2), float64) marks the voxels of
//    its corners, round(v / pitch) half to even; otherwise it forms its four midpoint children
//    ((a + b) / 2), marks the corners of those that already meet the limit (the test the next
//    level would apply, on the same coordinates) and appends only the others to the next level,
//    through one atomic per triangle.  The deepest level is thus never written out.  Marking is a
//    set operation, so the order in which triangles land in a level does not matter.
// 2. Hole fill: every empty voxel 6-connected to the grid's border is outside; the rest is
//    filled.  Outside status spreads along lines: each sweep kernel gives every line of one axis
//    to one thread, which walks it forwards and backwards; rounds of x, y and z sweeps repeat
//    until one changes nothing.
// 3. Points in x-major order of the grid (z fastest): per-chunk counts, an exclusive scan of the
//    chunk counts, and a second pass that writes (index - min) * pitch + min * pitch in float64.
//
// The grid spans the rounded corners of the input triangles: every sample of the subdivision is
// a convex combination of its triangle's corners, and rounding is monotonic.
#include <climits>
#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "ivp_internal.h"
#include "ivp_launch.h"

using namespace ivp;

namespace {

struct Tri {
    double v[9];   // corners a, b, c
};

struct VoxGrid {
    long long lo[3];
    long long dims[3];
    double pitch;
};

__device__ __forceinline__ double edge_len(const Tri& t, int p, int q) {
    double dx = t.v[3 * q + 0] - t.v[3 * p + 0];
    double dy = t.v[3 * q + 1] - t.v[3 * p + 1];
    double dz = t.v[3 * q + 2] - t.v[3 * p + 2];
    return sqrt((dx * dx + dy * dy) + dz * dz);
}

__device__ __forceinline__ bool is_long(const Tri& t, double max_edge) {
    return edge_len(t, 0, 1) > max_edge || edge_len(t, 1, 2) > max_edge || edge_len(t, 2, 0) > max_edge;
}

__device__ __forceinline__ void mark_corners(const Tri& t, const VoxGrid& g, uint8_t* __restrict__ grid,
                                             int* __restrict__ flags) {
    for (int c = 0; c < 3; ++c) {
        long long ix = (long long)rint(t.v[3 * c + 0] / g.pitch) - g.lo[0];
        long long iy = (long long)rint(t.v[3 * c + 1] / g.pitch) - g.lo[1];
        long long iz = (long long)rint(t.v[3 * c + 2] / g.pitch) - g.lo[2];
        if (ix < 0 || iy < 0 || iz < 0 || ix >= g.dims[0] || iy >= g.dims[1] || iz >= g.dims[2]) {
            atomicOr(flags, 1);
            continue;
        }
        grid[(ix * g.dims[1] + iy) * g.dims[2] + iz] = 1;
    }
}

// flags: bit 0 a sample outside the grid (cannot happen), bit 1 next-level capacity exceeded.
// mode: LV_SPLIT stores the children that need more levels; LV_LAST_CHILDREN (the children are
// on the last level allowed) and LV_NO_SPLIT (this level is the last allowed) only count the
// triangles that would need one more level, which the caller reports as "max_iter exceeded".
enum { LV_SPLIT = 0, LV_LAST_CHILDREN = 1, LV_NO_SPLIT = 2 };
__global__ __launch_bounds__(BLOCK) void k_vox_level(const Tri* __restrict__ in, long long n, double max_edge,
                                                     VoxGrid g, uint8_t* __restrict__ grid, Tri* __restrict__ next,
                                                     unsigned long long* __restrict__ next_count,
                                                     unsigned long long next_cap, int mode, int* __restrict__ flags) {
    long long t = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n) return;
    const Tri tr = in[t];
    if (!is_long(tr, max_edge)) {
        mark_corners(tr, g, grid, flags);
        return;
    }
    if (mode == LV_NO_SPLIT) {
        atomicAdd(next_count, 1ull);
        return;
    }
    double m[9];   // midpoints ab, bc, ca
    for (int a = 0; a < 3; ++a) {
        m[0 + a] = (tr.v[0 + a] + tr.v[3 + a]) / 2.0;
        m[3 + a] = (tr.v[3 + a] + tr.v[6 + a]) / 2.0;
        m[6 + a] = (tr.v[6 + a] + tr.v[0 + a]) / 2.0;
    }
    Tri ch[4];
    for (int a = 0; a < 3; ++a) {
        ch[0].v[a] = tr.v[a];     ch[0].v[3 + a] = m[a];       ch[0].v[6 + a] = m[6 + a];
        ch[1].v[a] = m[a];        ch[1].v[3 + a] = tr.v[3 + a]; ch[1].v[6 + a] = m[3 + a];
        ch[2].v[a] = m[6 + a];    ch[2].v[3 + a] = m[3 + a];   ch[2].v[6 + a] = tr.v[6 + a];
        ch[3].v[a] = m[a];        ch[3].v[3 + a] = m[3 + a];   ch[3].v[6 + a] = m[6 + a];
    }
    bool keep[4];
    int nk = 0;
    for (int k = 0; k < 4; ++k) {
        keep[k] = is_long(ch[k], max_edge);
        if (!keep[k]) mark_corners(ch[k], g, grid, flags);
        nk += keep[k] ? 1 : 0;
    }
    if (nk == 0) return;
    unsigned long long base = atomicAdd(next_count, (unsigned long long)nk);
    if (mode != LV_SPLIT) return;
    if (base + nk > next_cap) {
        atomicOr(flags, 2);
        return;
    }
    for (int k = 0; k < 4; ++k)
        if (keep[k]) next[base++] = ch[k];
}

// grid values: 0 empty, 1 occupied, 2 outside (empty and 6-connected to the border)
__global__ __launch_bounds__(BLOCK) void k_vox_border(uint8_t* __restrict__ grid, VoxGrid g, long long nvox) {
    long long f = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (f >= nvox) return;
    long long iz = f % g.dims[2];
    long long r = f / g.dims[2];
    long long iy = r % g.dims[1];
    long long ix = r / g.dims[1];
    bool border = ix == 0 || iy == 0 || iz == 0 || ix == g.dims[0] - 1 || iy == g.dims[1] - 1 || iz == g.dims[2] - 1;
    if (border && grid[f] == 0) grid[f] = 2;
}

// One thread per line along `axis`: spread "outside" forwards then backwards along the line.
__global__ __launch_bounds__(BLOCK) void k_vox_sweep(uint8_t* __restrict__ grid, VoxGrid g, int axis, long long nlines,
                                                     int* __restrict__ changed) {
    long long l = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (l >= nlines) return;
    const long long d0 = g.dims[0], d1 = g.dims[1], d2 = g.dims[2];
    long long start, step, len;
    if (axis == 2) {          // lines along z: l = ix * d1 + iy
        start = l * d2; step = 1; len = d2;
    } else if (axis == 1) {   // lines along y: l = ix * d2 + iz
        start = (l / d2) * d1 * d2 + (l % d2); step = d2; len = d1;
    } else {                  // lines along x: l = iy * d2 + iz
        start = l; step = d1 * d2; len = d0;
    }
    int any = 0;
    bool out = false;
    for (long long s = 0; s < len; ++s) {
        long long f = start + s * step;
        uint8_t v = grid[f];
        if (v == 0 && out) { grid[f] = 2; v = 2; any = 1; }
        out = v == 2;
    }
    out = false;
    for (long long s = len - 1; s >= 0; --s) {
        long long f = start + s * step;
        uint8_t v = grid[f];
        if (v == 0 && out) { grid[f] = 2; v = 2; any = 1; }
        out = v == 2;
    }
    if (any) atomicOr(changed, 1);
}

constexpr int VOX_PER_THREAD = 16;
constexpr long long VOX_CHUNK = (long long)BLOCK * VOX_PER_THREAD;

__device__ __forceinline__ int block_excl_scan_int(int v, int* total) {
    __shared__ int wsum[WAVES_PER_BLOCK];
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) wsum[w] = x;
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < WAVES_PER_BLOCK; ++q) {
        if (q < w) off += wsum[q];
        tot += wsum[q];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// Filled voxels (occupied or enclosed) per chunk of VOX_CHUNK voxels in flat x-major order
__global__ __launch_bounds__(BLOCK) void k_vox_count(const uint8_t* __restrict__ grid, long long nvox,
                                                     int* __restrict__ chunk_count) {
    long long base = (long long)blockIdx.x * VOX_CHUNK + (long long)threadIdx.x * VOX_PER_THREAD;
    int c = 0;
    for (int k = 0; k < VOX_PER_THREAD; ++k) {
        long long f = base + k;
        if (f < nvox && grid[f] != 2) c++;
    }
    int tot;
    (void)block_excl_scan_int(c, &tot);
    if (threadIdx.x == 0) chunk_count[blockIdx.x] = tot;
}

__global__ __launch_bounds__(BLOCK) void k_vox_emit(const uint8_t* __restrict__ grid, VoxGrid g, long long nvox,
                                                    const long long* __restrict__ chunk_off, double ox, double oy,
                                                    double oz, double* __restrict__ points) {
    long long base = (long long)blockIdx.x * VOX_CHUNK + (long long)threadIdx.x * VOX_PER_THREAD;
    int c = 0;
    for (int k = 0; k < VOX_PER_THREAD; ++k) {
        long long f = base + k;
        if (f < nvox && grid[f] != 2) c++;
    }
    int tot;
    long long o = chunk_off[blockIdx.x] + block_excl_scan_int(c, &tot);
    for (int k = 0; k < VOX_PER_THREAD; ++k) {
        long long f = base + k;
        if (f >= nvox || grid[f] == 2) continue;
        long long iz = f % g.dims[2];
        long long r = f / g.dims[2];
        long long iy = r % g.dims[1];
        long long ix = r / g.dims[1];
        points[3 * o + 0] = (double)ix * g.pitch + ox;
        points[3 * o + 1] = (double)iy * g.pitch + oy;
        points[3 * o + 2] = (double)iz * g.pitch + oz;
        o++;
    }
}

// ---- allocation-free variant: kernels over a fixed scratch array ----------------------------

// Level-0 triangles from device copies of the vertex and face arrays (faces checked on the host)
__global__ __launch_bounds__(BLOCK) void k_vox_tris(const double* __restrict__ verts, const int* __restrict__ faces,
                                                    long long nf, Tri* __restrict__ out) {
    long long f = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (f >= nf) return;
    Tri t;
    for (int c = 0; c < 3; ++c) {
        const long long vi = faces[3 * f + c];
        for (int a = 0; a < 3; ++a) t.v[3 * c + a] = verts[3 * vi + a];
    }
    out[f] = t;
}

// Zero `words` 16-byte words (clears the voxel grid and the counters with vector stores)
__global__ __launch_bounds__(BLOCK) void k_vox_zero(uint4* __restrict__ p, long long words) {
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < words) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Exclusive scan of the per-chunk counts in one block: off[k] = sum of count[0..k), *total = sum
__global__ __launch_bounds__(BLOCK) void k_vox_chunk_scan(const int* __restrict__ count, long long n,
                                                          long long* __restrict__ off, long long* __restrict__ total) {
    long long run = 0;
    for (long long base = 0; base < n; base += BLOCK) {
        const long long k = base + threadIdx.x;
        const int c = k < n ? count[k] : 0;
        int tile;
        const int ex = block_excl_scan_int(c, &tile);
        if (k < n) off[k] = run + ex;
        run += tile;
    }
    if (threadIdx.x == 0) *total = run;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t bytes) {
        if (p) { (void)hipFree(p); p = nullptr; }
        return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? IVP_OK : IVP_ERR_CAPACITY;
    }
};

inline unsigned grid_blocks(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

extern "C" int ivp_voxelize(const double* vertices, int n_vertices, const int* faces, int n_faces, double pitch,
                            int max_iter, double* points, long long capacity, long long* n_points, void* stream) {
    if (!n_points) return IVP_ERR_ARG;
    *n_points = 0;
    if (n_vertices < 0 || n_faces < 0 || !(pitch > 0.0) || !isfinite(pitch) || max_iter < 0 || max_iter > 30)
        return IVP_ERR_ARG;
    if (n_faces == 0) return IVP_OK;
    if (!vertices || !faces) return IVP_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    // level-0 triangles and the grid bounds (rounded corners), on the host
    std::vector<Tri> tri((size_t)n_faces);
    long long lo[3] = {LLONG_MAX, LLONG_MAX, LLONG_MAX}, hi[3] = {LLONG_MIN, LLONG_MIN, LLONG_MIN};
    for (int f = 0; f < n_faces; ++f) {
        for (int c = 0; c < 3; ++c) {
            int vi = faces[3 * f + c];
            if (vi < 0 || vi >= n_vertices) return IVP_ERR_ARG;
            for (int a = 0; a < 3; ++a) {
                double x = vertices[3 * vi + a];
                if (!isfinite(x)) return IVP_ERR_ARG;
                double q = rint(x / pitch);
                if (!(fabs(q) < 1e15)) return IVP_ERR_ARG;
                long long iq = (long long)q;
                lo[a] = std::min(lo[a], iq);
                hi[a] = std::max(hi[a], iq);
                tri[f].v[3 * c + a] = x;
            }
        }
    }
    VoxGrid g;
    g.pitch = pitch;
    long long nvox = 1;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo[a];
        g.dims[a] = hi[a] - lo[a] + 1;
        if (g.dims[a] > (1ll << 20)) return IVP_ERR_CAPACITY;
        nvox *= g.dims[a];
        if (nvox > (1ll << 34)) return IVP_ERR_CAPACITY;
    }
    DevBuf d_grid, d_cur, d_next, d_cnt, d_flags;
    IVP_TRY(d_grid.alloc((size_t)nvox));
    IVP_TRY(d_cur.alloc(sizeof(Tri) * (size_t)n_faces));
    IVP_TRY(d_cnt.alloc(sizeof(unsigned long long)));
    IVP_TRY(d_flags.alloc(2 * sizeof(int)));
    uint8_t* grid = (uint8_t*)d_grid.p;
    int* flags = (int*)d_flags.p;
    IVP_HIP(hipMemsetAsync(grid, 0, (size_t)nvox, st));
    IVP_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), st));
    IVP_HIP(hipMemcpyAsync(d_cur.p, tri.data(), sizeof(Tri) * (size_t)n_faces, hipMemcpyHostToDevice, st));
    const double max_edge = pitch / 2.0;
    long long n = n_faces;
    for (int level = 0; n > 0; ++level) {
        // levels 0 .. max_iter exist; triangles of level max_iter must meet the limit.  Children
        // are tested one level ahead, so the loop stores levels below max_iter only.
        const int mode = level == max_iter ? LV_NO_SPLIT : level + 1 == max_iter ? LV_LAST_CHILDREN : LV_SPLIT;
        const unsigned long long cap = mode == LV_SPLIT ? 4ull * (unsigned long long)n : 0ull;
        IVP_TRY(d_next.alloc(sizeof(Tri) * (size_t)(cap ? cap : 1)));
        IVP_HIP(hipMemsetAsync(d_cnt.p, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_vox_level, dim3(grid_blocks(n)), dim3(BLOCK), 0, st, (const Tri*)d_cur.p, n, max_edge, g,
                           grid, (Tri*)d_next.p, (unsigned long long*)d_cnt.p, cap, mode, flags);
        IVP_HIP(hipGetLastError());
        unsigned long long nn = 0;
        int fl = 0;
        IVP_HIP(hipMemcpyAsync(&nn, d_cnt.p, sizeof(nn), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipStreamSynchronize(st));
        if (fl & 1) return IVP_ERR_INTERNAL;
        if (nn > 0 && mode != LV_SPLIT) return IVP_ERR_ARG;   // trimesh: "max_iter exceeded"
        if (fl & 2) return IVP_ERR_INTERNAL;
        std::swap(d_cur.p, d_next.p);
        n = (long long)nn;
    }
    // hole fill
    hipLaunchKernelGGL(k_vox_border, dim3(grid_blocks(nvox)), dim3(BLOCK), 0, st, grid, g, nvox);
    IVP_HIP(hipGetLastError());
    int* changed = flags + 1;
    for (int round = 0;; ++round) {
        IVP_HIP(hipMemsetAsync(changed, 0, sizeof(int), st));
        for (int axis = 0; axis < 3; ++axis) {
            long long nl = nvox / g.dims[axis];
            hipLaunchKernelGGL(k_vox_sweep, dim3(grid_blocks(nl)), dim3(BLOCK), 0, st, grid, g, axis, nl, changed);
            IVP_HIP(hipGetLastError());
        }
        int ch = 0;
        IVP_HIP(hipMemcpyAsync(&ch, changed, sizeof(int), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipStreamSynchronize(st));
        if (!ch) break;
        if (round > nvox) return IVP_ERR_INTERNAL;
    }
    // points
    const long long nchunks = (nvox + VOX_CHUNK - 1) / VOX_CHUNK;
    DevBuf d_ccount, d_coff;
    IVP_TRY(d_ccount.alloc(sizeof(int) * (size_t)nchunks));
    IVP_TRY(d_coff.alloc(sizeof(long long) * (size_t)nchunks));
    hipLaunchKernelGGL(k_vox_count, dim3((unsigned)nchunks), dim3(BLOCK), 0, st, (const uint8_t*)grid, nvox,
                       (int*)d_ccount.p);
    IVP_HIP(hipGetLastError());
    std::vector<int> cc((size_t)nchunks);
    std::vector<long long> off((size_t)nchunks);
    IVP_HIP(hipMemcpyAsync(cc.data(), d_ccount.p, sizeof(int) * (size_t)nchunks, hipMemcpyDeviceToHost, st));
    IVP_HIP(hipStreamSynchronize(st));
    long long total = 0;
    for (long long k = 0; k < nchunks; ++k) {
        off[k] = total;
        total += cc[k];
    }
    *n_points = total;
    if (total > capacity) return IVP_ERR_CAPACITY;
    if (total == 0) return IVP_OK;
    if (!points) return IVP_ERR_ARG;
    IVP_HIP(hipMemcpyAsync(d_coff.p, off.data(), sizeof(long long) * (size_t)nchunks, hipMemcpyHostToDevice, st));
    const double ox = (double)g.lo[0] * pitch, oy = (double)g.lo[1] * pitch, oz = (double)g.lo[2] * pitch;
    hipLaunchKernelGGL(k_vox_emit, dim3((unsigned)nchunks), dim3(BLOCK), 0, st, (const uint8_t*)grid, g, nvox,
                       (const long long*)d_coff.p, ox, oy, oz, points);
    IVP_HIP(hipGetLastError());
    IVP_HIP(hipStreamSynchronize(st));
    return IVP_OK;
}

// ---- ivp_voxelize_noalloc ----------------------------------------------------------------------
// The same voxelisation with every device buffer carved from one file-scope array, `scratch`, of
// VOX_SCRATCH_BYTES: no hipMalloc / hipFree, for callers that cannot allocate.  The array's
// address on each device comes from hipGetSymbolAddress at the first call on that device (the
// host-side name of a __device__ array is not a device address); kernels, copies and fills only
// ever see that address.  Layout (256-byte aligned regions): counters, voxel grid, chunk counts
// and offsets, then the rest split into two halves that hold one subdivision level each (the
// level being read and the level being written; the vertex and face copies sit in the second
// half until the level-0 triangles are formed from them).
constexpr size_t VOX_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[VOX_SCRATCH_BYTES];

namespace {

constexpr int VOX_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[VOX_MAX_DEVICES] = {};

int scratch_base(unsigned char** out) {
    int dev = 0;
    IVP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= VOX_MAX_DEVICES) return IVP_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        IVP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return IVP_ERR_HIP;
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return IVP_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int zero_bytes(void* p, size_t bytes, hipStream_t st) {
    // regions are 256-byte aligned and sized in whole 16-byte words (align256)
    const long long words = (long long)(align256(bytes) / 16);
    if (words == 0) return IVP_OK;
    hipLaunchKernelGGL(k_vox_zero, dim3(grid_blocks(words)), dim3(BLOCK), 0, st, (uint4*)p, words);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

}  // namespace

extern "C" int ivp_voxelize_noalloc(const double* vertices, int n_vertices, const int* faces, int n_faces,
                                    double pitch, int max_iter, double* points, long long capacity,
                                    long long* n_points, void* stream) {
    if (!n_points) return IVP_ERR_ARG;
    *n_points = 0;
    if (n_vertices < 0 || n_faces < 0 || !(pitch > 0.0) || !isfinite(pitch) || max_iter < 0 || max_iter > 30)
        return IVP_ERR_ARG;
    if (n_faces == 0) return IVP_OK;
    if (!vertices || !faces) return IVP_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    // grid bounds: the rounded corners of the input triangles (host loop, no copies)
    long long lo[3] = {LLONG_MAX, LLONG_MAX, LLONG_MAX}, hi[3] = {LLONG_MIN, LLONG_MIN, LLONG_MIN};
    for (int f = 0; f < n_faces; ++f) {
        for (int c = 0; c < 3; ++c) {
            int vi = faces[3 * f + c];
            if (vi < 0 || vi >= n_vertices) return IVP_ERR_ARG;
            for (int a = 0; a < 3; ++a) {
                double x = vertices[3 * vi + a];
                if (!isfinite(x)) return IVP_ERR_ARG;
                double q = rint(x / pitch);
                if (!(fabs(q) < 1e15)) return IVP_ERR_ARG;
                long long iq = (long long)q;
                lo[a] = std::min(lo[a], iq);
                hi[a] = std::max(hi[a], iq);
            }
        }
    }
    VoxGrid g;
    g.pitch = pitch;
    long long nvox = 1;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo[a];
        g.dims[a] = hi[a] - lo[a] + 1;
        if (g.dims[a] > (1ll << 20)) return IVP_ERR_CAPACITY;
        nvox *= g.dims[a];
        if (nvox > (long long)VOX_SCRATCH_BYTES) return IVP_ERR_CAPACITY;
    }
    const long long nchunks = (nvox + VOX_CHUNK - 1) / VOX_CHUNK;
    // carve the scratch array
    const size_t hdr = 256;   // [0, 8) next-level count, [8, 16) flags, [16, 24) point total
    const size_t b_grid = align256((size_t)nvox);
    const size_t b_cnt = align256(sizeof(int) * (size_t)nchunks);
    const size_t b_off = align256(sizeof(long long) * (size_t)nchunks);
    const size_t fixed = hdr + b_grid + b_cnt + b_off;
    if (fixed >= VOX_SCRATCH_BYTES) return IVP_ERR_CAPACITY;
    const size_t half = ((VOX_SCRATCH_BYTES - fixed) / 2) & ~size_t(255);
    const unsigned long long tri_cap = half / sizeof(Tri);
    const size_t b_verts = align256(sizeof(double) * 3 * (size_t)n_vertices);
    const size_t b_faces = align256(sizeof(int) * 3 * (size_t)n_faces);
    if ((unsigned long long)n_faces > tri_cap || b_verts + b_faces > half) return IVP_ERR_CAPACITY;
    unsigned char* base = nullptr;
    IVP_TRY(scratch_base(&base));
    unsigned long long* d_count = (unsigned long long*)base;
    int* flags = (int*)(base + 8);
    long long* d_total = (long long*)(base + 16);
    uint8_t* grid = base + hdr;
    int* d_ccount = (int*)(base + hdr + b_grid);
    long long* d_coff = (long long*)(base + hdr + b_grid + b_cnt);
    unsigned char* half_a = base + fixed;
    unsigned char* half_b = half_a + half;
    // every call starts from zeroed counters and an empty grid (the array keeps the last call's
    // contents)
    IVP_TRY(zero_bytes(base, hdr + b_grid, st));
    double* d_verts = (double*)half_b;
    int* d_faces = (int*)(half_b + b_verts);
    IVP_HIP(hipMemcpyAsync(d_verts, vertices, sizeof(double) * 3 * (size_t)n_vertices, hipMemcpyHostToDevice, st));
    IVP_HIP(hipMemcpyAsync(d_faces, faces, sizeof(int) * 3 * (size_t)n_faces, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_vox_tris, dim3(grid_blocks(n_faces)), dim3(BLOCK), 0, st, (const double*)d_verts,
                       (const int*)d_faces, (long long)n_faces, (Tri*)half_a);
    IVP_HIP(hipGetLastError());
    const double max_edge = pitch / 2.0;
    Tri* cur = (Tri*)half_a;
    Tri* next = (Tri*)half_b;
    long long n = n_faces;
    for (int level = 0; n > 0; ++level) {
        const int mode = level == max_iter ? LV_NO_SPLIT : level + 1 == max_iter ? LV_LAST_CHILDREN : LV_SPLIT;
        const unsigned long long cap = mode == LV_SPLIT ? tri_cap : 0ull;
        IVP_TRY(zero_bytes(d_count, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_vox_level, dim3(grid_blocks(n)), dim3(BLOCK), 0, st, (const Tri*)cur, n, max_edge, g,
                           grid, next, d_count, cap, mode, flags);
        IVP_HIP(hipGetLastError());
        unsigned long long nn = 0;
        int fl = 0;
        IVP_HIP(hipMemcpyAsync(&nn, d_count, sizeof(nn), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipStreamSynchronize(st));
        if (fl & 1) return IVP_ERR_INTERNAL;
        if (nn > 0 && mode != LV_SPLIT) return IVP_ERR_ARG;   // trimesh: "max_iter exceeded"
        if (fl & 2) return IVP_ERR_CAPACITY;                   // a level larger than half the scratch
        std::swap(cur, next);
        n = (long long)nn;
    }
    // hole fill
    hipLaunchKernelGGL(k_vox_border, dim3(grid_blocks(nvox)), dim3(BLOCK), 0, st, grid, g, nvox);
    IVP_HIP(hipGetLastError());
    int* changed = flags + 1;
    for (int round = 0;; ++round) {
        IVP_TRY(zero_bytes(base, hdr, st));   // the counters (changed among them)
        for (int axis = 0; axis < 3; ++axis) {
            long long nl = nvox / g.dims[axis];
            hipLaunchKernelGGL(k_vox_sweep, dim3(grid_blocks(nl)), dim3(BLOCK), 0, st, grid, g, axis, nl, changed);
            IVP_HIP(hipGetLastError());
        }
        int ch = 0;
        IVP_HIP(hipMemcpyAsync(&ch, changed, sizeof(int), hipMemcpyDeviceToHost, st));
        IVP_HIP(hipStreamSynchronize(st));
        if (!ch) break;
        if (round > nvox) return IVP_ERR_INTERNAL;
    }
    // points: chunk counts, their scan on the device, then the emit pass
    hipLaunchKernelGGL(k_vox_count, dim3((unsigned)nchunks), dim3(BLOCK), 0, st, (const uint8_t*)grid, nvox,
                       d_ccount);
    IVP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_vox_chunk_scan, dim3(1), dim3(BLOCK), 0, st, (const int*)d_ccount, nchunks, d_coff, d_total);
    IVP_HIP(hipGetLastError());
    long long total = 0;
    IVP_HIP(hipMemcpyAsync(&total, d_total, sizeof(total), hipMemcpyDeviceToHost, st));
    IVP_HIP(hipStreamSynchronize(st));
    *n_points = total;
    if (total > capacity) return IVP_ERR_CAPACITY;
    if (total == 0) return IVP_OK;
    if (!points) return IVP_ERR_ARG;
    const double ox = (double)g.lo[0] * pitch, oy = (double)g.lo[1] * pitch, oz = (double)g.lo[2] * pitch;
    hipLaunchKernelGGL(k_vox_emit, dim3((unsigned)nchunks), dim3(BLOCK), 0, st, (const uint8_t*)grid, g, nvox,
                       (const long long*)d_coff, ox, oy, oz, points);
    IVP_HIP(hipGetLastError());
    IVP_HIP(hipStreamSynchronize(st));
    return IVP_OK;
}
