// ivp_grid.hip -- particle store, cell sort, neighbour rows, density and Aii.
//
// Replaces the reference's Warp HashGrid build + neighbour cache (scripts/sim/
// neighbour_search.py:10-155) and PressureBase.compute_density / compute_Aii
// (scripts/sim/forces/pressure/pressure_base.py:12-87).
//
// Sort: a deterministic counting sort over cell keys.  count (atomic), exclusive scan,
// scatter (atomic slot), then every particle ranks itself inside its cell by user id, so the
// resulting order -- (cell key, user id) -- depends on the positions only.  Cell keys are
// x-fastest linear, so the three cells (cx-1..cx+1, cy, cz) of a stencil row are one contiguous
// particle range: a query visits 9 ranges per population.
#include <cmath>

#include "ivp_internal.h"
#include "ivp_launch.h"
#include "ivp_reduce.h"

using namespace ivp;

namespace {

constexpr int SCAN_PER_THREAD = 8;
constexpr int SCAN_CHUNK = BLOCK * SCAN_PER_THREAD;

__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one int per thread over the block; returns the exclusive prefix and the
// block total.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
    __shared__ int wsum[WAVES_PER_BLOCK];
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    int incl = wave_incl_scan(v, lane);
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < WAVES_PER_BLOCK; ++q) {
        if (q < w) off += wsum[q];
        tot += wsum[q];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// start[i] = sum(cnt[0..i)) within the block's chunk; bsum[block] = chunk total
__global__ __launch_bounds__(BLOCK) void k_scan1(int n, const int* __restrict__ cnt, int* __restrict__ start,
                                                 int* __restrict__ bsum) {
    int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
    int v[SCAN_PER_THREAD];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) {
        int idx = base + q;
        v[q] = idx < n ? cnt[idx] : 0;
        sum += v[q];
    }
    int tot;
    int run = block_excl_scan(sum, &tot);
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) {
        int idx = base + q;
        if (idx < n) start[idx] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of the nb chunk totals, bsum[nb] = grand total
__global__ __launch_bounds__(BLOCK) void k_scan2(int nb, int* __restrict__ bsum) {
    __shared__ int carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += BLOCK) {
        int idx = base + threadIdx.x;
        int v = idx < nb ? bsum[idx] : 0;
        int tot;
        int ex = block_excl_scan(v, &tot);
        int carry = carry_s;
        if (idx < nb) bsum[idx] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry_s;
}

__global__ __launch_bounds__(BLOCK) void k_scan3(int n, int* __restrict__ start, const int* __restrict__ bsum, int nb) {
    int base = blockIdx.x * SCAN_CHUNK;
    int add = bsum[blockIdx.x];
    for (int q = threadIdx.x; q < SCAN_CHUNK; q += BLOCK) {
        int idx = base + q;
        if (idx < n) start[idx] += add;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[n] = bsum[nb];
}

// Runs of equal keys among the valid lanes of a wave (valid lanes form a prefix of the wave).
// Particles arrive in the previous step's cell order, so runs are ~ one cell long and one
// atomic per run replaces one per particle.
struct KeyRun {
    bool head;   // first lane of its run
    int len;     // run length (head lanes)
    int head_lane;
};

__device__ __forceinline__ KeyRun key_run(int k, bool valid) {
    const int lane = threadIdx.x & (WAVE - 1);
    int prev = __shfl_up(k, 1, WAVE);
    bool head = valid && (lane == 0 || prev != k);
    unsigned long long heads = __ballot(head);
    int nvalid = __popcll(__ballot(valid));
    unsigned long long above = lane == WAVE - 1 ? 0ull : (heads & (~0ull << (lane + 1)));
    int next = above ? __ffsll((long long)above) - 1 : WAVE;
    KeyRun r;
    r.head = head;
    r.len = (next < nvalid ? next : nvalid) - lane;
    unsigned long long upto = lane == WAVE - 1 ? heads : (heads & ((2ull << lane) - 1));
    r.head_lane = upto ? WAVE - 1 - __clzll((long long)upto) : lane;
    return r;
}

// SCENES: particle i uses the grid of its scene, grids[sid[i]].  clear_flag (if set) is zeroed:
// the build's overflow flags, set only by passes after this one.
template <bool SCENES>
__global__ __launch_bounds__(BLOCK) void k_keys(int count, const float4* __restrict__ pos, Grid g,
                                                const Grid* __restrict__ grids, const int* __restrict__ sid,
                                                int* __restrict__ key, int* __restrict__ ccount,
                                                int* __restrict__ clear_flag) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (clear_flag && i == 0) *clear_flag = 0;
    bool valid = i < count;
    int k = -1;
    if (valid) {
        if (SCENES) g = grids[sid[i]];
        float4 p = pos[i];
        k = cell_key(g, cell_coord(g, p.x, 0), cell_coord(g, p.y, 1), cell_coord(g, p.z, 2));
        key[i] = k;
    }
    KeyRun r = key_run(k, valid);
    if (r.head) atomicAdd(&ccount[k], r.len);
}

// slot = start[key] + (remaining count - 1); leaves ccount at zero for the next sort.  A run's
// head takes the run's slots with one atomic and its lanes take them in order.  The slot receives
// the particle's user id (the ranking below compares user ids within a cell).
__global__ __launch_bounds__(BLOCK) void k_scatter(int count, const int* __restrict__ key,
                                                   const int* __restrict__ start, int* __restrict__ ccount,
                                                   const int* __restrict__ uid, int* __restrict__ cell_uid) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    bool valid = i < count;
    int k = valid ? key[i] : -1;
    KeyRun r = key_run(k, valid);
    int old = 0;
    if (r.head) old = atomicSub(&ccount[k], r.len);
    old = __shfl(old, r.head_lane, WAVE);
    if (valid) {
        int lane = threadIdx.x & (WAVE - 1);
        int slot = start[k] + old - 1 - (lane - r.head_lane);
        cell_uid[slot] = uid[i];
    }
}

// final position = cell start + number of cell members with a smaller user id
__global__ __launch_bounds__(BLOCK) void k_rank(int count, const int* __restrict__ key,
                                                const int* __restrict__ start, const int* __restrict__ cell_uid,
                                                const int* __restrict__ uid, int* __restrict__ order) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    int k = key[i];
    int lo = start[k], hi = start[k + 1];
    int u = uid[i];
    int rank = 0;
#pragma unroll 4
    for (int s = lo; s < hi; ++s) rank += cell_uid[s] < u ? 1 : 0;
    order[lo + rank] = i;
}

struct GatherArgs {
    int count, offset;
    const int* order;
    const float4* pos; float4* pos_b;
    const float* vel; float* vel_b;
    const float* minv; float* minv_b;
    const float* rho0; float* rho0_b;
    const int* mat; int* mat_b;
    const int* uid; int* uid_b;
    int* inv;
    const float* divg; float* divg_b;
};

template <bool VEL>
__global__ __launch_bounds__(BLOCK) void k_gather(GatherArgs a) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= a.count) return;
    int src = a.offset + a.order[s];
    int dst = a.offset + s;
    a.pos_b[dst] = a.pos[src];
    if (VEL) {
        a.vel_b[3 * dst + 0] = a.vel[3 * src + 0];
        a.vel_b[3 * dst + 1] = a.vel[3 * src + 1];
        a.vel_b[3 * dst + 2] = a.vel[3 * src + 2];
        a.divg_b[dst] = a.divg[src];
    }
    a.minv_b[dst] = a.minv[src];
    a.rho0_b[dst] = a.rho0[src];
    a.mat_b[dst] = a.mat[src];
    int u = a.uid[src];
    a.uid_b[dst] = u;
    a.inv[u] = dst;
}

// ---- neighbour enumeration ---------------------------------------------------------------
// Visits, in order, every particle j of one population (dynamic or static) with j != self and
// sqrtf(dx*dx + dy*dy + dz*dz) < h -- the reference predicate, neighbour_search.py:34-44 --
// over the 9 contiguous (cx-1..cx+1) runs of the 3x3 (dz, dy) stencil.  Within a population the
// visit order is ascending (cell key, user id).
template <class F>
__device__ __forceinline__ void visit(const Grid& g, const int* __restrict__ start, const float4* __restrict__ ppos,
                                      int self, float4 pi, float h, F&& f) {
    int cx = cell_coord(g, pi.x, 0), cy = cell_coord(g, pi.y, 1), cz = cell_coord(g, pi.z, 2);
    int x0 = cx > 0 ? cx - 1 : 0;
    int x1 = cx < g.dims[0] - 1 ? cx + 1 : g.dims[0] - 1;
    for (int dz = -1; dz <= 1; ++dz) {
        int zz = cz + dz;
        if (zz < 0 || zz >= g.dims[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int yy = cy + dy;
            if (yy < 0 || yy >= g.dims[1]) continue;
            int s0 = start[cell_key(g, x0, yy, zz)];
            int s1 = start[cell_key(g, x1, yy, zz) + 1];
            for (int j = s0; j < s1; ++j) {
                if (j == self) continue;
                float4 pj = ppos[j];
                float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                float rn = len3(rx, ry, rz);
                if (rn < h) f(j, pj, rx, ry, rz, rn);
            }
        }
    }
}

// Candidate scan of one population: calls f(j) for every j != self with
// sqrtf(dx*dx + dy*dy + dz*dz) < h, in the order of `visit`.  The predicate is decided without
// the square root outside a few-ulp band around h^2: d^2 < h2_lo = h^2 (1 - 2^-20) implies
// sqrtf(d^2) < h, and d^2 >= h2_hi = h^2 (1 + 2^-20) implies sqrtf(d^2) >= h.  The correctly
// rounded root runs only when some lane of the wave has a candidate inside the band (a uniform
// branch, so the common case carries no root at all).  Candidates are loaded two at a time.
template <class F>
__device__ __forceinline__ void scan_ids(const Grid& g, const int* __restrict__ start, const float4* __restrict__ ppos,
                                         int self, float4 pi, float h, float h2_lo, float h2_hi, F&& f) {
    int cx = cell_coord(g, pi.x, 0), cy = cell_coord(g, pi.y, 1), cz = cell_coord(g, pi.z, 2);
    int x0 = cx > 0 ? cx - 1 : 0;
    int x1 = cx < g.dims[0] - 1 ? cx + 1 : g.dims[0] - 1;
    auto test = [&](int j, float4 pj) {
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float d2 = rx * rx + ry * ry + rz * rz;
        bool in = d2 < h2_lo;
        bool band = !in && d2 < h2_hi;
        if (__ballot(band)) {
            if (band) {
                // the empty volatile asm keeps the root inside this branch (it would otherwise be
                // evaluated speculatively for every candidate)
                float d = d2;
                asm volatile("" : "+v"(d));
                in = sqrtf(d) < h;
            }
        }
        if (in && j != self) f(j);
    };
    for (int dz = -1; dz <= 1; ++dz) {
        int zz = cz + dz;
        if (zz < 0 || zz >= g.dims[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int yy = cy + dy;
            if (yy < 0 || yy >= g.dims[1]) continue;
            int s0 = start[cell_key(g, x0, yy, zz)];
            int s1 = start[cell_key(g, x1, yy, zz) + 1];
            for (int j = s0; j < s1; j += 2) {
                int j1 = j + 1 < s1 ? j + 1 : j;
                float4 p0 = ppos[j];
                float4 p1 = ppos[j1];
                __builtin_amdgcn_sched_barrier(0);   // both loads in flight before the first test
                test(j, p0);
                if (j + 1 < s1) test(j1, p1);
            }
        }
    }
}

struct BuildArgs {
    int nd, ns, C, Cs;
    float h, h2_lo, h2_hi;
    float kc, ks;   // cubic and spiky kernel constants
    float inv_h;    // 1 / h rounded to nearest (div_rn)
    float inv_gs;   // 1 / scale of float16 gradW records (a power of two)
    Grid g;
    const Grid* grids;   // independent scenes: grid of each scene, indexed by sid
    const int* sid;
    const float4* pos;
    const float* minv;
    const int* dstart;
    const int* sstart;
    uint32_t* nb;
    uint32_t* g16;   // float16 format: gradW / scale as two half-pair words at twice each id's word
    int* tile_len;
    int* rowcnt;
    int* rowtot;
    float* density;
    float* kk;
    float* Ss;
    int* err_flag;
    double* partials;
    int stride;
};

// One thread per dynamic row, one wave per 64-row tile, two phases.
//  A: candidate scan; the ids of dynamic neighbours go to slots 0, 1, ... of the row and the
//     ids of static neighbours to slots Cs-1, Cs-2, ... (a row with more than C neighbours sets
//     the error flag).
//  B: the wave walks its slots in lockstep, computing gradW (stored next to the id, coalesced),
//     density (pressure_base.py:12-46), Aii (pressure_base.py:49-87) and S_i, accumulated in the
//     reference row order: dynamic neighbours, then static neighbours.  Dynamic slots from the
//     row's count up to the tile length -- the wave's longest row rounded up to ROW_PAD slots
//     -- are padded with (id = row, gradW = 0).
// FMT is the row format (include/ivp.h): ids only, ids + float32 gradW, or ids + float16 gradW
// scaled by 1/inv_gs, packed as (gx, gy) and (gz, 0) half pairs.
template <int FMT>
constexpr unsigned rec_words() { return FMT == IVP_ROWS_F32 ? 256u : 64u; }

// Phase B of the build for the wave's 64-row tile (one thread per row, lane = row's place in the
// tile); cnt / cs / tot are the row's stored dynamic and static slots and its neighbour total.
template <int FMT>
__device__ __forceinline__ void build_eval(const BuildArgs& a, int i, int lane, int tile, bool valid, float4 pi,
                                           int cnt, int cs, int tot);

template <int FMT, bool SCENES>
__global__ __launch_bounds__(BLOCK) void k_build(BuildArgs a) {
    constexpr unsigned REC = rec_words<FMT>();
    int blk = xcd_block(blockIdx.x, gridDim.x);
    int i = blk * BLOCK + threadIdx.x;
    int lane = threadIdx.x & (WAVE - 1);
    int tile = i >> 6;
    bool valid = i < a.nd;
    const float h = a.h;
    const int C = a.C, Cs = a.Cs;
    uint32_t* tbase = a.nb + (size_t)tile * (size_t)Cs * REC + row_word0<REC>(lane);
    int cnt = 0, cs = 0, tot = 0;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) {
        pi = a.pos[i];
        const Grid g = SCENES ? a.grids[a.sid[i]] : a.g;
        scan_ids(g, a.dstart, a.pos, i, pi, h, a.h2_lo, a.h2_hi, [&](int j) {
            if (tot < C) tbase[slot_word<REC>(cnt)] = (uint32_t)j;
            if (tot < C) cnt++;
            tot++;
        });
        scan_ids(g, a.sstart, a.pos + a.nd, -1, pi, h, a.h2_lo, a.h2_hi, [&](int j) {
            if (tot < C) tbase[slot_word<REC>(Cs - 1 - cs)] = (uint32_t)j;
            if (tot < C) cs++;
            tot++;
        });
        if (tot > C) atomicOr(a.err_flag, 1);
    }
    build_eval<FMT>(a, i, lane, tile, valid, pi, cnt, cs, tot);
}

// Predicate codes of the wave-mask compare builtins (LLVM's CmpInst numbering)
constexpr int FCMP_OLT = 4, ICMP_NE = 33, ICMP_ULT = 36;

// Candidate scan of one population with G lanes per row: lane g of the row's group tests the
// candidates s0 + g, s0 + g + G, ... of every run, U batches in flight; the group's ballot
// ranks the hits, so slots fill in the order of the one-lane scan (ascending candidate index).
// `put(j, rank_in_batch, hits_before_batch)` stores a hit; returns the number of hits.
// A batch loads its G * U candidates unconditionally: entries past the run's end (at most
// G * U - 1 <= POS_PAD - 1 past the last particle) lie inside the padded position buffer and are
// masked by j < s1.  The predicate is decided as in scan_ids, on d2 formed with fused
// multiply-adds: that d2 is within 5 * 2^-24 (relative) of the unfused d2 of the reference's
// expression, far inside the 2^-20 margins of h2_lo and h2_hi, so outside the band both give
// the same answer; inside the band the unfused d2 is formed again and the correctly rounded root
// decides, exactly as in scan_ids.
template <int G, int U, class P>
__device__ __forceinline__ int scan_group(const Grid& g, const int* __restrict__ start,
                                          const float4* __restrict__ ppos, int self, float4 pi, float h, float h2_lo,
                                          float h2_hi, int gl, P&& put) {
    const int cx = cell_coord(g, pi.x, 0), cy = cell_coord(g, pi.y, 1), cz = cell_coord(g, pi.z, 2);
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx < g.dims[0] - 1 ? cx + 1 : g.dims[0] - 1;
    const int wl = (int)(threadIdx.x & (WAVE - 1));
    const int base = wl & ~(G - 1);
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const unsigned long long below = (1ull << gl) - 1ull;
    int found = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        int zz = cz + dz;
        if (zz < 0 || zz >= g.dims[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int yy = cy + dy;
            if (yy < 0 || yy >= g.dims[1]) continue;
            const int s0 = start[cell_key(g, x0, yy, zz)];
            const int s1 = start[cell_key(g, x1, yy, zz) + 1];
            for (int b = s0; b < s1; b += G * U) {
                const float4* q = ppos + (b + gl);
                float4 p[U];
#pragma unroll
                for (int u = 0; u < U; ++u) p[u] = q[u * G];
                __builtin_amdgcn_sched_barrier(0);   // every load of the batch in flight before the tests
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int j = b + u * G + gl;
                    float rx = pi.x - p[u].x, ry = pi.y - p[u].y, rz = pi.z - p[u].z;
                    float d2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
                    // wave masks straight from the compares: valid candidates (inside the run, not
                    // the row itself), those inside h2_lo, and those in the band [h2_lo, h2_hi)
                    const unsigned long long valid = __builtin_amdgcn_uicmp((unsigned)j, (unsigned)s1, ICMP_ULT) &
                                                     __builtin_amdgcn_uicmp((unsigned)j, (unsigned)self, ICMP_NE);
                    unsigned long long hit = __builtin_amdgcn_fcmpf(d2, h2_lo, FCMP_OLT) & valid;
                    const unsigned long long band = (__builtin_amdgcn_fcmpf(d2, h2_hi, FCMP_OLT) & valid) & ~hit;
                    if (band) {
                        bool in = false;
                        if ((band >> wl) & 1ull) {
                            float d = rx * rx + ry * ry + rz * rz;
                            asm volatile("" : "+v"(d));
                            in = sqrtf(d) < h;
                        }
                        hit |= __builtin_amdgcn_ballot_w64(in);
                    }
                    const unsigned long long m = (hit >> base) & gmask;
                    if ((m >> gl) & 1ull) put(j, __popcll(m & below), found);
                    found += __popcll(m);
                }
            }
        }
    }
    return found;
}

// Phase A alone with G lanes per row (threads = rows x G): ids into the rows' slots exactly as
// k_build's scan stores them, the row's dynamic slot count into rowcnt and its neighbour total
// into rowtot.  For particle counts that leave few waves per SIMD, where the one-lane scan is a
// chain of dependent load rounds.
template <int FMT, bool SCENES, int G>
__global__ __launch_bounds__(BLOCK) void k_scan(BuildArgs a) {
    constexpr unsigned REC = rec_words<FMT>();
    constexpr int U = G == 2 ? 4 : G == 4 ? 2 : 1;   // batches in flight (measured best per G)
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = blk * (BLOCK / G) + (int)(threadIdx.x / G);
    const int gl = (int)(threadIdx.x % G);
    if (i >= a.nd) return;   // the G lanes of a row leave together
    const int C = a.C, Cs = a.Cs;
    // the wave's rows share one 64-row tile (64 / G rows from a multiple of 64 / G); slot k of the
    // row is word row_word0(lane) + slot_word(k) of the tile, a 32-bit index from the tile's
    // (uniform) base
    uint32_t* tnb = a.nb + (size_t)__builtin_amdgcn_readfirstlane(i >> 6) * (size_t)Cs * REC;
    const uint32_t w0 = row_word0<REC>(i & (WAVE - 1));
    const float4 pi = a.pos[i];
    const Grid g = SCENES ? a.grids[a.sid[i]] : a.g;
    int tot = 0, cnt = 0, cs = 0;
    // the k-th neighbour found (dynamic ones first) is stored when k < C
    int nd_found = scan_group<G, U>(g, a.dstart, a.pos, i, pi, a.h, a.h2_lo, a.h2_hi, gl,
                                    [&](int j, int rank, int before) {
                                        if (before + rank < C) tnb[slot_word<REC>(before + rank) + w0] = (uint32_t)j;
                                    });
    cnt = nd_found < C ? nd_found : C;
    tot = nd_found;
    int ns_found = scan_group<G, U>(g, a.sstart, a.pos + a.nd, -1, pi, a.h, a.h2_lo, a.h2_hi, gl,
                                    [&](int j, int rank, int before) {
                                        const int k = nd_found + before + rank;
                                        if (k < C) tnb[slot_word<REC>(Cs - 1 - (before + rank)) + w0] = (uint32_t)j;
                                    });
    tot += ns_found;
    cs = (tot < C ? tot : C) - cnt;
    (void)cs;
    if (gl == 0) {
        a.rowcnt[i] = cnt;
        a.rowtot[i] = tot;
        if (tot > C) atomicOr(a.err_flag, 1);
    }
}

// Candidate scan by cells (ids-only and float16 rows; one grid or a grid per scene).  Every row of
// one cell visits the same nine candidate runs (scan_ids: the runs of the 3x3 (y, z) neighbour
// cells, each over the cells x - 1 .. x + 1), so a wave takes SCAN_ROWS consecutive rows of its
// block's 64-row tile, cuts them into segments of equal cell, and for each segment loads the
// segment's candidates once -- the nine runs laid end to end in scan_ids' visiting order, 64 per
// batch, up to SCAN_BATCHES batches in registers -- and tests each row of the segment against them.
// The row's hits of a batch are ranked by the batch's wave mask, so slots fill in the one-lane
// order.  The slots are collected in LDS (per row Cs words and 4 spare against bank conflicts; each
// wave owns its rows, so no block barrier) and written out as whole 8-slot blocks of the paired
// layout, 512 contiguous bytes per wave and block; words of a written block between the row's
// dynamic [0, cnt) and static [Cs - cs, Cs) ranges get the row's own id, as the exact pass pads.
constexpr int SCAN_ROWS = 16;     // rows per wave: a block's four waves cover its tile
constexpr int SCAN_BATCHES = 4;   // candidate batches of 64 held in registers per chunk

// The candidate runs of a segment's cell, both populations at once: lane 16 p + k (k < 9) loads
// population p's run k = 3 (dz + 1) + (dy + 1), the cells x - 1 .. x + 1 of row y + dy, layer
// z + dz (empty outside the grid).  Returns, per lane, the exclusive prefix of the run lengths
// within its population (rp) and the run's first particle minus it (rb: candidate v of run k is
// particle v + rb); lane 16 p + 8 also returns the population's total in T.
struct SegRuns {
    int rp, rb, T;
};
__device__ __forceinline__ SegRuns seg_runs(const BuildArgs& a, const Grid& g, int cx, int cy, int cz) {
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const int pop = lane >> 4, k = lane & 15;
    const int zz = cz + k / 3 - 1, yy = cy + k % 3 - 1;
    const bool live = pop < 2 && k < 9 && zz >= 0 && zz < g.dims[2] && yy >= 0 && yy < g.dims[1];
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx < g.dims[0] - 1 ? cx + 1 : g.dims[0] - 1;
    int s0 = 0, n = 0;
    if (live) {
        const int* start = pop == 0 ? a.dstart : a.sstart;
        s0 = start[cell_key(g, x0, yy, zz)];
        n = start[cell_key(g, x1, yy, zz) + 1] - s0;
    }
    // inclusive prefix over the 16 lanes of the population (DPP row shifts stay inside a row of
    // 16 lanes and bring in 0 from outside it)
    int inc = n;
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x111, 0xf, 0xf, false);   // row_shr:1
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x112, 0xf, 0xf, false);   // row_shr:2
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x114, 0xf, 0xf, false);   // row_shr:4
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x118, 0xf, 0xf, false);   // row_shr:8
    return SegRuns{inc - n, s0 - (inc - n), inc};
}

// One row (position pr, global index row) against NBU batches of candidates (qx, qy, qz, qj; ok =
// lanes holding a candidate): hit k of the row goes to slot k (dynamic population; an overflowing
// row's hits past C all go to slot C, which no reader uses) or to slot Cs - 1 - k while k < lim
// (static population), lanes without a stored hit write their own word of dum.  Returns the
// row's hit count after these batches.  Without EXACT, candidates in the band [h2_lo, h2_hi) count
// as misses and are flagged in near (the row is then done again with EXACT, which decides them by
// the correctly rounded root as scan_group does).
template <bool DYN, int NBU, bool EXACT>
__device__ __forceinline__ int cell_row(const BuildArgs& a, const float (&qx)[SCAN_BATCHES],
                                        const float (&qy)[SCAN_BATCHES], const float (&qz)[SCAN_BATCHES],
                                        const int (&qj)[SCAN_BATCHES], const unsigned long long (&ok)[SCAN_BATCHES],
                                        float3 pr, int row, int before, int lim, uint32_t* srow, uint32_t* dum,
                                        unsigned long long& near) {
#pragma unroll
    for (int u = 0; u < NBU; ++u) {
        const float dx = pr.x - qx[u], dy = pr.y - qy[u], dz = pr.z - qz[u];
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        unsigned long long valid = ok[u];
        if (DYN) valid &= __builtin_amdgcn_uicmp((unsigned)qj[u], (unsigned)row, ICMP_NE);
        const unsigned long long lo = __builtin_amdgcn_fcmpf(d2, a.h2_lo, FCMP_OLT);
        const unsigned long long hi = __builtin_amdgcn_fcmpf(d2, a.h2_hi, FCMP_OLT);
        unsigned long long hit = lo & valid;
        if (EXACT) {
            const unsigned long long band = (hi & valid) & ~lo;
            if (band) {
                bool in = false;
                if (__builtin_amdgcn_inverse_ballot_w64(band)) {
                    float d = dx * dx + dy * dy + dz * dz;
                    asm volatile("" : "+v"(d));
                    in = sqrtf(d) < a.h;
                }
                hit |= __builtin_amdgcn_ballot_w64(in);
            }
        } else {
            near |= hi ^ lo;
        }
        const int k = before + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(hit >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)hit, 0u));
        uint32_t* p;
        if (DYN) p = srow + (k < a.C ? k : a.C);
        else p = k < lim ? srow + (a.Cs - 1 - k) : dum;
        if (!__builtin_amdgcn_inverse_ballot_w64(hit)) p = dum;
        *p = (uint32_t)qj[u];
        before += (int)__popcll(hit);
    }
    return before;
}

// Rows rs .. re - 1 of a segment (global row row0 + r) against the chunk's NBU batches.  found
// (lane r) counts row r's hits of the population; lim is C for the dynamic population and C - ndf
// (lane r's dynamic total) for the static one.
template <bool DYN, int NBU>
__device__ __forceinline__ void cell_rows(const BuildArgs& a, const float (&qx)[SCAN_BATCHES],
                                          const float (&qy)[SCAN_BATCHES], const float (&qz)[SCAN_BATCHES],
                                          const int (&qj)[SCAN_BATCHES], const unsigned long long (&ok)[SCAN_BATCHES],
                                          int cb, int rs, int re, int row0, float px, float py, float pz, int& found,
                                          int ndf, uint32_t* st, uint32_t* dum, int S) {
    for (int r = rs; r < re; ++r) {
        const float3 pr = make_float3(
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, px), r)),
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, py), r)),
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, pz), r)));
        // a row's first chunk of a population starts at 0 (each row is in one segment)
        const int before = cb == 0 ? 0 : __builtin_amdgcn_readlane(found, r);
        int lim = DYN ? a.C : a.C - __builtin_amdgcn_readlane(ndf, r);
        lim = lim > 0 ? lim : 0;
        uint32_t* srow = st + r * S;
        unsigned long long near = 0ull;
        int after = cell_row<DYN, NBU, false>(a, qx, qy, qz, qj, ok, pr, row0 + r, before, lim, srow, dum, near);
        if (near)   // a candidate in the band (rare): the row again, deciding the band exactly
            after = cell_row<DYN, NBU, true>(a, qx, qy, qz, qj, ok, pr, row0 + r, before, lim, srow, dum, near);
        found = (int)(threadIdx.x & (WAVE - 1)) == r ? after : found;
    }
}

// One population of a segment: rows rs .. re - 1 of the wave (global row row0 + r) in the cell
// (cx, cy, cz).  found (lane r) counts row r's hits of this population; the dynamic population
// stores hit k at slot k while k < C, the static one at slot Cs - 1 - k while ndf + k < C (ndf =
// lane r's dynamic total), as k_scan / scan_ids do.
template <bool DYN>
__device__ __forceinline__ void cell_pop_scan(const BuildArgs& a, const SegRuns& R, int T, int rs, int re, int row0,
                                              float px, float py, float pz, int& found, int ndf, uint32_t* st,
                                              uint32_t* dum, int S) {
    constexpr int NB = SCAN_BATCHES;
    constexpr int pl = DYN ? 0 : 16;
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const float4* __restrict__ ppos = DYN ? a.pos : a.pos + a.nd;
    int rb[9], rp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        rp[k] = __builtin_amdgcn_readlane(R.rp, pl + k);
        rb[k] = __builtin_amdgcn_readlane(R.rb, pl + k);
    }
    // a particle of the population for lanes without a candidate (masked by ok)
    const int jsafe = DYN ? row0 : a.ns - 1;
    for (int cb = 0; cb < T; cb += WAVE * NB) {
        float qx[NB], qy[NB], qz[NB];
        int qj[NB];
        unsigned long long ok[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            // candidate v = base + lane: the particle v + rb[k] of the last run k with rp[k] <= v
            const int v = cb + WAVE * u + lane;
            int off = rb[0];
#pragma unroll
            for (int k = 1; k < 9; ++k) off = v >= rp[k] ? rb[k] : off;
            const bool in = v < T;
            ok[u] = __builtin_amdgcn_ballot_w64(in);
            qj[u] = in ? v + off : jsafe;
        }
        // every batch's loads in flight together (a batch past the candidates reads jsafe)
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const float4 q = ppos[qj[u]];
            qx[u] = q.x; qy[u] = q.y; qz[u] = q.z;
        }
        const int nbu = (T - cb + WAVE - 1) / WAVE;
        if (nbu >= 4)
            cell_rows<DYN, 4>(a, qx, qy, qz, qj, ok, cb, rs, re, row0, px, py, pz, found, ndf, st, dum, S);
        else if (nbu == 3)
            cell_rows<DYN, 3>(a, qx, qy, qz, qj, ok, cb, rs, re, row0, px, py, pz, found, ndf, st, dum, S);
        else if (nbu == 2)
            cell_rows<DYN, 2>(a, qx, qy, qz, qj, ok, cb, rs, re, row0, px, py, pz, found, ndf, st, dum, S);
        else
            cell_rows<DYN, 1>(a, qx, qy, qz, qj, ok, cb, rs, re, row0, px, py, pz, found, ndf, st, dum, S);
    }
}

// LDS of one k_scan_cells block: its tile's 64 rows of Cs + 4 words, then a discard word per
// thread
inline size_t scan_cells_lds(int Cs) { return ((size_t)WAVE * (size_t)(Cs + 4) + BLOCK) * sizeof(uint32_t); }
constexpr size_t SCAN_CELLS_LDS_MAX = 64 * 1024;

template <int FMT, bool SCENES>
__global__ __launch_bounds__(BLOCK) void k_scan_cells(BuildArgs a) {
    static_assert(rec_words<FMT>() == 64u, "the cell scan writes the paired ids-only layout");
    extern __shared__ uint32_t scan_stage[];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE)), lane = (int)(threadIdx.x & (WAVE - 1));
    const int row0 = tile * WAVE + wv * SCAN_ROWS;
    if (row0 >= a.nd) return;   // whole waves; the block has no barrier
    const int nrw = a.nd - row0 < SCAN_ROWS ? a.nd - row0 : SCAN_ROWS;
    const int C = a.C, Cs = a.Cs, S = Cs + 4;
    uint32_t* st = scan_stage + wv * SCAN_ROWS * S;
    uint32_t* dum = scan_stage + WAVE * S + wv * WAVE + lane;   // the lane's discard word
    // lane r < nrw: row row0 + r, its position and cell (independent scenes: in its scene's
    // grid, whose keys follow the previous scene's, so a segment never spans two scenes)
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    int cx = 0, cy = 0, cz = 0, key = -1 - lane, sid = 0;
    if (lane < nrw) {
        const float4 p = a.pos[row0 + lane];
        px = p.x; py = p.y; pz = p.z;
        if (SCENES) sid = a.sid[row0 + lane];
        const Grid& g = SCENES ? a.grids[sid] : a.g;
        cx = cell_coord(g, px, 0); cy = cell_coord(g, py, 1); cz = cell_coord(g, pz, 2);
        key = cell_key(g, cx, cy, cz);
    }
    int fd = 0, fs = 0;   // lane r: row r's dynamic and static hits
    for (int s = 0; s < nrw;) {
        const int k0 = __builtin_amdgcn_readlane(key, s);
        const unsigned long long diff = __builtin_amdgcn_ballot_w64(lane > s && lane < nrw && key != k0);
        const int e = diff ? (int)__builtin_ctzll(diff) : nrw;
        const int scx = __builtin_amdgcn_readlane(cx, s), scy = __builtin_amdgcn_readlane(cy, s),
                  scz = __builtin_amdgcn_readlane(cz, s);
        const Grid& sg = SCENES ? a.grids[__builtin_amdgcn_readlane(sid, s)] : a.g;
        const SegRuns R = seg_runs(a, sg, scx, scy, scz);
        const int Td = __builtin_amdgcn_readlane(R.T, 8), Ts = __builtin_amdgcn_readlane(R.T, 24);
        if (Td > 0) cell_pop_scan<true>(a, R, Td, s, e, row0, px, py, pz, fd, 0, st, dum, S);
        if (Ts > 0) cell_pop_scan<false>(a, R, Ts, s, e, row0, px, py, pz, fs, fd, st, dum, S);
        s = e;
    }
    int cnt = 0, cs = 0;
    if (lane < nrw) {
        const int tot = fd + fs;
        cnt = fd < C ? fd : C;
        cs = (tot < C ? tot : C) - cnt;
        a.rowcnt[row0 + lane] = cnt;
        a.rowtot[row0 + lane] = tot;
        if (tot > C) atomicOr(a.err_flag, 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // write-out: lane (q, sub) writes slots 8 b + sub and 8 b + sub + 4 of row q (words
    // 512 b + 8 (row's place in the tile) + 2 sub), every block of [0, ceil8(cnt)) and of
    // [floor8(Cs - cs), Cs); the slots of those blocks between the two ranges get the row's own id
    // (the exact pass pads them so; until then it gathers them unused)
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const int q = lane >> 2, sub = lane & 3;
    const int qc = __shfl(cnt, q, WAVE), qs = __shfl(cs, q, WAVE);
    if (q < nrw) {
        const int nbd = (qc + 7) >> 3;
        const int bs0 = qs > 0 ? (Cs - qs) >> 3 : (Cs >> 3);
        const uint32_t* srow = st + q * S;
        u32x2* dst = reinterpret_cast<u32x2*>(a.nb + (size_t)tile * (size_t)Cs * 64u +
                                              8u * (uint32_t)(wv * SCAN_ROWS + q) + 2u * (uint32_t)sub);
        const uint32_t self = (uint32_t)(row0 + q);
        auto word = [&](int k) { return k < qc || k >= Cs - qs ? srow[k] : self; };
        for (int b = 0; b < (Cs >> 3); ++b)
            if (b < nbd || b >= bs0) dst[b * 256] = u32x2{word(8 * b + sub), word(8 * b + sub + 4)};
    }
}

// Phase B after k_scan: one thread per row, the counts from k_scan
template <int FMT>
__global__ __launch_bounds__(BLOCK) void k_eval(BuildArgs a) {
    int blk = xcd_block(blockIdx.x, gridDim.x);
    int i = blk * BLOCK + threadIdx.x;
    int lane = threadIdx.x & (WAVE - 1);
    bool valid = i < a.nd;
    int cnt = 0, tot = 0;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) {
        pi = a.pos[i];
        cnt = a.rowcnt[i];
        tot = a.rowtot[i];
    }
    int cs = (tot < a.C ? tot : a.C) - cnt;
    build_eval<FMT>(a, i, lane, i >> 6, valid, pi, cnt, cs, tot);
}

template <int FMT>
__device__ __forceinline__ void build_eval(const BuildArgs& a, int i, int lane, int tile, bool valid, float4 pi,
                                           int cnt, int cs, int tot) {
    constexpr unsigned REC = rec_words<FMT>();
    const float h = a.h;
    const int Cs = a.Cs;
    uint32_t* tbase = a.nb + (size_t)tile * (size_t)Cs * REC + row_word0<REC>(lane);
    int tl = wave_max(cnt);
    int ts = wave_max(cs);
    int tlp = (tl + ROW_PAD - 1) / ROW_PAD * ROW_PAD;   // <= Cs: tl <= C <= Cs, Cs a multiple of ROW_PAD
    float rho = 0.0f;
    rho = rho + pi.w * cubic_kernel(0.0f, h, a.kc);
    float aii = 0.0f;
    float jx = 0.0f, jy = 0.0f, jz = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    // one dynamic slot: gradW (stored next to the id in the float32 format; returned as the two
    // float16 side-array words h0, h1 in the float16 format), density, Aii, S
    auto dyn_slot = [&](uint32_t* rec, int j, float4 pj, float minv_j, uint32_t& h0, uint32_t& h1) {
        if ((unsigned)j >= (unsigned)a.nd) {  // never taken: phase A stored a dynamic id here
            atomicOr(a.err_flag, 4);
            return;
        }
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float rn = len3_rn(rx, ry, rz);
        float3 gw = spiky_grad_rn(rx, ry, rz, rn, h, a.ks);
        if (FMT == IVP_ROWS_F32) {
            rec[64] = __float_as_uint(gw.x);
            rec[128] = __float_as_uint(gw.y);
            rec[192] = __float_as_uint(gw.z);
        } else if (FMT == IVP_ROWS_F16) {
            h0 = pack_half2(gw.x * a.inv_gs, gw.y * a.inv_gs);
            h1 = pack_half2(gw.z * a.inv_gs, 0.0f);
        }
        rho += pj.w * cubic_kernel_rn(rn, h, a.inv_h, a.kc);
        float Jx = pj.w * gw.x, Jy = pj.w * gw.y, Jz = pj.w * gw.z;
        aii += (Jx * Jx + Jy * Jy + Jz * Jz) * minv_j;
        jx -= Jx; jy -= Jy; jz -= Jz;
    };
    // Two slots in flight (the cached-gradient formats): both ids, then both neighbours' (x, m)
    // and 1/m, then the two slots in order (the accumulation order of a one-slot loop).  An id
    // past a row's count is read but never used, and the gathers take a clamped index.
    auto dyn_pair = [&](uint32_t* r0, uint32_t* r1, int j0, int j1, int k, uint32_t (&hq)[4]) {
        // the empty asm statements keep each pair of loads together (the compiler would sink the
        // second slot's loads into its conditional use)
        asm volatile("" ::"v"(j0), "v"(j1));
        const int c0 = (unsigned)j0 < (unsigned)a.nd ? j0 : 0;
        const int c1 = (unsigned)j1 < (unsigned)a.nd ? j1 : 0;
        const float4 p0 = a.pos[c0];
        const float4 p1 = a.pos[c1];
        const float m0 = a.minv[c0];
        const float m1 = a.minv[c1];
        asm volatile("" ::"v"(p0.x), "v"(p0.y), "v"(p0.z), "v"(p0.w), "v"(m0), "v"(p1.x), "v"(p1.y), "v"(p1.z),
                     "v"(p1.w), "v"(m1));
        if (valid && k < cnt) dyn_slot(r0, j0, p0, m0, hq[0], hq[1]);
        if (valid && k + 1 < cnt) dyn_slot(r1, j1, p1, m1, hq[2], hq[3]);
    };
    // Four slots in flight the same way (ids-only rows, whose eight ids of a block come from two
    // loads)
    auto dyn_quad = [&](const int* j, int k, uint32_t (&hq)[8]) {
        asm volatile("" ::"v"(j[0]), "v"(j[1]), "v"(j[2]), "v"(j[3]));
        float4 p[4];
        float m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = (unsigned)j[u] < (unsigned)a.nd ? j[u] : 0;
            p[u] = a.pos[c];
            m[u] = a.minv[c];
        }
        asm volatile("" ::"v"(p[0].x), "v"(p[0].y), "v"(p[0].z), "v"(p[0].w), "v"(m[0]), "v"(p[1].x), "v"(p[1].y),
                     "v"(p[1].z), "v"(p[1].w), "v"(m[1]), "v"(p[2].x), "v"(p[2].y), "v"(p[2].z), "v"(p[2].w), "v"(m[2]),
                     "v"(p[3].x), "v"(p[3].y), "v"(p[3].z), "v"(p[3].w), "v"(m[3]));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (valid && k + u < cnt) dyn_slot(nullptr, j[u], p[u], m[u], hq[2 * u], hq[2 * u + 1]);
    };
    if constexpr (REC == 64u) {
        // ids-only rows: the row's 8 words of a block (slots 0, 4, 1, 5, 2, 6, 3, 7) with two
        // 16-byte loads; whole blocks stay inside the tile (Cs is a multiple of 8).  The float16
        // format collects the block's gradient words (slot s at word position
        // ((s & 3) << 1) | (s >> 2), zero past the row's count) and writes them as one 64-byte run
        // at twice the block's word index.
        for (int b = 0; b < tl; b += 8) {
            uint32_t* blk = tbase + slot_word<REC>(b);
            const uint4 w0 = *reinterpret_cast<const uint4*>(blk);
            const uint4 w1 = *reinterpret_cast<const uint4*>(blk + 4);
            const int ids[8] = {(int)w0.x, (int)w0.z, (int)w1.x, (int)w1.z,
                                (int)w0.y, (int)w0.w, (int)w1.y, (int)w1.w};
            uint32_t gq[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) gq[q] = 0u;
#pragma unroll
            for (int t = 0; t < 8; t += 4) {
                if (b + t >= tl) break;
                uint32_t hq[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
                dyn_quad(ids + t, b + t, hq);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int sl = t + u, pw = ((sl & 3) << 1) | (sl >> 2);
                    gq[2 * pw] = hq[2 * u]; gq[2 * pw + 1] = hq[2 * u + 1];
                }
            }
            if (FMT == IVP_ROWS_F16 && valid) {
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4* dst = reinterpret_cast<u32x4*>(a.g16 + 2 * (size_t)(blk - a.nb));
#pragma unroll
                for (int q = 0; q < 4; ++q) dst[q] = u32x4{gq[4 * q], gq[4 * q + 1], gq[4 * q + 2], gq[4 * q + 3]};
            }
        }
    } else {
        for (int k = 0; k < tl; k += 2) {
            uint32_t* r0 = tbase + slot_word<REC>(k);
            uint32_t* r1 = tbase + slot_word<REC>(k + 1 < tl ? k + 1 : k);
            uint32_t hq[4];
            dyn_pair(r0, r1, (int)r0[0], (int)r1[0], k, hq);
        }
    }
    // static slots Cs - 1, Cs - 2, ...: two ids, then both positions, in flight at once
    auto st_slot = [&](int j, float4 pj) {
        if ((unsigned)j >= (unsigned)a.ns) {  // never taken: phase A stored a static id here
            atomicOr(a.err_flag, 4);
            return;
        }
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float rn = len3_rn(rx, ry, rz);
        float3 gw = spiky_grad_rn(rx, ry, rz, rn, h, a.ks);
        rho += pj.w * cubic_kernel_rn(rn, h, a.inv_h, a.kc);
        float Jx = pj.w * gw.x, Jy = pj.w * gw.y, Jz = pj.w * gw.z;
        jx -= Jx; jy -= Jy; jz -= Jz;
        sx += Jx; sy += Jy; sz += Jz;
    };
    for (int k = 0; k < ts; k += 2) {
        const int j0 = (int)tbase[slot_word<REC>(Cs - 1 - k)];
        const int j1 = (int)tbase[slot_word<REC>(k + 1 < ts ? Cs - 2 - k : Cs - 1 - k)];
        asm volatile("" ::"v"(j0), "v"(j1));
        const float4 q0 = a.pos[a.nd + ((unsigned)j0 < (unsigned)a.ns ? j0 : 0)];
        const float4 q1 = a.pos[a.nd + ((unsigned)j1 < (unsigned)a.ns ? j1 : 0)];
        asm volatile("" ::"v"(q0.x), "v"(q0.y), "v"(q0.z), "v"(q0.w), "v"(q1.x), "v"(q1.y), "v"(q1.z), "v"(q1.w));
        if (valid && k < cs) st_slot(j0, q0);
        if (valid && k + 1 < cs) st_slot(j1, q1);
    }
    // Pad only after the static ids (slots Cs-cs .. Cs-1) have been consumed: the padded range
    // [cnt, tlp) follows the wave's longest row and can reach into them.
    for (int k = 0; k < tlp; ++k) {
        if (!valid || k < cnt) continue;
        uint32_t* rec = tbase + slot_word<REC>(k);
        rec[0] = (uint32_t)i;
        if (FMT == IVP_ROWS_F32) {
            rec[64] = 0u;
            rec[128] = 0u;
            rec[192] = 0u;
        }   // float16 gradient words past the count were written as zeros with their block
    }
    if (valid) {
        aii += (jx * jx + jy * jy + jz * jz) * a.minv[i];
        a.kk[i] = 1.0f / (aii + 1e-3f);
        a.density[i] = rho;
        a.Ss[3 * i + 0] = sx;
        a.Ss[3 * i + 1] = sy;
        a.Ss[3 * i + 2] = sz;
        a.rowcnt[i] = cnt;
        a.rowtot[i] = tot;
    }
    // row 1 starts at zero: the window pass adds its tiles' window sizes there
    double v[2] = {(double)cnt, 0.0};
    block_partials<2>(v, a.partials, a.stride);
    if (lane == 0) a.tile_len[tile] = tlp;
}

struct WinArgs {
    int nd, Cs;
    int cap;   // largest window kept, <= WIN_MAX
    const uint32_t* nb;
    const int* tile_len;
    uint32_t* loc;
    int* win_ids;
    int* win_n;
    double* partials;   // the build's partials (row 1: window sizes)
    int stride;
};

// Tile windows of ids-only rows, one block per 64-row tile.  The window is the ascending list of
// the distinct ids in the walked slots [0, tile_len) of the tile's valid rows (padding slots name
// the row itself); each such slot gets the index of its id in that list, 16 bits at the slot's id
// word (the 8-byte word pair of slots g, g + 4 becomes one 32-bit word).  The ids are marked in
// an LDS bitmap over [lo, hi] of the tile's ids, and an id's index is the number of marks below
// it; every slot then writes its id at its index into the list.  A tile whose range exceeds
// WIN_SPAN_WORDS words, or whose window exceeds the cap (at most WIN_MAX ids), gets win_n = -1
// and no indices.  Each thread keeps its PPT word pairs of the tile in registers from one
// coalesced load.  Measured on the 1M dam break after 20 steps: 69 us per build (107 us with an
// atomic for every slot; a bit loop writing the list, thread-contiguous pairs, per-word merged
// marks or the list staged in LDS were no faster).
template <int PPT>
__global__ __launch_bounds__(BLOCK) void k_window(WinArgs a) {
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    __shared__ uint32_t bm[WIN_SPAN_WORDS];
    __shared__ uint16_t pre[WIN_SPAN_WORDS];
    __shared__ int red[2 * WAVES_PER_BLOCK];
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int nrow = min(WAVE, a.nd - t * WAVE);
    const int npair = a.tile_len[t] * (WAVE / 2);
    const size_t base = (size_t)t * (size_t)a.Cs * WAVE;
    const u32x2* ids2 = reinterpret_cast<const u32x2*>(a.nb + base);
    // word pair q of a tile: block q >> 8, row (q >> 2) & 63; thread pairs q = thread + BLOCK k
    u32x2 v[PPT];
    bool on[PPT];
    int lo = 0x7fffffff, hi = -1;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int q = (int)threadIdx.x + BLOCK * k;
        on[k] = q < npair && ((q >> 2) & (WAVE - 1)) < nrow;
        v[k] = on[k] ? ids2[q] : u32x2{0u, 0u};
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k)
        if (on[k]) {
            lo = min(lo, (int)min(v[k].x, v[k].y));
            hi = max(hi, (int)max(v[k].x, v[k].y));
        }
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        red[w] = lo;
        red[WAVES_PER_BLOCK + w] = hi;
    }
    __syncthreads();
    lo = min(min(red[0], red[1]), min(red[2], red[3]));
    hi = max(max(red[4], red[5]), max(red[6], red[7]));
    if (hi < 0) {   // no walked slot
        if (threadIdx.x == 0) a.win_n[t] = 0;
        return;
    }
    const int words = ((hi - lo) >> 5) + 1;
    if (words > WIN_SPAN_WORDS) {
        if (threadIdx.x == 0) a.win_n[t] = -1;
        return;
    }
    for (int k = threadIdx.x; k < words; k += BLOCK) bm[k] = 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k)
        if (on[k]) {
            const uint32_t p0 = v[k].x - (uint32_t)lo, p1 = v[k].y - (uint32_t)lo;
            // an id repeats in many rows of a tile, and atomics on one LDS word serialise: mark
            // only what is not yet marked
            const uint32_t m0 = bm[p0 >> 5], m1 = bm[p1 >> 5];
            if (!(m0 & (1u << (p0 & 31)))) atomicOr(&bm[p0 >> 5], 1u << (p0 & 31));
            if (!(m1 & (1u << (p1 & 31)))) atomicOr(&bm[p1 >> 5], 1u << (p1 & 31));
        }
    __syncthreads();
    // thread k counts the marks of words [w0, w1); exclusive scan over the block
    const int per = (words + BLOCK - 1) / BLOCK;
    const int w0 = min((int)threadIdx.x * per, words), w1 = min(w0 + per, words);
    int cnt = 0;
    for (int k = w0; k < w1; ++k) cnt += __popc(bm[k]);
    int total;
    int off = block_excl_scan(cnt, &total);
    if (total > a.cap) {
        if (threadIdx.x == 0) a.win_n[t] = -1;
        return;
    }
    for (int k = w0; k < w1; ++k) {
        pre[k] = (uint16_t)off;
        off += __popc(bm[k]);
    }
    __syncthreads();
    auto index = [&](uint32_t j) {
        const uint32_t p = j - (uint32_t)lo;
        return (uint32_t)pre[p >> 5] + (uint32_t)__popc(bm[p >> 5] & ((1u << (p & 31)) - 1u));
    };
    // every window id is named by some walked slot: each slot writes its id at its index (slots
    // of one id write the same word)
    int* wid = a.win_ids + (size_t)t * WIN_MAX;
    uint32_t* l32 = a.loc + base / 2;
#pragma unroll
    for (int k = 0; k < PPT; ++k)
        if (on[k]) {
            const uint32_t i0 = index(v[k].x), i1 = index(v[k].y);
            wid[i0] = (int)v[k].x;
            wid[i1] = (int)v[k].y;
            l32[threadIdx.x + BLOCK * k] = i0 | (i1 << 16);
        }
    if (threadIdx.x == 0) {
        a.win_n[t] = total;
        // the build's partial row 1 (zeroed by the build's blocks of four tiles; the sums are
        // integers, exact in any order)
        atomicAdd(&a.partials[a.stride + t / WAVES_PER_BLOCK], (double)total);
    }
}

struct ExportArgs {
    int n, nd, C;
    float h, ks;
    Grid g;
    const Grid* grids;
    const int* sid;
    const float4* pos;
    const int* dstart;
    const int* sstart;
    const int* uid;
    int* count;
    int* ids;
    float* gradW;
};

// The reference cache layout (every row, user ids, count capped at C), same enumeration as
// k_build.
template <bool SCENES>
__global__ __launch_bounds__(BLOCK) void k_export(ExportArgs a) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= a.n) return;
    const Grid g = SCENES ? a.grids[a.sid[s]] : a.g;
    float4 pi = a.pos[s];
    int u = a.uid[s];
    size_t base = (size_t)u * (size_t)a.C;
    int cnt = 0;
    bool is_dyn = s < a.nd;
    auto put = [&](int jint, float rx, float ry, float rz) {
        if (cnt < a.C) {
            float3 gw = spiky_grad(rx, ry, rz, a.h, a.ks);
            a.ids[base + cnt] = a.uid[jint];
            a.gradW[3 * (base + cnt) + 0] = gw.x;
            a.gradW[3 * (base + cnt) + 1] = gw.y;
            a.gradW[3 * (base + cnt) + 2] = gw.z;
        }
        cnt++;
    };
    visit(g, a.dstart, a.pos, is_dyn ? s : -1, pi, a.h,
          [&](int j, float4, float rx, float ry, float rz, float) { put(j, rx, ry, rz); });
    visit(g, a.sstart, a.pos + a.nd, is_dyn ? -1 : s - a.nd, pi, a.h,
          [&](int j, float4, float rx, float ry, float rz, float) { put(a.nd + j, rx, ry, rz); });
    a.count[u] = cnt < a.C ? cnt : a.C;
}

// ---- particle store loading ----------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_dyn_flags(int n, const int* __restrict__ is_dynamic,
                                                     const int* __restrict__ material, int fluid,
                                                     int* __restrict__ flag, int* __restrict__ err) {
    int u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    int dyn = is_dynamic[u] != 0 ? 1 : 0;
    flag[u] = dyn;
    if (dyn && material[u] != fluid) atomicOr(err, 2);
    if (!dyn && material[u] == fluid) atomicOr(err, 8);
}

struct StageArgs {
    int n, nd;
    int S;                 // independent scenes (0: one scene)
    const int* ustart;     // S + 1 user-order scene starts
    int* sid;
    const int* dpos;
    const int* is_dynamic;
    const float* x; const float* v; const float* m; const float* m_inv;
    const float* density0; const int* material;
    float4* pos; float* vel; float* minv; float* rho0; int* mat; int* uid; int* inv;
    float* density; float* kk; float* divg;
};

__global__ __launch_bounds__(BLOCK) void k_stage(StageArgs a) {
    int u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= a.n) return;
    bool dyn = a.is_dynamic[u] != 0;
    int s = dyn ? a.dpos[u] : a.nd + (u - a.dpos[u]);
    a.pos[s] = make_float4(a.x[3 * u + 0], a.x[3 * u + 1], a.x[3 * u + 2], a.m[u]);
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (dyn) { vx = a.v[3 * u + 0]; vy = a.v[3 * u + 1]; vz = a.v[3 * u + 2]; }
    a.vel[3 * s + 0] = vx; a.vel[3 * s + 1] = vy; a.vel[3 * s + 2] = vz;
    a.minv[s] = a.m_inv[u];
    a.rho0[s] = a.density0[u];
    a.mat[s] = a.material[u];
    a.uid[s] = u;
    a.inv[u] = s;
    if (a.S > 0) {
        int lo = 0, hi = a.S;   // scene q with ustart[q] <= u < ustart[q + 1]
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (a.ustart[mid] <= u) lo = mid; else hi = mid;
        }
        a.sid[s] = lo;
    }
    a.density[s] = a.density0[u];
    a.kk[s] = 0.0f;
    a.divg[s] = 0.0f;
}

__global__ __launch_bounds__(BLOCK) void k_write_x(int nd, const float* __restrict__ x, const int* __restrict__ uid,
                                                   float4* __restrict__ pos) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= nd) return;
    int u = uid[s];
    float4 p = pos[s];
    p.x = x[3 * u + 0]; p.y = x[3 * u + 1]; p.z = x[3 * u + 2];
    pos[s] = p;
}

__global__ __launch_bounds__(BLOCK) void k_write_v(int nd, const float* __restrict__ v, const int* __restrict__ uid,
                                                   float* __restrict__ vel) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= nd) return;
    int u = uid[s];
    vel[3 * s + 0] = v[3 * u + 0];
    vel[3 * s + 1] = v[3 * u + 1];
    vel[3 * s + 2] = v[3 * u + 2];
}

__global__ __launch_bounds__(BLOCK) void k_read_x(int n, const float4* __restrict__ pos, const int* __restrict__ uid,
                                                  float* __restrict__ x) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    int u = uid[s];
    float4 p = pos[s];
    x[3 * u + 0] = p.x; x[3 * u + 1] = p.y; x[3 * u + 2] = p.z;
}

__global__ __launch_bounds__(BLOCK) void k_read_v(int n, int nd, const float* __restrict__ vel,
                                                  const int* __restrict__ uid, float* __restrict__ v) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    int u = uid[s];
    bool dyn = s < nd;
    v[3 * u + 0] = dyn ? vel[3 * s + 0] : 0.0f;
    v[3 * u + 1] = dyn ? vel[3 * s + 1] : 0.0f;
    v[3 * u + 2] = dyn ? vel[3 * s + 2] : 0.0f;
}

__global__ __launch_bounds__(BLOCK) void k_read_f(int n, const float* __restrict__ f, const int* __restrict__ uid,
                                                  float* __restrict__ out) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    out[uid[s]] = f[s];
}

int scan(ivp_ctx* c, const int* cnt, int* start, int n) {
    int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nb + 1 > c->nscan_blocks) return IVP_ERR_CAPACITY;
    if (nb == 0) {
        IVP_HIP(hipMemsetAsync(start, 0, sizeof(int), c->stream));
        return IVP_OK;
    }
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(BLOCK), 0, c->stream, n, cnt, start, c->scan_bsum);
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(BLOCK), 0, c->stream, nb, c->scan_bsum);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(BLOCK), 0, c->stream, n, start, c->scan_bsum, nb);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

template <class T>
void swap_ptr(T*& a, T*& b) {
    T* t = a;
    a = b;
    b = t;
}

}  // namespace

namespace ivp {

// Sort one population [offset, offset + count) of the store by (cell key, user id) and fill
// its cell table `start_table` (ncell + 1 entries, relative to offset).
int grid_sort_population(ivp_ctx* c, int offset, int count, int* start_table, bool dynamic) {
    if (count <= 0) {
        IVP_HIP(hipMemsetAsync(start_table, 0, sizeof(int) * (size_t)(c->ncell + 1), c->stream));
        return IVP_OK;
    }
    const dim3 nb(blocks_for(count));
    // bytes: keys (read x 16, write key 4), scatter (key 4 + slot 4), rank (key, uid, cell members
    // ~ 3 x 4), gather (index 4 + every per-particle array read and written)
    const double per = dynamic ? 4 + 2 * (16 + 12 + 4 + 4 + 4 + 4 + 4) + 4 : 4 + 2 * (16 + 4 + 4 + 4 + 4) + 4;
    // the dynamic population's sort starts a rebuild: its first pass clears the overflow flags
    int* clear = dynamic ? c->err_flag : nullptr;
    if (c->S > 0)
        IVP_TRY(launch(c, KC_SORT, -1, 24.0 * count, k_keys<true>, nb, count, (const float4*)(c->pos + offset), c->grid,
                       (const Grid*)c->grids, (const int*)(c->sid + offset), c->key, c->ccount, clear));
    else
        IVP_TRY(launch(c, KC_SORT, -1, 20.0 * count, k_keys<false>, nb, count, (const float4*)(c->pos + offset),
                       c->grid, (const Grid*)nullptr, (const int*)nullptr, c->key, c->ccount, clear));
    IVP_TRY(scan(c, c->ccount, start_table, c->ncell));
    IVP_TRY(launch(c, KC_SORT, -1, 16.0 * count, k_scatter, nb, count, (const int*)c->key, (const int*)start_table,
                   c->ccount, (const int*)(c->uid + offset), c->tmp_src));
    IVP_TRY(launch(c, KC_SORT, -1, 20.0 * count, k_rank, nb, count, (const int*)c->key, (const int*)start_table,
                   (const int*)c->tmp_src, (const int*)(c->uid + offset), c->order));
    GatherArgs ga{count, offset, c->order, c->pos, c->pos_b, c->vel, c->vel_b, c->minv, c->minv_b,
                  c->rho0, c->rho0_b, c->mat, c->mat_b, c->uid, c->uid_b, c->inv, c->divg, c->divg_b};
    if (dynamic)
        IVP_TRY(launch(c, KC_SORT, -1, per * count, k_gather<true>, nb, ga));
    else
        IVP_TRY(launch(c, KC_SORT, -1, per * count, k_gather<false>, nb, ga));
    swap_ptr(c->pos, c->pos_b);
    swap_ptr(c->minv, c->minv_b);
    swap_ptr(c->rho0, c->rho0_b);
    swap_ptr(c->mat, c->mat_b);
    swap_ptr(c->uid, c->uid_b);
    if (dynamic) {
        swap_ptr(c->vel, c->vel_b);
        swap_ptr(c->divg, c->divg_b);
    }
    return IVP_OK;
}

int grid_load(ivp_ctx* c, int n, const float* x, const float* v, const float* m, const float* m_inv,
              const float* density0, const int* is_dynamic, const int* material) {
    if (n < 0 || n > c->cap) return IVP_ERR_CAPACITY;
    if (n > 0 && (!x || !v || !m || !m_inv || !density0 || !is_dynamic || !material)) return IVP_ERR_ARG;
    c->loaded = false;
    c->built = false;
    c->n = n;
    if (n == 0) {
        c->nd = c->ns = 0;
        c->loaded = true;
        return IVP_OK;
    }
    int nb = blocks_for(n);
    IVP_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
    // flags -> exclusive prefix = rank of each dynamic particle
    hipLaunchKernelGGL(k_dyn_flags, dim3(nb), dim3(BLOCK), 0, c->stream, n, is_dynamic, material,
                       c->cfg.material_fluid, c->order, c->err_flag);
    int rc = scan(c, c->order, c->tmp_src, n);
    if (rc != IVP_OK) return rc;
    IVP_HIP(hipMemcpyAsync(c->count_host, c->tmp_src + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipMemcpyAsync(c->err_host, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    c->nd = *c->count_host;
    c->ns = n - c->nd;
    c->all_dyn_fluid = (*c->err_host & 2) == 0;
    c->static_fluid = (*c->err_host & 8) != 0;
    int* ustart = nullptr;
    if (c->S > 0) {
        // scene boundaries in user order -> dynamic / static starts in internal order
        int S = c->S;
        std::vector<int> us(S + 1, 0);
        for (int s = 0; s < S; ++s) us[s + 1] = us[s] + c->scenes[s].n_particles;
        if (us[S] != n) return IVP_ERR_ARG;
        std::vector<int> pre(S + 1, 0);
        for (int s = 0; s <= S; ++s)
            IVP_HIP(hipMemcpy(&pre[s], c->tmp_src + us[s], sizeof(int), hipMemcpyDeviceToHost));
        c->scene_row0.assign(S + 1, 0);
        c->scene_srow0.assign(S + 1, 0);
        for (int s = 0; s <= S; ++s) {
            c->scene_row0[s] = pre[s];
            c->scene_srow0[s] = us[s] - pre[s];
        }
        for (int s = 0; s < S; ++s)
            if (c->scene_row0[s + 1] - c->scene_row0[s] < MIN_SCENE_DYNAMIC) return IVP_ERR_ARG;
        ustart = c->key;   // sort scratch, free until the static sort below
        IVP_HIP(hipMemcpy(ustart, us.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice));
    }
    StageArgs sa{n, c->nd, c->S, ustart, c->sid, c->tmp_src, is_dynamic, x, v, m, m_inv, density0, material,
                 c->pos, c->vel, c->minv, c->rho0, c->mat, c->uid, c->inv, c->density, c->kk, c->divg};
    hipLaunchKernelGGL(k_stage, dim3(nb), dim3(BLOCK), 0, c->stream, sa);
    IVP_HIP(hipGetLastError());
    // static particles: sort once.  The sort gathers [nd, n) into the reorder buffers and swaps,
    // so afterwards the staged dynamic segment [0, nd) lives in the reorder buffers and the
    // sorted static segment in the primary ones: copy each segment across so both copies of
    // every per-particle array are complete and identical.
    rc = grid_sort_population(c, c->nd, c->ns, c->sstart, false);
    if (rc != IVP_OK) return rc;
    size_t off = (size_t)c->nd, cnt = (size_t)c->ns;
    if (cnt > 0) {
        size_t nd = off;
        auto mirror = [&](void* prim, void* alt, size_t elem) -> int {
            char* p = (char*)prim;
            char* b = (char*)alt;
            if (nd > 0) IVP_HIP(hipMemcpyAsync(p, b, nd * elem, hipMemcpyDeviceToDevice, c->stream));
            IVP_HIP(hipMemcpyAsync(b + nd * elem, p + nd * elem, cnt * elem, hipMemcpyDeviceToDevice, c->stream));
            return IVP_OK;
        };
        if ((rc = mirror(c->pos, c->pos_b, sizeof(float4))) != IVP_OK) return rc;
        if ((rc = mirror(c->minv, c->minv_b, sizeof(float))) != IVP_OK) return rc;
        if ((rc = mirror(c->rho0, c->rho0_b, sizeof(float))) != IVP_OK) return rc;
        if ((rc = mirror(c->mat, c->mat_b, sizeof(int))) != IVP_OK) return rc;
        if ((rc = mirror(c->uid, c->uid_b, sizeof(int))) != IVP_OK) return rc;
        // vel and divg were not swapped by the static sort: primary holds everything
        IVP_HIP(hipMemcpyAsync(c->vel_b, c->vel, 3 * (nd + cnt) * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        IVP_HIP(hipMemcpyAsync(c->divg_b, c->divg, (nd + cnt) * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        // static density stays rho0 (compute_density skips static rows, pressure_base.py:28-29);
        // static k and divergence are 0 in any order
        IVP_HIP(hipMemcpyAsync(c->density + off, c->rho0 + off, cnt * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    c->loaded = true;
    return IVP_OK;
}

// Lanes per row of the candidate scan: the option's value, or by the dynamic count.  Measured
// (build and sort per call): 8k rows 132 us with one fused lane, 92 with 16 lanes; 102k rows
// 145 -> 115 with 4 lanes; 1M rows 501 -> 404 with 2 lanes and 4 batches in flight.  Fewer rows
// leave fewer waves per SIMD, and the one-lane scan is then a chain of dependent load rounds.
// 64 is the scan by cells (k_scan_cells), for ids-only and float16 rows (one scene or many) whose
// staged rows fit the LDS: scan alone 1M rows 225 -> 186 us, 102k rows 32 -> 28 us, 8k rows 12-14 ->
// 15 us (16 lanes stay there); where it does not apply, the count's lanes per row.
static int lanes_by_count(int nd) { return nd >= (1 << 18) ? 2 : nd >= (1 << 15) ? 4 : nd >= (1 << 12) ? 8 : 16; }

int build_lanes_for(const ivp_ctx* c) {
    const int nd = c->nd;
    const bool cells = c->rec_words == 64 && scan_cells_lds(c->Cs) <= SCAN_CELLS_LDS_MAX;
    int G = c->build_lanes > 0 ? c->build_lanes : cells && nd >= (1 << 15) ? 64 : lanes_by_count(nd);
    if (G == 64 && !cells) G = lanes_by_count(nd);
    return G;
}

template <int FMT, bool SCENES>
int launch_scan(ivp_ctx* c, int G, const BuildArgs& ba) {
    if constexpr (rec_words<FMT>() == 64u) {
        if (G == 64)
            return launch_shm(c, KC_SCAN, -1, 0.0, k_scan_cells<FMT, SCENES>, dim3((c->nd + WAVE - 1) / WAVE),
                              (unsigned)scan_cells_lds(c->Cs), ba);
    }
    const dim3 grid(blocks_for(c->nd * G));
    switch (G) {
    case 2: return launch(c, KC_SCAN, -1, 0.0, k_scan<FMT, SCENES, 2>, grid, ba);
    case 4: return launch(c, KC_SCAN, -1, 0.0, k_scan<FMT, SCENES, 4>, grid, ba);
    case 8: return launch(c, KC_SCAN, -1, 0.0, k_scan<FMT, SCENES, 8>, grid, ba);
    case 16: return launch(c, KC_SCAN, -1, 0.0, k_scan<FMT, SCENES, 16>, grid, ba);
    default: return IVP_ERR_ARG;
    }
}

int grid_build(ivp_ctx* c) {
    if (!c->loaded) return IVP_ERR_STATE;
    if (c->nd > 0) {
        int rc = grid_sort_population(c, 0, c->nd, c->dstart, true);
        if (rc != IVP_OK) return rc;
    } else {
        IVP_HIP(hipMemsetAsync(c->dstart, 0, sizeof(int) * (size_t)(c->ncell + 1), c->stream));
        IVP_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));   // (else cleared by the sort)
    }
    if (c->nd > 0) {
        // float h2_lo <= h^2 (1 - 2^-20) and h2_hi >= h^2 (1 + 2^-20): the predicate bounds of scan_ids
        const double hh = (double)c->h * (double)c->h;
        float h2_lo = std::nextafter((float)(hh * (1.0 - 1.0 / 1048576.0)), 0.0f);
        float h2_hi = std::nextafter((float)(hh * (1.0 + 1.0 / 1048576.0)), 1.0f);
        BuildArgs ba{c->nd, c->ns, c->C, c->Cs, c->h, h2_lo, h2_hi, c->kcubic, c->kspiky, 1.0f / c->h, 1.0f / c->gscale, c->grid,
                     c->grids, c->sid, c->pos, c->minv, c->dstart, c->sstart, c->nb, c->g16, c->tile_len,
                     c->rowcnt, c->rowtot, c->density, c->kk, c->Ss, c->err_flag, c->partials, c->max_blocks};
        // bytes (pair count of the previous build): the pair records written, the ids read back
        // once, x and m of every particle once, the per-row outputs (rho, k, S, counts)
        double pairs = (double)c->state_host->total_pairs;
        double bytes = pairs * (4.0 * c->rec_words / 64 + 4.0 + (c->row_fmt == IVP_ROWS_F16 ? 8.0 : 0.0)) +
                       16.0 * c->n + 32.0 * c->nd;
        const dim3 nb(blocks_for(c->nd));
        const int G = build_lanes_for(c);
        if (G == 1) {
            if (c->S > 0)   // independent scenes use ids-only rows (ivp_set_scenes fixes the format)
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_build<IVP_ROWS_IDS, true>, nb, ba));
            else if (c->row_fmt == IVP_ROWS_F32)
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_build<IVP_ROWS_F32, false>, nb, ba));
            else if (c->row_fmt == IVP_ROWS_F16)
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_build<IVP_ROWS_F16, false>, nb, ba));
            else
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_build<IVP_ROWS_IDS, false>, nb, ba));
        } else {
            if (c->S > 0)
                IVP_TRY((launch_scan<IVP_ROWS_IDS, true>(c, G, ba)));
            else if (c->row_fmt == IVP_ROWS_F32)
                IVP_TRY((launch_scan<IVP_ROWS_F32, false>(c, G, ba)));
            else if (c->row_fmt == IVP_ROWS_F16)
                IVP_TRY((launch_scan<IVP_ROWS_F16, false>(c, G, ba)));
            else
                IVP_TRY((launch_scan<IVP_ROWS_IDS, false>(c, G, ba)));
            if (c->row_fmt == IVP_ROWS_F32 && c->S == 0)
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_eval<IVP_ROWS_F32>, nb, ba));
            else if (c->row_fmt == IVP_ROWS_F16 && c->S == 0)
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_eval<IVP_ROWS_F16>, nb, ba));
            else
                IVP_TRY(launch(c, KC_BUILD, -1, bytes, k_eval<IVP_ROWS_IDS>, nb, ba));
        }
    }
    c->win_built = false;
    // a tile walks at most Cs slots: Cs * 32 word pairs, PPT per thread
    const int ppt = (c->Cs * (WAVE / 2) + BLOCK - 1) / BLOCK;
    // (not for independent scenes: their passes run late in a batched solve on few active blocks,
    // and the window pass over every scene's rows cost more than it saved -- 4.28 -> 4.48 ms per
    // step on 256 columns)
    if (c->nd > 0 && c->windows && c->rec_words == 64 && c->S == 0 && ppt <= 32) {
        WinArgs wa{c->nd, c->Cs, c->win_cap, c->nb, c->tile_len, c->loc, c->win_ids, c->win_n, c->partials,
                   c->max_blocks};
        const dim3 grid((c->nd + WAVE - 1) / WAVE);
        if (ppt <= 8)
            IVP_TRY(launch(c, KC_WINDOW, -1, 0.0, k_window<8>, grid, wa));
        else if (ppt <= 16)
            IVP_TRY(launch(c, KC_WINDOW, -1, 0.0, k_window<16>, grid, wa));
        else
            IVP_TRY(launch(c, KC_WINDOW, -1, 0.0, k_window<32>, grid, wa));
        c->win_built = true;
    }
    // pair count (row 0) and window sizes (row 1) of the build's blocks
    if (c->nd > 0) IVP_TRY(finalize<FIN_BUILD>(c, (int)blocks_for(c->nd)));
    c->built = true;
    return IVP_OK;
}

int grid_export(ivp_ctx* c, int cache_size, int* count, int* ids, float* gradW) {
    if (!c->built) return IVP_ERR_STATE;
    if (cache_size <= 0 || !count || !ids || !gradW) return IVP_ERR_ARG;
    if (c->n == 0) return IVP_OK;
    ExportArgs ea{c->n, c->nd, cache_size, c->h, c->kspiky, c->grid, c->grids, c->sid, c->pos, c->dstart, c->sstart,
                  c->uid, count, ids, gradW};
    if (c->S > 0)
        hipLaunchKernelGGL(k_export<true>, dim3(blocks_for(c->n)), dim3(BLOCK), 0, c->stream, ea);
    else
        hipLaunchKernelGGL(k_export<false>, dim3(blocks_for(c->n)), dim3(BLOCK), 0, c->stream, ea);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

int store_write(ivp_ctx* c, int field, const void* src) {
    if (!c->loaded) return IVP_ERR_STATE;
    if (!src) return IVP_ERR_ARG;
    if (c->nd == 0) return IVP_OK;
    int nb = blocks_for(c->nd);
    if (field == IVP_FIELD_X) {
        hipLaunchKernelGGL(k_write_x, dim3(nb), dim3(BLOCK), 0, c->stream, c->nd, (const float*)src, c->uid, c->pos);
        c->built = false;
    } else if (field == IVP_FIELD_V) {
        hipLaunchKernelGGL(k_write_v, dim3(nb), dim3(BLOCK), 0, c->stream, c->nd, (const float*)src, c->uid, c->vel);
    } else {
        return IVP_ERR_ARG;
    }
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

int store_read(ivp_ctx* c, int field, void* dst) {
    if (!c->loaded) return IVP_ERR_STATE;
    if (!dst) return IVP_ERR_ARG;
    if (c->n == 0) return IVP_OK;
    int nb = blocks_for(c->n);
    switch (field) {
        case IVP_FIELD_X:
            hipLaunchKernelGGL(k_read_x, dim3(nb), dim3(BLOCK), 0, c->stream, c->n, c->pos, c->uid, (float*)dst);
            break;
        case IVP_FIELD_V:
            hipLaunchKernelGGL(k_read_v, dim3(nb), dim3(BLOCK), 0, c->stream, c->n, c->nd, c->vel, c->uid, (float*)dst);
            break;
        case IVP_FIELD_DENSITY:
            hipLaunchKernelGGL(k_read_f, dim3(nb), dim3(BLOCK), 0, c->stream, c->n, c->density, c->uid, (float*)dst);
            break;
        case IVP_FIELD_K:
            hipLaunchKernelGGL(k_read_f, dim3(nb), dim3(BLOCK), 0, c->stream, c->n, c->kk, c->uid, (float*)dst);
            break;
        case IVP_FIELD_DIVERGENCE:
            hipLaunchKernelGGL(k_read_f, dim3(nb), dim3(BLOCK), 0, c->stream, c->n, c->divg, c->uid, (float*)dst);
            break;
        default:
            return IVP_ERR_ARG;
    }
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

}  // namespace ivp
