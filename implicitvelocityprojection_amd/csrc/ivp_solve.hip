// ivp_solve.hip -- IP velocity projection, Weiler 2018 implicit viscosity, explicit viscosity
// and the step glue, on the tiled neighbour rows built by ivp_grid.hip.
//
// Reference: scripts/sim/forces/pressure/ip.py (IP.solve, IP.PCG, five kernels),
// scripts/sim/forces/pressure/pressure_base.py (J, J^T, errors, divergence),
// scripts/sim/math_utils.py (vector ops, dots), scripts/sim/forces/non_pressure/viscosity/
// {viscosity,weiler18}.py, scripts/framework.py and scripts/sim/boundary.py.
//
// Operator form.  The reference applies A = M + omega J^T diag(dfdt k) J with a scatter
// (compute_Ax_part2, ip.py:83-121: two float atomics per pair).  Neighbour rows here are
// complete (no truncation: an overflowing row is an error) and therefore symmetric, and
// gradW_ji = -gradW_ij exactly, so the scatter equals the gather
//     (A p)_i = m_i p_i + omega * ( sum_{j dyn} (m_j y_i + m_i y_j) gradW_ij + y_i S_i ),
//     y = diag(dfdt k) J p,  S_i = sum_{j static} m_j gradW_ij,
// i.e. one J pass and one J^T pass, both pure gathers with no atomics.  Rows whose closed
// neighbourhood holds no particle with dfdt > 0 contribute exactly 0 and skip their loop.
//
// Pair gradients.  Rows either cache float32 gradW per pair (16 B per pair and pass, the
// reference's layout) or hold ids only (4 B) with gradW evaluated from the positions in every
// pass: 1/|r| from the hardware reciprocal square root (v_rsq_f32, 1 ulp), then
// -45/(pi h^6) (h - |r|)^2 r/|r|.  The evaluation is a function of r_ij = x_i - x_j alone, so
// gradW_ji = -gradW_ij still holds bitwise; density and Aii always use the reference's exact
// expression (in the neighbour build).
//
// Control.  CG dot products are reduced in FP64 on the device; alpha, beta and the stop test
// (ip.py:302-327) are evaluated by a one-block finalize kernel after each reduction, and the
// host enqueues CG iterations in chunks, synchronising once per chunk instead of twice per
// iteration.
#include <math.h>

#include <algorithm>
#include <chrono>

#include "ivp_internal.h"
#include "ivp_launch.h"
#include "ivp_reduce.h"

using namespace ivp;

namespace {

// Row format seen by one pass: REC words per (tile, slot) record and the source of gradW --
// evaluated from positions (GS_EVAL), read from the record as float32 (GS_F32), or read as two
// packed float16 words scaled by Rows::gs from the side array Rows::g16 (GS_F16: the ids-only
// layout, REC = 64, the two words of a slot at twice the word index of its id).
enum { GS_EVAL = 0, GS_F32 = 1, GS_F16 = 2 };
template <unsigned REC_, int GS_>
struct Fmt {
    static constexpr unsigned REC = REC_;
    static constexpr int GS = GS_;
    static constexpr bool EVAL = GS_ == GS_EVAL;
};
using FmtIds = Fmt<64u, GS_EVAL>;
using FmtF32 = Fmt<256u, GS_F32>;
using FmtF16 = Fmt<64u, GS_F16>;
using FmtF16Eval = FmtIds;   // the float16 format's passes outside the IP CG: ids, gradW evaluated

__device__ __forceinline__ float half_lo(uint32_t w) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}

struct Rows {
    const uint32_t* nb;
    const uint32_t* g16;   // float16 gradW side array (float16 format): words 2w, 2w + 1 for id word w
    const int* tile_len;   // per tile: slots to walk, a multiple of ROW_PAD, padded with the row itself
    int Cs;                // slots per row in the tile layout
    int nd;
    const float4* pos;   // particle positions (x, y, z, m), for evaluated gradients
    float h;
    float ks;            // 45 / (pi h^6)
    float gs;            // scale of float16 gradW records
    const uint32_t* loc; // tile windows: the window index of each walked slot, 16 bits per slot
    const int* win_ids;  // WIN_MAX ids per tile
    const int* win_n;    // ids per tile window (-1: none); nullptr when the rows have no windows
};

// gradW_ij = s * r_ij in the evaluated form; returns s
__device__ __forceinline__ float grad_scale(float h, float ks, float rx, float ry, float rz) {
    float d2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
    float inv = __builtin_amdgcn_rsqf(fmaxf(d2, 1e-12f));   // 1 / max(|r|, 1e-6)
    float t = h - d2 * inv;
    return (t * t) * (-ks * inv);
}

// Independent scenes: the scene of each internal particle and the scene states.  A pass over
// scenes (template flag SC) runs on the rows whose scene is in the pass's phase.
struct SceneRef {
    const int* sid;
    const SceneState* ss;
};

// Row gate of a pass.  One scene: every row i < nd is on.  Scenes: the block's rows span at
// most two scenes (s0 = the first row's); gate_block returns false when neither is in `phase`,
// and the whole block then skips (writing no partials, which the per-scene finalize of a scene
// in the phase never reads).
struct Gate {
    int s0;      // scene of the block's first row
    bool two;    // the block's rows span a second scene
    bool on;     // this thread's row takes part
    int seg;     // 0: row of scene s0, 1: row of the second scene
    int s;       // scene of this thread's row
};

__device__ __forceinline__ Gate gate_all(int i, int nd) { return Gate{0, false, i < nd, 0, 0}; }

template <int RPB = BLOCK>
__device__ __forceinline__ bool gate_block(const SceneRef& sc, int blk, int i, int nd, int phase, Gate& g) {
    const int r0 = blk * RPB;
    const int rl = (r0 + RPB < nd ? r0 + RPB : nd) - 1;
    g.s0 = sc.sid[r0];
    const int s1 = sc.sid[rl];
    g.two = s1 != g.s0;
    if (sc.ss[g.s0].phase != phase && sc.ss[s1].phase != phase) return false;
    g.s = i < nd ? sc.sid[i] : g.s0;
    g.seg = g.s != g.s0 ? 1 : 0;
    g.on = i < nd && sc.ss[g.s].phase == phase;
    return true;
}

template <bool SC, int NV>
__device__ __forceinline__ void row_partials(const double (&v)[NV], const Gate& g, double* partials, int stride,
                                             int blk) {
    if (SC)
        block_partials_seg<NV>(v, g.seg, g.two, partials, stride, blk);
    else
        block_partials<NV>(v, partials, stride);
}

// Pair passes run G = PASS_LANES lanes per row: lane g of a row walks its slots g, g + G, ...,
// and the G partial sums are combined by a fixed xor tree.  Adjacent slots of one row name
// neighbours that are mostly adjacent in memory (rows list neighbours in (cell, id) order), so a
// gather instruction over 64 / G rows x G slots touches fewer distinct cache lines than one over
// 64 rows x 1 slot (measured on the dam break's rows: 11.8 against 25.4 lines per gather).
#ifndef IVP_PASS_LANES
#define IVP_PASS_LANES 4
#endif
#ifndef IVP_PASS_UNROLL
#define IVP_PASS_UNROLL 4
#endif
constexpr int PASS_LANES = IVP_PASS_LANES;
constexpr int PASS_UNROLL = IVP_PASS_UNROLL;    // slots in flight per lane
static_assert(PASS_LANES <= MAX_PASS_LANES, "partials are sized for MAX_PASS_LANES lanes per row");

inline dim3 pass_grid(int nd) { return dim3(blocks_for(nd * PASS_LANES)); }

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

template <int G>
__device__ __forceinline__ bool group_any(bool b) {
    int v = b ? 1 : 0;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v |= __shfl_xor(v, o, WAVE);
    return v != 0;
}

// Row and lane of this thread in a pair pass with G lanes per row, over row block blk
template <int G = PASS_LANES>
__device__ __forceinline__ int pass_row(int blk) { return blk * (BLOCK / G) + (int)(threadIdx.x / G); }
template <int G = PASS_LANES>
__device__ __forceinline__ int pass_lane() { return (int)(threadIdx.x % G); }

// Passes over independent scenes take their lanes per row from the context (1 by default: late
// in a solve most blocks belong to scenes in another phase and return at once, and four times
// as many blocks can make that exit the larger cost); the autotuner chooses.
template <bool SC>
constexpr int lanes_for() { return SC ? 1 : PASS_LANES; }

// Per-particle pass data of NB floats.
template <int NB>
struct Vec {
    float v[NB > 0 ? NB : 1];
};

// Walk the dynamic slots of row i in order, calling f(A_j, B_j, gx, gy, gz) with A_j = loadA(j) =
// (x, y, z, a_j) and B_j = loadB(j).  The tile length is a multiple of ROW_PAD, and slots past a
// row's count name the row itself, whose gradW is 0.
// With G lanes per row, this lane (g) walks slots g, g + G, ... of the row.  With ids-only rows and
// four lanes per row, slots g + 8 b and g + 4 + 8 b are adjacent words (row_word0 / slot_word),
// so the lane reads both ids with one 8-byte load; the slots are visited in the same order.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
#ifndef IVP_LOC_PF
#define IVP_LOC_PF 8
#endif
constexpr int LOC_PF = IVP_LOC_PF;   // window-index words requested at once per lane
// With LOC (tile windows, four lanes per row) the lane reads its two slots' window indices as
// one 32-bit word at half the word index of their ids, and loadA / loadB take window indices.
template <class FM, int NB, int G = 1, int U = 4, bool LOC = false, class LA, class LB, class F>
__device__ __forceinline__ void row_pairs(const Rows& R, int i, float4 xi, LA&& loadA, LB&& loadB, F&& f, int g = 0) {
    constexpr unsigned REC = FM::REC;
    static_assert(FM::GS != GS_F16 || REC == 64u, "float16 gradients live beside the ids-only layout");
    const size_t w_row = (size_t)(i >> 6) * (size_t)R.Cs * REC + row_word0<REC>(i & (WAVE - 1));
    const uint32_t* e = R.nb + w_row;
    const int tl = __builtin_amdgcn_readfirstlane(R.tile_len[i >> 6]);   // one tile per wave
    // gw: the slot's gradient words (float32 x, y, z at 64-word strides, or two float16 words)
    auto visit = [&](int j, const uint32_t* gw) {
        float4 Aj = loadA(j);
        Vec<NB> Bj = loadB(j);
        float gx, gy, gz;
        if (FM::GS == GS_EVAL) {
            float rx = xi.x - Aj.x, ry = xi.y - Aj.y, rz = xi.z - Aj.z;
            float sc = grad_scale(R.h, R.ks, rx, ry, rz);
            gx = sc * rx; gy = sc * ry; gz = sc * rz;
        } else if (FM::GS == GS_F32) {
            gx = __uint_as_float(__builtin_nontemporal_load(gw + 64));
            gy = __uint_as_float(__builtin_nontemporal_load(gw + 128));
            gz = __uint_as_float(__builtin_nontemporal_load(gw + 192));
        } else {
            const uint32_t w1 = gw[0], w2 = gw[1];
            gx = R.gs * half_lo(w1); gy = R.gs * half_hi(w1); gz = R.gs * half_lo(w2);
        }
        f(Aj, Bj, gx, gy, gz);
    };
    static_assert(!LOC || (REC == 64u && G == 4), "tile windows index the paired ids-only layout");
    if constexpr (LOC && FM::GS != GS_F16) {
        // all index words of up to LOC_PF blocks are requested before the first is used (the
        // words come from HBM; the neighbours from LDS)
        const uint32_t* l = R.loc + ((w_row + 2 * g) >> 1);
        for (int b0 = 0; b0 < tl; b0 += 8 * LOC_PF) {
            uint32_t q[LOC_PF];
#pragma unroll
            for (int k = 0; k < LOC_PF; ++k)
                q[k] = b0 + 8 * k < tl ? __builtin_nontemporal_load(l + ((b0 >> 3) + k) * 256) : 0u;
#pragma unroll
            for (int k = 0; k < LOC_PF; ++k)
                if (b0 + 8 * k < tl) {
                    visit((int)(q[k] & 0xffffu), nullptr);
                    visit((int)(q[k] >> 16), nullptr);
                }
        }
    } else if constexpr (LOC) {
        const uint32_t* l = R.loc + ((w_row + 2 * g) >> 1);
#pragma unroll((U + 1) / 2)
        for (int b = 0; b < tl; b += 8) {
            const uint32_t q = __builtin_nontemporal_load(l + (b >> 3) * 256);
            if constexpr (FM::GS == GS_F16) {
                const u32x4* g4 = reinterpret_cast<const u32x4*>(R.g16 + 2 * (w_row + 2 * g));
                const u32x4 h = __builtin_nontemporal_load(g4 + (b >> 3) * 256);
                const uint32_t w0[2] = {h.x, h.y}, w1[2] = {h.z, h.w};
                visit((int)(q & 0xffffu), w0);
                visit((int)(q >> 16), w1);
            } else {
                visit((int)(q & 0xffffu), nullptr);
                visit((int)(q >> 16), nullptr);
            }
        }
    } else if constexpr (REC == 64u && G == 4) {
        // tl is a multiple of ROW_PAD = 8: whole blocks; block b's pair for this lane is the
        // 8-byte word pair 2 g of the row's 8 words in block b (512 words per block), and its
        // float16 gradients the 16-byte quad at twice that word index
        const u32x2* e2 = reinterpret_cast<const u32x2*>(e + 2 * g);
#pragma unroll((U + 1) / 2)
        for (int b = 0; b < tl; b += 8) {
            const u32x2 ids = __builtin_nontemporal_load(e2 + (b >> 3) * 256);
            if constexpr (FM::GS == GS_F16) {
                const u32x4* g4 = reinterpret_cast<const u32x4*>(R.g16 + 2 * (w_row + 2 * g));
                const u32x4 q = __builtin_nontemporal_load(g4 + (b >> 3) * 256);
                const uint32_t w0[2] = {q.x, q.y}, w1[2] = {q.z, q.w};
                visit((int)ids.x, w0);
                visit((int)ids.y, w1);
            } else {
                visit((int)ids.x, nullptr);
                visit((int)ids.y, nullptr);
            }
        }
    } else {
#pragma unroll(U)
        for (int k = g; k < tl; k += G) {
            const uint32_t* r = e + slot_word<REC>(k);
            const uint32_t* gw = FM::GS == GS_F16 ? R.g16 + 2 * (size_t)(r - R.nb) : r;
            visit((int)__builtin_nontemporal_load(r), gw);
        }
    }
}

// Tile windows: the block of a four-lane pass over one scene is one 64-row tile.  stage_window
// loads the tile's window into LDS, A[k] = la(win[k]) and B[k] = lb(win[k]) (the pass's B record
// in the first NB components of a float4), and returns true; it returns false, staging nothing,
// when the rows have no windows or this tile has none.  Every thread of the block calls it (it
// synchronises the block); the result is block-uniform.  row_pairs_w then walks a row through
// the window, or through memory when there is none: the same values in the same order.
template <class FM, int G>
constexpr bool has_windows() { return FM::REC == 64u && G == 4; }

template <class FM, int G>
struct WinLds {
    static constexpr int N = has_windows<FM, G>() ? WIN_MAX : 1;
    float4 A[N];
    float4 B[N];
};

template <class FM, int G, class LA, class LB>
__device__ __forceinline__ bool stage_window(const Rows& R, int tile, WinLds<FM, G>& w, LA&& la, LB&& lb) {
    if constexpr (!has_windows<FM, G>()) {
        return false;
    } else {
        static_assert(BLOCK / G == WAVE, "a four-lane pass block is one tile");
        if (R.win_n == nullptr) return false;
        const int nw = __builtin_amdgcn_readfirstlane(R.win_n[tile]);
        if (nw < 0) return false;
        const int* wid = R.win_ids + (size_t)tile * WIN_MAX;
        for (int k = threadIdx.x; k < nw; k += BLOCK) {
            const int j = wid[k];
            w.A[k] = la(j);
            w.B[k] = lb(j);
        }
        __syncthreads();
        return true;
    }
}

template <int NB>
__device__ __forceinline__ Vec<NB> vec_of(const float4& b) {
    Vec<NB> B;
    const float c[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < (NB > 0 ? NB : 1); ++q) B.v[q] = c[q];
    return B;
}

template <class FM, int NB, int G, class LA, class LB, class F>
__device__ __forceinline__ void row_pairs_w(bool win, const WinLds<FM, G>& w, const Rows& R, int i, float4 xi,
                                            LA&& loadA, LB&& loadB, F&& f, int g) {
    if constexpr (has_windows<FM, G>()) {
        if (win) {
            row_pairs<FM, NB, G, PASS_UNROLL, true>(R, i, xi, [&](int k) { return w.A[k]; },
                                                    [&](int k) { return vec_of<NB>(w.B[k]); }, f, g);
            return;
        }
    }
    row_pairs<FM, NB, G, PASS_UNROLL>(R, i, xi, loadA, loadB, f, g);
}

// ---- IP ------------------------------------------------------------------------------------

// compute_constraint (pressure_base.py:169-184) and d = dt v (ip.py:343)
__global__ __launch_bounds__(BLOCK) void k_ip_init(int nd, float dt, const float* __restrict__ density,
                                                   const float* __restrict__ rho0, const float* __restrict__ vel,
                                                   const float4* __restrict__ pos, float* __restrict__ cc,
                                                   float4* __restrict__ d4) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    cc[i] = density[i] - rho0[i];
    d4[i] = make_float4(dt * vel[3 * i + 0], dt * vel[3 * i + 1], dt * vel[3 * i + 2], pos[i].w);
}

// Neighbour record (x, y, z, m) of the J passes: positions are needed only to evaluate gradW;
// with cached gradients only m_j is read, from the w of the pass vector (u4[j].w = m_j).
template <bool EVAL>
__device__ __forceinline__ float4 j_record(const float4* __restrict__ pos, const float4* __restrict__ u4, int j) {
    return EVAL ? pos[j] : make_float4(0.0f, 0.0f, 0.0f, u4[j].w);
}

// xyz of a float4 array entry as pass data
__device__ __forceinline__ Vec<3> vec3_of(const float4* __restrict__ u4, int j) {
    float4 u = u4[j];
    Vec<3> B;
    B.v[0] = u.x; B.v[1] = u.y; B.v[2] = u.z;
    return B;
}

struct PackArgs {
    Rows R;
    const float4* d4;
    const float* Ss;
    const float* cc;
    const float* kk;
    const float* rho0;
    const int* rowcnt;
    float4* q4;      // (f k, dfdt k, m, 0) gathered by the rhs pass
    float* sfac;     // dfdt k
    float* jd;       // J d, kept for the divergence of the final d
    double* partials;
    int stride;
    SceneRef sc;
    // when set: the pass runs only if the CG of this state has stopped without failing (a pass
    // queued behind a CG chunk's read-back)
    const SolverState* gate = nullptr;
};

// A CG of state st has stopped, and not on a non-positive p.Ap
__device__ __forceinline__ bool cg_stopped(const SolverState* st) { return st->active == 0 && st->status == 0; }

// ip_pack_fused (ip.py:124-171) + coef_wise_div + compute_avg_density_error
// (math_utils.py:76-91, pressure_base.py:187-203, 413-423).  J d in the form of
// compute_J_x_kernel (pressure_base.py:111-127): sum_j m_j dot(d_i - d_j, gradW_ij) + dot(d_i, S_i).
template <class FM, bool SC, int G = lanes_for<SC>()>
__global__ __launch_bounds__(BLOCK) void k_ip_pack(PackArgs a) {
    if (!SC && a.gate && !cg_stopped(a.gate)) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    Gate gt = gate_all(i, a.R.nd);
    if (SC && !gate_block<BLOCK / G>(a.sc, blk, i, a.R.nd, PH_PACK, gt)) return;
    // neighbour record: A = (x, y, z, m), B = d
    auto loadA = [&](int j) { return j_record<FM::EVAL>(a.R.pos, a.d4, j); };
    auto loadB = [&](int j) { return vec3_of(a.d4, j); };
    __shared__ WinLds<FM, G> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.d4[j]; });
    double err = 0.0, nact = 0.0, npairs = 0.0;
    float acc = 0.0f;
    if (gt.on) {
        float4 di = a.d4[i];
        row_pairs_w<FM, 3, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                             [&](const float4& Aj, const Vec<3>& Bj, float gx, float gy, float gz) {
                                 float dx = di.x - Bj.v[0], dy = di.y - Bj.v[1], dz = di.z - Bj.v[2];
                                 acc = fmaf(Aj.w, fmaf(dz, gz, fmaf(dy, gy, dx * gx)), acc);
                             }, g);
        acc = group_sum<G>(acc);
    }
    if (gt.on && g == 0) {
        float4 di = a.d4[i];
        float jd = acc + (di.x * a.Ss[3 * i + 0] + di.y * a.Ss[3 * i + 1] + di.z * a.Ss[3 * i + 2]);
        a.jd[i] = jd;
        float t = jd + a.cc[i];
        float f = fmaxf(t, 0.0f);
        float dfdt = t >= 1e-6f ? 1.0f : 0.0f;
        float k = a.kk[i];
        float sk = dfdt * k;
        a.sfac[i] = sk;
        a.q4[i] = make_float4(f * k, sk, di.w, 0.0f);
        float denom = a.rho0[i];
        if (denom < 1e-6f) denom = 1e-6f;
        err = (double)fmaxf(f / denom, 0.0f);
        if (sk != 0.0f) {
            nact = 1.0;
            npairs = (double)a.rowcnt[i];
        }
    }
    double v[3] = {err, nact, npairs};
    row_partials<SC, 3>(v, gt, a.partials, a.stride, blk);
}

struct RhsArgs {
    Rows R;
    const float4* q4;
    const float* Ss;
    float omega;
    float* b;
    int* nbact;
    const int* rowcnt;
    double* partials;
    int stride;
    SceneRef sc;
};

// ip_rhs_fused (ip.py:174-215): b = omega J^T (f k); also marks rows whose CG J^T loop can be
// non-zero (own dfdt > 0 or a dynamic neighbour with dfdt > 0).
template <class FM, bool SC, int G = lanes_for<SC>()>
__global__ __launch_bounds__(BLOCK) void k_ip_rhs(RhsArgs a) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    Gate gt = gate_all(i, a.R.nd);
    if (SC && !gate_block<BLOCK / G>(a.sc, blk, i, a.R.nd, PH_RHS, gt)) return;
    // neighbour record: A = (x, y, z, m), B = (f k, dfdt k); q4 = (f k, dfdt k, m, 0), so B is
    // one 8-byte gather
    auto loadA = [&](int j) {
        return FM::EVAL ? a.R.pos[j] : make_float4(0.0f, 0.0f, 0.0f, a.q4[j].z);
    };
    auto loadB = [&](int j) {
        float4 q = a.q4[j];
        Vec<2> B;
        B.v[0] = q.x; B.v[1] = q.y;
        return B;
    };
    __shared__ WinLds<FM, G> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.q4[j]; });
    double touched = 0.0, tpairs = 0.0;
    if (gt.on) {
        float4 qi = a.q4[i];
        float fki = qi.x, mi = qi.z;
        bool act = qi.y != 0.0f;
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
        row_pairs_w<FM, 2, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                             [&](const float4& Aj, const Vec<2>& Bj, float gx, float gy, float gz) {
                                 float sc = fmaf(Aj.w, fki, mi * Bj.v[0]);
                                 rx = fmaf(sc, gx, rx); ry = fmaf(sc, gy, ry); rz = fmaf(sc, gz, rz);
                                 act = act || (Bj.v[1] != 0.0f);
                             }, g);
        rx = group_sum<G>(rx);
        ry = group_sum<G>(ry);
        rz = group_sum<G>(rz);
        act = group_any<G>(act);
        if (g == 0) {
            rx += fki * a.Ss[3 * i + 0];
            ry += fki * a.Ss[3 * i + 1];
            rz += fki * a.Ss[3 * i + 2];
            a.b[3 * i + 0] = a.omega * rx;
            a.b[3 * i + 1] = a.omega * ry;
            a.b[3 * i + 2] = a.omega * rz;
            a.nbact[i] = act ? 1 : 0;
            if (act) {
                touched = 1.0;
                tpairs = (double)a.rowcnt[i];
            }
        }
    }
    double v[2] = {touched, tpairs};
    row_partials<SC, 2>(v, gt, a.partials, a.stride, blk);
}

struct CGInitArgs {
    int nd;
    const float* b;
    const float4* pos;
    float* sol;
    float* r;
    float4* p4;
    float2* y2;
    double* partials;
    int stride;
    SceneRef sc;
};

// IP.PCG prologue (ip.py:292-304): x = 0, r = b, rr = r.r, p = r
template <bool SC>
__global__ __launch_bounds__(BLOCK) void k_cg_init(CGInitArgs a) {
    const int blk = blockIdx.x;
    int i = blk * BLOCK + threadIdx.x;
    Gate gt = gate_all(i, a.nd);
    if (SC && !gate_block(a.sc, blk, i, a.nd, PH_RHS, gt)) return;
    double rr = 0.0;
    if (gt.on) {
        float bx = a.b[3 * i + 0], by = a.b[3 * i + 1], bz = a.b[3 * i + 2];
        float m = a.pos[i].w;
        a.sol[3 * i + 0] = 0.0f; a.sol[3 * i + 1] = 0.0f; a.sol[3 * i + 2] = 0.0f;
        a.r[3 * i + 0] = bx; a.r[3 * i + 1] = by; a.r[3 * i + 2] = bz;
        a.p4[i] = make_float4(bx, by, bz, m);
        a.y2[i] = make_float2(0.0f, m);
        rr = (double)(bx * bx + by * by + bz * bz);
    }
    double v[1] = {rr};
    row_partials<SC, 1>(v, gt, a.partials, a.stride, blk);
}

struct CGJArgs {
    Rows R;
    const float4* p4;
    const float* Ss;
    const float* sfac;
    float2* y2;
    const SolverState* st;
    SceneRef sc;
};

// compute_Jx_pcg (ip.py:19-60): y_i = dfdt_i k_i (J p)_i on rows with dfdt > 0, J p in the form
// sum_j dot(p_i - p_j, m_j gradW_ij) + dot(p_i, S_i) (ip.py:51-58)
template <class FM, bool SC, int G = lanes_for<SC>()>
__global__ __launch_bounds__(BLOCK) void k_cg_J(CGJArgs a) {
    if (!SC && a.st->active == 0) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    Gate gt = gate_all(i, a.R.nd);
    if (SC && !gate_block<BLOCK / G>(a.sc, blk, i, a.R.nd, PH_CG, gt)) return;
    const float sk = gt.on ? a.sfac[i] : 0.0f;
    // neighbour record: A = (x, y, z, m), B = p
    auto loadA = [&](int j) { return j_record<FM::EVAL>(a.R.pos, a.p4, j); };
    auto loadB = [&](int j) { return vec3_of(a.p4, j); };
    __shared__ WinLds<FM, G> w;
    bool win = false;
    if constexpr (has_windows<FM, G>()) {
        // a tile without active rows neither stages nor walks
        if (!__syncthreads_or(sk != 0.0f)) return;
        win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.p4[j]; });
    }
    if (sk == 0.0f) return;
    float4 pi = a.p4[i];
    float acc = 0.0f;
    row_pairs_w<FM, 3, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                         [&](const float4& Aj, const Vec<3>& Bj, float gx, float gy, float gz) {
                             float dx = pi.x - Bj.v[0], dy = pi.y - Bj.v[1], dz = pi.z - Bj.v[2];
                             acc = fmaf(Aj.w, fmaf(dz, gz, fmaf(dy, gy, dx * gx)), acc);
                         }, g);
    acc = group_sum<G>(acc);
    if (g != 0) return;
    acc = acc + (pi.x * a.Ss[3 * i + 0] + pi.y * a.Ss[3 * i + 1] + pi.z * a.Ss[3 * i + 2]);
    a.y2[i] = make_float2(acc * sk, pi.w);
}

struct CGJTArgs {
    Rows R;
    const float4* p4;
    const float2* y2;
    const float* Ss;
    const int* nbact;
    float omega;
    float* Ap;
    double* partials;
    int stride;
    const SolverState* st;
    SceneRef sc;
};

// compute_Ax_part1 + compute_Ax_part2 (ip.py:63-121) in gather form, fused with p.Ap
// (ip.py:307)
template <class FM, bool SC, int G = lanes_for<SC>()>
__global__ __launch_bounds__(BLOCK) void k_cg_JT(CGJTArgs a) {
    if (!SC && a.st->active == 0) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    Gate gt = gate_all(i, a.R.nd);
    if (SC && !gate_block<BLOCK / G>(a.sc, blk, i, a.R.nd, PH_CG, gt)) return;
    const bool touched = gt.on && a.nbact[i] != 0;
    // neighbour record: A = (x, y, z, m), B = y; y2 = (y, m)
    auto loadA = [&](int j) {
        return FM::EVAL ? a.R.pos[j] : make_float4(0.0f, 0.0f, 0.0f, a.y2[j].y);
    };
    auto loadB = [&](int j) {
        Vec<1> B;
        B.v[0] = a.y2[j].x;
        return B;
    };
    __shared__ WinLds<FM, G> w;
    bool win = false;
    if constexpr (has_windows<FM, G>()) {
        // a tile without touched rows stages nothing (its rows only write m p)
        if (__syncthreads_or(touched))
            win = stage_window(a.R, blk, w, loadA, [&](int j) {
                const float2 y = a.y2[j];
                return make_float4(y.x, y.y, 0.0f, 0.0f);
            });
    }
    double pap = 0.0;
    if (gt.on) {
        float4 pi = a.p4[i];
        float mi = pi.w;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (touched) {
            float yi = a.y2[i].x;
            row_pairs_w<FM, 1, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                                 [&](const float4& Aj, const Vec<1>& Bj, float gx, float gy, float gz) {
                                     float sc = fmaf(Aj.w, yi, mi * Bj.v[0]);
                                     ax = fmaf(sc, gx, ax); ay = fmaf(sc, gy, ay); az = fmaf(sc, gz, az);
                                 }, g);
            ax = group_sum<G>(ax);
            ay = group_sum<G>(ay);
            az = group_sum<G>(az);
            ax += yi * a.Ss[3 * i + 0];
            ay += yi * a.Ss[3 * i + 1];
            az += yi * a.Ss[3 * i + 2];
        }
        float Apx = mi * pi.x + a.omega * ax;
        float Apy = mi * pi.y + a.omega * ay;
        float Apz = mi * pi.z + a.omega * az;
        if (g == 0) {
            a.Ap[3 * i + 0] = Apx; a.Ap[3 * i + 1] = Apy; a.Ap[3 * i + 2] = Apz;
            pap = (double)(pi.x * Apx + pi.y * Apy + pi.z * Apz);
        }
    }
    double v[1] = {pap};
    row_partials<SC, 1>(v, gt, a.partials, a.stride, blk);
}

struct CGUpdArgs {
    int nd;
    float* sol;
    float* r;
    const float4* p4;
    const float* Ap;
    double* partials;
    int stride;
    const SolverState* st;
    SceneRef sc;
};

// pcg_update_x_r_scalar (math_utils.py:180-195) + r.r (ip.py:314-317)
template <bool SC>
__global__ __launch_bounds__(BLOCK) void k_cg_update(CGUpdArgs a) {
    if (!SC && a.st->active == 0) return;
    const int blk = blockIdx.x;
    int i = blk * BLOCK + threadIdx.x;
    Gate gt = gate_all(i, a.nd);
    if (SC && !gate_block(a.sc, blk, i, a.nd, PH_CG, gt)) return;
    float alpha = (float)(SC ? a.sc.ss[gt.s].cg.alpha : a.st->alpha);
    double rr = 0.0;
    if (gt.on) {
        float4 p = a.p4[i];
        float x0 = a.sol[3 * i + 0] + alpha * p.x;
        float x1 = a.sol[3 * i + 1] + alpha * p.y;
        float x2 = a.sol[3 * i + 2] + alpha * p.z;
        float r0 = a.r[3 * i + 0] - alpha * a.Ap[3 * i + 0];
        float r1 = a.r[3 * i + 1] - alpha * a.Ap[3 * i + 1];
        float r2 = a.r[3 * i + 2] - alpha * a.Ap[3 * i + 2];
        a.sol[3 * i + 0] = x0; a.sol[3 * i + 1] = x1; a.sol[3 * i + 2] = x2;
        a.r[3 * i + 0] = r0; a.r[3 * i + 1] = r1; a.r[3 * i + 2] = r2;
        rr = (double)(r0 * r0 + r1 * r1 + r2 * r2);
    }
    double v[1] = {rr};
    row_partials<SC, 1>(v, gt, a.partials, a.stride, blk);
}

// p = r + beta p (add_vec3, ip.py:323-326)
template <bool SC>
__global__ __launch_bounds__(BLOCK) void k_cg_p(int nd, const float* __restrict__ r, float4* __restrict__ p4,
                                                const SolverState* __restrict__ st, SceneRef sc) {
    if (!SC && st->active == 0) return;
    const int blk = blockIdx.x;
    int i = blk * BLOCK + threadIdx.x;
    Gate gt = gate_all(i, nd);
    if (SC && !gate_block(sc, blk, i, nd, PH_CG, gt)) return;
    if (!gt.on) return;
    float beta = (float)(SC ? sc.ss[gt.s].cg.beta : st->beta);
    float4 p = p4[i];
    p.x = r[3 * i + 0] + beta * p.x;
    p.y = r[3 * i + 1] + beta * p.y;
    p.z = r[3 * i + 2] + beta * p.z;
    p4[i] = p;
}

// (x, m) of the CG iterate, the input of the true-residual passes
__global__ __launch_bounds__(BLOCK) void k_x4(int nd, const float* __restrict__ sol, const float4* __restrict__ pos,
                                              float4* __restrict__ xs4, const SolverState* __restrict__ st) {
    if (st->active == 0) return;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    xs4[i] = make_float4(sol[3 * i + 0], sol[3 * i + 1], sol[3 * i + 2], pos[i].w);
}

// Residual replacement: r = b - A x with A applied in float32 from positions, and r.r
__global__ __launch_bounds__(BLOCK) void k_cg_refresh(int nd, const float* __restrict__ b,
                                                      const float* __restrict__ Ax, float* __restrict__ r,
                                                      double* __restrict__ partials, int stride,
                                                      const SolverState* __restrict__ st) {
    if (st->active == 0) return;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    double rr = 0.0;
    if (i < nd) {
        float r0 = b[3 * i + 0] - Ax[3 * i + 0];
        float r1 = b[3 * i + 1] - Ax[3 * i + 1];
        float r2 = b[3 * i + 2] - Ax[3 * i + 2];
        r[3 * i + 0] = r0; r[3 * i + 1] = r1; r[3 * i + 2] = r2;
        rr = (double)(r0 * r0 + r1 * r1 + r2 * r2);
    }
    double v[1] = {rr};
    block_partials<1>(v, partials, stride);
}

// d -= sol (add_vec3 with scale -1, ip.py:389)
template <bool SC>
__global__ __launch_bounds__(BLOCK) void k_ip_dupdate(int nd, float4* __restrict__ d4, const float* __restrict__ sol,
                                                      SceneRef sc, const SolverState* __restrict__ gate) {
    if (gate && !cg_stopped(gate)) return;   // (see PackArgs::gate)
    const int blk = blockIdx.x;
    int i = blk * BLOCK + threadIdx.x;
    Gate gt = gate_all(i, nd);
    if (SC && !gate_block(sc, blk, i, nd, PH_DUPDATE, gt)) return;
    if (!gt.on) return;
    float4 d = d4[i];
    d.x = d.x + -1.0f * sol[3 * i + 0];
    d.y = d.y + -1.0f * sol[3 * i + 1];
    d.z = d.z + -1.0f * sol[3 * i + 2];
    d4[i] = d;
}

// v = (1/dt) d (ip.py:398) and compute_divergence (pressure_base.py:449-457, 240-254):
// max(J v, 0).  The outer loop only ends right after a J pass over the final d, which left
// J d in jd; J is linear, so J v = (1/dt) J d (equal up to float rounding of the terms).
__global__ __launch_bounds__(BLOCK) void k_ip_finish(int nd, float inv_dt, const float4* __restrict__ d4,
                                                     const float* __restrict__ jd, float* __restrict__ vel,
                                                     float* __restrict__ divg) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    float4 d = d4[i];
    vel[3 * i + 0] = inv_dt * d.x; vel[3 * i + 1] = inv_dt * d.y; vel[3 * i + 2] = inv_dt * d.z;
    divg[i] = fmaxf(inv_dt * jd[i], 0.0f);
}

// ---- PRJ (prj.py) --------------------------------------------------------------------------
//
// Projected Jacobi on the constant-density constraint.  A pass is two gathers over the rows:
// the head dt v = dt v0 - J^T lambda / m (the J^T form of the right-hand side pass) and the tail
// r = c + J (dt v) with lambda = max(lambda + omega D r, 0) and the density-error partial.  The
// tail reads only its own row's lambda, so lambda is updated in place.  The stop test runs on the
// device after every pass; the host enqueues passes in chunks and reads the state once per chunk.

struct PrjArgs {
    Rows R;
    const float* Ss;
    const float* v0;      // dt v at the start of the loop, 3 per row
    float2* lam;          // (dt^2 lambda, m)
    float4* d4;           // dt v of the current pass, w = m
    const float* cc;      // constraint c = rho - rho0 (constant-density loop)
    const float* kk;      // D^-1 = k = 1 / (Aii + 1e-3)
    const float* rho0;
    const int* rowtot;    // neighbours of the row (dynamic and static)
    float* jd;            // J (dt v), kept for the divergence
    float omega;
    int C;                // neighborCacheSize: the reference's neighbour count is capped at it
    int df_threshold;     // divergence-free loop: rows with more neighbours than this take part
    double* partials;
    int stride;
    const SolverState* st;
};

// PRJ.solve prologue (prj.py:307-324): c = rho - rho0 (constant-density loop only),
// dt v0 = dt v, lambda = 0 or 0.5 x the previous solve's (warm start, add_scalar); prev is in
// user order.
template <bool DF>
__global__ __launch_bounds__(BLOCK) void k_prj_init(int nd, float dt, const float* __restrict__ density,
                                                    const float* __restrict__ rho0, const float* __restrict__ vel,
                                                    const float4* __restrict__ pos, const int* __restrict__ uid,
                                                    const float* __restrict__ prev, int warm, float* __restrict__ cc,
                                                    float* __restrict__ v0, float2* __restrict__ lam) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    if (!DF) cc[i] = density[i] - rho0[i];
    v0[3 * i + 0] = dt * vel[3 * i + 0];
    v0[3 * i + 1] = dt * vel[3 * i + 1];
    v0[3 * i + 2] = dt * vel[3 * i + 2];
    float l = 0.0f;
    if (warm) l = l + 0.5f * prev[uid[i]];
    lam[i] = make_float2(l, pos[i].w);
}

// pj_head_fused (prj.py:17-59); also the head of solve_df (prj.py:243-253)
template <class FM>
__global__ __launch_bounds__(BLOCK) void k_prj_head(PrjArgs a) {
    if (a.st->active == 0) return;
    constexpr int G = PASS_LANES;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    // neighbour record: A = (x, y, z, m), B = lambda; lam = (lambda, m)
    auto loadA = [&](int j) {
        return FM::EVAL ? a.R.pos[j] : make_float4(0.0f, 0.0f, 0.0f, a.lam[j].y);
    };
    auto loadB = [&](int j) {
        Vec<1> B;
        B.v[0] = a.lam[j].x;
        return B;
    };
    __shared__ WinLds<FM, G> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) {
        const float2 l = a.lam[j];
        return make_float4(l.x, l.y, 0.0f, 0.0f);
    });
    if (i >= a.R.nd) return;   // the G lanes of a row leave together
    const float2 li = a.lam[i];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    row_pairs_w<FM, 1, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                                     [&](const float4& Aj, const Vec<1>& Bj, float gx, float gy, float gz) {
                                         float s = fmaf(Aj.w, li.x, li.y * Bj.v[0]);
                                         ax = fmaf(s, gx, ax); ay = fmaf(s, gy, ay); az = fmaf(s, gz, az);
                                     }, g);
    ax = group_sum<G>(ax);
    ay = group_sum<G>(ay);
    az = group_sum<G>(az);
    if (g != 0) return;
    ax += li.x * a.Ss[3 * i + 0];
    ay += li.x * a.Ss[3 * i + 1];
    az += li.x * a.Ss[3 * i + 2];
    const float denom = li.y < 1e-6f ? 1e-6f : li.y;
    a.d4[i] = make_float4(a.v0[3 * i + 0] - ax / denom, a.v0[3 * i + 1] - ay / denom, a.v0[3 * i + 2] - az / denom,
                          li.y);
}

// pj_tail_fused (prj.py:62-115); DF: the pass of solve_df (prj.py:255-283: J v masked by the
// active rows, D^-1 J v, lambda + omega D^-1 J v, max(., 0), error of J v / rho0)
template <class FM, bool DF>
__global__ __launch_bounds__(BLOCK) void k_prj_tail(PrjArgs a) {
    if (a.st->active == 0) return;
    constexpr int G = PASS_LANES;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row<G>(blk), g = pass_lane<G>();
    const bool on = i < a.R.nd;
    // neighbour record: A = (x, y, z, m), B = dt v
    auto loadA = [&](int j) { return j_record<FM::EVAL>(a.R.pos, a.d4, j); };
    auto loadB = [&](int j) { return vec3_of(a.d4, j); };
    __shared__ WinLds<FM, G> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.d4[j]; });
    double err = 0.0;
    if (on) {
        const float4 di = a.d4[i];
        float acc = 0.0f;
        row_pairs_w<FM, 3, G>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                                         [&](const float4& Aj, const Vec<3>& Bj, float gx, float gy, float gz) {
                                             float dx = di.x - Bj.v[0], dy = di.y - Bj.v[1], dz = di.z - Bj.v[2];
                                             acc = fmaf(Aj.w, fmaf(dz, gz, fmaf(dy, gy, dx * gx)), acc);
                                         }, g);
        acc = group_sum<G>(acc);
        if (g == 0) {
            const float jd = acc + (di.x * a.Ss[3 * i + 0] + di.y * a.Ss[3 * i + 1] + di.z * a.Ss[3 * i + 2]);
            a.jd[i] = jd;
            float l = a.lam[i].x, e;
            if (!DF) {
                const float r = jd + a.cc[i];
                l = fmaxf(l + a.omega * r * a.kk[i], 0.0f);
                e = r;
            } else {
                const int cnt = a.rowtot[i] < a.C ? a.rowtot[i] : a.C;
                const float dJ = jd * (cnt > a.df_threshold ? 1.0f : 0.0f);
                const float Dr = dJ * a.kk[i];
                l = fmaxf(l + a.omega * Dr, 0.0f);
                e = dJ;
            }
            a.lam[i] = make_float2(l, di.w);
            float rho0 = a.rho0[i];
            if (rho0 < 1e-6f) rho0 = 1e-6f;
            err = (double)fmaxf(e / rho0, 0.0f);
        }
    }
    double v[1] = {err};
    block_partials<1>(v, a.partials, a.stride);
}

struct PrjFinArgs {
    const double* partials;
    int stride;
    int nblocks;
    SolverState* st;
    double tol;
    int max_pass;
    int fixed;
    int nd;
    double* hist;      // density error per pass (nullptr: not recorded)
    int hist_cap;
};

// The stop test of the pass loop (prj.py:197-206, 277-281): err = avg(max(r / rho0, 0)); stop when
// err < tol after at least one update, or at pass max_iteration_opt.  st->iter is the pass index
// (stats_outer_iters once the loop has stopped); rr0 / rr_last hold the first and the last err.
__global__ __launch_bounds__(FIN_BLOCK) void k_fin_prj(PrjFinArgs a) {
    SolverState* st = a.st;
    if (st->active == 0) return;
    double t[3] = {0.0, 0.0, 0.0};
    fold_partials<1>(a.partials, a.stride, a.nblocks, t);
    if (threadIdx.x != 0) return;
    const int pass = st->iter;
    const double err = t[0] / (double)a.nd;
    if (a.hist && pass < a.hist_cap) a.hist[pass] = err;
    if (pass == 0) st->rr0 = err;
    st->rr_last = err;
    const bool converged = err < a.tol && !a.fixed;
    if ((converged && pass >= 1) || pass >= a.max_pass)
        st->active = 0;
    else
        st->iter = pass + 1;
}

__global__ void k_prj_state_init(SolverState* st) {
    st->iter = 0;
    st->active = 1;
    st->status = 0;
    st->rr0 = 0.0;
    st->rr_last = 0.0;
}

// wp.copy(p_previous, dt2_lam) (prj.py:204, 280), into user order
__global__ __launch_bounds__(BLOCK) void k_prj_save(int nd, const float2* __restrict__ lam, const int* __restrict__ uid,
                                                    float* __restrict__ prev) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    prev[uid[i]] = lam[i].x;
}

// ---- explicit viscosity (viscosity.py:9-108), Jacobi reads ---------------------------------

__global__ __launch_bounds__(BLOCK) void k_v4(int nd, const float* __restrict__ vel, const float* __restrict__ density,
                                              float4* __restrict__ v4) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    v4[i] = make_float4(vel[3 * i + 0], vel[3 * i + 1], vel[3 * i + 2], density[i]);
}

// viscosity.py:62-108 with Jacobi reads.  Neighbour record: A = (x, y, z, m_j), m_j = 0 for a
// non-fluid neighbour (its term is then exactly 0, as the reference's skip), B = (v_j, rho_j).
template <class FM, bool ALLFLUID>
__global__ __launch_bounds__(BLOCK) void k_visc(Rows R, float coeff, float dt, int fluid,
                                                const float4* __restrict__ v4, const float* __restrict__ rho0,
                                                const int* __restrict__ mat, float* __restrict__ vel) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    auto loadA = [&](int j) {
        float4 A = R.pos[j];
        if (!ALLFLUID && mat[j] != fluid) A.w = 0.0f;
        return A;
    };
    auto loadB = [&](int j) {
        float4 u = v4[j];
        Vec<4> B;
        B.v[0] = u.x; B.v[1] = u.y; B.v[2] = u.z; B.v[3] = u.w;
        return B;
    };
    __shared__ WinLds<FM, PASS_LANES> w;
    const bool win = stage_window(R, blk, w, loadA, [&](int j) { return v4[j]; });
    if (i >= R.nd) return;
    if (!ALLFLUID && mat[i] != fluid) return;
    float4 xi = R.pos[i];
    float4 vi = v4[i];
    float rho_i = vi.w, rho0_i = rho0[i];
    float eps2 = (0.01f * R.h) * (0.01f * R.h);
    float k_visc = 2.0f * (3.0f + 2.0f);
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    row_pairs_w<FM, 4, PASS_LANES>(win, w, R, i, xi, loadA, loadB,
                         [&](const float4& xj, const Vec<4>& vj, float gx, float gy, float gz) {
                             float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                             float rn = len3(rx, ry, rz);
                             float vx = vi.x - vj.v[0], vy = vi.y - vj.v[1], vz = vi.z - vj.v[2];
                             float vxy = vx * rx + vy * ry + vz * rz;
                             float rho_j = vj.v[3];
                             float K = 2.0f * rho0_i / (rho_i + rho_j);
                             float sc = k_visc * coeff * (xj.w / rho_j) * (vxy / (rn * rn + eps2));
                             ax = ax + K * (sc * gx);
                             ay = ay + K * (sc * gy);
                             az = az + K * (sc * gz);
                         }, g);
    ax = group_sum<PASS_LANES>(ax);
    ay = group_sum<PASS_LANES>(ay);
    az = group_sum<PASS_LANES>(az);
    if (g != 0) return;
    vel[3 * i + 0] = vi.x + ax * dt;
    vel[3 * i + 1] = vi.y + ay * dt;
    vel[3 * i + 2] = vi.z + az * dt;
}

// ---- Weiler 2018 (weiler18.py) -------------------------------------------------------------

// (v, m / rho) for the operator gathers and factor = dt mu_i / rho_i, mu_i = coeff rho0_i
__global__ __launch_bounds__(BLOCK) void k_w18_prep(int nd, float coeff, float dt, const float* __restrict__ vel,
                                                    const float4* __restrict__ pos, const float* __restrict__ density,
                                                    const float* __restrict__ rho0, float4* __restrict__ v4,
                                                    float* __restrict__ fac) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    float rho = density[i];
    v4[i] = make_float4(vel[3 * i + 0], vel[3 * i + 1], vel[3 * i + 2], pos[i].w / rho);
    float mu = coeff * rho0[i];
    fac[i] = dt * mu / rho;
}

__device__ __forceinline__ void inverse33(const float a[9], float out[9]) {
    float c00 = a[4] * a[8] - a[5] * a[7];
    float c01 = a[5] * a[6] - a[3] * a[8];
    float c02 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.0f) {
        for (int q = 0; q < 9; ++q) out[q] = 0.0f;
        return;
    }
    float id = 1.0f / det;
    out[0] = c00 * id;
    out[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    out[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    out[3] = c01 * id;
    out[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    out[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    out[6] = c02 * id;
    out[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    out[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// Weiler18 neighbour record: A = (x, y, z, V_j), V_j = m_j / rho_j for a fluid neighbour and 0
// otherwise (a zero term, as the reference's skip of non-fluid neighbours); B = the operator input.
template <bool ALLFLUID>
__device__ __forceinline__ float4 w18_record(const float4* __restrict__ pos, const float4* __restrict__ u4,
                                             const int* __restrict__ mat, int fluid, int j) {
    float4 x = pos[j];
    float V = (ALLFLUID || mat[j] == fluid) ? u4[j].w : 0.0f;
    return make_float4(x.x, x.y, x.z, V);
}

struct W18Args {
    Rows R;
    int fluid;
    const float* fac;
    const int* mat;
    float* pinv;
    float4* v4;   // operator input (v for the rhs, p for CG), w = m / rho
    float* out;   // rhs or Ap
    float* dv; float* r; float* z;
    double* partials;
    int stride;
    const SolverState* st;
    int precond;            // use_preconditioner: block-Jacobi blocks, else identity (plain CG)
    const int* uid;         // user index of each internal row
    const float* dv_prev;   // the last solve's correction in user order (warm starts)
};

// The operator pass's neighbour record: w18_record, with w = -10 ks V_j when the gradients are
// evaluated (the constant of the folded pair scalar, w18_Av_row)
template <class FM, bool ALLFLUID>
__device__ __forceinline__ float4 w18_Av_record(const W18Args& a, int j) {
    float4 r = w18_record<ALLFLUID>(a.R.pos, a.v4, a.mat, a.fluid, j);
    if constexpr (FM::EVAL) r.w = r.w * (-10.0f * a.R.ks);
    return r;
}

// (A u)_i = u_i - factor_i * 10 sum_fluid (m_j/rho_j) (u_ij . x_ij) / (r^2 + 0.01 h^2) gradW_ij
// (compute_Av_weiler18, weiler18.py:103-172).  The caller stages the block's window.
template <class FM, bool ALLFLUID, int G, class LA, class LB>
__device__ __forceinline__ float3 w18_Av_row(const W18Args& a, int i, float4 ui, bool win, const WinLds<FM, G>& w,
                                             LA&& loadA, LB&& loadB, int g = 0) {
    if (!ALLFLUID && a.mat[i] != a.fluid) return make_float3(ui.x, ui.y, ui.z);
    float4 xi = a.R.pos[i];
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
    float eps2 = 0.01f * a.R.h * a.R.h;
    if constexpr (FM::EVAL) {
        // evaluated gradients: gradW_ij = -ks (h - |r|)^2 x_ij / |r| and the pair factor
        // 10 V_j (u_ij . x_ij) / (r^2 + 0.01 h^2) fold into one scalar per pair, the constant
        // -10 ks carried by the neighbour record's w (w18_Av_record); x_ij is scaled once
        const float h = a.R.h;
        row_pairs_w<FM, 3, G>(win, w, a.R, i, xi, loadA, loadB,
                             [&](const float4& xj, const Vec<3>& uj, float, float, float) {
                                 float x0 = xi.x - xj.x, x1 = xi.y - xj.y, x2 = xi.z - xj.z;
                                 float v0 = ui.x - uj.v[0], v1 = ui.y - uj.v[1], v2 = ui.z - uj.v[2];
                                 float r_sq = fmaf(x2, x2, fmaf(x1, x1, x0 * x0));
                                 float vx = fmaf(v2, x2, fmaf(v1, x1, v0 * x0));
                                 float inv = __builtin_amdgcn_rsqf(fmaxf(r_sq, 1e-12f));   // 1 / max(|r|, 1e-6)
                                 float t = fmaf(-r_sq, inv, h);
                                 float wp = xj.w * ((vx * __builtin_amdgcn_rcpf(r_sq + eps2)) * ((t * t) * inv));
                                 lx = fmaf(wp, x0, lx); ly = fmaf(wp, x1, ly); lz = fmaf(wp, x2, lz);
                             }, g);
        lx = group_sum<G>(lx);
        ly = group_sum<G>(ly);
        lz = group_sum<G>(lz);
        float f = a.fac[i];
        return make_float3(ui.x - f * lx, ui.y - f * ly, ui.z - f * lz);
    }
    float dd = 10.0f;
    row_pairs_w<FM, 3, G>(win, w, a.R, i, xi, loadA, loadB,
                         [&](const float4& xj, const Vec<3>& uj, float gx, float gy, float gz) {
                             float x0 = xi.x - xj.x, x1 = xi.y - xj.y, x2 = xi.z - xj.z;
                             float v0 = ui.x - uj.v[0], v1 = ui.y - uj.v[1], v2 = ui.z - uj.v[2];
                             float r_sq = fmaf(x2, x2, fmaf(x1, x1, x0 * x0));
                             float vx = fmaf(v2, x2, fmaf(v1, x1, v0 * x0));
                             float sc = dd * xj.w * (vx * __builtin_amdgcn_rcpf(r_sq + eps2));
                             lx = fmaf(sc, gx, lx); ly = fmaf(sc, gy, ly); lz = fmaf(sc, gz, lz);
                         }, g);
    lx = group_sum<G>(lx);
    ly = group_sum<G>(ly);
    lz = group_sum<G>(lz);
    float f = a.fac[i];
    return make_float3(ui.x - f * lx, ui.y - f * ly, ui.z - f * lz);
}

// The solve's set-up in one pass over the pairs: compute_precond_weiler18 (weiler18.py:21-85,
// the inverse of I - factor_i L_ii), the right-hand side v - A v (weiler18.py:390-393) and the
// PCG prologue (weiler18.py:320-335: dv = 0, r = rhs, z = P r, p = z, r.z).  The preconditioner
// block and the operator share each pair's gradient, (m_j / rho_j) and 1 / (r^2 + 0.01 h^2); every
// accumulator sums the same terms in the same order as it would in a pass of its own.
// With WARM (warm_start, weiler18.py:311-316) the pass stores the blocks and the right-hand side
// (a.out) and loads the initial correction dv = dv_prev into dv and p4; r = rhs - A dv follows
// from the operator pass on p4 and k_w18_warm_init.  Without the preconditioner the blocks are
// the identity, so z = P r is r exactly (apply_preconditioner's copy, weiler18.py:279-292).
template <class FM, bool ALLFLUID, bool WARM = false>
__global__ __launch_bounds__(BLOCK) void k_w18_setup(W18Args a, float4* __restrict__ p4) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    auto loadA = [&](int j) { return w18_record<ALLFLUID>(a.R.pos, a.v4, a.mat, a.fluid, j); };
    auto loadB = [&](int j) { return vec3_of(a.v4, j); };
    __shared__ WinLds<FM, PASS_LANES> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.v4[j]; });
    double rz = 0.0;
    if (i < a.R.nd) {
        const float4 vi = a.v4[i];
        const bool fl = ALLFLUID || a.mat[i] == a.fluid;
        float L[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) L[q] = 0.0f;
        float lx = 0.0f, ly = 0.0f, lz = 0.0f;
        if (fl) {
            const float4 xi = a.R.pos[i];
            const float eps2 = 0.01f * a.R.h * a.R.h;
            const float dd = 10.0f;
            row_pairs_w<FM, 3, PASS_LANES>(win, w, a.R, i, xi, loadA, loadB,
                [&](const float4& xj, const Vec<3>& uj, float gx, float gy, float gz) {
                    float x[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
                    float v0 = vi.x - uj.v[0], v1 = vi.y - uj.v[1], v2 = vi.z - uj.v[2];
                    float r_sq = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
                    float rc = __builtin_amdgcn_rcpf(r_sq + eps2);
                    float t = dd * xj.w;
                    float c = t * rc;
                    float cg[3] = {c * gx, c * gy, c * gz};
#pragma unroll
                    for (int p = 0; p < 3; ++p)
#pragma unroll
                        for (int q = 0; q < 3; ++q) L[3 * p + q] = fmaf(cg[p], x[q], L[3 * p + q]);
                    float vx = fmaf(v2, x[2], fmaf(v1, x[1], v0 * x[0]));
                    float sc = t * (vx * rc);
                    lx = fmaf(sc, gx, lx); ly = fmaf(sc, gy, ly); lz = fmaf(sc, gz, lz);
                }, g);
#pragma unroll
            for (int q = 0; q < 9; ++q) L[q] = group_sum<PASS_LANES>(L[q]);
            lx = group_sum<PASS_LANES>(lx);
            ly = group_sum<PASS_LANES>(ly);
            lz = group_sum<PASS_LANES>(lz);
        }
        if (g == 0) {
            float P[9];
            float Av0 = vi.x, Av1 = vi.y, Av2 = vi.z;
            if (fl) {
                const float f = a.fac[i];
                if (a.precond) {
                    float M[9];
#pragma unroll
                    for (int q = 0; q < 9; ++q) M[q] = ((q % 4) == 0 ? 1.0f : 0.0f) - f * L[q];
                    inverse33(M, P);
                } else {
#pragma unroll
                    for (int q = 0; q < 9; ++q) P[q] = (q % 4) == 0 ? 1.0f : 0.0f;
                }
                Av0 = vi.x - f * lx; Av1 = vi.y - f * ly; Av2 = vi.z - f * lz;
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) P[q] = (q % 4) == 0 ? 1.0f : 0.0f;
            }
            float* Pi = a.pinv + 9 * (size_t)i;
#pragma unroll
            for (int q = 0; q < 9; ++q) Pi[q] = P[q];
            const float r0 = vi.x + -1.0f * Av0, r1 = vi.y + -1.0f * Av1, r2 = vi.z + -1.0f * Av2;
            if constexpr (WARM) {
                // rhs = v - A v; dv = the last correction of this particle
                a.out[3 * i + 0] = r0; a.out[3 * i + 1] = r1; a.out[3 * i + 2] = r2;
                const float* d = a.dv_prev + 3 * (size_t)a.uid[i];
                const float d0 = d[0], d1 = d[1], d2 = d[2];
                a.dv[3 * i + 0] = d0; a.dv[3 * i + 1] = d1; a.dv[3 * i + 2] = d2;
                p4[i] = make_float4(d0, d1, d2, vi.w);
                return;
            }
            const float z0 = P[0] * r0 + P[1] * r1 + P[2] * r2;
            const float z1 = P[3] * r0 + P[4] * r1 + P[5] * r2;
            const float z2 = P[6] * r0 + P[7] * r1 + P[8] * r2;
            a.dv[3 * i + 0] = 0.0f; a.dv[3 * i + 1] = 0.0f; a.dv[3 * i + 2] = 0.0f;
            a.r[3 * i + 0] = r0; a.r[3 * i + 1] = r1; a.r[3 * i + 2] = r2;
            a.z[3 * i + 0] = z0; a.z[3 * i + 1] = z1; a.z[3 * i + 2] = z2;
            p4[i] = make_float4(z0, z1, z2, vi.w);
            rz = (double)(r0 * z0 + r1 * z1 + r2 * z2);
        }
    }
    if constexpr (!WARM) {
        double v[1] = {rz};
        block_partials<1>(v, a.partials, a.stride);
    }
}

// Warm-start prologue (weiler18.py:311-335): r = rhs - A dv (a.out = A dv), z = P r, p = z,
// r.z partials; dv keeps the loaded correction
__global__ __launch_bounds__(BLOCK) void k_w18_warm_init(W18Args a, const float* __restrict__ rhs,
                                                         float4* __restrict__ p4) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    double rz = 0.0;
    if (i < a.R.nd) {
        const float* P = a.pinv + 9 * (size_t)i;
        float r0 = rhs[3 * i + 0] + -1.0f * a.out[3 * i + 0];
        float r1 = rhs[3 * i + 1] + -1.0f * a.out[3 * i + 1];
        float r2 = rhs[3 * i + 2] + -1.0f * a.out[3 * i + 2];
        float z0 = P[0] * r0 + P[1] * r1 + P[2] * r2;
        float z1 = P[3] * r0 + P[4] * r1 + P[5] * r2;
        float z2 = P[6] * r0 + P[7] * r1 + P[8] * r2;
        a.r[3 * i + 0] = r0; a.r[3 * i + 1] = r1; a.r[3 * i + 2] = r2;
        a.z[3 * i + 0] = z0; a.z[3 * i + 1] = z1; a.z[3 * i + 2] = z2;
        p4[i] = make_float4(z0, z1, z2, a.v4[i].w);
        rz = (double)(r0 * z0 + r1 * z1 + r2 * z2);
    }
    double v[1] = {rz};
    block_partials<1>(v, a.partials, a.stride);
}

// The correction of this solve in user order, the next warm start's initial guess (the
// reference's self.dv, indexed by particle): one thread per user index, contiguous writes,
// the internal row found through inv (static particles keep their zero)
__global__ __launch_bounds__(BLOCK) void k_w18_save(int n, int nd, const float* __restrict__ dv,
                                                    const int* __restrict__ inv, float* __restrict__ dv_prev,
                                                    const SolverState* __restrict__ st) {
    if (st && st->active != 0) return;   // the solve continues: a later launch saves
    int u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    const int i = inv[u];
    if (i >= nd) return;
    dv_prev[3 * u + 0] = dv[3 * i + 0];
    dv_prev[3 * u + 1] = dv[3 * i + 1];
    dv_prev[3 * u + 2] = dv[3 * i + 2];
}

// PCG prologue (weiler18.py:320-335): dv = 0, r = rhs, z = P r, p = z, rz = r.z
__global__ __launch_bounds__(BLOCK) void k_w18_init(W18Args a, float4* __restrict__ p4) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    double rz = 0.0;
    if (i < a.R.nd) {
        const float* P = a.pinv + 9 * (size_t)i;
        float r0 = a.out[3 * i + 0], r1 = a.out[3 * i + 1], r2 = a.out[3 * i + 2];
        float z0 = P[0] * r0 + P[1] * r1 + P[2] * r2;
        float z1 = P[3] * r0 + P[4] * r1 + P[5] * r2;
        float z2 = P[6] * r0 + P[7] * r1 + P[8] * r2;
        a.dv[3 * i + 0] = 0.0f; a.dv[3 * i + 1] = 0.0f; a.dv[3 * i + 2] = 0.0f;
        a.r[3 * i + 0] = r0; a.r[3 * i + 1] = r1; a.r[3 * i + 2] = r2;
        a.z[3 * i + 0] = z0; a.z[3 * i + 1] = z1; a.z[3 * i + 2] = z2;
        p4[i] = make_float4(z0, z1, z2, a.v4[i].w);
        rz = (double)(r0 * z0 + r1 * z1 + r2 * z2);
    }
    double v[1] = {rz};
    block_partials<1>(v, a.partials, a.stride);
}

// Ap = A p, p.Ap (weiler18.py:337-338).  Without CHECK (the warm-start residual) it runs
// regardless of the solver state.
template <class FM, bool ALLFLUID, bool CHECK = true>
__global__ __launch_bounds__(BLOCK) void k_w18_Av(W18Args a) {
    if (CHECK && a.st->active == 0) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    auto loadA = [&](int j) { return w18_Av_record<FM, ALLFLUID>(a, j); };
    auto loadB = [&](int j) { return vec3_of(a.v4, j); };
    __shared__ WinLds<FM, PASS_LANES> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.v4[j]; });
    double pap = 0.0;
    if (i < a.R.nd) {
        float4 pi = a.v4[i];
        float3 Ap = w18_Av_row<FM, ALLFLUID, PASS_LANES>(a, i, pi, win, w, loadA, loadB, g);
        if (g == 0) {
            a.out[3 * i + 0] = Ap.x; a.out[3 * i + 1] = Ap.y; a.out[3 * i + 2] = Ap.z;
            pap = (double)(pi.x * Ap.x + pi.y * Ap.y + pi.z * Ap.z);
        }
    }
    double v[1] = {pap};
    block_partials<1>(v, a.partials, a.stride);
}

// r -= alpha Ap, r.r, z = P r, r.z (weiler18.py:344-360); dv += alpha p of the same lines
// follows in the p pass, which reads p anyway
__global__ __launch_bounds__(BLOCK) void k_w18_update(W18Args a) {
    if (a.st->active == 0) return;
    float nalpha = (float)(-a.st->alpha);
    int i = blockIdx.x * BLOCK + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < a.R.nd) {
        float r0 = a.r[3 * i + 0] + nalpha * a.out[3 * i + 0];
        float r1 = a.r[3 * i + 1] + nalpha * a.out[3 * i + 1];
        float r2 = a.r[3 * i + 2] + nalpha * a.out[3 * i + 2];
        a.r[3 * i + 0] = r0; a.r[3 * i + 1] = r1; a.r[3 * i + 2] = r2;
        const float* P = a.pinv + 9 * (size_t)i;
        float z0 = P[0] * r0 + P[1] * r1 + P[2] * r2;
        float z1 = P[3] * r0 + P[4] * r1 + P[5] * r2;
        float z2 = P[6] * r0 + P[7] * r1 + P[8] * r2;
        a.z[3 * i + 0] = z0; a.z[3 * i + 1] = z1; a.z[3 * i + 2] = z2;
        rr = (double)(r0 * r0 + r1 * r1 + r2 * r2);
        rz = (double)(r0 * z0 + r1 * z1 + r2 * z2);
    }
    double v[2] = {rr, rz};
    block_partials<2>(v, a.partials, a.stride);
}

// dv += alpha p (the iteration's update, weiler18.py:344-346), then p = z + beta p while the
// solve continues
__global__ __launch_bounds__(BLOCK) void k_w18_p(int nd, const float* __restrict__ z, float4* __restrict__ p4,
                                                 float* __restrict__ dv, const SolverState* __restrict__ st) {
    const int apply = st->dv_apply, active = st->active;
    if (apply == 0 && active == 0) return;
    float alpha = (float)st->alpha;
    float beta = (float)st->beta;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    float4 p = p4[i];
    if (apply) {
        dv[3 * i + 0] = dv[3 * i + 0] + alpha * p.x;
        dv[3 * i + 1] = dv[3 * i + 1] + alpha * p.y;
        dv[3 * i + 2] = dv[3 * i + 2] + alpha * p.z;
    }
    if (active) {
        p.x = z[3 * i + 0] + beta * p.x;
        p.y = z[3 * i + 1] + beta * p.y;
        p.z = z[3 * i + 2] + beta * p.z;
        p4[i] = p;
    }
}

// v += dv (weiler18.py:399-401), once the solve has stopped
__global__ __launch_bounds__(BLOCK) void k_w18_apply(int n3, const float* __restrict__ dv, float* __restrict__ vel,
                                                     const SolverState* __restrict__ st) {
    if (st && st->active != 0) return;   // nullptr: unconditional
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n3) return;
    vel[i] = vel[i] + 1.0f * dv[i];
}

// ---- Jeske23 implicit surface tension (jeske23.py) -----------------------------------------

// compute_Wst, jeske23.py:23-43: theta W_cubic(max(r, d); h_st), theta = 10/7 (same operation
// order as the reference; norm = 8 / (pi h_st^3) is passed in; q = rc / h_st by div_rn from
// inv_hst = 1 / h_st rounded to nearest, the correctly rounded quotient)
__device__ __forceinline__ float st_wst(float r, float d, float hst, float inv_hst, float tn) {
    float rc = r > d ? r : d;
    float q = div_rn(rc, hst, inv_hst);
    float fq = 0.0f;
    if (q <= 0.5f) {
        fq = 1.0f - 6.0f * q * q + 6.0f * q * q * q;
    } else if (q <= 1.0f) {
        float t = 1.0f - q;
        fq = 2.0f * t * t * t;
    }
    return tn * fq;
}

// compute_grad_Wst, jeske23.py:46-66: returns s with grad = s * (x_ij / r); 0 for r <= d
__device__ __forceinline__ float st_grad_scale(float r, float d, float hst, float inv_hst, float tnh) {
    if (r <= d || r < 1.0e-12f) return 0.0f;
    float q = div_rn(r, hst, inv_hst);
    if (q > 1.0f) return 0.0f;
    float fpq;
    if (q <= 0.5f) {
        fpq = -12.0f * q + 18.0f * q * q;
    } else {
        float t = 1.0f - q;
        fpq = -6.0f * t * t;
    }
    return tnh * fpq;
}

struct StArgs {
    Rows R;
    int fluid;
    const int* mat;
    const float* density;
    float4* u4;        // (v, m) of the pass input
    float* div;
    float* gt;         // g~ per slot, tile layout of the ids
    float* b;          // extra right-hand side
    float* pinv;       // 3x3 per row (diagonal)
    float* out;        // M u, or v - M v + b
    float sigma, dt, d, hst;
    float inv_hst;     // 1 / h_st rounded to nearest
    float tn, tnh;     // theta * 8/(pi h_st^3), and that times 1/h_st
    double* partials;
    int stride;
    const SolverState* st;
};

// compute_divergence_jeske23 (jeske23.py:73-112): div_i = sum_fluid m_j (v_i - v_j) . gradW_ij
template <class FM, bool ALLFLUID>
__global__ __launch_bounds__(BLOCK) void k_st_div(StArgs a) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    auto loadA = [&](int j) {
        float4 A = a.R.pos[j];
        if (!ALLFLUID && a.mat[j] != a.fluid) A.w = 0.0f;
        return A;
    };
    auto loadB = [&](int j) { return vec3_of(a.u4, j); };
    __shared__ WinLds<FM, PASS_LANES> w;
    const bool win = stage_window(a.R, blk, w, loadA, [&](int j) { return a.u4[j]; });
    if (i >= a.R.nd) return;
    if (!ALLFLUID && a.mat[i] != a.fluid) {
        if (g == 0) a.div[i] = 0.0f;
        return;
    }
    float4 ui = a.u4[i];
    float acc = 0.0f;
    row_pairs_w<FM, 3, PASS_LANES>(win, w, a.R, i, a.R.pos[i], loadA, loadB,
                         [&](const float4& Aj, const Vec<3>& Bj, float gx, float gy, float gz) {
                             float vx = ui.x - Bj.v[0], vy = ui.y - Bj.v[1], vz = ui.z - Bj.v[2];
                             acc = acc + Aj.w * (vx * gx + vy * gy + vz * gz);
                         }, g);
    acc = group_sum<PASS_LANES>(acc);
    if (g == 0) a.div[i] = acc;
}

// compute_g_tilde_jeske23 + compute_rhs_jeske23 + compute_precond_jeske23 (jeske23.py:115-350)
// in one pass over each row's slots: g~_ij (stored per slot), b_i and the diagonal
// preconditioner.  Padding slots (id = the row) and non-fluid neighbours get g~ = 0.
// The set-up's window records: A = (x, y, z, rho_j), B = (v_j, m_j), C = div_j, D = j, with bit
// 31 set for a non-fluid neighbour.
template <class FM>
struct StWinLds {
    static constexpr int N = has_windows<FM, PASS_LANES>() ? WIN_MAX : 1;
    float4 A[N];
    float4 B[N];
    float C[N];
    int D[N];
};

template <class FM, bool ALLFLUID>
__global__ __launch_bounds__(BLOCK) void k_st_setup(StArgs a) {
    constexpr unsigned REC = FM::REC;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    // the tile window (every thread of the block stages before any row leaves)
    __shared__ StWinLds<FM> w;
    bool win = false;
    if constexpr (has_windows<FM, PASS_LANES>()) {
        if (a.R.win_n != nullptr) {
            const int nw = __builtin_amdgcn_readfirstlane(a.R.win_n[blk]);
            if (nw >= 0) {
                const int* wid = a.R.win_ids + (size_t)blk * WIN_MAX;
                for (int k = threadIdx.x; k < nw; k += BLOCK) {
                    const int j = wid[k];
                    const float4 xj = a.R.pos[j];
                    w.A[k] = make_float4(xj.x, xj.y, xj.z, a.density[j]);
                    w.B[k] = a.u4[j];
                    w.C[k] = a.div[j];
                    w.D[k] = j | ((ALLFLUID || a.mat[j] == a.fluid) ? 0 : (int)0x80000000u);
                }
                __syncthreads();
                win = true;
            }
        }
    }
    if (i >= a.R.nd) return;
    float* P = a.pinv + 9 * (size_t)i;
    const bool fluid_i = ALLFLUID || a.mat[i] == a.fluid;
    if (!fluid_i) {
        if (g == 0) {
            for (int q = 0; q < 9; ++q) P[q] = (q % 4) == 0 ? 1.0f : 0.0f;
            a.b[3 * i + 0] = 0.0f; a.b[3 * i + 1] = 0.0f; a.b[3 * i + 2] = 0.0f;
        }
        return;
    }
    const uint32_t* e = a.R.nb + (size_t)(i >> 6) * (size_t)a.R.Cs * REC + row_word0<REC>(i & (WAVE - 1));
    // the per-pair g-tilde array has one word per slot in the layout of ids-only rows
    float* eg = a.gt + (size_t)(i >> 6) * (size_t)a.R.Cs * 64u + row_word0<64u>(i & (WAVE - 1));
    const int tl = __builtin_amdgcn_readfirstlane(a.R.tile_len[i >> 6]);
    const float4 xi = a.R.pos[i];
    const float4 ui = a.u4[i];
    const float m_i = ui.w, rho_i = a.density[i], div_i = a.div[i];
    float bx = 0.0f, by = 0.0f, bz = 0.0f, diag = 0.0f;
    // one slot: g~ of the pair (stored by the caller) and the sums; rho_j and div_j are read only
    // for a used pair within h_st
    auto pair = [&](float xjx, float xjy, float xjz, const float4& uj, bool use, auto&& rho_j, auto&& div_j) {
        float gt = 0.0f;
        float x0 = xi.x - xjx, x1 = xi.y - xjy, x2 = xi.z - xjz;
        float r = len3(x0, x1, x2);
        if (use && r < a.hst) {
            float v0 = ui.x - uj.x, v1 = ui.y - uj.y, v2 = ui.z - uj.z;
            float rho_bar = 0.5f * (rho_i + rho_j());
            float wst = st_wst(r, a.d, a.hst, a.inv_hst, a.tn);
            float s = st_grad_scale(r, a.d, a.hst, a.inv_hst, a.tnh);
            // x / r correctly rounded (div_rn; r = 0 gives NaN as the division does, zeroed below)
            const float yr = 1.0f / r;
            float g0 = s * div_rn(x0, r, yr), g1 = s * div_rn(x1, r, yr), g2 = s * div_rn(x2, r, yr);
            if (s == 0.0f) { g0 = 0.0f; g1 = 0.0f; g2 = 0.0f; }
            float g_ij = wst / rho_bar;
            float term_kernel = (g0 * v0 + g1 * v1 + g2 * v2) / rho_bar;
            float term_density = wst / (2.0f * rho_bar * rho_bar) * (div_i + div_j());
            gt = g_ij + a.dt * (term_kernel - term_density);
        }
        if (use) {
            float m_bar = 0.5f * (m_i + uj.w);
            diag = diag + m_bar * gt;
            if (gt != 0.0f) {
                bx = bx + m_bar * x0 * gt;
                by = by + m_bar * x1 * gt;
                bz = bz + m_bar * x2 * gt;
            }
        }
        return gt;
    };
    if (has_windows<FM, PASS_LANES>() && win) {
        // slots g + 8 b and g + 4 + 8 b (the lane's order of the loop below): one window-index
        // word and one 8-byte g~ store per block
        typedef float f32x2 __attribute__((ext_vector_type(2)));
        const uint32_t* l = a.R.loc + ((size_t)(i >> 6) * (size_t)a.R.Cs * 32u + 4u * (i & (WAVE - 1)) + g);
        f32x2* g2 = reinterpret_cast<f32x2*>(eg + 2 * g);
        auto from = [&](uint32_t k) {
            const float4 A = w.A[k];
            const int d = w.D[k];
            const bool use = (d & 0x7fffffff) != i && d >= 0;
            return pair(A.x, A.y, A.z, w.B[k], use, [&] { return A.w; }, [&] { return w.C[k]; });
        };
        for (int b = 0; b < tl; b += 8) {
            const uint32_t q = __builtin_nontemporal_load(l + (b >> 3) * 256);
            const float t0 = from(q & 0xffffu);
            const float t1 = from(q >> 16);
            g2[(b >> 3) * 256] = f32x2{t0, t1};
        }
    } else {
        for (int k = g; k < tl; k += PASS_LANES) {
            const int j = (int)__builtin_nontemporal_load(e + slot_word<REC>(k));
            const float4 xj = a.R.pos[j];
            const float4 uj = a.u4[j];
            const bool use = j != i && (ALLFLUID || a.mat[j] == a.fluid);
            eg[slot_word<64u>(k)] = pair(xj.x, xj.y, xj.z, uj, use, [&] { return a.density[j]; },
                                         [&] { return a.div[j]; });
        }
    }
    bx = group_sum<PASS_LANES>(bx);
    by = group_sum<PASS_LANES>(by);
    bz = group_sum<PASS_LANES>(bz);
    diag = group_sum<PASS_LANES>(diag);
    if (g != 0) return;
    float fb = -a.dt * a.sigma / m_i;
    a.b[3 * i + 0] = fb * bx; a.b[3 * i + 1] = fb * by; a.b[3 * i + 2] = fb * bz;
    float dt2 = a.dt * a.dt;
    float fm = dt2 * a.sigma / m_i;
    float dv = 1.0f + fm * diag;
    if (dv < 1.0e-12f) dv = 1.0e-12f;
    float inv = 1.0f / dv;
    for (int q = 0; q < 9; ++q) P[q] = (q % 4) == 0 ? inv : 0.0f;
}

// compute_Av_jeske23 (jeske23.py:245-295): (M u)_i = u_i + dt^2 (sigma/m_i) sum m_bar g~ (u_i - u_j).
// RHS: out = u - M u + b (jeske23.py:637-647); otherwise out = M u and the block's p.Mp partial.
// Only the ids and g~ are streamed and (u_j, m_j) gathered: positions are not needed.
template <class FM, bool ALLFLUID, bool RHS>
__global__ __launch_bounds__(BLOCK) void k_st_Av(StArgs a) {
    if (!RHS && a.st->active == 0) return;
    constexpr unsigned REC = FM::REC;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = pass_row(blk), g = pass_lane();
    // only (u_j, m_j) is read per neighbour: the window's B records
    __shared__ WinLds<FM, PASS_LANES> w;
    const bool win = stage_window(a.R, blk, w, [](int) { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); },
                                  [&](int j) { return a.u4[j]; });
    double pap = 0.0;
    if (i < a.R.nd) {
        const float4 ui = a.u4[i];
        float Mx = ui.x, My = ui.y, Mz = ui.z;
        if (ALLFLUID || a.mat[i] == a.fluid) {
            const uint32_t* e = a.R.nb + (size_t)(i >> 6) * (size_t)a.R.Cs * REC + row_word0<REC>(i & (WAVE - 1));
            const float* eg = a.gt + (size_t)(i >> 6) * (size_t)a.R.Cs * 64u + row_word0<64u>(i & (WAVE - 1));
            const int tl = __builtin_amdgcn_readfirstlane(a.R.tile_len[i >> 6]);
            const float m_i = ui.w;
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            auto visit_u = [&](const float4& uj, float gt) {
                float w = (0.5f * (m_i + uj.w)) * gt;
                sx = sx + w * (ui.x - uj.x);
                sy = sy + w * (ui.y - uj.y);
                sz = sz + w * (ui.z - uj.z);
            };
            auto visit = [&](int j, float gt) { visit_u(a.u4[j], gt); };
            if (has_windows<FM, PASS_LANES>() && win) {
                // window indices of slots g + 8 b and g + 4 + 8 b: one 32-bit word
                typedef float f32x2 __attribute__((ext_vector_type(2)));
                const uint32_t* l = a.R.loc + ((size_t)(i >> 6) * (size_t)a.R.Cs * 32u + 4u * (i & (WAVE - 1)) + g);
                const f32x2* g2 = reinterpret_cast<const f32x2*>(eg + 2 * g);
#pragma unroll((PASS_UNROLL + 1) / 2)
                for (int b = 0; b < tl; b += 8) {
                    const uint32_t q = __builtin_nontemporal_load(l + (b >> 3) * 256);
                    const f32x2 gts = __builtin_nontemporal_load(g2 + (b >> 3) * 256);
                    visit_u(w.B[q & 0xffffu], gts.x);
                    visit_u(w.B[q >> 16], gts.y);
                }
            } else if constexpr (REC == 64u && PASS_LANES == 4) {
                // slots g + 8 b and g + 4 + 8 b: one 8-byte load of ids and one of g-tilde
                typedef float f32x2 __attribute__((ext_vector_type(2)));
                const u32x2* e2 = reinterpret_cast<const u32x2*>(e + 2 * g);
                const f32x2* g2 = reinterpret_cast<const f32x2*>(eg + 2 * g);
#pragma unroll((PASS_UNROLL + 1) / 2)
                for (int b = 0; b < tl; b += 8) {
                    const u32x2 ids = __builtin_nontemporal_load(e2 + (b >> 3) * 256);
                    const f32x2 gts = __builtin_nontemporal_load(g2 + (b >> 3) * 256);
                    visit((int)ids.x, gts.x);
                    visit((int)ids.y, gts.y);
                }
            } else {
#pragma unroll(PASS_UNROLL)
                for (int k = g; k < tl; k += PASS_LANES)
                    visit((int)__builtin_nontemporal_load(e + slot_word<REC>(k)),
                          __builtin_nontemporal_load(eg + slot_word<64u>(k)));
            }
            sx = group_sum<PASS_LANES>(sx);
            sy = group_sum<PASS_LANES>(sy);
            sz = group_sum<PASS_LANES>(sz);
            float dt2 = a.dt * a.dt;
            float factor = dt2 * a.sigma / m_i;
            Mx = ui.x + factor * sx;
            My = ui.y + factor * sy;
            Mz = ui.z + factor * sz;
        }
        if (g == 0) {
            if (RHS) {
                a.out[3 * i + 0] = (ui.x + -1.0f * Mx) + 1.0f * a.b[3 * i + 0];
                a.out[3 * i + 1] = (ui.y + -1.0f * My) + 1.0f * a.b[3 * i + 1];
                a.out[3 * i + 2] = (ui.z + -1.0f * Mz) + 1.0f * a.b[3 * i + 2];
            } else {
                a.out[3 * i + 0] = Mx; a.out[3 * i + 1] = My; a.out[3 * i + 2] = Mz;
                pap = (double)(ui.x * Mx + ui.y * My + ui.z * Mz);
            }
        }
    }
    if (!RHS) {
        double v[1] = {pap};
        block_partials<1>(v, a.partials, a.stride);
    }
}

// (v, m) of the particles for the surface-tension passes
__global__ __launch_bounds__(BLOCK) void k_st_prep(int nd, const float* __restrict__ vel, const float4* __restrict__ pos,
                                                   float4* __restrict__ u4) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    u4[i] = make_float4(vel[3 * i + 0], vel[3 * i + 1], vel[3 * i + 2], pos[i].w);
}

// ---- glue ----------------------------------------------------------------------------------

// apply_gravity (framework.py:26-41)
__global__ __launch_bounds__(BLOCK) void k_gravity(int nd, float gx, float gy, float gz, const float* __restrict__ vin,
                                                   float* __restrict__ vout) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    vout[3 * i + 0] = vin[3 * i + 0] + gx;
    vout[3 * i + 1] = vin[3 * i + 1] + gy;
    vout[3 * i + 2] = vin[3 * i + 2] + gz;
}

// advect_position (framework.py:117-136) + enforce_boundary (boundary.py:27-91); with SC each
// particle is clamped to its scene's domain size sdom[3 sid[i] ..].
template <bool SC>
__global__ __launch_bounds__(BLOCK) void k_advect_boundary(int nd, float dt, float L0, float L1, float L2, float pad,
                                                           float4* __restrict__ pos, float* __restrict__ vel,
                                                           const float* __restrict__ sdom, const int* __restrict__ sid) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nd) return;
    float4 p = pos[i];
    float v[3] = {vel[3 * i + 0], vel[3 * i + 1], vel[3 * i + 2]};
    float x[3] = {p.x + dt * v[0], p.y + dt * v[1], p.z + dt * v[2]};
    float L[3] = {L0, L1, L2};
    if (SC) {
        const int s = sid[i];
        L[0] = sdom[3 * s + 0]; L[1] = sdom[3 * s + 1]; L[2] = sdom[3 * s + 2];
    }
    float nrm[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (x[q] > L[q] - pad) { nrm[q] = nrm[q] + 1.0f; x[q] = L[q] - pad; }
        if (x[q] <= pad) { nrm[q] = nrm[q] + -1.0f; x[q] = pad; }
    }
    float len = len3(nrm[0], nrm[1], nrm[2]);
    if (len > 1e-6f) {
        float c0 = nrm[0] / len, c1 = nrm[1] / len, c2 = nrm[2] / len;
        float s = (1.0f + 0.5f) * (v[0] * c0 + v[1] * c1 + v[2] * c2);
        vel[3 * i + 0] = v[0] - s * c0;
        vel[3 * i + 1] = v[1] - s * c1;
        vel[3 * i + 2] = v[2] - s * c2;
    }
    pos[i] = make_float4(x[0], x[1], x[2], p.w);
}

Rows rows_of(ivp_ctx* c) {
    const bool win = c->win_built && c->windows;
    return Rows{c->nb, c->g16, c->tile_len, c->Cs, c->nd, c->pos, c->h, c->kspiky, c->gscale,
                c->loc, c->win_ids, win ? c->win_n : nullptr};
}

// ---- independent scenes -------------------------------------------------------------------

// Start of a batched solve: every scene begins its outer loop (ip.py:345)
__global__ __launch_bounds__(BLOCK) void k_scene_init(int S, SceneState* __restrict__ ss) {
    int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= S) return;
    SceneState& q = ss[s];
    q.phase = PH_PACK;
    q.opt_iter = 0;
    q.outer_iters = 0;
    q.status = IVP_OK;
    q.cg_total = 0;
    q.n_err = 0;
    q.final_err = 0.0;
    q.cg.active = 0;
    q.cg.status = 0;
    q.cg.iter = 0;
}

// End of a replay (one block): scenes whose displacement was just updated start their next
// outer iteration (ip.py:389-391); *active = scenes not done.
__global__ __launch_bounds__(BLOCK) void k_scene_advance(int S, SceneState* __restrict__ ss, int* __restrict__ active) {
    __shared__ double lds[WAVES_PER_BLOCK];
    int cnt = 0;
    for (int s = threadIdx.x; s < S; s += BLOCK) {
        SceneState& q = ss[s];
        if (q.phase == PH_DUPDATE) {
            q.phase = PH_PACK;
            q.opt_iter = q.opt_iter + 1;
        }
        cnt += q.phase != PH_DONE ? 1 : 0;
    }
    double t = block_sum((double)cnt, lds);
    if (threadIdx.x == 0) *active = (int)t;
}

// One replay of the batched outer iteration on stream `st`: constraint pass for scenes starting
// an outer iteration, right-hand side and CG set-up for those that continue, K CG iterations
// for scenes in their CG, the displacement update for scenes whose CG ended.  Every launch
// covers all rows; rows of scenes in another phase skip (whole blocks when both of their
// scenes do).
template <int G>
int enqueue_scene_replay_g(ivp_ctx* c, hipStream_t st, float omega, int K) {
    const int nd = c->nd;
    const dim3 nb(blocks_for(nd)), np(blocks_for(nd * G)), blk(BLOCK);
    const dim3 nf((c->S + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK);
    const int stride = c->max_blocks;
    Rows R = rows_of(c);
    SceneRef sc{c->sid, c->ss};
    double* P = c->spartials;
    FinScenesArgs fa{P, stride, c->sid, c->ss, c->sprm, c->shist_err, c->shist_pcg, c->S, BLOCK};
    FinScenesArgs fp = fa;   // partials of the pair passes: BLOCK / G rows per block
    fp.rpb = BLOCK / G;
    PackArgs pa{R, c->d4, c->Ss, c->cc, c->kk, c->rho0, c->rowcnt, c->q4, c->sfac, c->jd, P, stride, sc};
    hipLaunchKernelGGL((k_ip_pack<FmtIds, true, G>), np, blk, 0, st, pa);
    hipLaunchKernelGGL(k_fin_scenes<FIN_PACK>, nf, blk, 0, st, fp);
    RhsArgs ra{R, c->q4, c->Ss, omega, c->b, c->nbact, c->rowcnt, P, stride, sc};
    hipLaunchKernelGGL((k_ip_rhs<FmtIds, true, G>), np, blk, 0, st, ra);
    hipLaunchKernelGGL(k_fin_scenes<FIN_RHS>, nf, blk, 0, st, fp);
    CGInitArgs ia{nd, c->b, c->pos, c->sol, c->r, c->p4, c->y2, P, stride, sc};
    hipLaunchKernelGGL(k_cg_init<true>, nb, blk, 0, st, ia);
    hipLaunchKernelGGL(k_fin_scenes<FIN_CG_INIT>, nf, blk, 0, st, fa);
    CGJArgs ja{R, c->p4, c->Ss, c->sfac, c->y2, c->state, sc};
    CGJTArgs ta{R, c->p4, c->y2, c->Ss, c->nbact, omega, c->Ap, P, stride, c->state, sc};
    CGUpdArgs ua{nd, c->sol, c->r, c->p4, c->Ap, P, stride, c->state, sc};
    for (int k = 0; k < K; ++k) {
        hipLaunchKernelGGL((k_cg_J<FmtIds, true, G>), np, blk, 0, st, ja);
        hipLaunchKernelGGL((k_cg_JT<FmtIds, true, G>), np, blk, 0, st, ta);
        hipLaunchKernelGGL(k_fin_scenes<FIN_CG_PAP>, nf, blk, 0, st, fp);
        hipLaunchKernelGGL(k_cg_update<true>, nb, blk, 0, st, ua);
        hipLaunchKernelGGL(k_fin_scenes<FIN_CG_UPDATE>, nf, blk, 0, st, fa);
        hipLaunchKernelGGL(k_cg_p<true>, nb, blk, 0, st, nd, (const float*)c->r, c->p4,
                           (const SolverState*)c->state, sc);
    }
    hipLaunchKernelGGL(k_ip_dupdate<true>, nb, blk, 0, st, nd, c->d4, (const float*)c->sol, sc,
                       (const SolverState*)nullptr);
    hipLaunchKernelGGL(k_scene_advance, dim3(1), blk, 0, st, c->S, c->ss, c->sactive);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

int enqueue_scene_replay(ivp_ctx* c, hipStream_t st, float omega, int K) {
    switch (c->scene_lanes) {
    case 2: return enqueue_scene_replay_g<2>(c, st, omega, K);
    case 4: return enqueue_scene_replay_g<4>(c, st, omega, K);
    default: return enqueue_scene_replay_g<1>(c, st, omega, K);
    }
}

// The instantiated graph of one replay for (K, omega) and the current store buffers (the sort
// alternates two buffers, so at most two captures are live).
int scene_graph(ivp_ctx* c, float omega, int K, hipGraphExec_t* out) {
    const bool win = rows_of(c).win_n != nullptr;
    for (auto& g : c->sgraphs) {
        if (g.exec && g.pos == (const void*)c->pos && g.chunk == K && g.omega == omega && g.nd == c->nd &&
            g.lanes == c->scene_lanes && g.win == win) {
            *out = g.exec;
            return IVP_OK;
        }
    }
    if (!c->cap_stream) IVP_HIP(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    IVP_HIP(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed));
    int rc = enqueue_scene_replay(c, c->cap_stream, omega, K);
    hipError_t e = hipStreamEndCapture(c->cap_stream, &graph);
    if (rc != IVP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    IVP_HIP(e);
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    IVP_HIP(e);
    SceneGraph& slot = c->sgraphs[c->sgraph_next];
    c->sgraph_next ^= 1;
    if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
    slot.exec = exec;
    slot.pos = (const void*)c->pos;
    slot.chunk = K;
    slot.omega = omega;
    slot.nd = c->nd;
    slot.lanes = c->scene_lanes;
    slot.win = win;
    *out = exec;
    return IVP_OK;
}

int read_state(ivp_ctx* c) {
    IVP_HIP(hipMemcpyAsync(c->state_host, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    return IVP_OK;
}

// Enqueue CG iterations in chunks until the device-side state says stop.  `launch_it(tag)`
// enqueues one iteration.  Returns the iteration count in *iters.
// The solver state copied back after the work enqueued so far, then `tail` enqueued (work that
// follows a finished solve and returns at once while it continues), then a wait for the copy
// alone: the device runs the tail while the host takes the state.
template <class T>
int read_state_then(ivp_ctx* c, T&& tail) {
    IVP_HIP(hipMemcpyAsync(c->state_host, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
    if (!c->wait_ev) IVP_HIP(hipEventCreateWithFlags(&c->wait_ev, hipEventDisableTiming));
    IVP_HIP(hipEventRecord(c->wait_ev, c->stream));
    IVP_TRY(tail());
    IVP_HIP(hipEventSynchronize(c->wait_ev));
    return IVP_OK;
}

// With a tail, every chunk's read-back is followed by the tail; it has run on a finished solve
// when the returned state is inactive.
template <class L, class T>
int run_chunks(ivp_ctx* c, int max_iter, int guess, L&& launch_it, int* iters, int* chunks, T&& tail) {
    int issued = 0;
    int chunk = std::max(2, std::min(guess, 64));
    for (;;) {
        int todo = std::min(chunk, max_iter - issued);
        if (todo <= 0) {
            // no iteration allowed at all: the set-up left the solve stopped; read it back (and
            // run the tail) all the same
            if (issued == 0) IVP_TRY(read_state_then(c, tail));
            break;
        }
        for (int q = 0; q < todo; ++q) IVP_TRY(launch_it(issued + q));
        issued += todo;
        IVP_TRY(read_state_then(c, tail));
        (*chunks)++;
        const SolverState& s = *c->state_host;
        if (s.active == 0) break;
        const double rr0 = s.rr0;
        // geometric extrapolation of r.r to the tolerance
        double rate = s.iter > 0 && rr0 > 0.0 && s.rr_last > 0.0 ? pow(s.rr_last / rr0, 1.0 / s.iter) : 0.5;
        int est = 8;
        if (rate > 0.0 && rate < 0.999 && s.rr_last > s.tol)
            est = (int)ceil(log(s.tol / s.rr_last) / log(rate)) + 1;
        chunk = std::max(1, std::min(est, 64));
    }
    *iters = c->state_host->iter;
    return IVP_OK;
}

template <class L>
int run_chunks(ivp_ctx* c, int max_iter, int guess, L&& launch_it, int* iters, int* chunks) {
    return run_chunks(c, max_iter, guess, launch_it, iters, chunks, []() { return (int)IVP_OK; });
}

// Bytes a pass of format FM reads from the pair stream: the id (4 B), plus 12 B of float32 or
// 8 B of float16 gradW when it reads cached gradients.
template <class FM>
constexpr double pair_bytes() { return FM::GS == GS_EVAL ? 4.0 : FM::GS == GS_F32 ? 16.0 : 12.0; }

// Bytes of a four-lane pass's pair stream over `pairs` pairs: with tile windows a 2-byte window
// index replaces each 4-byte id, and the window lists (4 bytes per entry) are read once.
template <class FM>
double stream_bytes(const ivp_ctx* c, double pairs) {
    if (has_windows<FM, PASS_LANES>() && c->win_built && c->windows)
        return (pair_bytes<FM>() - 2.0) * pairs + 4.0 * (double)c->state_host->win_entries;
    return pair_bytes<FM>() * pairs;
}

// CG iterations between true-residual replacements when the CG operator reads float16 gradW
constexpr int REFRESH_EVERY = 8;

// FP: format of the outer loop's passes (pack, rhs); FC: format of the CG operator passes.
template <class FP, class FC>
int ip_solve_body(ivp_ctx* c, float dt, const ivp_ip_params* prm, ivp_ip_stats* stats) {
    ivp_ip_stats& st = *stats;
    const int nd = c->nd;
    if (nd == 0) return IVP_OK;
    const dim3 nb(blocks_for(nd)), np = pass_grid(nd);
    const int stride = c->max_blocks;
    const double pb = pair_bytes<FP>(), pbc = pair_bytes<FC>();
    Rows R = rows_of(c);
    float4* d4 = c->d4;
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_init, nb, nd, dt, (const float*)c->density, (const float*)c->rho0,
                   (const float*)c->vel, (const float4*)c->pos, c->cc, d4));
    int opt_iter = 0;
    bool spec_finish = false;   // k_ip_finish enqueued before the last read-back
    bool pack_queued = false;   // this evaluation's passes were queued behind the CG's read-back
    bool finish_queued = false; // and k_ip_finish with them
    const float inv_dt = (float)(1.0 / (double)dt);
    double err = 0.0;
    IVP_TRY(iter_mark(c, 0));
    for (;;) {
        PackArgs pa{R, d4, c->Ss, c->cc, c->kk, c->rho0, c->rowcnt, c->q4, c->sfac, c->jd, c->partials, stride};
        double total_pairs = (double)c->state_host->total_pairs;
        // the constraint pass, its finalize and the flags' copy-back (queued in one piece)
        auto queue_pack = [&](const SolverState* gate) -> int {
            PackArgs pg = pa;
            pg.gate = gate;
            IVP_TRY(launch(c, KC_IP_PACK, -1, stream_bytes<FP>(c, total_pairs) + 68.0 * nd, k_ip_pack<FP, false>, np, pg));
            IVP_TRY(finalize<FIN_PACK>(c, np.x, 0.0, 0, gate != nullptr));
            IVP_HIP(hipMemcpyAsync(c->err_host, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            return IVP_OK;
        };
        if (!pack_queued) IVP_TRY(queue_pack(nullptr));
        pack_queued = false;
        // The first evaluation never ends the loop (the stop needs opt_iter >= 1, ip.py:374-376),
        // so the first CG follows without a wait: its first chunk's synchronisation brings back the
        // overflow flags and this pass's error sum and active counts (no later finalize writes them)
        // before anything reads them.
        auto record = [&]() -> int {
            if (*c->err_host & 4) return IVP_ERR_INTERNAL;
            if (*c->err_host & 1) return IVP_NEIGHBOR_OVERFLOW;
            err = c->state_host->err_sum / (double)nd;
            if (st.n_density_errors < st.capacity && st.density_errors) st.density_errors[st.n_density_errors] = err;
            st.n_density_errors++;
            return IVP_OK;
        };
        const bool deferred = opt_iter == 0 && prm->max_iteration_opt >= 1;
        long long act_pairs = c->state_host->active_pairs;   // (deferred: the last evaluation's, for
        int act_rows = c->state_host->n_active_rows;         //  the byte model until read back)
        if (!deferred) {
            // An evaluation below the iteration limit ends the loop only when it converged, and
            // then the solve closes with k_ip_finish on this d: enqueue it now, so that the
            // device runs it while the host reads the state.  Should the loop go on, nothing in
            // it reads the velocities or divergences it writes, and the closing pass writes them
            // again.
            spec_finish = opt_iter >= 1 && opt_iter < prm->max_iteration_opt && !prm->use_fixed_iters;
            if (spec_finish && !finish_queued)
                IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)d4, (const float*)c->jd,
                               c->vel, c->divg));
            finish_queued = false;
            IVP_TRY(read_state(c));
            IVP_TRY(record());
            bool converged = err < prm->tol_opt && !prm->use_fixed_iters;
            if ((converged && opt_iter >= 1) || opt_iter >= prm->max_iteration_opt) {
                IVP_TRY(iter_mark(c, opt_iter + 1));
                break;
            }
            spec_finish = false;
            act_pairs = c->state_host->active_pairs;
            act_rows = c->state_host->n_active_rows;
        }

        RhsArgs ra{R, c->q4, c->Ss, prm->omega, c->b, c->nbact, c->rowcnt, c->partials, stride};
        IVP_TRY(launch(c, KC_IP_RHS, -1, stream_bytes<FP>(c, total_pairs) + 48.0 * nd, k_ip_rhs<FP, false>, np, ra));
        IVP_TRY(finalize<FIN_RHS>(c, np.x));
        CGInitArgs ia{nd, c->b, c->pos, c->sol, c->r, c->p4, c->y2, c->partials, stride};
        IVP_TRY(launch(c, -1, -1, 0.0, k_cg_init<false>, nb, ia));
        IVP_TRY(finalize<FIN_CG_INIT>(c, nb.x, prm->tol_pcg, prm->max_iteration_pcg));
        // The first CG chunk is enqueued without waiting for the set-up: its kernels return at once
        // when there is nothing to solve.  The J^T byte model (touched rows, reported by the set-up)
        // is attached to the launches once the first chunk's state has been read.
        long long touch_pairs = 0;
        int touch_rows = 0;
        int iters = 0;
        bool fin_next = false;   // the tail below queues k_ip_finish
        {
            c->prof_epoch++;
            int epoch = c->prof_epoch;
            // algorithmic bytes per launch: the pair stream once per traversed pair, per-row
            // arrays once (gathered neighbour values counted once per particle)
            double bJ = stream_bytes<FC>(c, (double)act_pairs) + (double)act_rows * (16 + 12 + 8) + (double)nd * 4;
            double bJT = stream_bytes<FC>(c, (double)touch_pairs) + (double)touch_rows * (8 + 12) + (double)nd * (16 + 4 + 12);
            double bU = (double)nd * (12 + 12 + 16 + 12 + 12 + 12);
            double bP = (double)nd * (12 + 16 + 16);
            CGJArgs ja{R, c->p4, c->Ss, c->sfac, c->y2, c->state};
            CGJTArgs ta{R, c->p4, c->y2, c->Ss, c->nbact, prm->omega, c->Ap, c->partials, stride, c->state};
            CGUpdArgs ua{nd, c->sol, c->r, c->p4, c->Ap, c->partials, stride, c->state};
            // true-residual passes: the same J / J^T kernels on (x, m), gradW evaluated
            using FE = Fmt<FC::REC, GS_EVAL>;
            CGJArgs jx{R, c->xs4, c->Ss, c->sfac, c->y2, c->state};
            // (its x.Ax partials go to row 2, clear of the update's r.r in row 0)
            CGJTArgs tx{R, c->xs4, c->y2, c->Ss, c->nbact, prm->omega, c->Ap, c->partials + 2 * stride, stride,
                        c->state};
            auto one_iter = [&](int tag) -> int {
                IVP_TRY(launch(c, KC_CG_J, tag, bJ, k_cg_J<FC, false>, np, ja));
                IVP_TRY(launch(c, KC_CG_JT, tag, bJT, k_cg_JT<FC, false>, np, ta));
                IVP_TRY(finalize<FIN_CG_PAP>(c, np.x));
                IVP_TRY(launch(c, KC_CG_UPDATE, tag, bU, k_cg_update<false>, nb, ua));
                if (FC::GS == GS_F16 && (tag + 1) % REFRESH_EVERY == 0) {
                    // residual replacement: r = b - A x with A in float32 from the positions;
                    // its r.r (row 1) and the recursive r.r (row 0) go to the stop test and beta
                    IVP_TRY(launch(c, KC_CG_REFRESH, tag, 0.0, k_x4, nb, nd, (const float*)c->sol,
                                   (const float4*)c->pos, c->xs4, (const SolverState*)c->state));
                    IVP_TRY(launch(c, KC_CG_REFRESH, tag, 0.0, k_cg_J<FE, false>, np, jx));
                    IVP_TRY(launch(c, KC_CG_REFRESH, tag, 0.0, k_cg_JT<FE, false>, np, tx));
                    IVP_TRY(launch(c, KC_CG_REFRESH, tag, 0.0, k_cg_refresh, nb, nd, (const float*)c->b,
                                   (const float*)c->Ap, c->r, c->partials + stride, stride,
                                   (const SolverState*)c->state));
                    IVP_TRY(finalize<FIN_CG_REPLACE>(c, nb.x));
                } else {
                    IVP_TRY(finalize<FIN_CG_UPDATE>(c, nb.x));
                }
                return launch(c, KC_CG_P, tag, bP, k_cg_p<false>, nb, nd, (const float*)c->r, c->p4,
                              (const SolverState*)c->state, SceneRef{});
            };
            // Behind each chunk's state copy: d -= x, the outer-iteration mark and the next constraint
            // pass with its finalize, all returning at once unless the CG has stopped (and not
            // failed), so the device works while the host reads the state; after the first CG of
            // the solve (whose read-back also checks the rows for overflow) also the closing
            // k_ip_finish that the next evaluation may need (see above).
            fin_next = !deferred && opt_iter + 1 < prm->max_iteration_opt && !prm->use_fixed_iters;
            auto tail = [&]() -> int {
                const SolverState* gate = c->state;
                IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_dupdate<false>, nb, nd, d4, (const float*)c->sol, SceneRef{},
                               gate));
                IVP_TRY(iter_mark(c, opt_iter + 1));
                IVP_TRY(queue_pack(gate));
                if (fin_next)
                    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)d4,
                                   (const float*)c->jd, c->vel, c->divg));
                return IVP_OK;
            };
            IVP_TRY(run_chunks(c, prm->max_iteration_pcg, c->cg_guess, one_iter, &iters, &st.cg_launch_chunks, tail));
            if (deferred) {
                IVP_TRY(record());
                act_pairs = c->state_host->active_pairs;
                act_rows = c->state_host->n_active_rows;
                prof_set_bytes(c, epoch, KC_CG_J,
                               stream_bytes<FC>(c, (double)act_pairs) + (double)act_rows * (16 + 12 + 8) + (double)nd * 4);
            }
            touch_pairs = c->state_host->touched_pairs;
            touch_rows = c->state_host->n_touched_rows;
            prof_set_bytes(c, epoch, KC_CG_JT,
                           stream_bytes<FC>(c, (double)touch_pairs) + (double)touch_rows * (8 + 12) +
                               (double)nd * (16 + 4 + 12));
            prof_collect(c, iters, epoch);
            c->cg_guess = std::max(2, iters);
            st.pairs_traversed += (long long)iters * (act_pairs + touch_pairs);
            st.rows_traversed += (long long)iters * (act_rows + touch_rows);
            if (c->state_host->status == IVP_NOT_PD) {
                st.pAp_bad = c->state_host->pAp;
                st.pcg_iter_bad = iters;
                return IVP_NOT_PD;
            }
        }
        if (st.n_pcg < st.capacity && st.pcg_iters) st.pcg_iters[st.n_pcg] = iters;
        st.n_pcg++;
        // the CG stopped at the device's own limit at the latest, so its last read-back showed it
        // stopped and the tail behind it (d -= x, the mark, the next constraint pass) has run
        if (c->state_host->active != 0) return IVP_ERR_INTERNAL;
        pack_queued = true;
        finish_queued = fin_next;
        opt_iter++;
    }
    st.outer_iters = opt_iter;
    if (!prm->use_fixed_iters && err >= prm->tol_opt && opt_iter >= prm->max_iteration_opt) return IVP_NOT_CONVERGED;
    // final_density_error (pressure_base.py:438-447) evaluates avg(max((J d + c)/rho0, 0)) at the
    // d of the last loop evaluation, with the same per-row arithmetic: it equals that err.
    st.final_density_error = err;
    if (spec_finish) return IVP_OK;
    return launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)d4, (const float*)c->jd, c->vel,
                  c->divg);
}

template <class FM>
int visc_explicit_body(ivp_ctx* c, float coeff, float dt) {
    const int nd = c->nd;
    const dim3 nb(blocks_for(nd));
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_v4, nb, nd, (const float*)c->vel, (const float*)c->density, c->v4));
    double bytes = stream_bytes<FM>(c, (double)c->state_host->total_pairs) + 48.0 * nd;
    Rows R = rows_of(c);
    if (c->all_dyn_fluid)
        return launch(c, KC_VISC, -1, bytes, k_visc<FM, true>, pass_grid(nd), R, coeff, dt, c->cfg.material_fluid,
                      (const float4*)c->v4, (const float*)c->rho0, (const int*)c->mat, c->vel);
    return launch(c, KC_VISC, -1, bytes, k_visc<FM, false>, pass_grid(nd), R, coeff, dt, c->cfg.material_fluid,
                  (const float4*)c->v4, (const float*)c->rho0, (const int*)c->mat, c->vel);
}

template <class FM, bool ALLFLUID>
int weiler18_body(ivp_ctx* c, float coeff, float dt, int* iters_out) {
    const int nd = c->nd;
    const dim3 nb(blocks_for(nd));
    const int stride = c->max_blocks;
    Rows R = rows_of(c);
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_w18_prep, nb, nd, coeff, dt, (const float*)c->vel, (const float4*)c->pos,
                   (const float*)c->density, (const float*)c->rho0, c->w_v4, c->w_fac));
    const double pairs = (double)c->state_host->total_pairs;
    const double pb = pair_bytes<FM>();
    W18Args wa{R, c->cfg.material_fluid, c->w_fac, c->mat, c->w_pinv, c->w_v4, c->w_rhs,
               c->w_dv, c->w_r, c->w_z, c->partials, stride, c->state, c->w18_precond ? 1 : 0, c->uid,
               c->w_dv_prev};
    if (!c->w18_warm) {
        // one pair pass: preconditioner blocks, right-hand side and the PCG prologue
        IVP_TRY(launch(c, KC_W18_PRE, -1, stream_bytes<FM>(c, pairs) + (16.0 + 4.0 + 36.0 + 12.0 * 3 + 16.0) * nd,
                       k_w18_setup<FM, ALLFLUID>, pass_grid(nd), wa, c->w_p4));
    } else {
        // warm start: blocks, rhs and dv = the last correction; then r = rhs - A dv
        IVP_TRY(launch(c, KC_W18_PRE, -1, stream_bytes<FM>(c, pairs) + (16.0 + 4.0 + 36.0 + 12.0 * 3 + 16.0) * nd,
                       k_w18_setup<FM, ALLFLUID, true>, pass_grid(nd), wa, c->w_p4));
        W18Args res = wa;
        res.v4 = c->w_p4;
        res.out = c->w_Ap;
        IVP_TRY(launch(c, KC_W18_PRE, -1, stream_bytes<FM>(c, pairs) + 48.0 * nd, k_w18_Av<FM, ALLFLUID, false>, pass_grid(nd),
                       res));
        IVP_TRY(launch(c, KC_W18_PRE, -1, 0.0, k_w18_warm_init, nb, res, (const float*)c->w_rhs, c->w_p4));
    }
    IVP_TRY(finalize<FIN_W18_INIT>(c, c->w18_warm ? nb.x : pass_grid(nd).x, 1e-5, 1000));
    // the first CG chunk follows without waiting (its kernels return at once when r.z is tiny)
    int iters = 0, chunks = 0;
    {
        W18Args cg = wa;
        cg.v4 = c->w_p4;
        cg.out = c->w_Ap;
        const double bAv = stream_bytes<FM>(c, (double)c->state_host->total_pairs) + 48.0 * nd;
        auto one_iter = [&](int tag) -> int {
            IVP_TRY(launch(c, KC_W18_AV, tag, bAv, k_w18_Av<FM, ALLFLUID>, pass_grid(nd), cg));
            IVP_TRY(finalize<FIN_W18_PAP>(c, pass_grid(nd).x));
            IVP_TRY(launch(c, KC_W18_UPDATE, tag, 84.0 * nd, k_w18_update, nb, cg));
            IVP_TRY(finalize<FIN_W18_UPDATE>(c, nb.x));
            return launch(c, KC_W18_P, tag, 68.0 * nd, k_w18_p, nb, nd, (const float*)c->w_z, c->w_p4, c->w_dv,
                          (const SolverState*)c->state);
        };
        // the save of the correction and v += dv follow each chunk's state copy, running once the
        // solve has stopped (while the host reads the state)
        auto tail = [&]() -> int {
            IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_w18_save, dim3(blocks_for(c->n)), c->n, nd, (const float*)c->w_dv,
                           (const int*)c->inv, c->w_dv_prev, (const SolverState*)c->state));
            return launch(c, KC_GLUE, -1, 0.0, k_w18_apply, dim3(blocks_for(3 * nd)), 3 * nd, (const float*)c->w_dv,
                          c->vel, (const SolverState*)c->state);
        };
        c->prof_epoch++;
        int epoch = c->prof_epoch;
        IVP_TRY(run_chunks(c, 1000, c->w18_guess, one_iter, &iters, &chunks, tail));
        // the pair count of the last build, now read back with the solver state
        prof_set_bytes(c, epoch, KC_W18_AV, stream_bytes<FM>(c, (double)c->state_host->total_pairs) + 48.0 * nd);
        prof_collect(c, iters, epoch);
        c->w18_guess = std::max(2, iters);
        // the device stops the solve at the same 1000 iterations, so the last read-back shows it
        // stopped and the tail after it has run
        if (c->state_host->active != 0) return IVP_ERR_INTERNAL;
    }
    *iters_out = iters;
    return IVP_OK;
}

template <class FM, bool ALLFLUID>
int jeske23_body(ivp_ctx* c, float sigma, float dt, int* iters_out) {
    const int nd = c->nd;
    const dim3 nb(blocks_for(nd)), np = pass_grid(nd);
    const int stride = c->max_blocks;
    IVP_TRY(st_buffers(c));
    Rows R = rows_of(c);
    // d = 2 r and h_st = 2 d = 4 r = h (jeske23.py:391-392 with h = 4 r): exact in float32
    const float hst = c->h, d = 0.5f * c->h;
    const float theta = 10.0f / 7.0f;
    const float norm = 8.0f / (3.14159265358979f * hst * hst * hst);
    const float tn = theta * norm;
    const float tnh = tn * (1.0f / hst);
    StArgs sa{R, c->cfg.material_fluid, c->mat, c->density, c->w_v4, c->st_div, c->st_gt, c->st_b, c->w_pinv,
              c->w_rhs, sigma, dt, d, hst, 1.0f / hst, tn, tnh, c->partials, stride, c->state};
    const double pairs = (double)c->state_host->total_pairs;
    const double pb = pair_bytes<FM>();
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_st_prep, nb, nd, (const float*)c->vel, (const float4*)c->pos, c->w_v4));
    IVP_TRY(launch(c, KC_ST_SETUP, -1, stream_bytes<FM>(c, pairs) + 36.0 * nd, k_st_div<FM, ALLFLUID>, np, sa));
    IVP_TRY(launch(c, KC_ST_SETUP, -1, (pb + 4.0) * pairs + 72.0 * nd, k_st_setup<FM, ALLFLUID>, np, sa));
    IVP_TRY(launch(c, KC_ST_AV, -1, stream_bytes<FM>(c, pairs) + 4.0 * pairs + 40.0 * nd, k_st_Av<FM, ALLFLUID, true>, np, sa));
    W18Args wa{R, c->cfg.material_fluid, c->w_fac, c->mat, c->w_pinv, c->w_v4, c->w_rhs,
               c->w_dv, c->w_r, c->w_z, c->partials, stride, c->state};
    IVP_TRY(launch(c, -1, -1, 0.0, k_w18_init, nb, wa, c->w_p4));
    // Jeske23.PCG (jeske23.py:499-570): tol 1e-5 on r.r, at most 100 iterations
    IVP_TRY(finalize<FIN_W18_INIT>(c, nb.x, 1e-5, 100));
    int iters = 0, chunks = 0;
    {
        StArgs cga = sa;
        cga.u4 = c->w_p4;
        cga.out = c->w_Ap;
        W18Args cg = wa;
        cg.v4 = c->w_p4;
        cg.out = c->w_Ap;
        const double bAv = stream_bytes<FM>(c, pairs) + 4.0 * pairs + 28.0 * nd;
        auto one_iter = [&](int tag) -> int {
            IVP_TRY(launch(c, KC_ST_AV, tag, bAv, k_st_Av<FM, ALLFLUID, false>, np, cga));
            IVP_TRY(finalize<FIN_W18_PAP>(c, np.x));
            IVP_TRY(launch(c, KC_W18_UPDATE, tag, 84.0 * nd, k_w18_update, nb, cg));
            IVP_TRY(finalize<FIN_W18_UPDATE>(c, nb.x));
            return launch(c, KC_W18_P, tag, 68.0 * nd, k_w18_p, nb, nd, (const float*)c->w_z, c->w_p4, c->w_dv,
                          (const SolverState*)c->state);
        };
        c->prof_epoch++;
        int epoch = c->prof_epoch;
        IVP_TRY(run_chunks(c, 100, c->st_guess, one_iter, &iters, &chunks));
        prof_set_bytes(c, epoch, KC_ST_AV, stream_bytes<FM>(c, (double)c->state_host->total_pairs) + 4.0 * (double)c->state_host->total_pairs + 28.0 * nd);
        prof_collect(c, iters, epoch);
        c->st_guess = std::max(2, iters);
    }
    *iters_out = iters;
    return launch(c, KC_GLUE, -1, 0.0, k_w18_apply, dim3(blocks_for(3 * nd)), 3 * nd, (const float*)c->w_dv, c->vel,
                  (const SolverState*)nullptr);
}

// One pass loop of PRJ (solve_PJ, prj.py:162-211; DF: solve_df's loop, prj.py:238-281) on the
// multipliers and dt v0 set up by k_prj_init.  Passes are enqueued in chunks; each pass ends with
// the device-side stop test, after which the remaining passes of a chunk return at once.
// Returns the pass index the loop stopped at (stats_outer_iters) and its density error.
template <class FM, bool DF>
int prj_loop(ivp_ctx* c, const ivp_prj_params* prm, double tol, int* passes, double* err, int* chunks) {
    const int nd = c->nd;
    const dim3 np = pass_grid(nd);
    const double pb = pair_bytes<FM>(), pairs = (double)c->state_host->total_pairs;
    hipLaunchKernelGGL(k_prj_state_init, dim3(1), dim3(1), 0, c->stream, c->state);
    IVP_HIP(hipGetLastError());
    PrjArgs pa{rows_of(c), c->Ss, c->prj_v0, c->prj_lam, c->d4, c->cc, c->kk, c->rho0, c->rowtot, c->jd,
               prm->omega, c->C, prm->df_num_threshold, c->partials, c->max_blocks, c->state};
    PrjFinArgs fa{c->partials, c->max_blocks, (int)np.x, c->state, tol, prm->max_iteration_opt,
                  prm->use_fixed_iters ? 1 : 0, nd, DF ? nullptr : c->prj_hist, DF ? 0 : c->prj_hist_cap};
    // algorithmic bytes: the pair stream once, the gathered per-particle data once per row
    const double bH = stream_bytes<FM>(c, pairs) + (double)nd * (8 + 12 + 12 + 16);
    const double bT = stream_bytes<FM>(c, pairs) + (double)nd * (16 + 12 + 4 + 4 + 4 + 8 + 4 + 8 + (DF ? 4 : 0));
    c->prof_epoch++;
    const int epoch = c->prof_epoch;
    const int total = prm->max_iteration_opt + 1;
    int issued = 0;
    int chunk = std::max(2, std::min(c->prj_guess, 128));
    if (!DF) IVP_TRY(iter_mark(c, 0));
    for (;;) {
        const int todo = std::min(chunk, total - issued);
        if (todo <= 0) break;
        for (int q = 0; q < todo; ++q) {
            const int tag = issued + q;
            IVP_TRY(launch(c, KC_PRJ_HEAD, tag, bH, k_prj_head<FM>, np, pa));
            IVP_TRY(launch(c, KC_PRJ_TAIL, tag, bT, k_prj_tail<FM, DF>, np, pa));
            hipLaunchKernelGGL(k_fin_prj, dim3(1), dim3(FIN_BLOCK), 0, c->stream, fa);
            IVP_HIP(hipGetLastError());
            if (!DF) IVP_TRY(iter_mark(c, tag + 1));
        }
        issued += todo;
        IVP_TRY(read_state(c));
        (*chunks)++;
        const SolverState& s = *c->state_host;
        if (s.active == 0) break;
        // geometric extrapolation of the density error to the tolerance
        const int done = s.iter;   // passes completed
        double rate = done > 1 && s.rr0 > 0.0 && s.rr_last > 0.0 ? pow(s.rr_last / s.rr0, 1.0 / (done - 1)) : 0.0;
        int est = 16;
        if (rate > 0.0 && rate < 0.999 && s.rr_last > tol) est = (int)ceil(log(tol / s.rr_last) / log(rate)) + 1;
        chunk = std::max(1, std::min(est, 128));
    }
    *passes = c->state_host->iter;
    *err = c->state_host->rr_last;
    prof_collect(c, *passes + 1, epoch);
    return IVP_OK;
}

// PRJ.solve (prj.py:292-340)
template <class FM>
int prj_solve_body(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* stats) {
    ivp_ip_stats& st = *stats;
    const int nd = c->nd;
    if (nd == 0) return IVP_OK;
    const dim3 nb(blocks_for(nd));
    const float inv_dt = (float)(1.0 / (double)dt);
    auto t0 = std::chrono::steady_clock::now();
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_prj_init<false>, nb, nd, dt, (const float*)c->density,
                   (const float*)c->rho0, (const float*)c->vel, (const float4*)c->pos, (const int*)c->uid,
                   (const float*)c->prj_prev, prm->warm_start ? 1 : 0, c->cc, c->prj_v0, c->prj_lam));
    int passes = 0;
    double err = 0.0;
    IVP_TRY((prj_loop<FM, false>(c, prm, prm->tol_opt, &passes, &err, &st.cg_launch_chunks)));
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_prj_save, nb, nd, (const float2*)c->prj_lam, (const int*)c->uid,
                   c->prj_prev));
    const int n_err = passes + 1;
    const int n_copy = std::min(std::min(n_err, c->prj_hist_cap), st.density_errors ? st.capacity : 0);
    if (n_copy > 0)
        IVP_HIP(hipMemcpyAsync(st.density_errors, c->prj_hist, sizeof(double) * (size_t)n_copy,
                               hipMemcpyDeviceToHost, c->stream));
    st.n_density_errors = n_err;
    st.outer_iters = passes;
    const long long tp = c->state_host->total_pairs;
    st.pairs_traversed += 2ll * (passes + 1) * tp;
    st.rows_traversed += 2ll * (passes + 1) * nd;
    // final_density_error (pressure_base.py:438-447) evaluates avg(max((J dt v + c) / rho0, 0)) at
    // the dt v of the last pass with the tail's per-row arithmetic: it equals that pass's err.
    st.final_density_error = err;
    c->prj_guess = std::max(2, passes + 1);
    if (!prm->use_fixed_iters && err >= prm->tol_opt && passes >= prm->max_iteration_opt) {
        IVP_HIP(hipStreamSynchronize(c->stream));
        iter_costs(c, n_err, &st);
        return IVP_NOT_CONVERGED;
    }
    // v = dt v / dt and the divergence max(J v, 0) from J (dt v) of the last pass
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)c->d4, (const float*)c->jd,
                   c->vel, c->divg));
    IVP_HIP(hipStreamSynchronize(c->stream));
    iter_costs(c, n_err, &st);
    st.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return IVP_OK;
}

// PRJ.solve_df (prj.py:213-290) on the stored v
template <class FM>
int prj_df_body(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* stats) {
    ivp_ip_stats& st = *stats;
    const int nd = c->nd;
    if (nd == 0) return IVP_OK;
    const dim3 nb(blocks_for(nd));
    const float inv_dt = (float)(1.0 / (double)dt);
    const long long tp = c->state_host->total_pairs;
    auto t1 = std::chrono::steady_clock::now();
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_prj_init<true>, nb, nd, dt, (const float*)c->density,
                   (const float*)c->rho0, (const float*)c->vel, (const float4*)c->pos, (const int*)c->uid,
                   (const float*)c->prj_df_prev, prm->warm_start ? 1 : 0, c->cc, c->prj_v0, c->prj_lam));
    int df_passes = 0;
    double df_err = 0.0;
    IVP_TRY((prj_loop<FM, true>(c, prm, prm->df_tol_opt, &df_passes, &df_err, &st.cg_launch_chunks)));
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_prj_save, nb, nd, (const float2*)c->prj_lam, (const int*)c->uid,
                   c->prj_df_prev));
    st.df_outer_iters = df_passes;
    st.pairs_traversed += 2ll * (df_passes + 1) * tp;
    st.rows_traversed += 2ll * (df_passes + 1) * nd;
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)c->d4, (const float*)c->jd,
                   c->vel, c->divg));
    IVP_HIP(hipStreamSynchronize(c->stream));
    st.df_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    return IVP_OK;
}

}  // namespace

namespace ivp {

// Jeske23.solve (jeske23.py:590-655).  sigma = 0 leaves v unchanged (M = I, zero right-hand
// side) and is skipped.  Static neighbours never enter: the reference skips non-fluid
// neighbours, and a static particle with the fluid material is refused.
int st_jeske23(ivp_ctx* c, float sigma, float dt, ivp_cg_stats* out) {
    if (!c->built) return IVP_ERR_STATE;
    if (out) { out->cg_iters = 0; out->residual_norm2 = 0.0; }
    if (c->static_fluid) return IVP_ERR_ARG;
    if (c->nd == 0 || sigma == 0.0f) return IVP_OK;
    int iters = 0, rc;
    switch (c->row_fmt) {
    case IVP_ROWS_F32:
        rc = c->all_dyn_fluid ? jeske23_body<FmtF32, true>(c, sigma, dt, &iters)
                              : jeske23_body<FmtF32, false>(c, sigma, dt, &iters);
        break;
    case IVP_ROWS_F16:
        rc = c->all_dyn_fluid ? jeske23_body<FmtF16Eval, true>(c, sigma, dt, &iters)
                              : jeske23_body<FmtF16Eval, false>(c, sigma, dt, &iters);
        break;
    default:
        rc = c->all_dyn_fluid ? jeske23_body<FmtIds, true>(c, sigma, dt, &iters)
                              : jeske23_body<FmtIds, false>(c, sigma, dt, &iters);
        break;
    }
    if (rc != IVP_OK) return rc;
    if (out) {
        out->cg_iters = iters;
        out->residual_norm2 = c->state_host->rr_last;
    }
    return IVP_OK;
}

// IP.solve with its wall-clock cost (stats_total_cost_ms, ip.py:335, 403-404: the reference
// synchronises before reading the clock, and so does this).
int ip_solve(ivp_ctx* c, float dt, const ivp_ip_params* prm, ivp_ip_stats* stats, bool defer) {
    if (!c->built) return IVP_ERR_STATE;
    if (!prm || !stats) return IVP_ERR_ARG;
    ivp_ip_stats& st = *stats;
    st.n_density_errors = 0;
    st.n_pcg = 0;
    st.outer_iters = 0;
    st.final_density_error = 0.0;
    st.pAp_bad = 0.0;
    st.pcg_iter_bad = 0;
    st.pairs_traversed = 0;
    st.rows_traversed = 0;
    st.cg_launch_chunks = 0;
    st.df_outer_iters = 0;
    st.df_ms = 0.0;
    st.n_iter_costs = 0;
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    switch (c->row_fmt) {
    case IVP_ROWS_F32: rc = ip_solve_body<FmtF32, FmtF32>(c, dt, prm, stats); break;
    case IVP_ROWS_F16: rc = ip_solve_body<FmtF16Eval, FmtF16>(c, dt, prm, stats); break;
    default: rc = ip_solve_body<FmtIds, FmtIds>(c, dt, prm, stats); break;
    }
    if (rc == IVP_OK && defer && c->nd > 0) {
        if (!c->ip_end_ev) IVP_HIP(hipEventCreateWithFlags(&c->ip_end_ev, hipEventDisableTiming));
        IVP_HIP(hipEventRecord(c->ip_end_ev, c->stream));
        c->ip_wait_pending = true;
        c->ip_t0 = t0;
        c->ip_n_err = st.n_density_errors;
        return rc;
    }
    if (rc >= 0) {
        IVP_HIP(hipStreamSynchronize(c->stream));
        if (c->nd > 0) iter_costs(c, st.n_density_errors, stats);
    }
    stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// The closing wait of a deferred ip_solve: the solve's end mark, then stats_iter_costs and the
// solve's wall time as ip_solve would have filled them.
int ip_solve_wait(ivp_ctx* c, ivp_ip_stats* stats) {
    if (!c->ip_wait_pending) return IVP_OK;
    c->ip_wait_pending = false;
    IVP_HIP(hipEventSynchronize(c->ip_end_ev));
    iter_costs(c, c->ip_n_err, stats);
    stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c->ip_t0).count();
    return IVP_OK;
}

// PRJ.solve (prj.py:292-340), followed by solve_df (prj.py:213-290) when part == PRJ_BOTH and
// enable_df is set, or solve_df alone (part == PRJ_DF).  Rows are complete and symmetric, so the
// head's J^T and the tail's J are the gathers of the IP passes; the operator arithmetic follows
// the row format (FP16 rows evaluate gradW from the positions, as the IP outer-loop passes do).
int prj_solve(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* stats, int part) {
    if (!c->built) return IVP_ERR_STATE;
    if (!prm || !stats) return IVP_ERR_ARG;
    if (prm->max_iteration_opt < 0 || prm->df_num_threshold < 0) return IVP_ERR_ARG;
    ivp_ip_stats& st = *stats;
    st.n_density_errors = 0;
    st.n_pcg = 0;
    st.outer_iters = 0;
    st.final_density_error = 0.0;
    st.pAp_bad = 0.0;
    st.pcg_iter_bad = 0;
    st.pairs_traversed = 0;
    st.rows_traversed = 0;
    st.cg_launch_chunks = 0;
    st.solve_ms = 0.0;
    st.df_outer_iters = 0;
    st.df_ms = 0.0;
    st.n_iter_costs = 0;
    IVP_TRY(prj_buffers(c, prm->max_iteration_opt + 1));
    // a row that overflowed the neighbour capacity at the last build (the IP solve reads the
    // same flag at its first synchronisation)
    IVP_HIP(hipMemcpyAsync(c->err_host, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    if (*c->err_host & 4) return IVP_ERR_INTERNAL;
    if (*c->err_host & 1) return IVP_NEIGHBOR_OVERFLOW;
    const int fmt = c->row_fmt;
    if (part != PRJ_DF) {
        int rc = fmt == IVP_ROWS_F32   ? prj_solve_body<FmtF32>(c, dt, prm, stats)
                 : fmt == IVP_ROWS_F16 ? prj_solve_body<FmtF16Eval>(c, dt, prm, stats)
                                       : prj_solve_body<FmtIds>(c, dt, prm, stats);
        if (rc != IVP_OK || part == PRJ_CD || !prm->enable_df) return rc;
    }
    return fmt == IVP_ROWS_F32   ? prj_df_body<FmtF32>(c, dt, prm, stats)
           : fmt == IVP_ROWS_F16 ? prj_df_body<FmtF16Eval>(c, dt, prm, stats)
                                 : prj_df_body<FmtIds>(c, dt, prm, stats);
}

int visc_explicit(ivp_ctx* c, float coeff, float dt) {
    if (!c->built) return IVP_ERR_STATE;
    if (c->nd == 0) return IVP_OK;
    switch (c->row_fmt) {
    case IVP_ROWS_F32: return visc_explicit_body<FmtF32>(c, coeff, dt);
    case IVP_ROWS_F16: return visc_explicit_body<FmtF16Eval>(c, coeff, dt);
    default: return visc_explicit_body<FmtIds>(c, coeff, dt);
    }
}

int visc_weiler18(ivp_ctx* c, float coeff, float dt, ivp_cg_stats* out) {
    if (!c->built) return IVP_ERR_STATE;
    if (out) { out->cg_iters = 0; out->residual_norm2 = 0.0; }
    if (c->nd == 0) return IVP_OK;
    IVP_TRY(w18_buffers(c));
    int iters = 0, rc;
    switch (c->row_fmt) {
    case IVP_ROWS_F32:
        rc = c->all_dyn_fluid ? weiler18_body<FmtF32, true>(c, coeff, dt, &iters)
                              : weiler18_body<FmtF32, false>(c, coeff, dt, &iters);
        break;
    case IVP_ROWS_F16:
        rc = c->all_dyn_fluid ? weiler18_body<FmtF16Eval, true>(c, coeff, dt, &iters)
                              : weiler18_body<FmtF16Eval, false>(c, coeff, dt, &iters);
        break;
    default:
        rc = c->all_dyn_fluid ? weiler18_body<FmtIds, true>(c, coeff, dt, &iters)
                              : weiler18_body<FmtIds, false>(c, coeff, dt, &iters);
        break;
    }
    if (rc != IVP_OK) return rc;
    if (out) {
        out->cg_iters = iters;
        out->residual_norm2 = c->state_host->rr_last;
    }
    return IVP_OK;
}

// IP.solve (ip.py:331-404) for every scene of the context at once: per-scene outer loops and
// CGs advanced by replays of one captured launch sequence, one host synchronisation per replay.
int ip_solve_scenes(ivp_ctx* c, float dt, const ivp_ip_params* prm) {
    if (!c->built || c->S == 0) return IVP_ERR_STATE;
    if (!prm) return IVP_ERR_ARG;
    const int nd = c->nd;
    const dim3 nb(blocks_for(nd));
    SceneParams& hp = *c->sprm_host;
    hp.tol_opt = prm->tol_opt;
    hp.tol_pcg = prm->tol_pcg;
    hp.max_opt = prm->max_iteration_opt;
    hp.max_pcg = prm->max_iteration_pcg;
    hp.fixed = prm->use_fixed_iters ? 1 : 0;
    IVP_HIP(hipMemcpyAsync(c->sprm, c->sprm_host, sizeof(SceneParams), hipMemcpyHostToDevice, c->stream));
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_init, nb, nd, dt, (const float*)c->density, (const float*)c->rho0,
                   (const float*)c->vel, (const float4*)c->pos, c->cc, c->d4));
    hipLaunchKernelGGL(k_scene_init, dim3((c->S + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream, c->S, c->ss);
    IVP_HIP(hipGetLastError());
    const int K = std::max(1, c->scene_chunk);
    hipGraphExec_t exec = nullptr;
    if (c->scene_graph) IVP_TRY(scene_graph(c, prm->omega, K, &exec));
    // every replay advances each unfinished scene by a phase or by K CG iterations
    const long long cap = (long long)(std::max(prm->max_iteration_opt, 0) + 1) *
                          ((long long)std::max(prm->max_iteration_pcg, 0) / K + 3) + 2;
    c->scene_replays = 0;
    for (;;) {
        if (exec)
            IVP_HIP(hipGraphLaunch(exec, c->stream));
        else
            IVP_TRY(enqueue_scene_replay(c, c->stream, prm->omega, K));
        c->scene_replays++;
        IVP_HIP(hipMemcpyAsync(c->sactive_host, c->sactive, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        if (c->scene_replays == 1)   // the flags of the rebuild that produced these rows
            IVP_HIP(hipMemcpyAsync(c->err_host, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        IVP_HIP(hipStreamSynchronize(c->stream));
        if (c->scene_replays == 1) {
            if (*c->err_host & 4) return IVP_ERR_INTERNAL;
            if (*c->err_host & 1) return IVP_NEIGHBOR_OVERFLOW;
        }
        if (*c->sactive_host == 0) break;
        if (c->scene_replays >= cap) return IVP_ERR_INTERNAL;
    }
    float inv_dt = (float)(1.0 / (double)dt);
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_ip_finish, nb, nd, inv_dt, (const float4*)c->d4, (const float*)c->jd, c->vel,
                   c->divg));
    IVP_HIP(hipMemcpyAsync(c->ss_host, c->ss, sizeof(SceneState) * (size_t)c->S, hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    return IVP_OK;
}

// v + g dt into the second velocity buffer, which becomes the first; the previous velocities stay
// in the second until the next re-sort, so a step that must not start can restore them
int glue_gravity(ivp_ctx* c, const float g[3], float dt) {
    if (!c->loaded) return IVP_ERR_STATE;
    if (c->nd == 0) return IVP_OK;
    IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_gravity, dim3(blocks_for(c->nd)), c->nd, g[0] * dt, g[1] * dt, g[2] * dt,
                   (const float*)c->vel, c->vel_b));
    std::swap(c->vel, c->vel_b);
    return IVP_OK;
}

int glue_advect_boundary(ivp_ctx* c, float dt) {
    if (!c->loaded) return IVP_ERR_STATE;
    if (c->nd == 0) return IVP_OK;
    if (c->S > 0)
        IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_advect_boundary<true>, dim3(blocks_for(c->nd)), c->nd, dt, 0.0f, 0.0f,
                       0.0f, c->h, c->pos, c->vel, (const float*)c->sdomain, (const int*)c->sid));
    else
        IVP_TRY(launch(c, KC_GLUE, -1, 0.0, k_advect_boundary<false>, dim3(blocks_for(c->nd)), c->nd, dt,
                       c->cfg.domain_size[0], c->cfg.domain_size[1], c->cfg.domain_size[2], c->h, c->pos, c->vel,
                       (const float*)nullptr, (const int*)nullptr));
    c->built = false;
    return IVP_OK;
}

}  // namespace ivp
