// ivp_device.h -- device-side helpers shared by the libivp kernels (gfx950, wave64).
//
// Floating-point policy: every kernel file is compiled with -ffp-contract=off, and HIP's
// float division and sqrtf are correctly rounded, so the per-pair quantities below (kernel
// values, gradients, distances) are bit-identical to a float32 CPU evaluation of the same
// expression order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ivp {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;            // 4 waves; every row kernel maps wave w of a block to tile
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

// Tile windows: the pair passes stage the distinct neighbours of a 64-row tile in LDS, at most
// WIN_MAX of them, found by the build in a bitmap of WIN_SPAN_WORDS words over the tile's id
// range; a tile outside either bound gathers its neighbours from memory.
constexpr int WIN_MAX = 608;
constexpr int WIN_SPAN_WORDS = 3072;

struct Grid {
    float origin[3];
    float inv_w;
    int dims[3];
    int base;      // first key of this grid (independent scenes: one grid per scene, keys consecutive)
};

// The kernel constants of sph_kernel.py:14-15 and :55, evaluated in float32 with the reference's
// operation order (once per launch: a division inside a loop is not hoisted on this target).
__host__ __device__ inline float cubic_constant(float h) {
    float k = 8.0f / 3.14159265359f;
    return k / (h * h * h);
}
__host__ __device__ inline float spiky_constant(float h) {
    return 45.0f / (3.14159265359f * h * h * h * h * h * h);
}

// cubic_kernel, scripts/sim/sph_kernel.py:8-26 (same operation order); k = cubic_constant(h)
__device__ __forceinline__ float cubic_kernel(float r_norm, float h, float k) {
    float res = 0.0f;
    float q = r_norm / h;
    if (q <= 1.0f) {
        if (q <= 0.5f) {
            float q2 = q * q;
            float q3 = q2 * q;
            res = k * (6.0f * q3 - 6.0f * q2 + 1.0f);
        } else {
            float factor = 1.0f - q;
            res = k * 2.0f * factor * factor * factor;
        }
    }
    return res;
}

// x / b rounded to nearest, given y = 1 / b rounded to nearest: q = RN(x y) is within an ulp of
// x / b, and one correction with the exact residual x - b q gives the correctly rounded quotient
// (Markstein's theorem; it needs y correctly rounded and no overflow or underflow, which holds for
// the build's quotients: components of a vector over its length, a length over h).  Three
// quotients by one divisor then cost one division and three corrections.
__device__ __forceinline__ float div_rn(float x, float b, float y) {
    const float q = x * y;
    return fmaf(fmaf(-b, q, x), y, q);
}

// cubic_kernel with q = r_norm / h formed by div_rn from inv_h = 1 / h (rounded to nearest)
__device__ __forceinline__ float cubic_kernel_rn(float r_norm, float h, float inv_h, float k) {
    float res = 0.0f;
    float q = div_rn(r_norm, h, inv_h);
    if (q <= 1.0f) {
        if (q <= 0.5f) {
            float q2 = q * q;
            float q3 = q2 * q;
            res = k * (6.0f * q3 - 6.0f * q2 + 1.0f);
        } else {
            float factor = 1.0f - q;
            res = k * 2.0f * factor * factor * factor;
        }
    }
    return res;
}

__device__ __forceinline__ float len3(float x, float y, float z) {
    return sqrtf(x * x + y * y + z * z);
}

// sqrt(x) rounded to nearest from the hardware root s (v_sqrt_f32, within one ulp): s moves to a
// neighbour when the neighbour's residual x - n s (one fused multiply-add) shows the root beyond
// the midpoint.  Exact for every x >= 2^-96 and any s within one ulp (checked over every float of
// [1, 4) and its three candidate roots); below 2^-96 (|r| < 2^-48) the result stays below 2^-48,
// where the build's kernels do not depend on it (spiky clamps |r| to 1e-6; the cubic kernel is
// exactly its constant for |r| / h < 7e-5).  The same as sqrtf for the build's pairs, without
// sqrtf's scaling of small arguments.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
    float r = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : r;
}

__device__ __forceinline__ float len3_rn(float x, float y, float z) {
    return sqrt_rn(x * x + y * y + z * z);
}

// 1/b rounded to nearest for normal b > 0 with a normal reciprocal: the hardware reciprocal y0
// (v_rcp_f32, within one ulp) and one Newton step with fused multiply-adds, which rounds
// correctly from any y0 within two ulps except when b's significand is all ones; there
// 1/b = 2^-e-1 (1 + 2^-24 + ...) rounds to 2^-e-1 (1 + 2^-23), formed from b's exponent (checked
// over every significand and five candidate y0).
__device__ __forceinline__ float rcp_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float y1 = fmaf(fmaf(-b, y0, 1.0f), y0, y0);
    const uint32_t u = __float_as_uint(b);
    return (u & 0x7fffffu) == 0x7fffffu ? __uint_as_float((0x7e800000u - (u & 0x7f800000u)) | 1u) : y1;
}

// spiky_kernel_derivative, scripts/sim/sph_kernel.py:52-61 (same operation order);
// k = spiky_constant(h)
__device__ __forceinline__ float3 spiky_grad(float rx, float ry, float rz, float h, float k) {
    float r_norm = len3(rx, ry, rz);
    if (r_norm < 1e-6f) r_norm = 1e-6f;
    float gx = rx / r_norm, gy = ry / r_norm, gz = rz / r_norm;
    float t = h - r_norm;
    float s = -k * (t * t);
    return make_float3(s * gx, s * gy, s * gz);
}

// spiky_grad given r_norm = len3(rx, ry, rz), the three quotients by div_rn
__device__ __forceinline__ float3 spiky_grad_rn(float rx, float ry, float rz, float r_norm, float h, float k) {
    if (r_norm < 1e-6f) r_norm = 1e-6f;
    const float y = rcp_rn(r_norm);
    float gx = div_rn(rx, r_norm, y), gy = div_rn(ry, r_norm, y), gz = div_rn(rz, r_norm, y);
    float t = h - r_norm;
    float s = -k * (t * t);
    return make_float3(s * gx, s * gy, s * gz);
}

// Two floats rounded to float16 (round to nearest even), packed low | high << 16
__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
    uint32_t a = __builtin_bit_cast(uint16_t, (_Float16)lo);
    uint32_t b = __builtin_bit_cast(uint16_t, (_Float16)hi);
    return a | (b << 16);
}

// Cell coordinate along one axis: floor((x - origin) * inv_w), clamped into the grid.
// The cell width w = h * (1 + 2^-10) keeps the 3x3x3 stencil complete under rounding.
__device__ __forceinline__ int cell_coord(const Grid& g, float x, int a) {
    float t = (x - g.origin[a]) * g.inv_w;
    t = fminf(fmaxf(t, -1.0f), (float)g.dims[a]);
    int c = (int)floorf(t);
    c = c < 0 ? 0 : c;
    c = c > g.dims[a] - 1 ? g.dims[a] - 1 : c;
    return c;
}

// Neighbour records of a 64-row tile: the word of slot k of a row, relative to the tile's first
// word, is row_word0(lane) + slot_word(k).  Ids-only rows (REC = 64 words per slot) keep slots
// g and g + 4 of each 8-slot block of a row in adjacent words: block b holds, for row `lane`,
// words 512 b + 8 lane + [slot 0, 4, 1, 5, 2, 6, 3, 7 of the block], so a lane that walks slots
// g, g + 4, ... (four lanes per row) reads two ids with one 8-byte load, and the 16 rows x 4
// lanes of a wave read 512 contiguous bytes.  Formats with per-pair gradient words (REC = 192,
// 256) keep slot-major records: word k REC + lane, the gradient words 64 and 128 (and 192)
// words further.
template <unsigned REC>
__device__ __forceinline__ uint32_t row_word0(int lane) {
    return REC == 64u ? 8u * (uint32_t)lane : (uint32_t)lane;
}
template <unsigned REC>
__device__ __forceinline__ uint32_t slot_word(int k) {
    if (REC == 64u) return ((uint32_t)(k >> 3) << 9) | (uint32_t)(((k & 3) << 1) | ((k >> 2) & 1));
    return (uint32_t)k * REC;
}

__device__ __forceinline__ int cell_key(const Grid& g, int cx, int cy, int cz) {
    return g.base + (cz * g.dims[1] + cy) * g.dims[0] + cx;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8).  Remap so
// each XCD walks one contiguous range of particle tiles: spatial neighbours (and the particle
// data they gather) then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    int xcd = b & 7, idx = b >> 3;
    int per = nblocks >> 3, rem = nblocks & 7;
    int start = xcd * per + (xcd < rem ? xcd : rem);
    return start + idx;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, WAVE);
        v = u < v ? u : v;
    }
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, WAVE);
        v = u > v ? u : v;
    }
    return v;
}

// Sum of one double per thread over a 256-thread block, in a fixed order.  Result valid in
// thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4) {
    v = wave_sum(v);
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
    if (lane == 0) lds4[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

}  // namespace ivp
