// ivp_kinematic.hip -- keyframe-animated rigid bodies and the CFL speed bound.
//
// ivp_rigid_transform: apply_transform_kernel (animation.py:7-34) on user-order device arrays:
// for every particle of the target object, x = (pivot + T) + R (x0 - pivot) in float32, the
// matrix-vector product summed ((R_i0 d_0 + R_i1 d_1) + R_i2 d_2).
// ivp_max_speed2: compute_max_velocity_sq (framework.py:44-58), max |v|^2 = (vx vx + vy vy) + vz vz
// over the dynamic particles, as one atomic max per block on the float's bit pattern (squares
// are non-negative, so the bit patterns order like the values).
#include <string.h>

#include "ivp_internal.h"
#include "ivp_launch.h"

using namespace ivp;

namespace {

struct RigidXf {
    float pivot[3];
    float R[9];
    float T[3];
};

__global__ __launch_bounds__(BLOCK) void k_rigid_transform(const float* __restrict__ x0, float* __restrict__ x,
                                                           const int* __restrict__ object_id, int n, int target,
                                                           RigidXf f, int* __restrict__ moved) {
    int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n || object_id[p] != target) return;
    float d0 = x0[3 * p + 0] - f.pivot[0];
    float d1 = x0[3 * p + 1] - f.pivot[1];
    float d2 = x0[3 * p + 2] - f.pivot[2];
    float r0 = (f.R[0] * d0 + f.R[1] * d1) + f.R[2] * d2;
    float r1 = (f.R[3] * d0 + f.R[4] * d1) + f.R[5] * d2;
    float r2 = (f.R[6] * d0 + f.R[7] * d1) + f.R[8] * d2;
    x[3 * p + 0] = (f.pivot[0] + f.T[0]) + r0;
    x[3 * p + 1] = (f.pivot[1] + f.T[1]) + r1;
    x[3 * p + 2] = (f.pivot[2] + f.T[2]) + r2;
    if (moved) atomicOr(moved, 1);
}

__global__ __launch_bounds__(BLOCK) void k_max_speed2(int nd, const float* __restrict__ vel,
                                                      unsigned* __restrict__ out) {
    __shared__ unsigned lds[WAVES_PER_BLOCK];
    int i = blockIdx.x * BLOCK + threadIdx.x;
    unsigned b = 0u;
    if (i < nd) {
        float vx = vel[3 * i + 0], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        float s = (vx * vx + vy * vy) + vz * vz;
        b = isnan(s) ? 0x7fc00000u : __float_as_uint(s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned u = __shfl_xor(b, o, WAVE);
        b = u > b ? u : b;
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = lds[0];
        for (int w = 1; w < WAVES_PER_BLOCK; ++w) m = lds[w] > m ? lds[w] : m;
        atomicMax(out, m);
    }
}

}  // namespace

extern "C" int ivp_rigid_transform(const float* x0, float* x, const int* object_id, int n, int object,
                                   const float pivot[3], const float R[9], const float T[3], void* stream) {
    if (n < 0 || !pivot || !R || !T) return IVP_ERR_ARG;
    if (n == 0) return IVP_OK;
    if (!x0 || !x || !object_id) return IVP_ERR_ARG;
    RigidXf f;
    for (int a = 0; a < 3; ++a) {
        f.pivot[a] = pivot[a];
        f.T[a] = T[a];
    }
    for (int a = 0; a < 9; ++a) f.R[a] = R[a];
    hipLaunchKernelGGL(k_rigid_transform, dim3(blocks_for(n)), dim3(BLOCK), 0, (hipStream_t)stream, x0, x, object_id, n,
                       object, f, (int*)nullptr);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

extern "C" int ivp_max_speed2(ivp_ctx* c, double* out) {
    if (!c || !out) return IVP_ERR_ARG;
    if (!c->loaded) return IVP_ERR_STATE;
    *out = 0.0;
    if (c->nd == 0) return IVP_OK;
    unsigned* d = c->speed_max;
    IVP_HIP(hipMemsetAsync(d, 0, sizeof(unsigned), c->stream));
    hipLaunchKernelGGL(k_max_speed2, dim3(blocks_for(c->nd)), dim3(BLOCK), 0, c->stream, c->nd, (const float*)c->vel, d);
    IVP_HIP(hipGetLastError());
    unsigned h = 0;
    IVP_HIP(hipMemcpyAsync(&h, d, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    float f;
    ::memcpy(&f, &h, sizeof(float));
    *out = (double)f;
    return IVP_OK;
}
