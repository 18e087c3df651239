// ivp_launch.h -- kernel launch with optional per-launch timing.
#pragma once

#include <hip/hip_ext.h>

#include "ivp_internal.h"

namespace ivp {

// Launch `kernel` (256-thread blocks) on the context's stream.  When timing is enabled for its
// class the launch goes through hipExtLaunchKernelGGL with two events that the dispatch itself
// stamps at kernel start and end, so the recorded time is the kernel's own execution time.
// kclass < 0 marks helper launches that are never timed.
// `shmem` bytes of dynamic LDS per block.
template <typename... KArgs, typename... Args>
inline int launch_shm(ivp_ctx* c, int kclass, int tag, double bytes, void (*kernel)(KArgs...), dim3 grid,
                      unsigned shmem, Args... args) {
    if (kclass >= 0 && c->prof && (c->prof_mask & (1u << kclass))) {
        ProfEvent ev;
        int rc = prof_take(c, kclass, tag, bytes, &ev);
        if (rc != IVP_OK) return rc;
        hipExtLaunchKernelGGL(kernel, grid, dim3(BLOCK), shmem, c->stream, ev.start, ev.stop, 0, args...);
        rc = prof_push(c, &ev);
        if (rc != IVP_OK) return rc;
    } else {
        hipLaunchKernelGGL(kernel, grid, dim3(BLOCK), shmem, c->stream, args...);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? IVP_OK : ivp_hip_fail(e, "kernel launch", __FILE__, __LINE__);
}

template <typename... KArgs, typename... Args>
inline int launch(ivp_ctx* c, int kclass, int tag, double bytes, void (*kernel)(KArgs...), dim3 grid,
                  Args... args) {
    return launch_shm(c, kclass, tag, bytes, kernel, grid, 0u, args...);
}

}  // namespace ivp

#define IVP_TRY(expr)                      \
    do {                                   \
        int rc_ = (expr);                  \
        if (rc_ != IVP_OK) return rc_;     \
    } while (0)
