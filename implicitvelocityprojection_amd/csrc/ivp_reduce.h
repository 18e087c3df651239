// ivp_reduce.h -- grid reductions: per-block partials plus a one-block finalize kernel that
// folds them in a fixed order (deterministic) and advances the device-resident solver state.
#pragma once

#include "ivp_internal.h"

namespace ivp {

// Per-block partial sums of up to 3 doubles (fixed order inside the block).  The grid totals are
// formed by k_finalize, launched next on the stream: the kernel boundary publishes the partials,
// so no device-scope fence (an L2 write-back on a multi-XCD part) is issued per block.
template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], double* partials, int stride) {
    __shared__ double lds[WAVES_PER_BLOCK];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = block_sum(v[q], lds);
        if (threadIdx.x == 0) partials[q * stride + blockIdx.x] = s;
    }
}

enum FinOp {
    FIN_BUILD, FIN_PACK, FIN_RHS, FIN_CG_INIT, FIN_CG_PAP, FIN_CG_UPDATE, FIN_W18_INIT, FIN_W18_PAP, FIN_W18_UPDATE,
    FIN_CG_REPLACE
};

struct FinArgs {
    const double* partials;
    int stride;
    int nblocks;
    SolverState* st;
    double tol;
    int max_iter;
    int gated;   // FIN_PACK queued behind a CG read-back: only once that CG stopped without failing
};

constexpr int FIN_BLOCK = 1024;
constexpr int FIN_UNROLL = 4;  // the fold below adds exactly four loads per row

// Fixed-order fold of NV rows of n partials by one FIN_BLOCK-thread block (valid in thread 0).
// Each thread issues NV * FIN_UNROLL independent loads before adding, so the fold costs a few
// memory latencies instead of one per loop trip.
template <int NV>
__device__ __forceinline__ void fold_partials(const double* __restrict__ partials, int stride, int n,
                                              double (&t)[3]) {
    __shared__ double lds[NV][FIN_BLOCK / WAVE];
    double s[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = 0.0;
    for (int base = 0; base < n; base += FIN_UNROLL * FIN_BLOCK) {
        double v[NV][FIN_UNROLL];
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int u = 0; u < FIN_UNROLL; ++u) {
                int idx = base + u * FIN_BLOCK + (int)threadIdx.x;
                v[q][u] = idx < n ? partials[q * stride + idx] : 0.0;
            }
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] += (v[q][0] + v[q][1]) + (v[q][2] + v[q][3]);
    }
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = wave_sum(s[q]);
        if (lane == 0) lds[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double x = 0.0;
            for (int k = 0; k < FIN_BLOCK / WAVE; ++k) x += lds[q][k];
            t[q] = x;
        }
    }
}

// Scalar state update after a reduction (the host logic of ip.py:296-327,
// pressure_base.py:405-423, weiler18.py:330-365), applied to one solver state.
template <int OP>
__device__ __forceinline__ void fin_apply(SolverState* st, const double (&t)[3], double tol, int max_iter) {
    if (OP == FIN_BUILD) {
        st->total_pairs = (long long)t[0];
        st->win_entries = (long long)t[1];
    } else if (OP == FIN_PACK) {
        st->err_sum = t[0];
        st->n_active_rows = (int)t[1];
        st->active_pairs = (long long)t[2];
    } else if (OP == FIN_RHS) {
        st->n_touched_rows = (int)t[0];
        st->touched_pairs = (long long)t[1];
    } else if (OP == FIN_CG_INIT || OP == FIN_W18_INIT) {
        // IP: CG runs when r.r > 1e-30 (ip.py:302); Weiler18: when r.z > 1e-12 (weiler18.py:337)
        double floor_ = OP == FIN_CG_INIT ? 1e-30 : 1e-12;
        st->rr_old = t[0];
        st->rr_new = t[0];
        st->rr_last = t[0];
        st->rr0 = t[0];
        st->rr_best = t[0];
        st->stalls = 0;
        st->tol = tol;
        st->max_iter = max_iter;
        st->iter = 0;
        st->status = 0;
        st->dv_apply = 0;
        st->active = (t[0] > floor_ && max_iter > 0) ? 1 : 0;
    } else if (OP == FIN_CG_PAP) {
        // ip.py:307-311
        st->pAp = t[0];
        if (t[0] < 0.0) {
            st->status = IVP_NOT_PD;
            st->active = 0;
        } else {
            st->alpha = st->rr_old / (t[0] > 1e-30 ? t[0] : 1e-30);
        }
    } else if (OP == FIN_W18_PAP) {
        // weiler18.py:338-342: pAp <= 0 ends the iteration
        st->pAp = t[0];
        if (t[0] <= 0.0) {
            st->active = 0;
        } else {
            st->alpha = st->rr_old / (t[0] > 1e-12 ? t[0] : 1e-12);
        }
    } else if (OP == FIN_CG_UPDATE) {
        // ip.py:317-327
        int it = st->iter + 1;
        st->iter = it;
        st->rr_new = t[0];
        st->rr_last = t[0];
        bool stop = (t[0] < st->tol && it >= 2) || it >= st->max_iter;
        if (stop) {
            st->active = 0;
        } else {
            double ro = st->rr_old;
            st->beta = t[0] / (ro > 1e-30 ? ro : 1e-30);
            st->rr_old = t[0];
        }
    } else if (OP == FIN_CG_REPLACE) {
        // CG iteration whose recursive residual (t[0] = its r.r) was replaced by the true
        // residual b - A x (t[1] = its r.r).  The stop test reads the true r.r; beta keeps the
        // recursion's ratio, so a true residual far above the recursive one (an operator error
        // floor) turns p into about r instead of amplifying p.  Two replacements in a row that
        // do not lower the smallest true r.r end the solve: that floor has been reached.
        int it = st->iter + 1;
        st->iter = it;
        st->rr_new = t[1];
        st->rr_last = t[1];
        if (t[1] < st->rr_best) {
            st->rr_best = t[1];
            st->stalls = 0;
        } else {
            st->stalls = st->stalls + 1;
        }
        bool stop = (t[1] < st->tol && it >= 2) || it >= st->max_iter || st->stalls >= 2;
        if (stop) {
            st->active = 0;
        } else {
            double ro = st->rr_old;
            st->beta = t[0] / (ro > 1e-30 ? ro : 1e-30);
            st->rr_old = t[1];
        }
    } else if (OP == FIN_W18_UPDATE) {
        // weiler18.py:350-365 (t[0] = r.r, t[1] = r.z); the iteration's dv += alpha p follows in
        // the p pass, also when this iteration ends the solve
        st->dv_apply = 1;
        int it = st->iter + 1;
        st->iter = it;
        st->rr_last = t[0];
        bool stop = (t[0] < st->tol && it >= 2) || it >= st->max_iter;
        if (stop) {
            st->active = 0;
        } else {
            double ro = st->rr_old;
            st->beta = t[1] / (ro > 1e-12 ? ro : 1e-12);
            st->rr_old = t[1];
        }
    }
}

template <int OP>
constexpr int fin_nv() {
    return OP == FIN_PACK ? 3 : (OP == FIN_BUILD || OP == FIN_RHS || OP == FIN_W18_UPDATE || OP == FIN_CG_REPLACE) ? 2 : 1;
}

// One block: fold the partials of the previous kernel in a fixed order and advance the solver
// state.
template <int OP>
__global__ __launch_bounds__(FIN_BLOCK) void k_finalize(FinArgs a) {
    constexpr int NV = fin_nv<OP>();
    SolverState* st = a.st;
    // the p pass before this one applied the last iteration's dv update, if any (iterations
    // enqueued past the stop must not apply it again)
    if (OP == FIN_W18_PAP && threadIdx.x == 0) st->dv_apply = 0;
    if (OP == FIN_PACK && a.gated && (st->active != 0 || st->status != 0)) return;
    if (OP == FIN_CG_PAP || OP == FIN_CG_UPDATE || OP == FIN_W18_PAP || OP == FIN_W18_UPDATE || OP == FIN_CG_REPLACE) {
        if (st->active == 0) return;
    }
    double t[3] = {0.0, 0.0, 0.0};
    fold_partials<NV>(a.partials, a.stride, a.nblocks, t);
    if (threadIdx.x != 0) return;
    fin_apply<OP>(st, t, a.tol, a.max_iter);
}


template <int OP>
inline int finalize(ivp_ctx* c, int nblocks, double tol = 0.0, int max_iter = 0, bool gated = false) {
    FinArgs fa{c->partials, c->max_blocks, nblocks, c->state, tol, max_iter, gated ? 1 : 0};
    hipLaunchKernelGGL(k_finalize<OP>, dim3(1), dim3(FIN_BLOCK), 0, c->stream, fa);
    IVP_HIP(hipGetLastError());
    return IVP_OK;
}

// ---- independent scenes ----------------------------------------------------------------------

// Per-block partial sums split by scene.  A block's rows span at most two scenes (every scene
// holds at least MIN_SCENE_DYNAMIC >= rows of a block): seg = 0 for rows of the block's first
// scene, 1 for the second.  Value q of segment g goes to row 2 q + g, column blk (the block's
// row block, so the blocks of one scene are one contiguous range).  `two` is block-uniform.
template <int NV>
__device__ __forceinline__ void block_partials_seg(const double (&v)[NV], int seg, bool two, double* partials,
                                                   int stride, int blk) {
    __shared__ double lds[WAVES_PER_BLOCK];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s0 = block_sum(seg == 0 ? v[q] : 0.0, lds);
        if (threadIdx.x == 0) partials[(2 * q) * stride + blk] = s0;
        if (two) {
            double s1 = block_sum(seg == 1 ? v[q] : 0.0, lds);
            if (threadIdx.x == 0) partials[(2 * q + 1) * stride + blk] = s1;
        }
    }
}

struct FinScenesArgs {
    const double* partials;
    int stride;
    const int* sid;
    SceneState* ss;
    const SceneParams* prm;
    double* hist_err;     // [S * SCENE_HIST]
    int* hist_pcg;        // [S * SCENE_HIST]
    int S;
    int rpb;              // rows per block of the kernel that wrote the partials
};

// One wave per scene: fold the scene's block partials in a fixed order and advance its state
// with the single-scene logic (fin_apply), then its outer-loop phase (the host logic of
// ip.py:345-395 for one scene).  Scenes not in the phase this reduction belongs to are skipped.
template <int OP>
__global__ __launch_bounds__(BLOCK) void k_fin_scenes(FinScenesArgs a) {
    constexpr int NV = fin_nv<OP>();
    constexpr int want = OP == FIN_PACK ? PH_PACK : (OP == FIN_RHS || OP == FIN_CG_INIT) ? PH_RHS : PH_CG;
    const int s = blockIdx.x * WAVES_PER_BLOCK + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & (WAVE - 1);
    if (s >= a.S) return;
    SceneState* sc = a.ss + s;
    if (sc->phase != want) return;
    const int b0 = sc->row0 / a.rpb, b1 = (sc->row0 + sc->nd - 1) / a.rpb;
    double t[3] = {0.0, 0.0, 0.0};
    for (int b = b0 + lane; b <= b1; b += WAVE) {
        const int seg = a.sid[b * a.rpb] == s ? 0 : 1;
#pragma unroll
        for (int q = 0; q < NV; ++q) t[q] += a.partials[(2 * q + seg) * a.stride + b];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = wave_sum(t[q]);
    if (lane != 0) return;
    const SceneParams& p = *a.prm;
    double* herr = a.hist_err + (size_t)s * SCENE_HIST;
    int* hpcg = a.hist_pcg + (size_t)s * SCENE_HIST;
    if (OP == FIN_PACK) {
        fin_apply<OP>(&sc->cg, t, 0.0, 0);
        double err = sc->cg.err_sum / (double)sc->nd;
        if (sc->n_err < SCENE_HIST) herr[sc->n_err] = err;
        sc->n_err = sc->n_err + 1;
        bool converged = err < p.tol_opt && !p.fixed;
        if ((converged && sc->opt_iter >= 1) || sc->opt_iter >= p.max_opt) {
            sc->phase = PH_DONE;
            sc->final_err = err;
            sc->outer_iters = sc->opt_iter;
            if (!p.fixed && err >= p.tol_opt && sc->opt_iter >= p.max_opt) sc->status = IVP_NOT_CONVERGED;
        } else {
            sc->phase = PH_RHS;
        }
    } else if (OP == FIN_RHS) {
        fin_apply<OP>(&sc->cg, t, 0.0, 0);
    } else if (OP == FIN_CG_INIT) {
        fin_apply<OP>(&sc->cg, t, p.tol_pcg, p.max_pcg);
        if (sc->cg.active) {
            sc->phase = PH_CG;
        } else {
            sc->phase = PH_DUPDATE;
            if (sc->opt_iter < SCENE_HIST) hpcg[sc->opt_iter] = 0;
        }
    } else if (OP == FIN_CG_PAP) {
        fin_apply<OP>(&sc->cg, t, 0.0, 0);
        if (sc->cg.status == IVP_NOT_PD) {
            sc->status = IVP_NOT_PD;
            sc->phase = PH_DONE;
            sc->outer_iters = sc->opt_iter;
            sc->cg_total = sc->cg_total + sc->cg.iter;
            if (sc->opt_iter < SCENE_HIST) hpcg[sc->opt_iter] = sc->cg.iter;
        }
    } else if (OP == FIN_CG_UPDATE) {
        fin_apply<OP>(&sc->cg, t, 0.0, 0);
        if (!sc->cg.active) {
            sc->phase = PH_DUPDATE;
            sc->cg_total = sc->cg_total + sc->cg.iter;
            if (sc->opt_iter < SCENE_HIST) hpcg[sc->opt_iter] = sc->cg.iter;
        }
    }
}

}  // namespace ivp
