// ivp_api.hip -- the extern "C" entry points of libivp.so (declared in include/ivp.h):
// context lifecycle, device memory, the particle store interface and instrumentation.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "ivp_internal.h"
#include "ivp_launch.h"

using namespace ivp;

int ivp_hip_fail(hipError_t e, const char* what, const char* file, int line) {
    if (getenv("IVP_VERBOSE")) fprintf(stderr, "libivp: %s failed at %s:%d: %s\n", what, file, line, hipGetErrorString(e));
    return IVP_ERR_HIP;
}

namespace {

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    IVP_HIP(hipMalloc((void**)p, count * sizeof(T)));
    return IVP_OK;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <class T>
void hfree(T*& p) {
    if (p) (void)hipHostFree((void*)p);
    p = nullptr;
}

// Captured replay graphs hold the device addresses of the rows and scene buffers: drop them
// whenever those are freed.
void scene_graphs_clear(ivp_ctx* c) {
    bool any = false;
    for (auto& g : c->sgraphs) any = any || g.exec != nullptr;
    if (any) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->sgraphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = SceneGraph{};
    }
}

void scenes_free(ivp_ctx* c) {
    scene_graphs_clear(c);
    dfree(c->grids); dfree(c->sid); dfree(c->sdomain); dfree(c->ss); dfree(c->sprm);
    dfree(c->shist_err); dfree(c->shist_pcg); dfree(c->spartials); dfree(c->sactive);
    hfree(c->ss_host); hfree(c->sprm_host); hfree(c->sactive_host);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    c->cap_stream = nullptr;
    c->S = 0;
    c->scenes.clear();
    c->grids_host.clear();
}

void free_all(ivp_ctx* c) {
    scenes_free(c);
    dfree(c->pos); dfree(c->pos_b); dfree(c->vel); dfree(c->vel_b); dfree(c->minv); dfree(c->minv_b);
    dfree(c->rho0); dfree(c->rho0_b); dfree(c->mat); dfree(c->mat_b); dfree(c->uid); dfree(c->uid_b);
    dfree(c->inv); dfree(c->density); dfree(c->kk); dfree(c->Ss); dfree(c->divg); dfree(c->divg_b);
    dfree(c->rowcnt); dfree(c->rowtot); dfree(c->nb); dfree(c->tile_len); dfree(c->dstart); dfree(c->sstart);
    dfree(c->ccount); dfree(c->key); dfree(c->tmp_src); dfree(c->order); dfree(c->scan_bsum);
    dfree(c->d4); dfree(c->q4); dfree(c->b); dfree(c->sol); dfree(c->r); dfree(c->Ap); dfree(c->p4); dfree(c->y2);
    dfree(c->cc); dfree(c->sfac); dfree(c->jd); dfree(c->xs4); dfree(c->nbact);
    dfree(c->w_pinv); dfree(c->w_p4); dfree(c->w_v4); dfree(c->w_dv); dfree(c->w_r); dfree(c->w_z);
    dfree(c->w_Ap); dfree(c->w_rhs); dfree(c->w_fac); dfree(c->v4); dfree(c->w_dv_prev);
    dfree(c->st_gt); dfree(c->st_div); dfree(c->st_b);
    dfree(c->prj_v0); dfree(c->prj_lam); dfree(c->prj_prev); dfree(c->prj_df_prev); dfree(c->prj_hist);
    c->prj_hist_cap = 0;
    dfree(c->partials); dfree(c->state); dfree(c->err_flag); dfree(c->speed_max);
    if (c->state_host) (void)hipHostFree(c->state_host);
    if (c->err_host) (void)hipHostFree(c->err_host);
    if (c->count_host) (void)hipHostFree(c->count_host);

    c->state_host = nullptr;
    c->err_host = nullptr;
    c->count_host = nullptr;
    for (auto& e : c->prof_pending) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    c->prof_pending.clear();
    for (auto e : c->prof_free) (void)hipEventDestroy(e);
    c->prof_free.clear();
    for (auto e : c->iter_ev) (void)hipEventDestroy(e);
    c->iter_ev.clear();
    if (c->ip_end_ev) (void)hipEventDestroy(c->ip_end_ev);
    c->ip_end_ev = nullptr;
    if (c->wait_ev) (void)hipEventDestroy(c->wait_ev);
    c->wait_ev = nullptr;
    c->ip_wait_pending = false;
}

int alloc_all(ivp_ctx* c) {
    const size_t cap = (size_t)c->cap;
    const size_t tiles = (size_t)blocks_for(c->cap) * WAVES_PER_BLOCK;
    c->nd_alloc_tiles = (int)tiles;
    int rc = IVP_OK;
#define A(p, n) if ((rc = dalloc(&c->p, (n))) != IVP_OK) return rc
    A(pos, cap + POS_PAD); A(pos_b, cap + POS_PAD); A(vel, 3 * cap); A(vel_b, 3 * cap); A(minv, cap); A(minv_b, cap);
    A(rho0, cap); A(rho0_b, cap); A(mat, cap); A(mat_b, cap); A(uid, cap); A(uid_b, cap); A(inv, cap);
    A(density, cap); A(kk, cap); A(Ss, 3 * cap); A(divg, cap); A(divg_b, cap); A(rowcnt, cap); A(rowtot, cap);
    A(tile_len, tiles);
    A(dstart, (size_t)c->ncell + 1); A(sstart, (size_t)c->ncell + 1); A(ccount, (size_t)c->ncell);
    A(key, cap); A(tmp_src, cap + 1); A(order, cap);
    size_t scan_n = std::max((size_t)c->ncell, cap) + 1;
    c->nscan_blocks = (int)((scan_n + 2047) / 2048) + 2;
    A(scan_bsum, (size_t)c->nscan_blocks + 1);
    A(d4, cap); A(q4, cap); A(b, 3 * cap); A(sol, 3 * cap); A(r, 3 * cap); A(Ap, 3 * cap); A(p4, cap); A(y2, cap);
    A(cc, cap); A(sfac, cap); A(jd, cap); A(xs4, cap); A(nbact, cap);
    A(w_pinv, 9 * cap); A(w_p4, cap); A(w_v4, cap); A(w_dv, 3 * cap); A(w_r, 3 * cap); A(w_z, 3 * cap);
    A(w_Ap, 3 * cap); A(w_rhs, 3 * cap); A(w_fac, cap); A(v4, cap);
    c->max_blocks = blocks_for(c->cap) * MAX_PASS_LANES + 1;   // blocks of a pair pass at most
    A(partials, 4 * (size_t)c->max_blocks);
    A(state, 1); A(err_flag, 1); A(speed_max, 1);
#undef A
    IVP_HIP(hipHostMalloc((void**)&c->state_host, sizeof(SolverState), hipHostMallocDefault));
    IVP_HIP(hipHostMalloc((void**)&c->err_host, sizeof(int), hipHostMallocDefault));
    IVP_HIP(hipHostMalloc((void**)&c->count_host, 4 * sizeof(int), hipHostMallocDefault));

    memset(c->state_host, 0, sizeof(SolverState));
    IVP_HIP(hipMemsetAsync(c->ccount, 0, sizeof(int) * (size_t)c->ncell, c->stream));
    IVP_HIP(hipMemsetAsync(c->state, 0, sizeof(SolverState), c->stream));
    IVP_HIP(hipMemsetAsync(c->tile_len, 0, sizeof(int) * tiles, c->stream));
    // the pads past cap are read by the candidate scan (and masked); keep them defined
    IVP_HIP(hipMemsetAsync(c->pos, 0, sizeof(float4) * (cap + POS_PAD), c->stream));
    IVP_HIP(hipMemsetAsync(c->pos_b, 0, sizeof(float4) * (cap + POS_PAD), c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    return IVP_OK;
}

// Neighbour rows sized for the loaded dynamic count.
int ensure_rows(ivp_ctx* c) {
    size_t tiles = (size_t)blocks_for(std::max(c->nd, 1)) * WAVES_PER_BLOCK;
    static_assert(sizeof(uint32_t) == 4, "");
    size_t slots = tiles * (size_t)c->Cs * 64u;
    if (c->nb) return IVP_OK;
    IVP_TRY(dalloc(&c->nb, slots * (size_t)(c->rec_words / 64)));
    if (c->row_fmt == IVP_ROWS_F16) IVP_TRY(dalloc(&c->g16, 2 * slots));
    if (c->rec_words == 64) {
        // tile windows: 16 bits per slot, WIN_MAX ids and a count per tile
        IVP_TRY(dalloc(&c->loc, slots / 2));
        IVP_TRY(dalloc(&c->win_ids, tiles * (size_t)WIN_MAX));
        IVP_TRY(dalloc(&c->win_n, tiles));
    }
    c->win_built = false;
    return IVP_OK;
}

void free_rows(ivp_ctx* c) {
    scene_graphs_clear(c);
    dfree(c->nb);
    dfree(c->g16);
    dfree(c->loc);
    dfree(c->win_ids);
    dfree(c->win_n);
    c->win_built = false;
    dfree(c->st_gt);
}

}  // namespace

namespace ivp {

// Jeske23 buffers: g~ per row slot (sized like the ids of the current rows), divergence and the
// extra right-hand side per particle.
int st_buffers(ivp_ctx* c) {
    if (!c->st_div) IVP_TRY(dalloc(&c->st_div, (size_t)c->cap));
    if (!c->st_b) IVP_TRY(dalloc(&c->st_b, 3 * (size_t)c->cap));
    if (!c->st_gt) {
        size_t tiles = (size_t)blocks_for(std::max(c->nd, 1)) * WAVES_PER_BLOCK;
        IVP_TRY(dalloc(&c->st_gt, tiles * (size_t)c->Cs * 64u));
    }
    return IVP_OK;
}

// The Weiler18 correction kept for warm starts, in user order; zero until a solve writes it (the
// reference allocates dv with wp.zeros, weiler18.py:193).
int w18_buffers(ivp_ctx* c) {
    if (!c->w_dv_prev) {
        IVP_TRY(dalloc(&c->w_dv_prev, 3 * (size_t)c->cap));
        IVP_HIP(hipMemsetAsync(c->w_dv_prev, 0, 3 * (size_t)c->cap * sizeof(float), c->stream));
    }
    return IVP_OK;
}

// PRJ buffers.  The multipliers kept for warm starts start at zero (the reference allocates
// p_previous and dt2_lam_df_previous with wp.zeros, prj.py:163-164, pressure_base.py:302).
int prj_buffers(ivp_ctx* c, int passes) {
    const size_t cap = (size_t)c->cap;
    if (!c->prj_v0) IVP_TRY(dalloc(&c->prj_v0, 3 * cap));
    if (!c->prj_lam) IVP_TRY(dalloc(&c->prj_lam, cap));
    if (!c->prj_prev) {
        IVP_TRY(dalloc(&c->prj_prev, cap));
        IVP_HIP(hipMemsetAsync(c->prj_prev, 0, cap * sizeof(float), c->stream));
    }
    if (!c->prj_df_prev) {
        IVP_TRY(dalloc(&c->prj_df_prev, cap));
        IVP_HIP(hipMemsetAsync(c->prj_df_prev, 0, cap * sizeof(float), c->stream));
    }
    if (passes > c->prj_hist_cap) {
        IVP_HIP(hipStreamSynchronize(c->stream));
        dfree(c->prj_hist);
        c->prj_hist_cap = 0;
        IVP_TRY(dalloc(&c->prj_hist, (size_t)passes));
        c->prj_hist_cap = passes;
    }
    return IVP_OK;
}

}  // namespace ivp

namespace {

// Independent scenes: one cell grid per scene (same cell width and margins as the single-scene
// grid), keys consecutive in scene order; cell tables and scene buffers reallocated.
int scenes_set(ivp_ctx* c, int S, const ivp_scene* sc) {
    if (S <= 0 || !sc) return IVP_ERR_ARG;
    long long total = 0;
    for (int s = 0; s < S; ++s) {
        if (sc[s].n_particles <= 0) return IVP_ERR_ARG;
        total += sc[s].n_particles;
    }
    if (total > c->cap) return IVP_ERR_CAPACITY;
    IVP_HIP(hipStreamSynchronize(c->stream));
    const float w = c->h * 1.0009765625f;
    std::vector<Grid> g((size_t)S);
    long long base = 0;
    for (int s = 0; s < S; ++s) {
        g[s].inv_w = c->grid.inv_w;
        g[s].base = (int)base;
        long long cells_s = 1;
        for (int a = 0; a < 3; ++a) {
            g[s].origin[a] = sc[s].domain_start[a] - 2.0f * w;
            float cells = ceilf(sc[s].domain_size[a] / w);
            if (!(cells >= 0.0f) || cells > 2044.0f) return IVP_ERR_GRID;
            g[s].dims[a] = (int)cells + 4;
            cells_s *= g[s].dims[a];
        }
        base += cells_s;
        if (base > (1ll << 28)) return IVP_ERR_GRID;
    }
    scenes_free(c);
    free_rows(c);
    c->loaded = false;
    c->built = false;
    const size_t cap = (size_t)c->cap;
    dfree(c->dstart); dfree(c->sstart); dfree(c->ccount); dfree(c->scan_bsum);
    c->ncell = (int)base;
    size_t scan_n = std::max((size_t)c->ncell, cap) + 1;
    c->nscan_blocks = (int)((scan_n + 2047) / 2048) + 2;
    int rc = IVP_OK;
#define A(p, n) if ((rc = dalloc(&c->p, (n))) != IVP_OK) return rc
    A(dstart, (size_t)c->ncell + 1); A(sstart, (size_t)c->ncell + 1); A(ccount, (size_t)c->ncell);
    A(scan_bsum, (size_t)c->nscan_blocks + 1);
    A(grids, (size_t)S); A(sid, cap); A(sdomain, 3 * (size_t)S); A(ss, (size_t)S); A(sprm, 1);
    A(shist_err, (size_t)S * SCENE_HIST); A(shist_pcg, (size_t)S * SCENE_HIST);
    A(spartials, 8 * (size_t)c->max_blocks); A(sactive, 1);
#undef A
    IVP_HIP(hipHostMalloc((void**)&c->ss_host, sizeof(SceneState) * (size_t)S, hipHostMallocDefault));
    IVP_HIP(hipHostMalloc((void**)&c->sprm_host, sizeof(SceneParams), hipHostMallocDefault));
    IVP_HIP(hipHostMalloc((void**)&c->sactive_host, sizeof(int), hipHostMallocDefault));
    memset(c->ss_host, 0, sizeof(SceneState) * (size_t)S);
    std::vector<float> dom(3 * (size_t)S);
    for (int s = 0; s < S; ++s)
        for (int a = 0; a < 3; ++a) dom[3 * s + a] = sc[s].domain_size[a];
    IVP_HIP(hipMemsetAsync(c->ccount, 0, sizeof(int) * (size_t)c->ncell, c->stream));
    IVP_HIP(hipMemsetAsync(c->shist_err, 0, sizeof(double) * (size_t)S * SCENE_HIST, c->stream));
    IVP_HIP(hipMemsetAsync(c->shist_pcg, 0, sizeof(int) * (size_t)S * SCENE_HIST, c->stream));
    IVP_HIP(hipMemsetAsync(c->ss, 0, sizeof(SceneState) * (size_t)S, c->stream));
    IVP_HIP(hipMemcpyAsync(c->grids, g.data(), sizeof(Grid) * (size_t)S, hipMemcpyHostToDevice, c->stream));
    IVP_HIP(hipMemcpyAsync(c->sdomain, dom.data(), sizeof(float) * dom.size(), hipMemcpyHostToDevice, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    c->S = S;
    c->scenes.assign(sc, sc + S);
    c->grids_host = g;
    c->row_fmt = IVP_ROWS_IDS;
    c->rec_words = 64;
    return IVP_OK;
}

}  // namespace

namespace ivp {

// Event timing of selected kernel classes.  Launches tagged with a CG iteration index are kept
// only if that iteration ran (prof_collect), so trailing no-op launches of the last CG chunk are
// not counted.
int prof_take(ivp_ctx* c, int kclass, int tag, double bytes, ProfEvent* ev) {
    ev->start = nullptr;
    auto take = [&](hipEvent_t* e) -> int {
        if (!c->prof_free.empty()) {
            *e = c->prof_free.back();
            c->prof_free.pop_back();
            return IVP_OK;
        }
        // timing-only events: no system-scope fence (cache write-back) around the timed kernel;
        // they are read only after the stream has been synchronised
        IVP_HIP(hipEventCreateWithFlags(e, hipEventDisableSystemFence));
        return IVP_OK;
    };
    if (take(&ev->start) != IVP_OK || take(&ev->stop) != IVP_OK) return IVP_ERR_HIP;
    ev->kclass = kclass;
    ev->tag = tag;
    ev->epoch = c->prof_epoch;
    ev->bytes = bytes;
    return IVP_OK;
}

int prof_push(ivp_ctx* c, ProfEvent* ev) {
    c->prof_pending.push_back(*ev);
    if (c->prof_pending.size() > 8192) {
        IVP_HIP(hipStreamSynchronize(c->stream));
        return prof_collect(c, -1, -1);
    }
    return IVP_OK;
}

// Fold finished events.  Events of `epoch` with tag >= cg_iters_done are dropped.
// Set the byte figure of the pending launches of one class in one solve (for byte models that
// depend on counts the device reports only after the launches were enqueued).
void prof_set_bytes(ivp_ctx* c, int epoch, int kclass, double bytes) {
    for (auto& e : c->prof_pending)
        if (e.epoch == epoch && e.kclass == kclass) e.bytes = bytes;
}

int iter_mark(ivp_ctx* c, int k) {
    while ((int)c->iter_ev.size() <= k) {
        hipEvent_t e = nullptr;
        // read after the stream is synchronised: no system-scope fence needed
        IVP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        c->iter_ev.push_back(e);
    }
    IVP_HIP(hipEventRecord(c->iter_ev[(size_t)k], c->stream));
    return IVP_OK;
}

void iter_costs(ivp_ctx* c, int n, ivp_ip_stats* st) {
    st->n_iter_costs = n;
    if (!st->iter_costs) return;
    for (int k = 0; k < n && k < st->capacity && k + 1 < (int)c->iter_ev.size(); ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->iter_ev[(size_t)k], c->iter_ev[(size_t)k + 1]) != hipSuccess) ms = 0.0f;
        st->iter_costs[k] = (double)ms;
    }
}

int prof_collect(ivp_ctx* c, int cg_iters_done, int epoch) {
    std::vector<ProfEvent> keep;
    for (auto& e : c->prof_pending) {
        if (hipEventQuery(e.stop) != hipSuccess) {
            keep.push_back(e);
            continue;
        }
        bool counted = e.tag < 0 || cg_iters_done < 0 || e.epoch != epoch || e.tag < cg_iters_done;
        if (counted) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e.start, e.stop) == hipSuccess) {
                c->prof_launches[e.kclass]++;
                c->prof_ms[e.kclass] += ms;
                c->prof_bytes[e.kclass] += e.bytes;
            }
        }
        c->prof_free.push_back(e.start);
        c->prof_free.push_back(e.stop);
    }
    c->prof_pending.swap(keep);
    return IVP_OK;
}

}  // namespace ivp

static int build_code(int err) {
    if (err & 4) return IVP_ERR_INTERNAL;
    return (err & 1) ? IVP_NEIGHBOR_OVERFLOW : IVP_OK;
}

// Status of the neighbour build last enqueued, read with one synchronisation: its error flags
// (and the solver state, whose pair count the byte figures of the following passes use).
static int build_status(ivp_ctx* c) {
    IVP_HIP(hipMemcpyAsync(c->err_host, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipMemcpyAsync(c->state_host, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    return build_code(*c->err_host);
}

static const char* kclass_name(int k) {
    static const char* names[KC_COUNT] = {"sort", "build", "ip_pack", "ip_rhs", "cg_J", "cg_JT", "cg_update", "cg_p",
                                          "visc", "w18_Av", "w18_update", "w18_p", "w18_pre", "cg_refresh", "st_setup", "st_Av", "prj_head", "prj_tail",
                                          "glue", "window", "scan"};
    return (k >= 0 && k < KC_COUNT) ? names[k] : "?";
}

extern "C" {

const char* ivp_version(void) { return "ivp 0.1 (gfx950)"; }

const char* ivp_status_string(int s) {
    switch (s) {
        case IVP_OK: return "ok";
        case IVP_NOT_CONVERGED: return "IP outer loop did not reach densityErrorTol within maxOuterIters";
        case IVP_NOT_PD: return "non-positive definite matrix (p.Ap < 0)";
        case IVP_NEIGHBOR_OVERFLOW: return "a particle has more neighbours than neighborCacheSize";
        case IVP_ERR_ARG: return "invalid argument";
        case IVP_ERR_HIP: return "HIP runtime error";
        case IVP_ERR_CAPACITY: return "capacity exceeded";
        case IVP_ERR_STATE: return "call out of order (particles not set or neighbours not built)";
        case IVP_ERR_GRID: return "grid too large (more than 2048 cells per axis or 2^28 cells)";
        case IVP_ERR_INTERNAL: return "internal consistency check failed in a kernel";
        default: return "unknown status";
    }
}

int ivp_create(const ivp_config* cfg, int max_particles, void* stream, ivp_ctx** out) {
    if (!cfg || !out || max_particles < 0) return IVP_ERR_ARG;
    *out = nullptr;
    if (!(cfg->support_radius > 0.0f) || cfg->cache_size <= 0) return IVP_ERR_ARG;
    ivp_ctx* c = new ivp_ctx();
    c->cfg = *cfg;
    c->stream = (hipStream_t)stream;
    c->h = cfg->support_radius;
    c->C = cfg->cache_size;
    c->Cs = (c->C + ROW_PAD - 1) / ROW_PAD * ROW_PAD;
    c->cap = std::max(max_particles, 1);
    // uniform grid: cell width w = h (1 + 2^-10), origin two cells below the domain start,
    // ceil(L / w) + 4 cells per axis
    float w = c->h * 1.0009765625f;
    c->grid.inv_w = 1.0f / w;
    {
        float h = c->h;
        c->kspiky = spiky_constant(h);
        c->kcubic = cubic_constant(h);
    }
    c->row_fmt = IVP_ROWS_IDS;
    c->rec_words = 64;
    {
        // power-of-two scale at or above the largest |gradW| = 45/(pi h^6) h^2: float16 records
        // then hold gradW / scale in [-1, 1] and decode by an exact multiplication
        int e = 0;
        double gmax = (double)c->kspiky * (double)c->h * (double)c->h;
        double m = std::frexp(gmax, &e);   // gmax = m 2^e, m in [0.5, 1)
        c->gscale = (float)std::ldexp(1.0, m == 0.5 ? e - 1 : e);
    }
    long long ncell = 1;
    for (int a = 0; a < 3; ++a) {
        c->grid.origin[a] = cfg->domain_start[a] - 2.0f * w;
        float cells = ceilf(cfg->domain_size[a] / w);
        if (!(cells >= 0.0f) || cells > 2044.0f) {
            delete c;
            return IVP_ERR_GRID;
        }
        c->grid.dims[a] = (int)cells + 4;
        ncell *= c->grid.dims[a];
    }
    if (ncell > (1ll << 28)) {
        delete c;
        return IVP_ERR_GRID;
    }
    c->ncell = (int)ncell;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return IVP_ERR_HIP;
    }
    int rc = alloc_all(c);
    if (rc != IVP_OK) {
        free_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return IVP_OK;
}

int ivp_destroy(ivp_ctx* c) {
    if (!c) return IVP_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    free_rows(c);
    free_all(c);
    delete c;
    return IVP_OK;
}

int ivp_set_particles(ivp_ctx* c, int n, const float* x, const float* v, const float* m, const float* m_inv,
                      const float* density0, const int* is_dynamic, const int* material) {
    if (!c) return IVP_ERR_ARG;
    free_rows(c);
    c->built = false;   // the rows describe the previous particle set
    int rc = grid_load(c, n, x, v, m, m_inv, density0, is_dynamic, material);
    if (rc != IVP_OK) return rc;
    if (c->S > 0) {
        // the dynamic rows of every scene, for the per-scene reductions
        for (int s = 0; s < c->S; ++s) {
            c->ss_host[s].row0 = c->scene_row0[s];
            c->ss_host[s].nd = c->scene_row0[s + 1] - c->scene_row0[s];
            c->ss_host[s].phase = PH_DONE;
        }
        IVP_HIP(hipMemcpyAsync(c->ss, c->ss_host, sizeof(SceneState) * (size_t)c->S, hipMemcpyHostToDevice,
                               c->stream));
        IVP_HIP(hipStreamSynchronize(c->stream));
    }
    return ensure_rows(c);
}

int ivp_reset(ivp_ctx* c) {
    if (!c) return IVP_ERR_ARG;
    if (!c->loaded) return IVP_ERR_STATE;
    if (c->n > 0) {
        const size_t n = (size_t)c->n;
        IVP_HIP(hipMemcpyAsync(c->density, c->rho0, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        IVP_HIP(hipMemsetAsync(c->kk, 0, n * sizeof(float), c->stream));
        IVP_HIP(hipMemsetAsync(c->divg, 0, n * sizeof(float), c->stream));
        IVP_HIP(hipMemsetAsync(c->divg_b, 0, n * sizeof(float), c->stream));
    }
    const size_t cap = (size_t)c->cap;
    if (c->prj_prev) IVP_HIP(hipMemsetAsync(c->prj_prev, 0, cap * sizeof(float), c->stream));
    if (c->prj_df_prev) IVP_HIP(hipMemsetAsync(c->prj_df_prev, 0, cap * sizeof(float), c->stream));
    // Weiler18.reset (weiler18.py:213-220), called by Framework.reset: the kept correction
    if (c->w_dv_prev) IVP_HIP(hipMemsetAsync(c->w_dv_prev, 0, 3 * cap * sizeof(float), c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    c->built = false;
    return IVP_OK;
}

int ivp_set_option(ivp_ctx* c, int option, int value) {
    if (!c) return IVP_ERR_ARG;
    if (option == IVP_OPT_SCENE_CG_CHUNK) {
        if (value < 1 || value > 64) return IVP_ERR_ARG;
        c->scene_chunk = value;
        return IVP_OK;
    }
    if (option == IVP_OPT_SCENE_GRAPH) {
        c->scene_graph = value != 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_W18_WARM_START) {
        c->w18_warm = value != 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_W18_PRECONDITIONER) {
        c->w18_precond = value != 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_TILE_WINDOWS) {
        if (value < 0) return IVP_ERR_ARG;
        c->windows = value != 0;
        c->win_cap = value > 1 ? std::min(value, WIN_MAX) : WIN_MAX;
        return IVP_OK;
    }
    if (option == IVP_OPT_BUILD_LANES) {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 64)
            return IVP_ERR_ARG;
        c->build_lanes = value;
        return IVP_OK;
    }
    if (option == IVP_OPT_SCENE_LANES) {
        if (value != 1 && value != 2 && value != 4) return IVP_ERR_ARG;
        c->scene_lanes = value;
        return IVP_OK;
    }
    if (option == IVP_OPT_GRADIENT_CACHE) {
        if (value != IVP_ROWS_IDS && value != IVP_ROWS_F32 && value != IVP_ROWS_F16) return IVP_ERR_ARG;
        if (value == c->row_fmt) return IVP_OK;
        if (c->S > 0) return IVP_ERR_ARG;   // independent scenes use ids-only rows
        IVP_HIP(hipStreamSynchronize(c->stream));
        free_rows(c);
        c->row_fmt = value;
        c->rec_words = value == IVP_ROWS_F32 ? 256 : 64;
        c->built = false;
        return c->loaded ? ensure_rows(c) : IVP_OK;
    }
    return IVP_ERR_ARG;
}

int ivp_get_option(ivp_ctx* c, int option, int* value) {
    if (!c || !value) return IVP_ERR_ARG;
    if (option == IVP_OPT_GRADIENT_CACHE) {
        *value = c->row_fmt;
        return IVP_OK;
    }
    if (option == IVP_OPT_SCENE_CG_CHUNK) {
        *value = c->scene_chunk;
        return IVP_OK;
    }
    if (option == IVP_OPT_SCENE_GRAPH) {
        *value = c->scene_graph ? 1 : 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_SCENE_LANES) {
        *value = c->scene_lanes;
        return IVP_OK;
    }
    if (option == IVP_OPT_BUILD_LANES) {
        *value = c->build_lanes;
        return IVP_OK;
    }
    if (option == IVP_OPT_W18_WARM_START) {
        *value = c->w18_warm ? 1 : 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_W18_PRECONDITIONER) {
        *value = c->w18_precond ? 1 : 0;
        return IVP_OK;
    }
    if (option == IVP_OPT_TILE_WINDOWS) {
        *value = !c->windows ? 0 : c->win_cap < WIN_MAX ? c->win_cap : 1;
        return IVP_OK;
    }
    return IVP_ERR_ARG;
}

int ivp_write(ivp_ctx* c, int field, const void* src) {
    if (!c) return IVP_ERR_ARG;
    return store_write(c, field, src);
}

int ivp_read(ivp_ctx* c, int field, void* dst) {
    if (!c) return IVP_ERR_ARG;
    return store_read(c, field, dst);
}

int ivp_neighbors_build(ivp_ctx* c) {
    if (!c) return IVP_ERR_ARG;
    int rc = grid_build(c);
    if (rc != IVP_OK) return rc;
    return build_status(c);
}

int ivp_neighbors_export(ivp_ctx* c, int cache_size, int* count, int* ids, float* gradW) {
    if (!c) return IVP_ERR_ARG;
    return grid_export(c, cache_size, count, ids, gradW);
}

int ivp_ip_solve(ivp_ctx* c, float dt, const ivp_ip_params* prm, ivp_ip_stats* stats) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    return ip_solve(c, dt, prm, stats);
}

int ivp_viscosity_explicit(ivp_ctx* c, float coeff, float dt) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    // rows that the last rebuild truncated are refused before v changes
    if (c->built) IVP_TRY(build_status(c));
    return visc_explicit(c, coeff, dt);
}

int ivp_viscosity_weiler18(ivp_ctx* c, float coeff, float dt, ivp_cg_stats* stats) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    // rows that the last rebuild truncated are refused before v changes
    if (c->built) IVP_TRY(build_status(c));
    return visc_weiler18(c, coeff, dt, stats);
}

int ivp_prj_solve(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* stats) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    return prj_solve(c, dt, prm, stats, PRJ_CD);
}

int ivp_prj_solve_df(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* stats) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    return prj_solve(c, dt, prm, stats, PRJ_DF);
}

int ivp_surface_tension_jeske23(ivp_ctx* c, float sigma, float dt, ivp_cg_stats* stats) {
    if (!c) return IVP_ERR_ARG;
    if (c->S > 0) return IVP_ERR_STATE;
    // rows that the last rebuild truncated are refused before v changes
    if (c->built) IVP_TRY(build_status(c));
    return st_jeske23(c, sigma, dt, stats);
}

int ivp_apply_gravity(ivp_ctx* c, const float gravity[3], float dt) {
    if (!c || !gravity) return IVP_ERR_ARG;
    return glue_gravity(c, gravity, dt);
}

int ivp_advect_boundary(ivp_ctx* c, float dt) {
    if (!c) return IVP_ERR_ARG;
    return glue_advect_boundary(c, dt);
}

int ivp_step(ivp_ctx* c, const ivp_step_params* prm, ivp_ip_stats* stats, ivp_cg_stats* visc, ivp_cg_stats* st) {
    if (!c || !prm || !stats) return IVP_ERR_ARG;
    if (!c->built || c->S > 0) return IVP_ERR_STATE;
    if (prm->surface_tension_coeff != 0.0f && prm->surface_tension_method != 1) return IVP_ERR_ARG;
    if (prm->pressure_method != 0 && prm->pressure_method != 1) return IVP_ERR_ARG;
    // gravity writes the step's velocities into the second buffer; a stage that finds the rows
    // truncated by the last rebuild (the pressure solve's first synchronisation reads its flags)
    // swaps the first one back, so such a step leaves v and x as they were
    int rc = glue_gravity(c, prm->gravity, prm->dt);
    if (rc != IVP_OK) return rc;
    auto truncated = [&](int r) {
        if (r == IVP_NEIGHBOR_OVERFLOW || r == IVP_ERR_INTERNAL) std::swap(c->vel, c->vel_b);
        return r;
    };
    if (visc) { visc->cg_iters = 0; visc->residual_norm2 = 0.0; }
    if (st) { st->cg_iters = 0; st->residual_norm2 = 0.0; }
    // compute_non_pressure_forces order (framework.py:400-421): gravity, surface tension, viscosity
    if (prm->surface_tension_coeff != 0.0f) {
        rc = st_jeske23(c, prm->surface_tension_coeff, prm->dt, st);
        if (rc != IVP_OK) return rc;
    }
    // coefficient 0 makes either viscosity model an exact no-op (the rhs v - A v is 0, and the
    // explicit force is scaled by the coefficient)
    if (prm->viscosity_coeff != 0.0f) {
        if (prm->viscosity_method == 1)
            rc = visc_weiler18(c, prm->viscosity_coeff, prm->dt, visc);
        else
            rc = visc_explicit(c, prm->viscosity_coeff, prm->dt);
        if (rc != IVP_OK) return rc;
    }
    // pressure solve (framework.py:481-487; PRJ's divergence-free pass runs inside prj_solve).
    // The IP solve's closing wait (its wall time and stats_iter_costs) comes after the rest of
    // the step is enqueued, so the device does not idle while the host reads the clock.
    rc = prm->pressure_method == 0 ? prj_solve(c, prm->dt, &prm->prj, stats, PRJ_BOTH)
                                   : ip_solve(c, prm->dt, &prm->ip, stats, true);
    if (rc != IVP_OK) return truncated(rc);
    rc = glue_advect_boundary(c, prm->dt);
    // the rebuild that ends the step, not waited for: the step's results come from the rows built
    // before it, and a row of the new build over the capacity stops the next step (above)
    if (rc == IVP_OK) rc = grid_build(c);
    const int rw = ip_solve_wait(c, stats);
    return rc != IVP_OK ? rc : rw;
}

int ivp_set_scenes(ivp_ctx* c, int n_scenes, const ivp_scene* scenes) {
    if (!c) return IVP_ERR_ARG;
    return scenes_set(c, n_scenes, scenes);
}

int ivp_step_scenes(ivp_ctx* c, const ivp_step_params* prm, ivp_scene_stats* stats) {
    if (!c || !prm || !stats) return IVP_ERR_ARG;
    if (!c->built || c->S == 0) return IVP_ERR_STATE;
    // viscosity and surface tension are not part of the scene step, which solves with IP
    if (prm->viscosity_coeff != 0.0f || prm->surface_tension_coeff != 0.0f) return IVP_ERR_ARG;
    if (prm->pressure_method != 1) return IVP_ERR_ARG;
    int rc = glue_gravity(c, prm->gravity, prm->dt);
    if (rc != IVP_OK) return rc;
    rc = ip_solve_scenes(c, prm->dt, &prm->ip);
    if (rc == IVP_NEIGHBOR_OVERFLOW || rc == IVP_ERR_INTERNAL) std::swap(c->vel, c->vel_b);   // v as it was
    if (rc != IVP_OK) return rc;
    int status = IVP_OK;
    for (int s = 0; s < c->S; ++s) {
        const SceneState& q = c->ss_host[s];
        stats[s].status = q.status;
        stats[s].outer_iters = q.outer_iters;
        stats[s].cg_iters = q.cg_total;
        stats[s].final_density_error = q.final_err;
        if (status == IVP_OK && q.status != IVP_OK) status = q.status;
    }
    rc = glue_advect_boundary(c, prm->dt);
    if (rc != IVP_OK) return rc;
    rc = grid_build(c);
    if (rc != IVP_OK) return rc;
    return status;
}

int ivp_scene_history(ivp_ctx* c, int scene, int capacity, double* density_errors, int* pcg_iters, int* n_errors,
                      int* n_pcg) {
    if (!c) return IVP_ERR_ARG;
    if (c->S == 0) return IVP_ERR_STATE;
    if (scene < 0 || scene >= c->S || capacity < 0) return IVP_ERR_ARG;
    const SceneState& q = c->ss_host[scene];
    int ne = q.n_err, np = q.outer_iters;
    if (n_errors) *n_errors = ne;
    if (n_pcg) *n_pcg = np;
    int ce = std::min(std::min(ne, capacity), SCENE_HIST), cp = std::min(std::min(np, capacity), SCENE_HIST);
    const size_t off = (size_t)scene * SCENE_HIST;
    if (density_errors && ce > 0)
        IVP_HIP(hipMemcpyAsync(density_errors, c->shist_err + off, sizeof(double) * ce, hipMemcpyDeviceToHost,
                               c->stream));
    if (pcg_iters && cp > 0)
        IVP_HIP(hipMemcpyAsync(pcg_iters, c->shist_pcg + off, sizeof(int) * cp, hipMemcpyDeviceToHost, c->stream));
    IVP_HIP(hipStreamSynchronize(c->stream));
    return IVP_OK;
}

int ivp_profile_enable(ivp_ctx* c, int enable) {
    if (!c) return IVP_ERR_ARG;
    c->prof = enable != 0;
    c->prof_mask = (unsigned)enable;
    return IVP_OK;
}

int ivp_profile_reset(ivp_ctx* c) {
    if (!c) return IVP_ERR_ARG;
    IVP_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c, -1, -1);
    for (int k = 0; k < KC_COUNT; ++k) {
        c->prof_launches[k] = 0;
        c->prof_ms[k] = 0.0;
        c->prof_bytes[k] = 0.0;
    }
    return IVP_OK;
}

int ivp_profile_read(ivp_ctx* c, ivp_kernel_timing* out, int capacity) {
    if (!c || (!out && capacity > 0)) return IVP_ERR_ARG;
    IVP_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c, -1, -1);
    int n = 0;
    for (int k = 0; k < KC_COUNT && n < capacity; ++k, ++n) {
        memset(out[n].name, 0, sizeof(out[n].name));
        strncpy(out[n].name, kclass_name(k), sizeof(out[n].name) - 1);
        out[n].launches = c->prof_launches[k];
        out[n].total_ms = c->prof_ms[k];
        out[n].bytes = c->prof_bytes[k];
    }
    return n;
}

int ivp_counts(ivp_ctx* c, int* n_total, int* n_dynamic, long long* dyn_pairs) {
    if (!c) return IVP_ERR_ARG;
    if (n_total) *n_total = c->n;
    if (n_dynamic) *n_dynamic = c->nd;
    if (dyn_pairs) {
        *dyn_pairs = 0;
        if (c->built && c->nd > 0) {
            IVP_HIP(hipMemcpyAsync(c->state_host, c->state, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
            IVP_HIP(hipStreamSynchronize(c->stream));
            *dyn_pairs = c->state_host->total_pairs;
        }
    }
    return IVP_OK;
}

int ivp_tile_windows(ivp_ctx* c, int* tiles, int* windowed, long long* window_ids) {
    if (!c) return IVP_ERR_ARG;
    int t = 0, w = 0;
    long long ids = 0;
    if (c->win_built && c->nd > 0) {
        t = (c->nd + WAVE - 1) / WAVE;
        std::vector<int> n((size_t)t);
        IVP_HIP(hipMemcpyAsync(n.data(), c->win_n, sizeof(int) * (size_t)t, hipMemcpyDeviceToHost, c->stream));
        IVP_HIP(hipStreamSynchronize(c->stream));
        for (int k : n)
            if (k >= 0) {
                w++;
                ids += k;
            }
    }
    if (tiles) *tiles = t;
    if (windowed) *windowed = w;
    if (window_ids) *window_ids = ids;
    return IVP_OK;
}

}  // extern "C"
