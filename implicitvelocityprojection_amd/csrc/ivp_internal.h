// ivp_internal.h -- the library context and the launch functions each kernel file provides.
//
// Particle store ("internal order"): dynamic particles occupy [0, nd) sorted by
// (cell key, user id) -- re-sorted at every neighbour rebuild -- and static particles occupy
// [nd, n), sorted once by ivp_set_particles (they never move: framework.py:130-133 zeroes
// their velocity and advects them by 0).  uid[s] is the user index of internal particle s and
// inv[u] its inverse.
//
// Neighbour rows exist for dynamic particles only and hold their DYNAMIC neighbours; the
// static neighbours of a row enter the operators only through S_i = sum_static m_j gradW_ij,
// density and Aii, all accumulated during the build.  Rows are stored in wave tiles: tile t =
// rows [64t, 64t+64), Cs slots each (Cs = cache size rounded up to ROW_PAD), lane = row - 64t.
// The word of (tile t, slot k, row lane) is t * Cs * REC + row_word0<REC>(lane) + slot_word<REC>(k)
// (ivp_device.h), REC = words per slot of a tile:
//   * ids only (REC = 64): 8-slot blocks of 512 words; block b holds, for row `lane`, the words
//     512 b + 8 lane + [slot 0, 4, 1, 5, 2, 6, 3, 7 of the block], so a pass lane walking slots
//     g, g + 4, ... (four lanes per row) reads two ids with one 8-byte load and a wave reads 512
//     contiguous bytes;
//   * cached gradW (REC = 256 float32: id, gx, gy, gz; REC = 192 float16: id, (gx, gy), (gz, 0)):
//     slot-major records, word k REC + c 64 + lane for component c, one contiguous record per
//     (tile, slot) read by one wave.
// The tile length is the tile's longest row rounded up to ROW_PAD; slots past a row's count hold
// (id = the row itself, gradW = 0): r_ii = 0 gives gradW = 0 either way, which adds exactly 0 to
// every operator, so the hot loops run a wave-uniform trip count.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/ivp.h"
#include "ivp_device.h"

#define IVP_HIP(call)                                   \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return ivp_hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

int ivp_hip_fail(hipError_t e, const char* what, const char* file, int line);

namespace ivp {

// Device-resident solver state.  Reducing kernels write per-block partials; a one-block
// finalize kernel folds them in a fixed order and updates these fields.
struct SolverState {
    double rr_old;      // r.r (IP) / r.z (Weiler18) of the current iteration
    double rr_new;
    double pAp;
    double alpha;
    double beta;
    double err_sum;     // sum of max(f / rho0, 0) over dynamic particles
    double tol;         // CG stop threshold on r.r
    double rr_last;     // r.r at the last completed iteration
    double rr0;         // r.r before the first iteration
    double rr_best;     // smallest true r.r so far (r0.r0, then residual replacements)
    int stalls;         // consecutive residual replacements that did not lower rr_best
    int iter;           // CG iterations completed
    int max_iter;
    int active;         // 1 while CG continues
    int dv_apply;       // Weiler18 / Jeske23: this iteration's dv += alpha p is left to the p pass
    int status;         // 0 ok, IVP_NOT_PD
    int n_active_rows;  // rows with dfdt > 0 (per outer iteration)
    int n_touched_rows; // rows whose J^T loop runs (active or with an active neighbour)
    long long touched_pairs;
    long long active_pairs;
    long long total_pairs;  // stored (dynamic-dynamic) neighbour pairs of the last build
    long long win_entries;  // neighbour ids in the tile windows of the last build
};

// Independent scenes in one context.  Each scene has its own cell grid (keys of scene s follow
// those of scene s-1), so the cell sort keeps every scene's particles contiguous: dynamic
// particles of scene s at internal [row0, row0 + nd), static ones in the same scene order after
// all dynamic particles.  Each scene runs its own IP outer loop and CG, advanced by a phase:
enum ScenePhase { PH_PACK = 0, PH_RHS = 1, PH_CG = 2, PH_DUPDATE = 3, PH_DONE = 4 };

constexpr int MIN_SCENE_DYNAMIC = 256;   // a block (at most 256 rows) then spans at most two scenes
constexpr int MAX_PASS_LANES = 4;        // lanes per row of a pair pass, at most
constexpr int SCENE_HIST = 1026;         // per-scene outer-loop history entries kept

struct SceneState {
    SolverState cg;      // the scene's CG and error sums (single-scene semantics)
    int phase;
    int opt_iter;        // outer iterations done
    int outer_iters;     // final count (stats_outer_iters)
    int status;          // IVP_OK, IVP_NOT_CONVERGED, IVP_NOT_PD
    int row0, nd;        // dynamic rows of the scene
    int cg_total;        // CG iterations over the solve
    int n_err;           // density errors recorded
    double final_err;
};

// An instantiated replay graph and what its capture depended on
struct SceneGraph {
    hipGraphExec_t exec = nullptr;
    const void* pos = nullptr;   // particle store buffer read by the captured launches
    int chunk = 0;
    float omega = 0.0f;
    int nd = 0;
    int lanes = 0;
    bool win = false;            // the captured passes read the tile windows
};

// IP parameters of the current batched solve (device copy, read by the finalize kernels)
struct SceneParams {
    double tol_opt, tol_pcg;
    int max_opt, max_pcg, fixed;
};

constexpr int ROW_PAD = 8;        // tile lengths are padded to a multiple of this many slots
// slot_word's 8-slot blocks and the paired (g, g + 4) id loads of the pair passes assume both
static_assert(ROW_PAD % 8 == 0, "ids-only rows are stored in 8-slot blocks");
constexpr int POS_PAD = 16;       // float4 entries past the last particle of the position buffers: a
                                  // candidate batch of the scan (at most 16 entries) may read past its run

enum KernelClass {
    KC_SORT = 0,
    KC_BUILD,
    KC_IP_PACK,
    KC_IP_RHS,
    KC_CG_J,
    KC_CG_JT,
    KC_CG_UPDATE,
    KC_CG_P,
    KC_VISC,
    KC_W18_AV,
    KC_W18_UPDATE,
    KC_W18_P,
    KC_W18_PRE,
    KC_CG_REFRESH,
    KC_ST_SETUP,
    KC_ST_AV,
    KC_PRJ_HEAD,
    KC_PRJ_TAIL,
    KC_GLUE,
    KC_WINDOW,
    KC_SCAN,
    KC_COUNT
};

struct ProfEvent {
    hipEvent_t start, stop;
    int kclass;
    int tag;       // CG iteration index the launch belongs to (-1: always counted)
    int epoch;     // CG solve the launch belongs to
    double bytes;
};

}  // namespace ivp

struct ivp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    ivp_config cfg{};
    float h = 0.0f;
    int C = 60;             // neighborCacheSize: the row capacity (overflow beyond it)
    int Cs = 64;            // slots per row in the tile layout: C rounded up to ROW_PAD
    ivp::Grid grid{};
    int ncell = 0;
    int cap = 0;
    int n = 0, nd = 0, ns = 0;
    int nd_alloc_tiles = 0;
    bool loaded = false;
    bool built = false;
    bool all_dyn_fluid = true;
    bool static_fluid = false;   // some static particle has the fluid material
    // neighbour-row format (IVP_ROWS_*): ids only, the operator passes evaluating gradW from
    // positions; or ids with cached float32 / float16 gradW.  rec_words = words per (tile, slot)
    // record; gscale = the float16 records' scale, the power of two at or above 45/(pi h^6) h^2,
    // the largest |gradW|.
    int row_fmt = IVP_ROWS_IDS;
    int rec_words = 64;
    float gscale = 1.0f;
    float kspiky = 0.0f;    // 45 / (pi h^6), the spiky gradient constant
    float kcubic = 0.0f;    // 8 / pi / h^3, the cubic kernel constant

    // particle store (internal order) and reorder buffers
    float4* pos = nullptr;  float4* pos_b = nullptr;    // x, y, z, m
    float* vel = nullptr;   float* vel_b = nullptr;     // 3 * cap
    float* minv = nullptr;  float* minv_b = nullptr;
    float* rho0 = nullptr;  float* rho0_b = nullptr;
    int* mat = nullptr;     int* mat_b = nullptr;
    int* uid = nullptr;     int* uid_b = nullptr;
    int* inv = nullptr;
    float* density = nullptr;
    float* kk = nullptr;
    float* Ss = nullptr;      // 3 * cap
    float* divg = nullptr;  float* divg_b = nullptr;
    int* rowcnt = nullptr;
    int* rowtot = nullptr;

    // neighbour rows
    uint32_t* nb = nullptr;
    uint32_t* g16 = nullptr;   // float16 format: gradW / gscale, two half-pair words per slot at
                               // twice the word index of the slot's id
    int* tile_len = nullptr;
    // tile windows (ids-only layout, one scene): per 64-row tile its distinct neighbour ids in
    // ascending order (win_ids, WIN_MAX per tile) and their count (win_n, -1: the tile has no
    // window); per walked slot of a valid row the index of its id in the window, 16 bits at the
    // slot's id word (loc).  win_built: the last build wrote them.
    uint32_t* loc = nullptr;
    int* win_ids = nullptr;
    int* win_n = nullptr;
    bool windows = true;
    int win_cap = ivp::WIN_MAX;   // largest window kept (a smaller cap leaves more tiles without)
    bool win_built = false;

    // cell tables and sort scratch
    int* dstart = nullptr;  // ncell + 1
    int* sstart = nullptr;  // ncell + 1
    int* ccount = nullptr;  // ncell, all zero between sorts
    int* key = nullptr;
    int* tmp_src = nullptr;
    int* order = nullptr;
    int* scan_bsum = nullptr;
    int nscan_blocks = 0;

    // IP solver vectors (dynamic rows)
    float4* d4 = nullptr;   // displacement d, w = m
    float4* q4 = nullptr;   // (f k, dfdt k, m, 0)
    float* b = nullptr;  float* sol = nullptr;  float* r = nullptr;  float* Ap = nullptr;  // 3 * cap
    float4* p4 = nullptr;   // CG direction, w = m
    float2* y2 = nullptr;   // (dfdt k J p, m)
    float* cc = nullptr;    // constraint c = rho - rho0
    float* sfac = nullptr;  // dfdt * k
    float* jd = nullptr;    // J d of the last outer iteration
    float4* xs4 = nullptr;  // (x, m) of the CG iterate, input of the true-residual passes
    int* nbact = nullptr;   // row has an active particle in its closed neighbourhood

    // Weiler18 vectors
    float* w_pinv = nullptr;  // 9 * cap
    float4* w_p4 = nullptr;   // (p, m/rho)
    float4* w_v4 = nullptr;   // (v, m/rho)
    float* w_dv = nullptr; float* w_r = nullptr; float* w_z = nullptr;
    float* w_Ap = nullptr; float* w_rhs = nullptr;
    float* w_fac = nullptr;   // dt * mu_i / rho_i
    float* w_dv_prev = nullptr;  // the last solve's correction in USER order (allocated at first
                                 // use, zeroed by ivp_reset): the warm start's initial guess
    bool w18_warm = false;       // Weiler18.warm_start
    bool w18_precond = true;     // Weiler18.use_preconditioner

    // explicit viscosity
    float4* v4 = nullptr;     // (v, rho)

    // Jeske23 surface tension (allocated at first use)
    float* st_gt = nullptr;   // g~ per stored pair, at the word of its id in the ids-only layout
    float* st_div = nullptr;  // velocity divergence per row
    float* st_b = nullptr;    // extra right-hand side, 3 * cap

    // PRJ (allocated at first use).  prj_prev / prj_df_prev hold the last solve's multipliers
    // of both loops in USER order, so that they follow the particles through every re-sort.
    float* prj_v0 = nullptr;      // dt v at the start of the pass loop, 3 * cap
    float2* prj_lam = nullptr;    // (dt^2 lambda, m) per dynamic row
    float* prj_prev = nullptr;    // p_previous, user order
    float* prj_df_prev = nullptr; // dt2_lam_df_previous, user order
    double* prj_hist = nullptr;   // density error per pass
    int prj_hist_cap = 0;
    int prj_guess = 8;            // first launch chunk of the next pass loop

    // reductions
    double* partials = nullptr;   // 4 * max_blocks
    int max_blocks = 0;
    ivp::SolverState* state = nullptr;      // device
    ivp::SolverState* state_host = nullptr; // pinned
    int* err_flag = nullptr;                // device: bit 0 neighbour overflow, bit 2 a kernel's
                                            // consistency guard (set by the build)
    unsigned* speed_max = nullptr;          // device: bits of max |v|^2 (ivp_max_speed2)
    int* err_host = nullptr;                // pinned
    int* count_host = nullptr;              // pinned scratch

    // independent scenes (S > 0 after ivp_set_scenes; S == 0: one scene, the grid above)
    int S = 0;
    std::vector<ivp_scene> scenes;
    std::vector<ivp::Grid> grids_host;
    std::vector<int> scene_row0;          // S + 1 dynamic row starts (internal order)
    std::vector<int> scene_srow0;         // S + 1 static starts, relative to nd
    ivp::Grid* grids = nullptr;           // device [S]
    int* sid = nullptr;                   // device [cap]: scene of each internal particle
    float* sdomain = nullptr;             // device [3 S]: domain size per scene (boundary clamp)
    ivp::SceneState* ss = nullptr;        // device [S]
    ivp::SceneState* ss_host = nullptr;   // pinned [S]
    ivp::SceneParams* sprm = nullptr;     // device
    ivp::SceneParams* sprm_host = nullptr;  // pinned
    double* shist_err = nullptr;          // device [S * SCENE_HIST]
    int* shist_pcg = nullptr;             // device [S * SCENE_HIST]
    double* spartials = nullptr;          // device [8 * max_blocks]: per block, two scene segments
    int* sactive = nullptr;               // device: scenes not done
    int* sactive_host = nullptr;          // pinned
    // first CG launch chunk of the next IP / Weiler18 solve: the last solve's iteration count
    int cg_guess = 4;
    int w18_guess = 4;
    int st_guess = 4;
    int build_lanes = 0;                  // lanes per row of the build's candidate scan (0: by count)
    int scene_chunk = 4;                  // CG iterations per replay
    int scene_lanes = 1;                  // lanes per row in the scene passes (1, 2 or 4)
    bool scene_graph = true;              // replay a captured hipGraph
    hipStream_t cap_stream = nullptr;     // private stream the replay graph is captured on
    ivp::SceneGraph sgraphs[2];
    int sgraph_next = 0;
    long long scene_replays = 0;          // replays of the last batched solve

    // stream marks at outer-iteration / pass boundaries (stats_iter_costs)
    std::vector<hipEvent_t> iter_ev;
    // an IP solve whose closing wait was left to the caller (ivp_step enqueues the rest of the
    // step first): its end mark, start time and outer-loop evaluation count
    hipEvent_t ip_end_ev = nullptr;
    hipEvent_t wait_ev = nullptr;   // a chunk's state copy, when work follows it before the wait
    bool ip_wait_pending = false;
    std::chrono::steady_clock::time_point ip_t0;
    int ip_n_err = 0;

    // profiling
    bool prof = false;
    unsigned prof_mask = 0xffffffffu;
    std::vector<ivp::ProfEvent> prof_pending;
    std::vector<hipEvent_t> prof_free;
    long long prof_launches[ivp::KC_COUNT] = {};
    double prof_ms[ivp::KC_COUNT] = {};
    double prof_bytes[ivp::KC_COUNT] = {};
    int prof_epoch = 0;
};

namespace ivp {

// profiling helpers (ivp_api.hip): take a pair of timing events, queue them after the launch,
// fold finished ones into the per-class totals
int prof_take(ivp_ctx* c, int kclass, int tag, double bytes, ProfEvent* ev);
int prof_push(ivp_ctx* c, ProfEvent* ev);
int prof_collect(ivp_ctx* c, int cg_iters_done, int epoch);
void prof_set_bytes(ivp_ctx* c, int epoch, int kclass, double bytes);
// stats_iter_costs: mark k on the stream at the start of outer-loop evaluation k (mark n after the
// last); iter_costs fills the n intervals once the stream has been synchronised
int iter_mark(ivp_ctx* c, int k);
void iter_costs(ivp_ctx* c, int n, ivp_ip_stats* st);

// ivp_grid.hip
int grid_sort_population(ivp_ctx* c, int offset, int count, int* start_table, bool dynamic);
int grid_build(ivp_ctx* c);
int grid_export(ivp_ctx* c, int cache_size, int* count, int* ids, float* gradW);
int grid_load(ivp_ctx* c, int n, const float* x, const float* v, const float* m, const float* m_inv,
              const float* density0, const int* is_dynamic, const int* material);
int store_write(ivp_ctx* c, int field, const void* src);
int store_read(ivp_ctx* c, int field, void* dst);

// ivp_solve.hip, independent scenes
int ip_solve_scenes(ivp_ctx* c, float dt, const ivp_ip_params* prm);

// ivp_solve.hip
// defer: leave the closing synchronisation (and with it stats_iter_costs and solve_ms) to
// ip_solve_wait, so that the caller can enqueue more work first
int ip_solve(ivp_ctx* c, float dt, const ivp_ip_params* prm, ivp_ip_stats* st, bool defer = false);
int ip_solve_wait(ivp_ctx* c, ivp_ip_stats* st);
int visc_explicit(ivp_ctx* c, float coeff, float dt);
int visc_weiler18(ivp_ctx* c, float coeff, float dt, ivp_cg_stats* st);
int st_jeske23(ivp_ctx* c, float sigma, float dt, ivp_cg_stats* st);
int st_buffers(ivp_ctx* c);   // ivp_api.hip: allocate the Jeske23 buffers on first use
int w18_buffers(ivp_ctx* c);  // ivp_api.hip: the Weiler18 warm-start buffer, zeroed, on first use
enum PrjPart { PRJ_BOTH = 0, PRJ_CD = 1, PRJ_DF = 2 };   // solve (+ solve_df if enabled), solve, solve_df
int prj_solve(ivp_ctx* c, float dt, const ivp_prj_params* prm, ivp_ip_stats* st, int part);
int prj_buffers(ivp_ctx* c, int passes);   // ivp_api.hip: PRJ buffers, history for `passes`
int glue_gravity(ivp_ctx* c, const float g[3], float dt);
int glue_advect_boundary(ivp_ctx* c, float dt);

inline int blocks_for(int n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace ivp
