/*
 * ivp.h -- C ABI of libivp.so, the MI355X (gfx950) implementation of the reference's
 * Implicit Velocity Projection hot path (medialab-ku/ImplicitVelocityProjection).
 *
 * Every pointer argument named x, v, m, ... is a DEVICE pointer to a contiguous array in the
 * caller's particle order ("user order"), laid out like the reference's Warp arrays:
 * vec3 fields are float[3*n], scalar fields float[n] / int[n].  Host structs are plain PODs.
 * All work is enqueued on the stream given to ivp_create; entry points that return solver
 * statistics synchronise that stream once (the reference synchronises on every dot product).
 *
 * Return values: IVP_OK, a positive solver status, or a negative error code.
 * ivp_status_string() turns either into text.
 */
#ifndef IVP_H
#define IVP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define IVP_OK                     0
#define IVP_NOT_CONVERGED          1  /* pressure_base.py:487-495 check_converged raise      */
#define IVP_NOT_PD                 2  /* ip.py:308-309 "non-positive definite matrix"         */
#define IVP_NEIGHBOR_OVERFLOW      3  /* a row has more than cache_size neighbours; the
                                         reference truncates silently (neighbour_search.py:45) */
#define IVP_ERR_ARG               -1
#define IVP_ERR_HIP               -2
#define IVP_ERR_CAPACITY          -3
#define IVP_ERR_STATE             -4
#define IVP_ERR_GRID              -5
#define IVP_ERR_INTERNAL          -6  /* a kernel's own consistency guard fired              */

/* ---- configuration ----------------------------------------------------------------------- */
typedef struct ivp_ctx ivp_ctx;

/* SimConfig fields the hot path reads (scripts/io/config_builder.py:10-100) */
typedef struct {
    float support_radius;    /* h = float32(4 * particleRadius) (config_builder.py:52)     */
    float domain_start[3];   /* domainStart                                                */
    float domain_size[3];    /* domainEnd - domainStart, float32 (simulation_data.py:124)  */
    int cache_size;          /* neighborCacheSize (simulation_data.py:157), default 60     */
    int material_fluid;      /* SimulationData.material_fluid (simulation_data.py:128) = 1 */
} ivp_config;

/* IP solver parameters, read at every ivp_ip_solve call (the GUI edits them live):
 * pressure_base.py:314-317, ip.py:228-229 */
typedef struct {
    float omega;             /* omega              */
    double tol_opt;          /* densityErrorTol    */
    int max_iteration_opt;   /* maxOuterIters      */
    int use_fixed_iters;     /* fixedOuterIters    */
    double tol_pcg;          /* tol_pcg            */
    int max_iteration_pcg;   /* max_iteration_pcg  */
} ivp_ip_params;

/* Per-solve statistics (pressure_base.py:332-339, ip.py:365-404).  The caller owns the two
 * host arrays; entries beyond `capacity` are counted but not stored. */
typedef struct {
    int capacity;
    double *density_errors;      /* stats_density_errors, one per outer-loop evaluation   */
    int *pcg_iters;              /* stats_pcg_iters, one per update                        */
    int n_density_errors;
    int n_pcg;
    int outer_iters;             /* stats_outer_iters                                      */
    double final_density_error;  /* stats_final_density_error                              */
    double pAp_bad;              /* p^T A p when IVP_NOT_PD is returned                    */
    int pcg_iter_bad;            /* CG iteration at which IVP_NOT_PD was detected          */
    long long pairs_traversed;   /* neighbour pairs read by the CG operator passes         */
    long long rows_traversed;    /* rows whose neighbour loop ran in the CG operator passes */
    int cg_launch_chunks;        /* host synchronisations spent on CG termination (IP) or
                                    on the pass loop's stop test (PRJ)                       */
    double solve_ms;             /* host wall time of the solve, synchronisations included */
    int df_outer_iters;          /* PRJ: stats_df_outer_iters of the divergence-free pass   */
    double df_ms;                /* PRJ: stats_df_total_cost_ms                             */
    double *iter_costs;          /* stats_iter_costs (ip.py:369, 392; prj.py:189): ms per
                                    outer-loop evaluation (IP) or pass (PRJ), one per density
                                    error, timed by events on the stream; caller-owned, at most
                                    `capacity` entries, NULL to skip                          */
    int n_iter_costs;
} ivp_ip_stats;
/* ivp_ip_stats holds PressureBase's statistics (pressure_base.py:332-339), shared by both
 * solvers: PRJ leaves pcg_iters empty (n_pcg = 0) and fills the df_ fields. */
/* PRJ solver parameters (prj.py:150-161, pressure_base.py:314-317), read at every call */
typedef struct {
    double tol_opt;          /* densityErrorTol                                          */
    int max_iteration_opt;   /* maxOuterIters: passes 0 .. max_iteration_opt              */
    int use_fixed_iters;     /* fixedOuterIters                                          */
    float omega;             /* PRJ.omega, the Jacobi relaxation (0.5)                    */
    int warm_start;          /* PRJ.warm_start: start from 0.5 x the last solve's lambda  */
    int enable_df;           /* PRJ.enable_df: divergence-free pass after the solve       */
    int df_num_threshold;    /* PRJ.df_num_threshold (25): rows of the divergence-free
                                pass have more neighbours than this                       */
    double df_tol_opt;       /* PRJ.df_tol_opt (1e-2)                                     */
} ivp_prj_params;

/* Weiler18 solve statistics */
typedef struct {
    int cg_iters;
    double residual_norm2;
} ivp_cg_stats;

/* One Framework.forward step (framework.py:423-544), velocity-based branch */
typedef struct {
    float dt;                /* timeStepSize                                       */
    float gravity[3];        /* gravitation                                        */
    int viscosity_method;    /* 0 explicit (Jacobi), 1 Weiler18 implicit           */
    float viscosity_coeff;   /* viscosityCoeff; 0 with method 0 skips the pass     */
    ivp_ip_params ip;
    int surface_tension_method;   /* surfaceTensionMethod: 1 Jeske23 implicit; 0 (the explicit
                                     model) only with coefficient 0                 */
    float surface_tension_coeff;  /* surfaceTensionCoeff; 0 skips the pass          */
    int pressure_method;     /* pressureMethod, the reference's numbering: 1 IP (ip above),
                                0 PRJ (prj below).  A zero-filled struct selects PRJ.  */
    ivp_prj_params prj;
} ivp_step_params;

/* Fields readable with ivp_read / writable with ivp_write */
enum {
    IVP_FIELD_X = 0,          /* float[3n] positions            */
    IVP_FIELD_V = 1,          /* float[3n] velocities           */
    IVP_FIELD_DENSITY = 2,    /* float[n]  density              */
    IVP_FIELD_K = 3,          /* float[n]  1/(Aii + eps)        */
    IVP_FIELD_DIVERGENCE = 4  /* float[n]  max(J v, 0)          */
};

/* ---- lifecycle ---------------------------------------------------------------------------- */
const char *ivp_version(void);
const char *ivp_status_string(int status);

/* SimulationData.allocate + NeighbourSearch.__init__ (simulation_data.py:165-212,
 * neighbour_search.py:93-114).  `stream` is a hipStream_t (NULL: the null stream). */
int ivp_create(const ivp_config *cfg, int max_particles, void *stream, ivp_ctx **out);
int ivp_destroy(ivp_ctx *ctx);

/* SimulationData.add_particles (simulation_data.py:214-287) for the whole particle set at once.
 * m and m_inv follow add_particle (simulation_data.py:40-57).  Dynamic particles are stored
 * first, sorted by cell; static particles are sorted once here and never move. */
int ivp_set_particles(ivp_ctx *ctx, int n, const float *x, const float *v, const float *m,
                      const float *m_inv, const float *density0, const int *is_dynamic,
                      const int *material);
/* particle_num growth (ScheduledSpawns, scheduled_spawn.py:68-117) is a new ivp_set_particles
 * call with the grown set, up to the capacity given to ivp_create. */

/* The solver part of Framework.reset (framework.py:254-270): density back to density0, k
 * (pressure.k), divergence and the PRJ multipliers kept for warm starts (p_previous,
 * dt2_lam_df_previous) zeroed, neighbour rows dropped.  The caller restores x = x0 and
 * v = v0 with ivp_write and calls ivp_neighbors_build, as Framework.reset re-initialises. */
int ivp_reset(ivp_ctx *ctx);

/* Options.  IVP_OPT_GRADIENT_CACHE selects the neighbour-row format:
 *   IVP_ROWS_IDS (0, default): ids only (4 B per pair and pass); every operator pass evaluates
 *     gradW from the positions.
 *   IVP_ROWS_F32 (1): float32 gradW cached next to every id (the reference's neighbor_gradW
 *     cache, 16 B per pair and pass) and read by every operator pass.
 *   IVP_ROWS_F16 (2): the ids-only rows plus float16 gradW per pair in a side array (8 B per
 *     pair) with one power-of-two scale per context; the IP CG's J and J^T passes read it, and every 8th CG
 *     iteration replaces the recursive residual by r = b - A x evaluated in float32 from the
 *     positions.  The outer-loop passes (constraint, right-hand side) and viscosity evaluate
 *     gradW from the positions as in IVP_ROWS_IDS.
 * Density and Aii use the reference's exact expression in every format.  Changing the option
 * drops the neighbour rows: call ivp_neighbors_build again.  No counterpart in the reference. */
#define IVP_OPT_GRADIENT_CACHE 1
#define IVP_ROWS_IDS 0
#define IVP_ROWS_F32 1
#define IVP_ROWS_F16 2
int ivp_set_option(ivp_ctx *ctx, int option, int value);
int ivp_get_option(ivp_ctx *ctx, int option, int *value);

/* Copy a user-order field in or out of the library's particle store.  ivp_write accepts
 * IVP_FIELD_X (dynamic particles only; static particles keep the positions given to
 * ivp_set_particles) and IVP_FIELD_V. */
int ivp_write(ivp_ctx *ctx, int field, const void *src);
int ivp_read(ivp_ctx *ctx, int field, void *dst);

/* ---- neighbour search + density + Aii ---------------------------------------------------- */
/* NeighbourSearch.build + build_cache (neighbour_search.py:116-155) fused with
 * PressureBase.compute_density and compute_Aii (pressure_base.py:12-87, 341-365):
 * cell keys, counting sort, reorder, neighbour rows, density, k = 1/(Aii + 1e-3). */
int ivp_neighbors_build(ivp_ctx *ctx);

/* The reference's neighbour cache layout (simulation_data.py:206-209): count[n] (capped at
 * cache_size), ids[n*cache_size], gradW[3*n*cache_size], user particle ids, every row
 * (static rows included).  Row order: dynamic neighbours by ascending (cell key, id), then
 * static neighbours by ascending (cell key, id). */
int ivp_neighbors_export(ivp_ctx *ctx, int cache_size, int *count, int *ids, float *gradW);

/* ---- solvers ----------------------------------------------------------------------------- */
/* IP.solve (ip.py:331-404): reads the stored v, density, k; writes v and divergence. */
int ivp_ip_solve(ivp_ctx *ctx, float dt, const ivp_ip_params *prm, ivp_ip_stats *stats);

/* PRJ.solve (prj.py:292-340): projected Jacobi on the constant-density constraint,
 *   pass l:  dt v = dt v0 - J^T lambda / m;  r = c + J (dt v);  lambda = max(lambda + omega D r, 0)
 * with D = k = 1/(Aii + 1e-3), stopping when avg(max(r / rho0, 0)) < densityErrorTol after at
 * least one update, or after maxOuterIters.  Reads the stored v, density, k; writes v and
 * divergence; keeps lambda (per particle) for warm starts.  IVP_NOT_CONVERGED leaves v
 * unchanged, as check_converged raises before the velocity update.  enable_df is not read. */
int ivp_prj_solve(ivp_ctx *ctx, float dt, const ivp_prj_params *prm, ivp_ip_stats *stats);
/* PRJ.solve_df (prj.py:213-290): the same loop on J v <= 0 over rows with more than
 * df_num_threshold neighbours (tolerance df_tol_opt), on the stored v; fills the df_ fields of
 * stats.  ivp_step runs it after ivp_prj_solve's loop when enable_df is set. */
int ivp_prj_solve_df(ivp_ctx *ctx, float dt, const ivp_prj_params *prm, ivp_ip_stats *stats);
/* Viscosity.solve (viscosity.py:134-164) with out-of-place (Jacobi) reads. */
int ivp_viscosity_explicit(ivp_ctx *ctx, float coeff, float dt);

/* Weiler18.solve (weiler18.py:371-401): block-Jacobi PCG from dv = 0, or as the options
 * IVP_OPT_W18_WARM_START / IVP_OPT_W18_PRECONDITIONER select. */
int ivp_viscosity_weiler18(ivp_ctx *ctx, float coeff, float dt, ivp_cg_stats *stats);

/* Jeske23.solve (jeske23.py:590-655): implicit surface tension, diagonal-preconditioned CG on
 * M dv = v - M v + b (tol 1e-5 on r.r, at most 100 iterations), v += dv.  h_st = 2 d = h.
 * IVP_ERR_ARG when a static particle has the fluid material (static neighbours are assumed to
 * be boundary, which the reference skips as non-fluid). */
int ivp_surface_tension_jeske23(ivp_ctx *ctx, float sigma, float dt, ivp_cg_stats *stats);

/* apply_gravity (framework.py:26-41) */
int ivp_apply_gravity(ivp_ctx *ctx, const float gravity[3], float dt);

/* advect_position + enforce_boundary (framework.py:117-136, boundary.py:27-91) */
int ivp_advect_boundary(ivp_ctx *ctx, float dt);

/* Framework.forward (framework.py:423-544), velocity-based branch: gravity, surface tension,
 * viscosity, the pressure solve of prm->pressure_method (IP, or PRJ with its optional
 * divergence-free pass), advect, boundary, neighbour rebuild, density, Aii.  stats receives the
 * pressure solver's statistics; visc and st (either may be NULL) the viscosity and
 * surface-tension CG statistics.  The closing rebuild is not waited for; if a row of it exceeds
 * cache_size, the next step (or solve) returns IVP_NEIGHBOR_OVERFLOW with v and x unchanged. */
int ivp_step(ivp_ctx *ctx, const ivp_step_params *prm, ivp_ip_stats *stats, ivp_cg_stats *visc,
             ivp_cg_stats *st);

/* ---- independent scenes in one context ------------------------------------------------------
 * No counterpart in the reference (one Framework per scene there).  For callers that simulate
 * many small scenes at once: every scene keeps its own coordinates, cell grid, IP outer loop,
 * CG and boundary clamp, exactly as if it ran alone in its own context, while each pass of the
 * step runs once over all scenes. */
typedef struct {
    int n_particles;         /* particles of the scene; scenes are consecutive in user order */
    float domain_start[3];   /* the scene's domainStart, in its own coordinates             */
    float domain_size[3];    /* the scene's domainEnd - domainStart                         */
} ivp_scene;

/* Declare the particle set of the next ivp_set_particles as n_scenes consecutive scenes of the
 * given sizes (their sum must equal that call's n).  support_radius, cache_size and
 * material_fluid come from ivp_create; ivp_create's domain is not used afterwards.  Every scene
 * needs at least 256 dynamic particles.  Entry points for one scene (ivp_ip_solve, ivp_step,
 * the viscosity solvers) return IVP_ERR_STATE on such a context; ivp_neighbors_build,
 * ivp_read / ivp_write, ivp_neighbors_export (ids are user indices over all scenes) and
 * ivp_apply_gravity work on all scenes. */
int ivp_set_scenes(ivp_ctx *ctx, int n_scenes, const ivp_scene *scenes);

typedef struct {
    int status;                  /* IVP_OK, IVP_NOT_CONVERGED or IVP_NOT_PD for this scene  */
    int outer_iters;             /* stats_outer_iters                                       */
    int cg_iters;                /* CG iterations summed over the outer iterations          */
    double final_density_error;  /* stats_final_density_error                               */
} ivp_scene_stats;

/* Framework.forward (framework.py:423-544, velocity-based IP branch, viscosity off:
 * viscosity_coeff must be 0) for every scene: gravity, IP solve, advect, per-scene boundary
 * clamp.  The IP outer iterations of all scenes advance together: each replay of one captured
 * launch sequence runs the constraint pass for scenes that start an outer iteration, the
 * right-hand side and CG set-up, a fixed number of CG iterations for scenes in their CG, and
 * the displacement update for scenes whose CG ended; each scene stops on its own density error
 * (densityErrorTol / maxOuterIters semantics of ip.py:345-395).  stats has n_scenes entries.
 * Returns IVP_OK or the first scene's non-OK status. */
int ivp_step_scenes(ivp_ctx *ctx, const ivp_step_params *prm, ivp_scene_stats *stats);

/* Outer-loop history of one scene in the last ivp_step_scenes: stats_density_errors and
 * stats_pcg_iters (at most `capacity` entries each; *n_errors and *n_pcg give the counts). */
int ivp_scene_history(ivp_ctx *ctx, int scene, int capacity, double *density_errors, int *pcg_iters,
                      int *n_errors, int *n_pcg);

/* Options of the scene solve: CG iterations per replay of the launch sequence (default 4);
 * whether the sequence is captured once into a hipGraph and replayed (1, default) or enqueued
 * launch by launch (0); lanes per neighbour row in its pair passes (1 default, 2 or 4). */
#define IVP_OPT_SCENE_CG_CHUNK 2
#define IVP_OPT_SCENE_GRAPH 3
#define IVP_OPT_SCENE_LANES 4
/* Lanes per row of the neighbour build's candidate scan: 0 (default) chooses by the dynamic
 * count; 1 one lane per row, fused with the gradient/density/Aii pass; 2, 4, 8 or 16 lanes per row
 * in a separate scan (slots filled in the same order); 64 the scan by cells (a wave loads the
 * candidates of a cell once, 64 at a time, and tests every row of the cell against them; ids-only
 * and float16 rows, independent scenes included, otherwise the count's choice).  The default takes the scan by
 * cells from 32768 dynamic rows where it applies.  The rows are identical either way. */
#define IVP_OPT_BUILD_LANES 5
/* Weiler18's mutable solver attributes, read at every Weiler18 solve (the reference's GUI toggles
 * them live, gui.py:429-432): IVP_OPT_W18_WARM_START 1 starts the PCG from the previous solve's
 * correction dv, r_0 = b - A dv (warm_start, weiler18.py:311-316, 385-388; default 0: dv = 0);
 * the correction is kept per particle through re-sorts and zeroed by ivp_reset (Weiler18.reset).
 * IVP_OPT_W18_PRECONDITIONER 0 runs plain CG, z = r (use_preconditioner, weiler18.py:279-292,
 * 382-383; default 1: the 3x3 block-Jacobi preconditioner). */
#define IVP_OPT_W18_WARM_START 6
#define IVP_OPT_W18_PRECONDITIONER 7
/* Tile windows (1, default; ids-only and float16 rows, one scene): each neighbour build also
 * lists, per 64-row tile, the distinct neighbours of its rows, and the pair passes load them into
 * on-chip memory once per tile instead of gathering each pair's neighbour from memory.  Results
 * are bitwise the same either way; 0 turns them off from the next build on, and a value v > 1
 * keeps windows of at most v neighbours only (tiles with more gather from memory). */
#define IVP_OPT_TILE_WINDOWS 8

/* ---- scene loading -------------------------------------------------------------------------- */
/* SceneLoader.load_mesh (scene_loader.py:547-580): the voxel centres of the solid bounded by a
 * placed triangle mesh, as the reference's trimesh call mesh.voxelized(pitch, max_iter).fill()
 * .points computes them: triangles split at edge midpoints until no edge is longer than
 * pitch / 2 (at most max_iter levels; more is IVP_ERR_ARG, trimesh's "max_iter exceeded"), every
 * corner rounded to its voxel, round(v / pitch) half to even, and every empty voxel that is not
 * 6-connected to the border of the occupied index box filled.  vertices (double[3 * n_vertices])
 * and faces (int[3 * n_faces], 0-based) are HOST arrays; points is a DEVICE double[3 * capacity]
 * that receives the centres in x-major voxel order (z fastest), (index - min) * pitch +
 * min * pitch in float64 (the reference rounds them to float32 when it adds the particles).  *n_points is the point count; when it exceeds
 * capacity nothing is written and IVP_ERR_CAPACITY is returned.  Runs on `stream` (a
 * hipStream_t) and synchronises it; needs no context. */
int ivp_voxelize(const double *vertices, int n_vertices, const int *faces, int n_faces, double pitch,
                 int max_iter, double *points, long long capacity, long long *n_points, void *stream);
/* ivp_voxelize without any device allocation, for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments, result and errors as ivp_voxelize; runs on the current device.  Every device buffer
 * (voxel grid, subdivision levels, counters, chunk counts) lives in one 64 MiB array that the
 * library holds on each device it runs on until the process ends: meshes whose voxel grid, or
 * one subdivision level of triangles still too long (72 B each), do not fit in it return
 * IVP_ERR_CAPACITY.  All calls on one device share that array, so calls that overlap in time on
 * one device (other threads, other streams) race: serialise them. */
int ivp_voxelize_noalloc(const double *vertices, int n_vertices, const int *faces, int n_faces, double pitch,
                         int max_iter, double *points, long long capacity, long long *n_points, void *stream);
/* ---- time step and animation -------------------------------------------------------------- */
/* compute_max_velocity_sq (framework.py:44-58): *out = max over dynamic particles of
 * (vx vx + vy vy) + vz vz in float32 (0 without dynamic particles), from the stored velocities;
 * Framework's CFL step dt = min(dt_default, cflFactor h / sqrt(out)) (framework.py:427-440)
 * reads it.  Synchronises the stream. */
int ivp_max_speed2(ivp_ctx *ctx, double *out);
/* apply_transform_kernel (animation.py:7-34) on user-order DEVICE arrays of n particles: every
 * particle with object_id == object gets x = (pivot + T) + R (x0 - pivot) in float32 (R row
 * major, products summed left to right).  pivot, R, T are host arrays.  Needs no context; runs
 * on `stream` without synchronising it.  The kinematic flags and velocities of the reference
 * kernel are not touched (an animated static rigid body's velocity is zeroed by the advection). */
int ivp_rigid_transform(const float *x0, float *x, const int *object_id, int n, int object,
                        const float pivot[3], const float R[9], const float T[3], void *stream);
/* ---- instrumentation ---------------------------------------------------------------------- */
/* Per-kernel-class event timing on the context's stream.  Only launches that did work are
 * counted (CG launches after convergence inside a launch chunk are excluded). */
typedef struct {
    char name[32];
    long long launches;
    double total_ms;
    double bytes;            /* algorithmic bytes moved by the counted launches */
} ivp_kernel_timing;

/* enable: 0 off, otherwise a bit mask of kernel classes (bit k = class k of ivp_profile_read;
 * -1 selects all) */
int ivp_profile_enable(ivp_ctx *ctx, int enable);
int ivp_profile_read(ivp_ctx *ctx, ivp_kernel_timing *out, int capacity);
int ivp_profile_reset(ivp_ctx *ctx);

/* Particle counts of the loaded set (n_total, n_dynamic) and the number of dynamic-dynamic
 * neighbour pairs stored by the last build (synchronises the stream when requested). */
int ivp_counts(ivp_ctx *ctx, int *n_total, int *n_dynamic, long long *dyn_pairs);

/* Tile windows of the last build (IVP_OPT_TILE_WINDOWS): 64-row tiles, tiles with a window, and
 * the neighbour ids staged per pass over all windows (synchronises the stream).  All zero when
 * the last build made no windows. */
int ivp_tile_windows(ivp_ctx *ctx, int *tiles, int *windowed, long long *window_ids);

#ifdef __cplusplus
}
#endif
#endif
