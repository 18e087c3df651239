/*
 * ivp_oracle.h -- CPU restatement of the reference IP hot path (test infrastructure only).
 * Arrays are in the reference's layout: vec3 fields are float[3n], the neighbour cache is
 * row-major [i*C + k] with C = neighborCacheSize (scripts/sim/simulation_data.py:157, 207-209).
 */
#ifndef IVP_ORACLE_H
#define IVP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_NOT_CONVERGED = 1, ORC_NOT_PD = 2 };
enum { ORC_VISC_INPLACE = 0, ORC_VISC_JACOBI = 1 };

typedef struct {
    float origin[3];
    float inv_w;
    int dims[3];
} orc_grid;

typedef struct {
    float omega;            /* IP.omega                (pressure_base.py:317)  */
    double tol_opt;         /* densityErrorTol         (pressure_base.py:316)  */
    int max_iteration_opt;  /* maxOuterIters           (pressure_base.py:314)  */
    int use_fixed_iters;    /* fixedOuterIters         (pressure_base.py:315)  */
    double tol_pcg;         /* tol_pcg                 (ip.py:228)             */
    int max_iteration_pcg;  /* max_iteration_pcg       (ip.py:229)             */
} orc_ip_params;

typedef struct {
    int capacity;              /* length of density_errors / pcg_iters */
    double *density_errors;    /* stats_density_errors */
    int *pcg_iters;            /* stats_pcg_iters */
    int n_density_errors;
    int n_pcg;
    int outer_iters;           /* stats_outer_iters */
    double final_density_error;/* stats_final_density_error */
    double final_err_loop;
    double pAp_bad;
    int pcg_iter_bad;
} orc_ip_stats;

typedef struct {
    double tol_opt;         /* densityErrorTol         (pressure_base.py:316)  */
    int max_iteration_opt;  /* maxOuterIters           (pressure_base.py:314)  */
    int use_fixed_iters;    /* fixedOuterIters         (pressure_base.py:315)  */
    float omega;            /* PRJ.omega = 0.5         (prj.py:155)            */
    int warm_start;         /* PRJ.warm_start          (prj.py:154)            */
    int enable_df;          /* PRJ.enable_df           (prj.py:158)            */
    int df_num_threshold;   /* PRJ.df_num_threshold    (prj.py:159)            */
    double df_tol_opt;      /* PRJ.df_tol_opt          (prj.py:160)            */
} orc_prj_params;

typedef struct {
    int capacity;              /* length of density_errors */
    double *density_errors;    /* stats_density_errors, one per pass */
    int n_density_errors;
    int outer_iters;           /* stats_outer_iters */
    double final_density_error;/* stats_final_density_error */
    double final_err_loop;     /* err of the last pass */
    int df_outer_iters;        /* stats_df_outer_iters */
    double df_final_err;       /* err of the last divergence-free pass */
} orc_prj_stats;

float orc_cubic_kernel(float r_norm, float h);
void orc_spiky_grad(const float r[3], float h, float out[3]);
void orc_grid_init(const float domain_start[3], const float domain_size[3], float h, orc_grid *g);
void orc_cell_keys(const orc_grid *g, int n, const float *x, int *key);
int orc_build_neighbors(const orc_grid *g, int n, const float *x, const int *is_dynamic, float h,
                        int cache_size, int *count, int *ids, float *gradW, int *max_total);
int orc_brute_neighbors(int n, const float *x, float h, int cap, int *count, int *ids);
void orc_density(int n, const float *x, const float *m, const int *is_dynamic, float h, int C,
                 const int *count, const int *ids, float *density);
void orc_aii(int n, const float *m, const float *m_inv, const int *is_dynamic, int C,
             const int *count, const int *ids, const float *gradW, float *k_out);
void orc_J(int n, const float *xv, const float *m, const int *is_dynamic, int C, const int *count,
           const int *ids, const float *gradW, float *ret);
void orc_JT(int n, const float *y, const float *m, const int *is_dynamic, int C, const int *count,
            const int *ids, const float *gradW, float *ret);
void orc_ip_apply_A(int n, const float *m, const int *is_dynamic, int C, const int *count,
                    const int *ids, const float *gradW, const float *k, const float *dfdt,
                    float omega, const float *p, float *Ap);
int orc_ip_solve(int n, float *v, const float *m, const float *density, const float *density0,
                 const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                 const float *k, float dt, const orc_ip_params *prm, float *divergence,
                 orc_ip_stats *st);
int orc_prj_solve(int n, float *v, const float *m, const float *density, const float *density0,
                  const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                  const float *k, float dt, const orc_prj_params *prm, float *lam_prev,
                  float *lam_df_prev, float *divergence, orc_prj_stats *st);
void orc_viscosity(int n, const float *x, float *v, const float *m, const float *density,
                   const float *density0, const int *material, const int *is_dynamic, float h,
                   int C, const int *count, const int *ids, const float *gradW, float coeff,
                   float dt, int material_fluid, int mode);
void orc_inverse33(const float a[9], float out[9]);
void orc_w18_apply_A(int n, const float *x, const float *m, const float *density,
                     const float *density0, const int *material, const int *is_dynamic, float h,
                     int C, const int *count, const int *ids, const float *gradW, float coeff,
                     float dt, int material_fluid, const float *vin, float *Av, float *Pinv);
int orc_weiler18(int n, const float *x, float *v, const float *m, const float *density,
                 const float *density0, const int *material, const int *is_dynamic, float h, int C,
                 const int *count, const int *ids, const float *gradW, float coeff, float dt,
                 int material_fluid);
int orc_weiler18_ex(int n, const float *x, float *v, const float *m, const float *density,
                    const float *density0, const int *material, const int *is_dynamic, float h, int C,
                    const int *count, const int *ids, const float *gradW, float coeff, float dt,
                    int material_fluid, int warm, int precond, float *dv);
float orc_wst(float r_ij, float d, float h_st);
void orc_grad_wst(const float x_ij[3], float r_ij, float d, float h_st, float out[3]);
void orc_jeske23_setup(int n, const float *x, const float *v, const float *m, const float *density,
                       const int *material, const int *is_dynamic, int C, const int *count, const int *ids,
                       const float *gradW, float sigma, float dt, float d, float h_st, int fluid,
                       float *div, float *g_tilde, float *b);
void orc_jeske23_apply_M(int n, const float *m, const int *material, const int *is_dynamic, int C,
                         const int *count, const int *ids, const float *g_tilde, float sigma, float dt,
                         int fluid, const float *vin, float *Mv, float *Pinv);
int orc_jeske23(int n, const float *x, float *v, const float *m, const float *density, const int *material,
                const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                float sigma, float dt, float d, float h_st, int fluid);
void orc_gravity(int n, float *v, const int *is_dynamic, const float g[3], float dt);
void orc_advect(int n, float *x, float *x_old, float *v, const int *is_dynamic, float dt);
void orc_enforce_boundary(int n, float *x, float *v, const int *is_dynamic,
                          const float domain_size[3], float padding);

#ifdef __cplusplus
}
#endif
#endif
