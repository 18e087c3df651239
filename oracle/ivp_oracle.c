/*
 * ivp_oracle.c -- CPU restatement of the reference's Implicit Velocity Projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/libivp_oracle.so and loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  The product (libivp.so) never links, loads or calls it.
 *
 * What it restates (all citations relative to the reference repository root):
 *   - scripts/sim/sph_kernel.py:8-26, 52-61           cubic_kernel, spiky_kernel_derivative
 *   - scripts/sim/neighbour_search.py:10-51, 93-155    neighbour cache (row-major [i*C+k])
 *   - scripts/sim/forces/pressure/pressure_base.py:12-203, 394-457  density, Aii, J, J^T, errors
 *   - scripts/sim/forces/pressure/ip.py:19-404        IP kernels, PCG, outer loop
 *   - scripts/sim/math_utils.py:10-195                 vector kernels, tiled dot (block 128)
 *   - scripts/sim/forces/non_pressure/viscosity/viscosity.py:9-108   explicit viscosity
 *   - scripts/sim/forces/non_pressure/viscosity/weiler18.py:21-401   implicit viscosity PCG
 *   - scripts/sim/forces/non_pressure/surface_tension/jeske23.py:23-655  implicit surface tension
 *   - scripts/framework.py:26-41, 117-136, 400-544     gravity, advect, step order
 *   - scripts/sim/boundary.py:10-91                    domain clamp + reflection
 *
 * Execution model: every reference "launch" is a serial loop over the particle index (the
 * Warp device="cpu" model), arithmetic is float32 with no FMA contraction (build with
 * -ffp-contract=off), and host scalars (alpha, beta, errors) are double, as in the
 * reference's Python host code.  Atomic accumulations are performed in index order.
 *
 * Points the reference leaves to its third-party runtime (Warp, not vendored) and which are
 * therefore fixed here by a documented choice ("parity unpinned" against Warp itself):
 *   - neighbour ORDER inside a row.  Warp's HashGrid query order is replaced by: dynamic
 *     neighbours in ascending (cell key, particle id), then static neighbours in ascending
 *     (cell key, particle id).  The neighbour SET is the reference's exact predicate
 *     (j != i and sqrtf(dx*dx + dy*dy + dz*dz) < h, float32, uncontracted) and is pinned
 *     against brute force.
 *   - tile_sum reduction tree of dot_vec3_tiled: sequential inside each 128-wide tile.
 *   - wp.inverse(mat33) singular case: returns the zero matrix.
 *
 * Built a second time with -fopenmp (oracle/_build/libivp_oracle_omp.so) for the all-cores CPU
 * baseline only: the row loops then run in parallel, the scatter of compute_Ax_part2 uses
 * atomics (as the reference's wp.atomic_add / atomic_sub do), and the float sums reduce in
 * per-thread partials.  The in-place (serial-order) explicit viscosity stays serial.  Only the
 * serial build is the checker; the parallel one changes summation order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ivp_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#define ORC_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_FOR_ROWS _Pragma("omp parallel for schedule(dynamic, 256)")
#define ORC_FOR_ROWS_MAX_MX _Pragma("omp parallel for schedule(dynamic, 256) reduction(max : mx)")
#define ORC_FOR_SUM_TOTAL _Pragma("omp parallel for schedule(static) reduction(+ : total)")
#define ORC_FOR_SUM_TOTAL_COUNT _Pragma("omp parallel for schedule(static) reduction(+ : total, count)")
#define ORC_ATOMIC _Pragma("omp atomic")
#define ORC_FOR_ROWS_IF_JACOBI _Pragma("omp parallel for schedule(dynamic, 256) if (jacobi)")
#else
#define ORC_FOR
#define ORC_FOR_ROWS
#define ORC_FOR_ROWS_MAX_MX
#define ORC_FOR_SUM_TOTAL
#define ORC_FOR_SUM_TOTAL_COUNT
#define ORC_ATOMIC
#define ORC_FOR_ROWS_IF_JACOBI
#endif

/* threads the row loops use: 1 in the serial build */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* SPH kernels: scripts/sim/sph_kernel.py                                                      */
/* ------------------------------------------------------------------------------------------ */

/* cubic_kernel, sph_kernel.py:8-26 */
float orc_cubic_kernel(float r_norm, float h)
{
    float res = 0.0f;
    float k = 8.0f / 3.14159265359f;
    k = k / (h * h * h);
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

static float len3(const float r[3])
{
    return sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

static float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* spiky_kernel_derivative, sph_kernel.py:52-61 */
void orc_spiky_grad(const float r[3], float h, float out[3])
{
    float k = 45.0f / (3.14159265359f * h * h * h * h * h * h);
    float r_norm = len3(r);
    if (r_norm < 1e-6f) r_norm = 1e-6f;
    float gq0 = r[0] / r_norm, gq1 = r[1] / r_norm, gq2 = r[2] / r_norm;
    float t = h - r_norm;
    float s = -k * (t * t);
    out[0] = s * gq0;
    out[1] = s * gq1;
    out[2] = s * gq2;
}

/* ------------------------------------------------------------------------------------------ */
/* Uniform grid (documented formula, replaces Warp HashGrid; neighbour_search.py:93-129)       */
/* ------------------------------------------------------------------------------------------ */

void orc_grid_init(const float domain_start[3], const float domain_size[3], float h, orc_grid *g)
{
    float w = h * 1.0009765625f; /* h * (1 + 2^-10): makes the 3x3x3 stencil provably complete */
    g->inv_w = 1.0f / w;
    for (int a = 0; a < 3; ++a) {
        g->origin[a] = domain_start[a] - 2.0f * w;
        g->dims[a] = (int)ceilf(domain_size[a] / w) + 4;
    }
}

static int cell_coord(const orc_grid *g, float x, int a)
{
    float t = (x - g->origin[a]) * g->inv_w;
    t = fminf(fmaxf(t, -1.0f), (float)g->dims[a]);
    int c = (int)floorf(t);
    if (c < 0) c = 0;
    if (c > g->dims[a] - 1) c = g->dims[a] - 1;
    return c;
}

static int cell_key(const orc_grid *g, int cx, int cy, int cz)
{
    return (cz * g->dims[1] + cy) * g->dims[0] + cx;
}

void orc_cell_keys(const orc_grid *g, int n, const float *x, int *key)
{
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        int cx = cell_coord(g, x[3 * i + 0], 0);
        int cy = cell_coord(g, x[3 * i + 1], 1);
        int cz = cell_coord(g, x[3 * i + 2], 2);
        key[i] = cell_key(g, cx, cy, cz);
    }
}

/* Stable counting sort of the particles with is_dynamic == want by cell key.
 * out: sorted particle ids (ascending key, then ascending id); start: [ncell+1] offsets. */
static int sort_population(int n, const int *key, const int *is_dynamic, int want, int ncell,
                           int *out, int *start)
{
    memset(start, 0, sizeof(int) * (size_t)(ncell + 1));
    int m = 0;
    for (int i = 0; i < n; ++i)
        if ((is_dynamic[i] != 0) == want) { start[key[i] + 1]++; m++; }
    for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int *cur = (int *)malloc(sizeof(int) * (size_t)ncell);
    memcpy(cur, start, sizeof(int) * (size_t)ncell);
    for (int i = 0; i < n; ++i)
        if ((is_dynamic[i] != 0) == want) out[cur[key[i]]++] = i;
    free(cur);
    return m;
}

/* build_neighbor_cache, neighbour_search.py:10-51 (every row, static rows included, first
 * cache_size hits kept as in :45, count capped as in :51).  max_total returns the largest
 * uncapped row length so callers can detect the silent truncation. */
int orc_build_neighbors(const orc_grid *g, int n, const float *x, const int *is_dynamic, float h,
                        int cache_size, int *count, int *ids, float *gradW, int *max_total)
{
    int ncell = g->dims[0] * g->dims[1] * g->dims[2];
    int *key = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *dyn = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *sta = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *dstart = (int *)malloc(sizeof(int) * (size_t)(ncell + 1));
    int *sstart = (int *)malloc(sizeof(int) * (size_t)(ncell + 1));
    if (!key || !dyn || !sta || !dstart || !sstart) return -1;
    orc_cell_keys(g, n, x, key);
    sort_population(n, key, is_dynamic, 1, ncell, dyn, dstart);
    sort_population(n, key, is_dynamic, 0, ncell, sta, sstart);

    int mx = 0;
    ORC_FOR_ROWS_MAX_MX
    for (int i = 0; i < n; ++i) {
        const float *xi = x + 3 * i;
        int cx = cell_coord(g, xi[0], 0), cy = cell_coord(g, xi[1], 1), cz = cell_coord(g, xi[2], 2);
        int x0 = cx > 0 ? cx - 1 : 0;
        int x1 = cx < g->dims[0] - 1 ? cx + 1 : g->dims[0] - 1;
        int cnt = 0, total = 0;
        size_t base = (size_t)i * (size_t)cache_size;
        for (int pop = 0; pop < 2; ++pop) {
            const int *list = pop == 0 ? dyn : sta;
            const int *st = pop == 0 ? dstart : sstart;
            for (int dz = -1; dz <= 1; ++dz) {
                int zz = cz + dz;
                if (zz < 0 || zz >= g->dims[2]) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    int yy = cy + dy;
                    if (yy < 0 || yy >= g->dims[1]) continue;
                    int k0 = cell_key(g, x0, yy, zz), k1 = cell_key(g, x1, yy, zz);
                    for (int s = st[k0]; s < st[k1 + 1]; ++s) {
                        int j = list[s];
                        if (j == i) continue;
                        float r[3] = {xi[0] - x[3 * j + 0], xi[1] - x[3 * j + 1], xi[2] - x[3 * j + 2]};
                        float rn = len3(r);
                        if (rn < h) {
                            total++;
                            if (cnt < cache_size) {
                                ids[base + (size_t)cnt] = j;
                                orc_spiky_grad(r, h, gradW + 3 * (base + (size_t)cnt));
                                cnt++;
                            }
                        }
                    }
                }
            }
        }
        count[i] = cnt;
        if (total > mx) mx = total;
    }
    if (max_total) *max_total = mx;
    free(key); free(dyn); free(sta); free(dstart); free(sstart);
    return 0;
}

/* Brute-force O(n^2) neighbour sets (ascending id), the pin for orc_build_neighbors. */
int orc_brute_neighbors(int n, const float *x, float h, int cap, int *count, int *ids)
{
    for (int i = 0; i < n; ++i) {
        int cnt = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            float r[3] = {x[3 * i + 0] - x[3 * j + 0], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
            if (len3(r) < h) {
                if (cnt < cap) ids[(size_t)i * (size_t)cap + (size_t)cnt] = j;
                cnt++;
            }
        }
        count[i] = cnt;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Density, Aii, constraint: pressure_base.py                                                  */
/* ------------------------------------------------------------------------------------------ */

static int capped(const int *count, int i, int C)
{
    int c = count[i];
    return c > C ? C : c;
}

/* compute_density_kernel, pressure_base.py:12-46 */
void orc_density(int n, const float *x, const float *m, const int *is_dynamic, float h, int C,
                 const int *count, const int *ids, float *density)
{
    ORC_FOR_ROWS
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        const float *xi = x + 3 * i;
        float rho = 0.0f;
        rho = rho + m[i] * orc_cubic_kernel(0.0f, h);
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int k = 0; k < cnt; ++k) {
            int j = ids[base + (size_t)k];
            float r[3] = {xi[0] - x[3 * j + 0], xi[1] - x[3 * j + 1], xi[2] - x[3 * j + 2]};
            float rn = len3(r);
            rho += m[j] * orc_cubic_kernel(rn, h);
        }
        density[i] = rho;
    }
}

/* compute_Aii_kernel, pressure_base.py:49-87 */
void orc_aii(int n, const float *m, const float *m_inv, const int *is_dynamic, int C,
             const int *count, const int *ids, const float *gradW, float *k_out)
{
    ORC_FOR_ROWS
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) { k_out[i] = 0.0f; continue; }
        float eps = 1e-3f;
        float Aii = 0.0f;
        float Jii[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            const float *g = gradW + 3 * (base + (size_t)kk);
            float Jij[3] = {m[j] * g[0], m[j] * g[1], m[j] * g[2]};
            if (is_dynamic[j] == 1) Aii += dot3(Jij, Jij) * m_inv[j];
            Jii[0] -= Jij[0]; Jii[1] -= Jij[1]; Jii[2] -= Jij[2];
        }
        Aii += dot3(Jii, Jii) * m_inv[i];
        k_out[i] = 1.0f / (Aii + eps);
    }
}

/* compute_constraint_kernel, pressure_base.py:169-184 */
static void constraint(int n, const float *density, const float *density0, const int *is_dynamic,
                       float *c)
{
    for (int i = 0; i < n; ++i) c[i] = is_dynamic[i] == 0 ? 0.0f : density[i] - density0[i];
}

/* compute_J_x_kernel, pressure_base.py:91-127 (also the J d part of ip_pack_fused, ip.py:150-163) */
void orc_J(int n, const float *xv, const float *m, const int *is_dynamic, int C, const int *count,
           const int *ids, const float *gradW, float *ret)
{
    ORC_FOR_ROWS
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) { ret[i] = 0.0f; continue; }
        const float *vi = xv + 3 * i;
        float result = 0.0f;
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int k = 0; k < cnt; ++k) {
            int j = ids[base + (size_t)k];
            const float *g = gradW + 3 * (base + (size_t)k);
            if (is_dynamic[j] == 1) {
                float dlt[3] = {vi[0] - xv[3 * j + 0], vi[1] - xv[3 * j + 1], vi[2] - xv[3 * j + 2]};
                result += m[j] * dot3(dlt, g);
            } else {
                result += m[j] * dot3(vi, g);
            }
        }
        ret[i] = result;
    }
}

/* compute_J_tr_x_kernel, pressure_base.py:130-166 (ip_rhs_fused without the omega factor) */
void orc_JT(int n, const float *y, const float *m, const int *is_dynamic, int C, const int *count,
            const int *ids, const float *gradW, float *ret)
{
    ORC_FOR_ROWS
    for (int i = 0; i < n; ++i) {
        float *o = ret + 3 * i;
        if (is_dynamic[i] == 0) { o[0] = o[1] = o[2] = 0.0f; continue; }
        float yi = y[i];
        float res[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int k = 0; k < cnt; ++k) {
            int j = ids[base + (size_t)k];
            const float *g = gradW + 3 * (base + (size_t)k);
            float s = is_dynamic[j] == 1 ? (m[j] * yi + m[i] * y[j]) : (m[j] * yi);
            res[0] += s * g[0]; res[1] += s * g[1]; res[2] += s * g[2];
        }
        o[0] = res[0]; o[1] = res[1]; o[2] = res[2];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Reductions: math_utils.py:143-173, pressure_base.py:187-203, 405-436                        */
/* ------------------------------------------------------------------------------------------ */

/* dot_vec3_tiled (block 128, tile_sum, FP32 atomic across tiles), read back as double */
static double dot_tiled(int n, const float *a, const float *b, const int *is_dynamic)
{
    float total = 0.0f;
    ORC_FOR_SUM_TOTAL
    for (int t0 = 0; t0 < n; t0 += 128) {
        float s = 0.0f;
        int t1 = t0 + 128 < n ? t0 + 128 : n;
        for (int i = t0; i < t1; ++i)
            if (is_dynamic[i] == 1) s += dot3(a + 3 * i, b + 3 * i);
        total += s;
    }
    return (double)total;
}

/* dot_vec3 (plain FP32 atomic), used by Weiler18.dot, weiler18.py:226-236 */
static double dot_atomic(int n, const float *a, const float *b, const int *is_dynamic)
{
    float total = 0.0f;
    ORC_FOR_SUM_TOTAL
    for (int i = 0; i < n; ++i)
        if (is_dynamic[i] != 0) total += dot3(a + 3 * i, b + 3 * i);
    return (double)total;
}

/* coef_wise_div (math_utils.py:76-91) + compute_avg_density_error (pressure_base.py:405-423) */
static double avg_density_error(int n, float *t, const float *num, const float *density0,
                                const int *is_dynamic)
{
    float total = 0.0f, count = 0.0f;
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        float denom = density0[i];
        if (denom < 1e-6f) denom = 1e-6f;
        t[i] = num[i] / denom;
    }
    ORC_FOR_SUM_TOTAL_COUNT
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        total += fmaxf(t[i], 0.0f);
        count += 1.0f;
    }
    return count > 0.0f ? (double)total / (double)count : 0.0;
}

/* ------------------------------------------------------------------------------------------ */
/* IP solver: ip.py                                                                            */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int n, C;
    const float *m;
    const int *is_dynamic, *count, *ids;
    const float *gradW, *k;
    float omega;
    float *dfdt, *Jx;
} ip_op;

/* compute_Jx_pcg, ip.py:19-60 */
static void ip_Jx(const ip_op *o, const float *xv)
{
    ORC_FOR_ROWS
    for (int i = 0; i < o->n; ++i) {
        o->Jx[i] = 0.0f;
        if (o->is_dynamic[i] == 0) continue;
        if (o->dfdt[i] > 0.0f) {
            const float *vi = xv + 3 * i;
            float result = 0.0f;
            size_t base = (size_t)i * (size_t)o->C;
            int cnt = capped(o->count, i, o->C);
            for (int k = 0; k < cnt; ++k) {
                int j = o->ids[base + (size_t)k];
                const float *g = o->gradW + 3 * (base + (size_t)k);
                float val[3] = {o->m[j] * g[0], o->m[j] * g[1], o->m[j] * g[2]};
                if (o->is_dynamic[j] == 1) {
                    float dlt[3] = {vi[0] - xv[3 * j + 0], vi[1] - xv[3 * j + 1], vi[2] - xv[3 * j + 2]};
                    result += dot3(dlt, val);
                } else {
                    result += dot3(vi, val);
                }
            }
            o->Jx[i] = result * o->dfdt[i] * o->k[i];
        }
    }
}

/* compute_Ax_part1 + compute_Ax_part2, ip.py:63-121 (scatter form, atomics in index order) */
static void ip_Ax(const ip_op *o, const float *p, float *Ax)
{
    ip_Jx(o, p);
    ORC_FOR
    for (int i = 0; i < o->n; ++i) {
        float *a = Ax + 3 * i;
        if (o->is_dynamic[i] == 0) { a[0] = a[1] = a[2] = 0.0f; continue; }
        a[0] = o->m[i] * p[3 * i + 0]; a[1] = o->m[i] * p[3 * i + 1]; a[2] = o->m[i] * p[3 * i + 2];
    }
    ORC_FOR_ROWS
    for (int i = 0; i < o->n; ++i) {
        if (o->is_dynamic[i] == 0) continue;
        if (o->dfdt[i] <= 0.0f) continue;
        float Jxi = o->Jx[i];
        size_t base = (size_t)i * (size_t)o->C;
        int cnt = capped(o->count, i, o->C);
        for (int k = 0; k < cnt; ++k) {
            int j = o->ids[base + (size_t)k];
            const float *g = o->gradW + 3 * (base + (size_t)k);
            float val[3] = {o->m[j] * g[0], o->m[j] * g[1], o->m[j] * g[2]};
            float s = o->omega * Jxi;
            float c[3] = {s * val[0], s * val[1], s * val[2]};
            ORC_ATOMIC
            Ax[3 * i + 0] += c[0];
            ORC_ATOMIC
            Ax[3 * i + 1] += c[1];
            ORC_ATOMIC
            Ax[3 * i + 2] += c[2];
            if (o->is_dynamic[j] == 1) {
                ORC_ATOMIC
                Ax[3 * j + 0] -= c[0];
                ORC_ATOMIC
                Ax[3 * j + 1] -= c[1];
                ORC_ATOMIC
                Ax[3 * j + 2] -= c[2];
            }
        }
    }
}

void orc_ip_apply_A(int n, const float *m, const int *is_dynamic, int C, const int *count,
                    const int *ids, const float *gradW, const float *k, const float *dfdt,
                    float omega, const float *p, float *Ap)
{
    float *Jx = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    ip_op o = {n, C, m, is_dynamic, count, ids, gradW, k, omega, (float *)dfdt, Jx};
    ip_Ax(&o, p, Ap);
    free(Jx);
}

/* IP.PCG, ip.py:283-329.  Returns the iteration count, or -1 on pAp < 0 (ip.py:308-309). */
static int ip_pcg(const ip_op *o, float *x, const float *b, float *r, float *p, float *Ap,
                  double tol_pcg, int max_it, double *pAp_bad)
{
    int n = o->n;
    memcpy(r, b, sizeof(float) * 3 * (size_t)n);
    double rr_old = dot_tiled(n, r, r, o->is_dynamic);
    int it = 0;
    if (rr_old > 1e-30) {
        memcpy(p, r, sizeof(float) * 3 * (size_t)n);
        for (int s = 0; s < max_it; ++s) {
            ip_Ax(o, p, Ap);
            double pAp = dot_tiled(n, p, Ap, o->is_dynamic);
            if (pAp < 0.0) { *pAp_bad = pAp; return -1 - it; }
            double alpha = rr_old / (pAp > 1e-30 ? pAp : 1e-30);
            float af = (float)alpha;
            ORC_FOR
            for (int i = 0; i < n; ++i) {
                if (o->is_dynamic[i] == 0) continue;
                for (int c = 0; c < 3; ++c) {
                    x[3 * i + c] = x[3 * i + c] + af * p[3 * i + c];
                    r[3 * i + c] = r[3 * i + c] - af * Ap[3 * i + c];
                }
            }
            double rr_new = dot_tiled(n, r, r, o->is_dynamic);
            it++;
            if ((rr_new < tol_pcg && it >= 2) || it >= max_it) break;
            double beta = rr_new / (rr_old > 1e-30 ? rr_old : 1e-30);
            float bf = (float)beta;
            ORC_FOR
            for (int i = 0; i < n; ++i) {
                if (o->is_dynamic[i] == 0) continue;
                for (int c = 0; c < 3; ++c) p[3 * i + c] = r[3 * i + c] + bf * p[3 * i + c];
            }
            rr_old = rr_new;
        }
    }
    return it;
}

/* IP.solve, ip.py:331-404.  Writes v (dynamic rows) and divergence.  Returns ORC_OK,
 * ORC_NOT_CONVERGED (check_converged, pressure_base.py:487-495) or ORC_NOT_PD (ip.py:308). */
int orc_ip_solve(int n, float *v, const float *m, const float *density, const float *density0,
                 const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                 const float *k, float dt, const orc_ip_params *prm, float *divergence,
                 orc_ip_stats *st)
{
    size_t n3 = 3 * (size_t)(n > 0 ? n : 1);
    float *c = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *t = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *f = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *dfdt = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *Jx = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *fk = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    float *d = (float *)calloc(n3, sizeof(float));
    float *rhs = (float *)calloc(n3, sizeof(float));
    float *sol = (float *)calloc(n3, sizeof(float));
    float *r = (float *)calloc(n3, sizeof(float));
    float *p = (float *)calloc(n3, sizeof(float));
    float *Ap = (float *)calloc(n3, sizeof(float));
    int status = ORC_OK;
    ip_op o = {n, C, m, is_dynamic, count, ids, gradW, k, prm->omega, dfdt, Jx};

    constraint(n, density, density0, is_dynamic, c);
    st->n_density_errors = 0;
    st->n_pcg = 0;
    float dtf = dt;
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) d[3 * i + a] = dtf * v[3 * i + a];
    }
    int opt_iter = 0;
    double err = 0.0;
    for (;;) {
        /* ip_pack_fused, ip.py:124-171 */
        orc_J(n, d, m, is_dynamic, C, count, ids, gradW, Jx);
        ORC_FOR
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) { t[i] = 0.0f; f[i] = 0.0f; dfdt[i] = 0.0f; continue; }
            float ti = Jx[i] + c[i];
            t[i] = ti;
            f[i] = fmaxf(ti, 0.0f);
            dfdt[i] = ti >= 1e-6f ? 1.0f : 0.0f;
        }
        err = avg_density_error(n, t, f, density0, is_dynamic);
        if (st->n_density_errors < st->capacity) st->density_errors[st->n_density_errors] = err;
        st->n_density_errors++;
        int converged = err < prm->tol_opt && !prm->use_fixed_iters;
        if ((converged && opt_iter >= 1) || opt_iter >= prm->max_iteration_opt) break;
        /* ip_rhs_fused, ip.py:174-215 */
        ORC_FOR
        for (int i = 0; i < n; ++i) fk[i] = f[i] * k[i];
        ORC_FOR
        for (int i = 0; i < n; ++i) {
            float *o3 = rhs + 3 * i;
            if (is_dynamic[i] == 0) { o3[0] = o3[1] = o3[2] = 0.0f; continue; }
            float fki = f[i] * k[i], mi = m[i];
            float res[3] = {0.0f, 0.0f, 0.0f};
            size_t base = (size_t)i * (size_t)C;
            int cnt = capped(count, i, C);
            for (int kk = 0; kk < cnt; ++kk) {
                int j = ids[base + (size_t)kk];
                const float *g = gradW + 3 * (base + (size_t)kk);
                float s;
                if (is_dynamic[j] == 1) {
                    float fkj = f[j] * k[j];
                    s = m[j] * fki + mi * fkj;
                } else {
                    s = m[j] * fki;
                }
                res[0] += s * g[0]; res[1] += s * g[1]; res[2] += s * g[2];
            }
            o3[0] = prm->omega * res[0]; o3[1] = prm->omega * res[1]; o3[2] = prm->omega * res[2];
        }
        ORC_FOR
        for (int i = 0; i < n; ++i)
            if (is_dynamic[i] != 0) sol[3 * i] = sol[3 * i + 1] = sol[3 * i + 2] = 0.0f;
        double bad = 0.0;
        int it = ip_pcg(&o, sol, rhs, r, p, Ap, prm->tol_pcg, prm->max_iteration_pcg, &bad);
        if (it < 0) {
            st->pAp_bad = bad;
            st->pcg_iter_bad = -1 - it;
            status = ORC_NOT_PD;
            goto done;
        }
        if (st->n_pcg < st->capacity) st->pcg_iters[st->n_pcg] = it;
        st->n_pcg++;
        ORC_FOR
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) continue;
            for (int a = 0; a < 3; ++a) d[3 * i + a] = d[3 * i + a] + -1.0f * sol[3 * i + a];
        }
        opt_iter++;
    }
    st->outer_iters = opt_iter;
    if (!prm->use_fixed_iters && err >= prm->tol_opt && opt_iter >= prm->max_iteration_opt)
        status = ORC_NOT_CONVERGED;
    if (status == ORC_OK) {
        /* v = d / dt: scale_vec3 with 1.0/dt computed in double and passed as float32 */
        float inv_dt = (float)(1.0 / (double)dt);
        ORC_FOR
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) continue;
            for (int a = 0; a < 3; ++a) v[3 * i + a] = inv_dt * d[3 * i + a];
        }
        /* final_density_error, pressure_base.py:438-447 */
        orc_J(n, d, m, is_dynamic, C, count, ids, gradW, Jx);
        ORC_FOR
        for (int i = 0; i < n; ++i)
            if (is_dynamic[i] != 0) t[i] = Jx[i] + 1.0f * c[i];
        st->final_density_error = avg_density_error(n, t, t, density0, is_dynamic);
        /* compute_divergence, pressure_base.py:449-457 */
        orc_J(n, v, m, is_dynamic, C, count, ids, gradW, Jx);
        if (divergence)
            ORC_FOR
            for (int i = 0; i < n; ++i) divergence[i] = is_dynamic[i] == 1 ? fmaxf(Jx[i], 0.0f) : 0.0f;
    }
done:
    st->final_err_loop = err;
    free(c); free(t); free(f); free(dfdt); free(Jx); free(fk);
    free(d); free(rhs); free(sol); free(r); free(p); free(Ap);
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* PRJ solver (projected Jacobi, constant density, optional divergence-free pass): prj.py     */
/* ------------------------------------------------------------------------------------------ */

/* pj_head_fused, prj.py:17-59: dt_v_i = dt_v_0_i - sum_j (m_j lam_i + [dyn j] m_i lam_j) gradW_ij
 * / max(m_i, 1e-6).  Also the head of the divergence-free pass (prj.py:243-253:
 * compute_J_tr_x, div_vec3_by_scalar, add_vec3 with scale -1 compute the same values). */
static void prj_head(int n, const float *lam, const float *m, const float *dt_v_0, const int *is_dynamic,
                     int C, const int *count, const int *ids, const float *gradW, float *dt_v)
{
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        float li = lam[i], mi = m[i];
        float acc[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            const float *g = gradW + 3 * (base + (size_t)kk);
            float s = is_dynamic[j] == 1 ? (m[j] * li + mi * lam[j]) : (m[j] * li);
            acc[0] += s * g[0]; acc[1] += s * g[1]; acc[2] += s * g[2];
        }
        float denom = mi < 1e-6f ? 1e-6f : mi;
        for (int a = 0; a < 3; ++a) dt_v[3 * i + a] = dt_v_0[3 * i + a] - acc[a] / denom;
    }
}

/* One PRJ solve_PJ pass loop, prj.py:162-211.  Returns the pass count (stats_outer_iters). */
static int prj_cd(int n, const float *m, const float *c, const float *D_inv, const float *density0,
                  const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                  const orc_prj_params *prm, const float *dt_v_0, float *dt_v, float *lam, float *Jx,
                  orc_prj_stats *st, double *err_out)
{
    memcpy(dt_v, dt_v_0, sizeof(float) * 3 * (size_t)n);
    st->n_density_errors = 0;
    double err = 0.0;
    int n_pass = 0;
    for (n_pass = 0; n_pass <= prm->max_iteration_opt; ++n_pass) {
        prj_head(n, lam, m, dt_v_0, is_dynamic, C, count, ids, gradW, dt_v);
        /* pj_tail_fused, prj.py:62-115 (float32 atomics, here in index order) */
        orc_J(n, dt_v, m, is_dynamic, C, count, ids, gradW, Jx);
        float total = 0.0f, cnt = 0.0f;
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) continue;
            float r = Jx[i] + c[i];
            lam[i] = fmaxf(lam[i] + prm->omega * r * D_inv[i], 0.0f);
            float rho0 = density0[i] < 1e-6f ? 1e-6f : density0[i];
            total += fmaxf(r / rho0, 0.0f);
            cnt += 1.0f;
        }
        err = cnt > 0.0f ? (double)total / (double)cnt : 0.0;   /* read_avg_error, pressure_base.py:394-400 */
        if (st->n_density_errors < st->capacity) st->density_errors[st->n_density_errors] = err;
        st->n_density_errors++;
        int converged = err < prm->tol_opt && !prm->use_fixed_iters;
        if ((converged && n_pass >= 1) || n_pass >= prm->max_iteration_opt) break;
    }
    if (n_pass > prm->max_iteration_opt) n_pass = prm->max_iteration_opt;   /* max_iteration_opt < 0 */
    *err_out = err;
    return n_pass;
}

/* PRJ.solve_df, prj.py:213-290: projected Jacobi on J v <= 0 over rows with more than
 * df_num_threshold neighbours (compute_active_mask, pressure_base.py:227-240), tol df_tol_opt. */
static int prj_df(int n, float *v, const float *m, const float *D_inv, const float *density0,
                  const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                  float dt, const orc_prj_params *prm, float *lam_df_prev, float *dt_v_0, float *dt_v,
                  float *lam, float *Jx, float *t, double *err_out)
{
    float *active = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    for (int i = 0; i < n; ++i)
        active[i] = (is_dynamic[i] == 1 && count[i] > prm->df_num_threshold) ? 1.0f : 0.0f;
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) dt_v_0[3 * i + a] = dt * v[3 * i + a];
        lam[i] = 0.0f;
        if (prm->warm_start) lam[i] = lam[i] + 0.5f * lam_df_prev[i];
    }
    memcpy(dt_v, dt_v_0, sizeof(float) * 3 * (size_t)n);
    double err = 0.0;
    int df_iter;
    for (df_iter = 0; df_iter <= prm->max_iteration_opt; ++df_iter) {
        prj_head(n, lam, m, dt_v_0, is_dynamic, C, count, ids, gradW, dt_v);
        orc_J(n, dt_v, m, is_dynamic, C, count, ids, gradW, Jx);
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) continue;
            float dJ = Jx[i] * active[i];                    /* coef_wise_mul */
            float Dr = dJ * D_inv[i];                        /* coef_wise_mul */
            lam[i] = fmaxf(lam[i] + prm->omega * Dr, 0.0f);  /* add_scalar, max_zero */
            float rho0 = density0[i] < 1e-6f ? 1e-6f : density0[i];
            t[i] = dJ / rho0;                                /* coef_wise_div */
        }
        float total = 0.0f, cnt = 0.0f;
        for (int i = 0; i < n; ++i) {
            if (is_dynamic[i] == 0) continue;
            total += fmaxf(t[i], 0.0f);
            cnt += 1.0f;
        }
        err = cnt > 0.0f ? (double)total / (double)cnt : 0.0;
        int converged = err < prm->df_tol_opt && !prm->use_fixed_iters;
        if ((converged && df_iter >= 1) || df_iter >= prm->max_iteration_opt) {
            memcpy(lam_df_prev, lam, sizeof(float) * (size_t)n);
            break;
        }
    }
    if (df_iter > prm->max_iteration_opt) df_iter = prm->max_iteration_opt;
    float inv_dt = (float)(1.0 / (double)dt);
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) v[3 * i + a] = inv_dt * dt_v[3 * i + a];
    }
    free(active);
    *err_out = err;
    return df_iter;
}

/* PRJ.solve, prj.py:292-340 (+ solve_df when enabled, framework.py:483-487).  Writes v (dynamic
 * rows), divergence, lam_prev (p_previous) and lam_df_prev.  Returns ORC_OK or
 * ORC_NOT_CONVERGED (check_converged, pressure_base.py:487-495: v is then left unchanged). */
int orc_prj_solve(int n, float *v, const float *m, const float *density, const float *density0,
                  const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                  const float *k, float dt, const orc_prj_params *prm, float *lam_prev,
                  float *lam_df_prev, float *divergence, orc_prj_stats *st)
{
    size_t n1 = (size_t)(n > 0 ? n : 1), n3 = 3 * n1;
    float *c = (float *)calloc(n1, sizeof(float));
    float *t = (float *)calloc(n1, sizeof(float));
    float *Jx = (float *)calloc(n1, sizeof(float));
    float *lam = (float *)calloc(n1, sizeof(float));
    float *dt_v_0 = (float *)calloc(n3, sizeof(float));
    float *dt_v = (float *)calloc(n3, sizeof(float));
    int status = ORC_OK;

    constraint(n, density, density0, is_dynamic, c);
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) dt_v_0[3 * i + a] = dt * v[3 * i + a];   /* scale_vec3 */
        lam[i] = 0.0f;
        if (prm->warm_start) lam[i] = lam[i] + 0.5f * lam_prev[i];          /* add_scalar */
    }
    double err = 0.0;
    st->outer_iters = prj_cd(n, m, c, k, density0, is_dynamic, C, count, ids, gradW, prm, dt_v_0, dt_v, lam,
                             Jx, st, &err);
    memcpy(lam_prev, lam, sizeof(float) * n1);
    st->final_err_loop = err;
    st->final_density_error = 0.0;
    st->df_outer_iters = 0;
    st->df_final_err = 0.0;
    if (!prm->use_fixed_iters && err >= prm->tol_opt && st->outer_iters >= prm->max_iteration_opt) {
        status = ORC_NOT_CONVERGED;
        goto done;
    }
    float inv_dt = (float)(1.0 / (double)dt);
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) v[3 * i + a] = inv_dt * dt_v[3 * i + a];
    }
    /* final_density_error, pressure_base.py:438-447 */
    orc_J(n, dt_v, m, is_dynamic, C, count, ids, gradW, Jx);
    for (int i = 0; i < n; ++i)
        if (is_dynamic[i] != 0) t[i] = Jx[i] + 1.0f * c[i];
    st->final_density_error = avg_density_error(n, t, t, density0, is_dynamic);
    if (prm->enable_df) {
        double derr = 0.0;
        st->df_outer_iters = prj_df(n, v, m, k, density0, is_dynamic, C, count, ids, gradW, dt, prm, lam_df_prev,
                                    dt_v_0, dt_v, lam, Jx, t, &derr);
        st->df_final_err = derr;
    }
    /* compute_divergence, pressure_base.py:449-457 */
    orc_J(n, v, m, is_dynamic, C, count, ids, gradW, Jx);
    if (divergence)
        for (int i = 0; i < n; ++i) divergence[i] = is_dynamic[i] == 1 ? fmaxf(Jx[i], 0.0f) : 0.0f;
done:
    free(c); free(t); free(Jx); free(lam); free(dt_v_0); free(dt_v);
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Explicit viscosity: viscosity.py:9-108                                                      */
/* ------------------------------------------------------------------------------------------ */

/* mode ORC_VISC_INPLACE: the reference kernel as written, run serially (Warp-CPU order, later
 * rows read already-updated v_j).  mode ORC_VISC_JACOBI: every row reads the input v only. */
void orc_viscosity(int n, const float *x, float *v, const float *m, const float *density,
                   const float *density0, const int *material, const int *is_dynamic, float h,
                   int C, const int *count, const int *ids, const float *gradW, float coeff,
                   float dt, int material_fluid, int mode)
{
    float *vin = v;
    float *copy = NULL;
    if (mode == ORC_VISC_JACOBI) {
        copy = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
        memcpy(copy, v, sizeof(float) * 3 * (size_t)n);
        vin = copy;
    }
    int jacobi = mode == ORC_VISC_JACOBI;
    (void)jacobi;
    ORC_FOR_ROWS_IF_JACOBI
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        if (material[i] != material_fluid) continue;
        const float *xi = x + 3 * i;
        float vi[3] = {vin[3 * i], vin[3 * i + 1], vin[3 * i + 2]};
        float rho_i = density[i], rho0_i = density0[i];
        float eps2 = (0.01f * h) * (0.01f * h);
        float k_visc = 2.0f * (3.0f + 2.0f);
        float acc[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            if (material[j] != material_fluid) continue;
            float r[3] = {xi[0] - x[3 * j], xi[1] - x[3 * j + 1], xi[2] - x[3 * j + 2]};
            float rn = len3(r);
            float vj[3] = {vin[3 * j], vin[3 * j + 1], vin[3 * j + 2]};
            float rho_j = density[j];
            float vij[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
            float vxy = dot3(vij, r);
            float K = 2.0f * rho0_i / (rho_i + rho_j);
            const float *g = gradW + 3 * (base + (size_t)kk);
            float s = k_visc * coeff * (m[j] / rho_j) * (vxy / (rn * rn + eps2));
            float fv[3] = {s * g[0], s * g[1], s * g[2]};
            acc[0] = acc[0] + K * fv[0]; acc[1] = acc[1] + K * fv[1]; acc[2] = acc[2] + K * fv[2];
        }
        for (int a = 0; a < 3; ++a) v[3 * i + a] = vi[a] + acc[a] * dt;
    }
    free(copy);
}

/* ------------------------------------------------------------------------------------------ */
/* Implicit viscosity, Weiler et al. 2018: weiler18.py                                        */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int n, C;
    const float *x, *m, *density, *density0;
    const int *material, *is_dynamic, *count, *ids;
    const float *gradW;
    float h, coeff, dt;
    int fluid;
} w18_op;

/* wp.inverse for mat33 via the adjugate; the zero matrix when det == 0 */
void orc_inverse33(const float a[9], float out[9])
{
    float c00 = a[4] * a[8] - a[5] * a[7];
    float c01 = a[5] * a[6] - a[3] * a[8];
    float c02 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.0f) { memset(out, 0, sizeof(float) * 9); return; }
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

/* compute_precond_weiler18, weiler18.py:21-85 */
static void w18_precond(const w18_op *o, float *Pinv)
{
    ORC_FOR_ROWS
    for (int i = 0; i < o->n; ++i) {
        float *P = Pinv + 9 * (size_t)i;
        if (o->is_dynamic[i] == 0 || o->material[i] != o->fluid) {
            memset(P, 0, sizeof(float) * 9);
            P[0] = P[4] = P[8] = 1.0f;
            continue;
        }
        const float *xi = o->x + 3 * i;
        float rho_i = o->density[i];
        float eps2 = 0.01f * o->h * o->h;
        float dd = 10.0f;
        float mu_i = o->coeff * o->density0[i];
        float factor = o->dt * mu_i / rho_i;
        float L[9] = {0};
        size_t base = (size_t)i * (size_t)o->C;
        int cnt = capped(o->count, i, o->C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = o->ids[base + (size_t)kk];
            if (o->material[j] != o->fluid) continue;
            float xij[3] = {xi[0] - o->x[3 * j], xi[1] - o->x[3 * j + 1], xi[2] - o->x[3 * j + 2]};
            float r_sq = dot3(xij, xij);
            const float *g = o->gradW + 3 * (base + (size_t)kk);
            float c = dd * (o->m[j] / o->density[j]) / (r_sq + eps2);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) L[3 * a + b] = L[3 * a + b] + c * (g[a] * xij[b]);
        }
        float M[9];
        for (int a = 0; a < 9; ++a) M[a] = ((a % 4) == 0 ? 1.0f : 0.0f) - factor * L[a];
        orc_inverse33(M, P);
    }
}

/* compute_Av_weiler18, weiler18.py:103-172 */
static void w18_Av(const w18_op *o, const float *vin, float *Av)
{
    ORC_FOR_ROWS
    for (int i = 0; i < o->n; ++i) {
        float *out = Av + 3 * (size_t)i;
        if (o->is_dynamic[i] == 0 || o->material[i] != o->fluid) {
            out[0] = vin[3 * i]; out[1] = vin[3 * i + 1]; out[2] = vin[3 * i + 2];
            continue;
        }
        const float *vi = vin + 3 * i, *xi = o->x + 3 * i;
        float rho_i = o->density[i];
        float eps2 = 0.01f * o->h * o->h;
        float dd = 10.0f;
        float mu_i = o->coeff * o->density0[i];
        float factor = o->dt * mu_i / rho_i;
        float lap[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)o->C;
        int cnt = capped(o->count, i, o->C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = o->ids[base + (size_t)kk];
            if (o->material[j] != o->fluid) continue;
            float xij[3] = {xi[0] - o->x[3 * j], xi[1] - o->x[3 * j + 1], xi[2] - o->x[3 * j + 2]};
            float vij[3] = {vi[0] - vin[3 * j], vi[1] - vin[3 * j + 1], vi[2] - vin[3 * j + 2]};
            float r_sq = dot3(xij, xij);
            const float *g = o->gradW + 3 * (base + (size_t)kk);
            float vx = dot3(vij, xij);
            float s = dd * (o->m[j] / o->density[j]) * (vx / (r_sq + eps2));
            lap[0] += s * g[0]; lap[1] += s * g[1]; lap[2] += s * g[2];
        }
        for (int a = 0; a < 3; ++a) out[a] = vi[a] - factor * lap[a];
    }
}

void orc_w18_apply_A(int n, const float *x, const float *m, const float *density,
                     const float *density0, const int *material, const int *is_dynamic, float h,
                     int C, const int *count, const int *ids, const float *gradW, float coeff,
                     float dt, int material_fluid, const float *vin, float *Av, float *Pinv)
{
    w18_op o = {n, C, x, m, density, density0, material, is_dynamic, count, ids, gradW, h, coeff, dt, material_fluid};
    if (vin && Av) w18_Av(&o, vin, Av);
    if (Pinv) w18_precond(&o, Pinv);
}

static void add_vec3(int n, float *ret, const float *v0, float scale, const float *v1,
                     const int *is_dynamic)
{
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) ret[3 * i + a] = v0[3 * i + a] + scale * v1[3 * i + a];
    }
}

/* z = M^-1 r (apply_precond_weiler18), or z = r (wp.copy) without the preconditioner:
 * weiler18.py:279-292 */
static void w18_apply_precond(int n, const float *Pinv, const float *r, float *z, const int *is_dynamic,
                              int precond)
{
    if (!precond) {
        memcpy(z, r, sizeof(float) * 3 * (size_t)n);
        return;
    }
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        const float *P = Pinv + 9 * (size_t)i;
        const float *ri = r + 3 * i;
        for (int a = 0; a < 3; ++a) z[3 * i + a] = P[3 * a] * ri[0] + P[3 * a + 1] * ri[1] + P[3 * a + 2] * ri[2];
    }
}

/* Weiler18.solve + Weiler18.PCG, weiler18.py:298-401.  warm: start from dv (the previous solve's
 * correction, weiler18.py:311-316, r_0 = b - A dv), else from dv = 0 (weiler18.py:387-388,
 * r_0 = b); precond: block-Jacobi preconditioner (use_preconditioner) or plain CG.  dv is in/out
 * (3n floats, the solver's self.dv).  Returns the CG iteration count. */
int orc_weiler18_ex(int n, const float *x, float *v, const float *m, const float *density,
                    const float *density0, const int *material, const int *is_dynamic, float h, int C,
                    const int *count, const int *ids, const float *gradW, float coeff, float dt,
                    int material_fluid, int warm, int precond, float *dv)
{
    size_t n3 = 3 * (size_t)(n > 0 ? n : 1);
    w18_op o = {n, C, x, m, density, density0, material, is_dynamic, count, ids, gradW, h, coeff, dt, material_fluid};
    float *Pinv = (float *)calloc(9 * (size_t)(n > 0 ? n : 1), sizeof(float));
    float *r = (float *)calloc(n3, sizeof(float));
    float *z = (float *)calloc(n3, sizeof(float));
    float *p = (float *)calloc(n3, sizeof(float));
    float *Ap = (float *)calloc(n3, sizeof(float));
    float *rhs = (float *)calloc(n3, sizeof(float));
    if (precond) w18_precond(&o, Pinv);
    if (!warm) memset(dv, 0, sizeof(float) * 3 * (size_t)n);
    w18_Av(&o, v, Ap);
    add_vec3(n, rhs, v, -1.0f, Ap, is_dynamic);
    /* PCG, weiler18.py:298-365 */
    if (warm) {
        w18_Av(&o, dv, Ap);
        add_vec3(n, r, rhs, -1.0f, Ap, is_dynamic);
    } else {
        memcpy(r, rhs, sizeof(float) * 3 * (size_t)n);
    }
    w18_apply_precond(n, Pinv, r, z, is_dynamic, precond);
    memcpy(p, z, sizeof(float) * 3 * (size_t)n);
    double rz_old = dot_atomic(n, r, z, is_dynamic);
    int it = 0;
    if (rz_old > 1e-12) {
        for (int s = 0; s < 1000; ++s) {
            w18_Av(&o, p, Ap);
            double pAp = dot_atomic(n, p, Ap, is_dynamic);
            if (pAp <= 0.0) break;
            double alpha = rz_old / (pAp > 1e-12 ? pAp : 1e-12);
            add_vec3(n, dv, dv, (float)alpha, p, is_dynamic);
            add_vec3(n, r, r, (float)(-alpha), Ap, is_dynamic);
            double rn = dot_atomic(n, r, r, is_dynamic);
            it++;
            if ((rn < 1e-5 && it >= 2) || it >= 1000) break;
            w18_apply_precond(n, Pinv, r, z, is_dynamic, precond);
            double rz_new = dot_atomic(n, r, z, is_dynamic);
            double beta = rz_new / (rz_old > 1e-12 ? rz_old : 1e-12);
            add_vec3(n, p, z, (float)beta, p, is_dynamic);
            rz_old = rz_new;
        }
    }
    add_vec3(n, v, v, 1.0f, dv, is_dynamic);
    free(Pinv); free(r); free(z); free(p); free(Ap); free(rhs);
    return it;
}

/* the reference's defaults: cold start, block-Jacobi preconditioner */
int orc_weiler18(int n, const float *x, float *v, const float *m, const float *density,
                 const float *density0, const int *material, const int *is_dynamic, float h, int C,
                 const int *count, const int *ids, const float *gradW, float coeff, float dt,
                 int material_fluid)
{
    float *dv = (float *)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(float));
    int it = orc_weiler18_ex(n, x, v, m, density, density0, material, is_dynamic, h, C, count, ids, gradW, coeff,
                             dt, material_fluid, 0, 1, dv);
    free(dv);
    return it;
}

/* ------------------------------------------------------------------------------------------ */
/* Jeske23 implicit surface tension, jeske23.py                                                */
/* ------------------------------------------------------------------------------------------ */

/* compute_Wst, jeske23.py:23-43: theta * W_cubic(max(r, d); h_st), theta = 10/7 */
float orc_wst(float r_ij, float d, float h_st)
{
    float theta = 10.0f / 7.0f;
    float r_c = r_ij > d ? r_ij : d;
    float q = r_c / h_st;
    float norm = 8.0f / (3.14159265358979f * h_st * h_st * h_st);
    float fq = 0.0f;
    if (q <= 0.5f) {
        fq = 1.0f - 6.0f * q * q + 6.0f * q * q * q;
    } else if (q <= 1.0f) {
        float t = 1.0f - q;
        fq = 2.0f * t * t * t;
    }
    return theta * norm * fq;
}

/* compute_grad_Wst, jeske23.py:46-66: 0 for r <= d (the kernel is flat there) and r > h_st */
void orc_grad_wst(const float x_ij[3], float r_ij, float d, float h_st, float out[3])
{
    float theta = 10.0f / 7.0f;
    out[0] = out[1] = out[2] = 0.0f;
    if (r_ij <= d || r_ij < 1.0e-12f) return;
    float q = r_ij / h_st;
    if (q > 1.0f) return;
    float norm = 8.0f / (3.14159265358979f * h_st * h_st * h_st);
    float fpq;
    if (q <= 0.5f) {
        fpq = -12.0f * q + 18.0f * q * q;
    } else {
        float t = 1.0f - q;
        fpq = -6.0f * t * t;
    }
    float s = theta * norm * (1.0f / h_st) * fpq;
    out[0] = s * (x_ij[0] / r_ij);
    out[1] = s * (x_ij[1] / r_ij);
    out[2] = s * (x_ij[2] / r_ij);
}

typedef struct {
    int n, C;
    const float *m;
    const int *material, *is_dynamic, *count, *ids;
    const float *g_tilde;
    float sigma, dt;
    int fluid;
} st_op;

/* compute_Av_jeske23, jeske23.py:245-295: (M v)_i = v_i + dt^2 (sigma / m_i) sum_j m_bar g~ v_ij */
static void st_Av(const st_op *o, const float *vin, float *Av)
{
    for (int i = 0; i < o->n; ++i) {
        float *out = Av + 3 * (size_t)i;
        const float *vi = vin + 3 * i;
        if (o->is_dynamic[i] == 0 || o->material[i] != o->fluid) {
            out[0] = vi[0]; out[1] = vi[1]; out[2] = vi[2];
            continue;
        }
        float m_i = o->m[i];
        float dt2 = o->dt * o->dt;
        float factor = dt2 * o->sigma / m_i;
        float sum[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)o->C;
        int cnt = capped(o->count, i, o->C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = o->ids[base + (size_t)kk];
            if (o->material[j] != o->fluid) continue;
            float g_t = o->g_tilde[base + (size_t)kk];
            if (g_t == 0.0f) continue;
            float m_bar = 0.5f * (m_i + o->m[j]);
            float w = m_bar * g_t;
            for (int a = 0; a < 3; ++a) sum[a] = sum[a] + w * (vi[a] - vin[3 * j + a]);
        }
        for (int a = 0; a < 3; ++a) out[a] = vi[a] + factor * sum[a];
    }
}

/* compute_precond_jeske23, jeske23.py:298-350: diagonal, 1 / max(1 + dt^2 sigma/m_i sum m_bar g~, 1e-12) */
static void st_precond(const st_op *o, float *Pinv)
{
    for (int i = 0; i < o->n; ++i) {
        float *P = Pinv + 9 * (size_t)i;
        memset(P, 0, sizeof(float) * 9);
        if (o->is_dynamic[i] == 0 || o->material[i] != o->fluid) {
            P[0] = P[4] = P[8] = 1.0f;
            continue;
        }
        float m_i = o->m[i];
        float dt2 = o->dt * o->dt;
        float factor = dt2 * o->sigma / m_i;
        float diag_sum = 0.0f;
        size_t base = (size_t)i * (size_t)o->C;
        int cnt = capped(o->count, i, o->C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = o->ids[base + (size_t)kk];
            if (o->material[j] != o->fluid) continue;
            float m_bar = 0.5f * (m_i + o->m[j]);
            diag_sum = diag_sum + m_bar * o->g_tilde[base + (size_t)kk];
        }
        float diag_val = 1.0f + factor * diag_sum;
        if (diag_val < 1.0e-12f) diag_val = 1.0e-12f;
        float inv_val = 1.0f / diag_val;
        P[0] = P[4] = P[8] = inv_val;
    }
}

/* compute_divergence_jeske23 + compute_g_tilde_jeske23 + compute_rhs_jeske23
 * (jeske23.py:73-242): div (cached spiky gradient), g~ per pair, b */
void orc_jeske23_setup(int n, const float *x, const float *v, const float *m, const float *density,
                       const int *material, const int *is_dynamic, int C, const int *count, const int *ids,
                       const float *gradW, float sigma, float dt, float d, float h_st, int fluid,
                       float *div, float *g_tilde, float *b)
{
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0 || material[i] != fluid) { div[i] = 0.0f; continue; }
        const float *vi = v + 3 * i;
        float div_val = 0.0f;
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            if (material[j] != fluid) continue;
            float v_ij[3] = {vi[0] - v[3 * j], vi[1] - v[3 * j + 1], vi[2] - v[3 * j + 2]};
            div_val = div_val + m[j] * dot3(v_ij, gradW + 3 * (base + (size_t)kk));
        }
        div[i] = div_val;
    }
    for (int i = 0; i < n; ++i) {
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        if (is_dynamic[i] == 0 || material[i] != fluid) {
            for (int kk = 0; kk < cnt; ++kk) g_tilde[base + (size_t)kk] = 0.0f;
            continue;
        }
        const float *xi = x + 3 * i, *vi = v + 3 * i;
        float rho_i = density[i], div_i = div[i];
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            float *gt = g_tilde + base + (size_t)kk;
            if (material[j] != fluid) { *gt = 0.0f; continue; }
            float x_ij[3] = {xi[0] - x[3 * j], xi[1] - x[3 * j + 1], xi[2] - x[3 * j + 2]};
            float r_ij = len3(x_ij);
            if (r_ij >= h_st) { *gt = 0.0f; continue; }
            float v_ij[3] = {vi[0] - v[3 * j], vi[1] - v[3 * j + 1], vi[2] - v[3 * j + 2]};
            float rho_bar = 0.5f * (rho_i + density[j]);
            float wst = orc_wst(r_ij, d, h_st);
            float gwst[3];
            orc_grad_wst(x_ij, r_ij, d, h_st, gwst);
            float g_ij = wst / rho_bar;
            float term_kernel = dot3(gwst, v_ij) / rho_bar;
            float term_density = wst / (2.0f * rho_bar * rho_bar) * (div_i + div[j]);
            *gt = g_ij + dt * (term_kernel - term_density);
        }
    }
    for (int i = 0; i < n; ++i) {
        float *bi = b + 3 * (size_t)i;
        bi[0] = bi[1] = bi[2] = 0.0f;
        if (is_dynamic[i] == 0 || material[i] != fluid) continue;
        const float *xi = x + 3 * i;
        float m_i = m[i];
        float factor = -dt * sigma / m_i;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        size_t base = (size_t)i * (size_t)C;
        int cnt = capped(count, i, C);
        for (int kk = 0; kk < cnt; ++kk) {
            int j = ids[base + (size_t)kk];
            if (material[j] != fluid) continue;
            float g_t = g_tilde[base + (size_t)kk];
            if (g_t == 0.0f) continue;
            float m_bar = 0.5f * (m_i + m[j]);
            for (int a = 0; a < 3; ++a) acc[a] = acc[a] + m_bar * (xi[a] - x[3 * j + a]) * g_t;
        }
        for (int a = 0; a < 3; ++a) bi[a] = factor * acc[a];
    }
}

void orc_jeske23_apply_M(int n, const float *m, const int *material, const int *is_dynamic, int C,
                         const int *count, const int *ids, const float *g_tilde, float sigma, float dt,
                         int fluid, const float *vin, float *Mv, float *Pinv)
{
    st_op o = {n, C, m, material, is_dynamic, count, ids, g_tilde, sigma, dt, fluid};
    if (vin && Mv) st_Av(&o, vin, Mv);
    if (Pinv) st_precond(&o, Pinv);
}

/* Jeske23.solve + Jeske23.PCG (jeske23.py:458-655): correction form M dv = v - M v + b from
 * dv = 0, diagonal preconditioner, tol 1e-5 on r.r, at most 100 iterations; v += dv.
 * Returns the CG iteration count. */
int orc_jeske23(int n, const float *x, float *v, const float *m, const float *density, const int *material,
                const int *is_dynamic, int C, const int *count, const int *ids, const float *gradW,
                float sigma, float dt, float d, float h_st, int fluid)
{
    size_t n1 = (size_t)(n > 0 ? n : 1), n3 = 3 * n1;
    float *div = (float *)calloc(n1, sizeof(float));
    float *gt = (float *)calloc(n1 * (size_t)(C > 0 ? C : 1), sizeof(float));
    float *bst = (float *)calloc(n3, sizeof(float));
    float *Pinv = (float *)calloc(9 * n1, sizeof(float));
    float *dv = (float *)calloc(n3, sizeof(float));
    float *r = (float *)calloc(n3, sizeof(float));
    float *z = (float *)calloc(n3, sizeof(float));
    float *p = (float *)calloc(n3, sizeof(float));
    float *Ap = (float *)calloc(n3, sizeof(float));
    float *rhs = (float *)calloc(n3, sizeof(float));
    orc_jeske23_setup(n, x, v, m, density, material, is_dynamic, C, count, ids, gradW, sigma, dt, d, h_st, fluid,
                      div, gt, bst);
    st_op o = {n, C, m, material, is_dynamic, count, ids, gt, sigma, dt, fluid};
    st_precond(&o, Pinv);
    st_Av(&o, v, Ap);
    add_vec3(n, rhs, v, -1.0f, Ap, is_dynamic);
    add_vec3(n, rhs, rhs, 1.0f, bst, is_dynamic);
    memcpy(r, rhs, sizeof(float) * 3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        const float *P = Pinv + 9 * (size_t)i;
        const float *ri = r + 3 * i;
        for (int a = 0; a < 3; ++a) z[3 * i + a] = P[3 * a] * ri[0] + P[3 * a + 1] * ri[1] + P[3 * a + 2] * ri[2];
    }
    memcpy(p, z, sizeof(float) * 3 * (size_t)n);
    double rz_old = dot_atomic(n, r, z, is_dynamic);
    int it = 0;
    if (rz_old > 1e-12) {
        for (int s = 0; s < 100; ++s) {
            st_Av(&o, p, Ap);
            double pAp = dot_atomic(n, p, Ap, is_dynamic);
            if (pAp <= 0.0) break;
            double alpha = rz_old / (pAp > 1e-12 ? pAp : 1e-12);
            add_vec3(n, dv, dv, (float)alpha, p, is_dynamic);
            add_vec3(n, r, r, (float)(-alpha), Ap, is_dynamic);
            double rn = dot_atomic(n, r, r, is_dynamic);
            it++;
            if ((rn < 1e-5 && it >= 2) || it >= 100) break;
            for (int i = 0; i < n; ++i) {
                if (is_dynamic[i] == 0) continue;
                const float *P = Pinv + 9 * (size_t)i;
                const float *ri = r + 3 * i;
                for (int a = 0; a < 3; ++a) z[3 * i + a] = P[3 * a] * ri[0] + P[3 * a + 1] * ri[1] + P[3 * a + 2] * ri[2];
            }
            double rz_new = dot_atomic(n, r, z, is_dynamic);
            double beta = rz_new / (rz_old > 1e-12 ? rz_old : 1e-12);
            add_vec3(n, p, z, (float)beta, p, is_dynamic);
            rz_old = rz_new;
        }
    }
    add_vec3(n, v, v, 1.0f, dv, is_dynamic);
    free(div); free(gt); free(bst); free(Pinv); free(dv); free(r); free(z); free(p); free(Ap); free(rhs);
    return it;
}

/* ------------------------------------------------------------------------------------------ */
/* Step glue: framework.py, boundary.py                                                        */
/* ------------------------------------------------------------------------------------------ */

/* apply_gravity, framework.py:26-41 */
void orc_gravity(int n, float *v, const int *is_dynamic, const float g[3], float dt)
{
    float gd[3] = {g[0] * dt, g[1] * dt, g[2] * dt};
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        for (int a = 0; a < 3; ++a) v[3 * i + a] = v[3 * i + a] + gd[a];
    }
}

/* advect_position, framework.py:117-136 */
void orc_advect(int n, float *x, float *x_old, float *v, const int *is_dynamic, float dt)
{
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (x_old) for (int a = 0; a < 3; ++a) x_old[3 * i + a] = x[3 * i + a];
        if (is_dynamic[i] == 0) v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0f;
        for (int a = 0; a < 3; ++a) x[3 * i + a] = x[3 * i + a] + dt * v[3 * i + a];
    }
}

/* enforce_boundary, boundary.py:27-91 (domain assumed to start at the origin, as in :59-86) */
void orc_enforce_boundary(int n, float *x, float *v, const int *is_dynamic,
                          const float domain_size[3], float padding)
{
    ORC_FOR
    for (int i = 0; i < n; ++i) {
        if (is_dynamic[i] == 0) continue;
        float *p = x + 3 * i;
        float nrm[3] = {0.0f, 0.0f, 0.0f};
        for (int a = 0; a < 3; ++a) {
            if (p[a] > domain_size[a] - padding) { nrm[a] = nrm[a] + 1.0f; p[a] = domain_size[a] - padding; }
            if (p[a] <= padding) { nrm[a] = nrm[a] + -1.0f; p[a] = padding; }
        }
        float L = len3(nrm);
        if (L > 1e-6f) {
            float cn[3] = {nrm[0] / L, nrm[1] / L, nrm[2] / L};
            float *vi = v + 3 * i;
            float c_f = 0.5f;
            float s = (1.0f + c_f) * dot3(vi, cn);
            float nv[3] = {vi[0] - s * cn[0], vi[1] - s * cn[1], vi[2] - s * cn[2]};
            vi[0] = nv[0]; vi[1] = nv[1]; vi[2] = nv[2];
        }
    }
}
