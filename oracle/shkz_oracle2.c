/*
 * shkz_oracle2.c -- CPU restatement of Shiokaze's 2D MAC pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY (see shkz_oracle.h).  It restates
 *   macpressuresolver2::project   src/projection/macpressuresolver2.cpp:48-232
 *   macutility2 face fractions    src/utility/macutility2.cpp:89-150
 * on dense arrays (y fastest): cells (w,h) idx = i*h + j; u (w+1,h); v (w,h+1);
 * nodal solid (w+1,h+1).  The linear solve is oracle_pcg (the same pcg module).
 */
#include "shkz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now2_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static inline size_t c2(const oracle_grid2 *g, int i, int j) { return (size_t)i * g->h + j; }
static inline size_t f2(const oracle_grid2 *g, int dim, int i, int j) { return (size_t)i * (g->h + (dim == 1)) + j; }
static inline size_t n2(const oracle_grid2 *g, int i, int j) { return (size_t)i * (g->h + 1) + j; }
static inline size_t fcount2(const oracle_grid2 *g, int dim) { return (size_t)(g->w + (dim == 0)) * (g->h + (dim == 1)); }
static inline double phi2(const oracle_grid2 *g, size_t c)
{
    return g->fluid_f64 ? ((const double *)g->fluid)[c] : (double)((const float *)g->fluid)[c];
}
static inline double solid2(const oracle_grid2 *g, size_t n)
{
    return g->solid_f64 ? ((const double *)g->solid)[n] : (double)((const float *)g->solid)[n];
}
static inline int clamp2(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int oob2(const oracle_grid2 *g, int i, int j) { return i < 0 || j < 0 || i >= g->w || j >= g->h; }

/* array_utility2 levelset_exist: a negative value somewhere */
static int exist_fluid2(const oracle_grid2 *g)
{
    const size_t n = (size_t)g->w * g->h;
    for (size_t c = 0; c < n; ++c)
        if (phi2(g, c) < 0.0) return 1;
    return 0;
}
static int exist_solid2(const oracle_grid2 *g)
{
    if (!g->solid) return 0;
    const size_t n = (size_t)(g->w + 1) * (g->h + 1);
    for (size_t c = 0; c < n; ++c)
        if (solid2(g, c) < 0.0) return 1;
    return 0;
}

/* macutility2.cpp:89-150 (area: 1 - fraction of the two nodal solid samples of the face, 0 on the
 * domain boundary; rho: fraction of the two clamped cell samples) and the first-order overrides
 * of macpressuresolver2.cpp:62-73. */
void oracle2_face_fractions(const oracle_grid2 *g, const oracle_params *p, float *area[2], float *rho[2])
{
    const int N[2] = {g->w, g->h};
    const int has_solid = exist_solid2(g), has_fluid = exist_fluid2(g);
    for (int dim = 0; dim < 2; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1);
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j) {
                const int pi[2] = {i, j};
                const size_t f = f2(g, dim, i, j);
                double a;
                if (has_solid) {
                    if (pi[dim] == 0 || pi[dim] == N[dim]) a = 0.0;
                    else a = 1.0 - oracle_fraction(solid2(g, n2(g, i, j)), solid2(g, n2(g, i + (dim != 0), j + (dim != 1))));
                    if (a && a < p->eps_solid) a = p->eps_solid;
                } else {
                    a = (pi[dim] == 0 || pi[dim] == N[dim]) ? 0.0 : 1.0;
                }
                float af = (float)a;
                if (!p->second_order_solid && af) af = 1.0f;
                area[dim][f] = af;
                double r;
                if (has_fluid) {
                    r = oracle_fraction(phi2(g, c2(g, clamp2(i, 0, g->w - 1), clamp2(j, 0, g->h - 1))),
                                        phi2(g, c2(g, clamp2(i - (dim == 0), 0, g->w - 1), clamp2(j - (dim == 1), 0, g->h - 1))));
                    if (r && r < p->eps_fluid) r = p->eps_fluid;
                } else {
                    r = 1.0;
                }
                float rf = (float)r;
                if (!p->second_order_fluid && rf) rf = 1.0f;
                rho[dim][f] = rf;
            }
    }
}

/* macpressuresolver2.cpp:75-103: 5-point Laplacian curvature (Real sums), face impulse. */
void oracle2_surface_tension(const oracle_grid2 *g, double kappa, float *rho[2])
{
    const size_t M = (size_t)g->w * g->h;
    float *curv = (float *)malloc((M ? M : 1) * sizeof(float));
    const double dx = g->dx;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j) {
            const int im = clamp2(i - 1, 0, g->w - 1), ip = clamp2(i + 1, 0, g->w - 1);
            const int jm = clamp2(j - 1, 0, g->h - 1), jp = clamp2(j + 1, 0, g->h - 1);
            double value;
            if (g->fluid_f64) {
                const double *f = (const double *)g->fluid;
                double s = +f[c2(g, im, j)] + f[c2(g, ip, j)] + f[c2(g, i, jm)] + f[c2(g, i, jp)];
                value = (s - 4.0 * f[c2(g, i, j)]) / (dx * dx);
            } else {
                const float *f = (const float *)g->fluid;
                float s = +f[c2(g, im, j)] + f[c2(g, ip, j)] + f[c2(g, i, jm)] + f[c2(g, i, jp)];
                value = (s - 4.0 * f[c2(g, i, j)]) / (dx * dx);
            }
            curv[c2(g, i, j)] = (float)value;
        }
    for (int dim = 0; dim < 2; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1);
        float *U = g->vel[dim];
        const uint8_t *act = g->active[dim];
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j) {
                const size_t f = f2(g, dim, i, j);
                if (act && !act[f]) continue;
                const double r = rho[dim][f];
                if (r && r < 1.0) {
                    const int i0 = clamp2(i, 0, g->w - 1), j0 = clamp2(j, 0, g->h - 1);
                    const int i1 = clamp2(i - (dim == 0), 0, g->w - 1), j1 = clamp2(j - (dim == 1), 0, g->h - 1);
                    const double sgn = phi2(g, c2(g, i0, j0)) < 0.0 ? -1.0 : 1.0;
                    const double theta = sgn < 0 ? 1.0 - r : r;
                    const double face_c = theta * curv[c2(g, i0, j0)] + (1.0 - theta) * curv[c2(g, i1, j1)];
                    U[f] += (float)(-sgn * g->dt / (dx * r) * kappa * face_c);
                }
            }
    }
    free(curv);
}

/* Neighbour order of macpressuresolver2.cpp:108-110,152-155: +x,-x,+y,-y; the face of a
 * +direction neighbour is indexed at the neighbour, of a -direction one at the cell. */
static const int D2I[4] = {1, -1, 0, 0};
static const int D2J[4] = {0, 0, 1, -1};
static const int D2DIM[4] = {0, 0, 1, 1};
static const int D2SGN[4] = {1, -1, 1, -1};

static inline size_t nbf2(const oracle_grid2 *g, int nq, int i, int j)
{
    if (D2SGN[nq] > 0) return f2(g, D2DIM[nq], i + D2I[nq], j + D2J[nq]);
    return f2(g, D2DIM[nq], i, j);
}

/* macpressuresolver2.cpp:105-143 (serial numbering, y fastest here). */
int64_t oracle2_label(const oracle_grid2 *g, float *area[2], float *rho[2], int64_t *index_map)
{
    int64_t index = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j) {
            const size_t c = c2(g, i, j);
            int inside = 0;
            if (phi2(g, c) < 0.0) {
                for (int nq = 0; nq < 4; ++nq) {
                    const int qi = i + D2I[nq], qj = j + D2J[nq];
                    if (oob2(g, qi, qj)) continue;
                    if (phi2(g, c2(g, qi, qj)) < 0.0) {
                        const size_t f = nbf2(g, nq, i, j);
                        const int dim = D2DIM[nq];
                        if (area[dim][f] && rho[dim][f]) {
                            inside = 1;
                            break;
                        }
                    }
                }
            }
            index_map[c] = inside ? index++ : -1;
        }
    return index;
}

static int row_entries2(const oracle_grid2 *g, float *area[2], float *rho[2], int i, int j)
{
    int cnt = 1;
    for (int nq = 0; nq < 4; ++nq) {
        const int qi = i + D2I[nq], qj = j + D2J[nq];
        if (oob2(g, qi, qj)) continue;
        const size_t f = nbf2(g, nq, i, j);
        const int dim = D2DIM[nq];
        if (area[dim][f] && rho[dim][f] && phi2(g, c2(g, qi, qj)) < 0.0) ++cnt;
    }
    return cnt;
}

int64_t oracle2_assemble_count(const oracle_grid2 *g, const int64_t *index_map, float *area[2], float *rho[2])
{
    int64_t nnz = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            if (index_map[c2(g, i, j)] >= 0) nnz += row_entries2(g, area, rho, i, j);
    return nnz;
}

/* macpressuresolver2.cpp:145-179.  Unlike the 3D module, the rhs term -sgn*area*u/dx is added
 * for every in-bounds neighbour, also where the area is 0 (:176). */
void oracle2_assemble(const oracle_grid2 *g, const int64_t *index_map, float *area[2], float *rho[2], int64_t *rowptr,
                      int64_t *col, double *val, double *b)
{
    const double dt = g->dt, dx = g->dx;
    int64_t n = 0, nnz = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j) {
            if (index_map[c2(g, i, j)] < 0) continue;
            rowptr[n++] = nnz;
            nnz += row_entries2(g, area, rho, i, j);
        }
    rowptr[n] = nnz;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j) {
            const int64_t row = index_map[c2(g, i, j)];
            if (row < 0) continue;
            int64_t cols[5];
            double vals[5];
            int cnt = 0;
            double rhs = 0.0, diagonal = 0.0;
            for (int nq = 0; nq < 4; ++nq) {
                const int dim = D2DIM[nq];
                const int qi = i + D2I[nq], qj = j + D2J[nq];
                if (oob2(g, qi, qj)) continue;
                const size_t f = nbf2(g, nq, i, j);
                const double a = area[dim][f];
                if (a) {
                    const double r = rho[dim][f];
                    if (r) {
                        const double value = dt * a / (dx * dx * r);
                        const size_t q = c2(g, qi, qj);
                        if (phi2(g, q) < 0.0) {
                            cols[cnt] = index_map[q];
                            vals[cnt] = -value;
                            ++cnt;
                        }
                        diagonal += value;
                    }
                }
                rhs += -D2SGN[nq] * a * (double)g->vel[dim][f] / dx;
            }
            cols[cnt] = row;
            vals[cnt] = diagonal;
            ++cnt;
            for (int a = 1; a < cnt; ++a) { /* rows kept sorted by column (RCMatrix.cpp:249-272) */
                const int64_t cc = cols[a];
                const double vv = vals[a];
                int bb = a - 1;
                while (bb >= 0 && cols[bb] > cc) {
                    cols[bb + 1] = cols[bb];
                    vals[bb + 1] = vals[bb];
                    --bb;
                }
                cols[bb + 1] = cc;
                vals[bb + 1] = vv;
            }
            const int64_t base = rowptr[row];
            for (int a = 0; a < cnt; ++a) {
                col[base + a] = cols[a];
                val[base + a] = vals[a];
            }
            b[row] = rhs;
        }
}

/* macpressuresolver2.cpp:212-230.  Open faces (area and rho nonzero) on the domain boundary
 * become 0, interior ones lose dt*(P(c)-P(c-e))/(rho*dx) (float difference of float pressures);
 * a closed boundary face with liquid inside stays active at 0 only while its velocity points
 * out of the domain; every other face is switched off (reads 0).  active_out (optional)
 * receives the activity after the update for the faces that were active. */
void oracle2_velocity_update(const oracle_grid2 *g, const float *P, float *area[2], float *rho[2],
                             uint8_t *active_out[2])
{
    const int N[2] = {g->w, g->h};
    const double dt = g->dt, dx = g->dx;
    for (int dim = 0; dim < 2; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1);
        float *U = g->vel[dim];
        const uint8_t *act = g->active[dim];
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j) {
                const int pi[2] = {i, j};
                const size_t f = f2(g, dim, i, j);
                if (act && !act[f]) continue;
                const double r = rho[dim][f];
                int keep = 1;
                if (area[dim][f] && r) {
                    if (pi[dim] == 0 || pi[dim] == N[dim]) U[f] = 0.0f;
                    else {
                        const float diff = P[c2(g, i, j)] - P[c2(g, i - (dim == 0), j - (dim == 1))];
                        U[f] -= (float)(dt * diff / (r * dx));
                    }
                } else {
                    if (pi[dim] == 0 && phi2(g, c2(g, i, j)) < 0.0 && U[f] < 0.0f) U[f] = 0.0f;
                    else if (pi[dim] == N[dim] && phi2(g, c2(g, i - (dim == 0), j - (dim == 1))) < 0.0 && U[f] > 0.0f)
                        U[f] = 0.0f;
                    else {
                        U[f] = 0.0f;
                        keep = 0;
                    }
                }
                if (active_out) active_out[dim][f] = (uint8_t)keep;
            }
    }
}

double oracle2_divergence(const oracle_grid2 *g, const int64_t *index_map, float *area[2], double *div)
{
    double mx = 0.0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j) {
            const size_t c = c2(g, i, j);
            if (index_map[c] < 0) {
                if (div) div[c] = 0.0;
                continue;
            }
            double s = 0.0;
            for (int nq = 0; nq < 4; ++nq) {
                const int dim = D2DIM[nq];
                if (oob2(g, i + D2I[nq], j + D2J[nq])) continue;
                const size_t f = nbf2(g, nq, i, j);
                const double a = area[dim][f];
                if (a) s += D2SGN[nq] * a * (double)g->vel[dim][f] / g->dx;
            }
            if (div) div[c] = s;
            if (fabs(s) > mx) mx = fabs(s);
        }
    return mx;
}

/* macpressuresolver2::project (macpressuresolver2.cpp:48-232). */
int oracle2_project(const oracle_grid2 *g, const oracle_params *p, float *pressure, double *pressure64,
                    oracle_result *res)
{
    memset(res, 0, sizeof(*res));
    const double t0 = now2_s();
    const size_t M = (size_t)g->w * g->h;
    float *area[2], *rho[2];
    for (int dim = 0; dim < 2; ++dim) {
        area[dim] = (float *)malloc(fcount2(g, dim) * sizeof(float));
        rho[dim] = (float *)malloc(fcount2(g, dim) * sizeof(float));
    }
    oracle2_face_fractions(g, p, area, rho);
    if (p->surface_tension) oracle2_surface_tension(g, p->surface_tension, rho);
    const double t1 = now2_s();
    int64_t *index_map = (int64_t *)malloc((M ? M : 1) * sizeof(int64_t));
    const int64_t n = oracle2_label(g, area, rho, index_map);
    const int64_t nnz = oracle2_assemble_count(g, index_map, area, rho);
    int64_t *rowptr = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t *col = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    double *val = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    double *b = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *x = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    oracle2_assemble(g, index_map, area, rho, rowptr, col, val, b);
    if (p->rhs_correct)
        for (int64_t r = 0; r < n; ++r) b[r] += p->rhs_correct;
    double *prev = NULL;
    if (p->warm_start && n) { /* :193-203, cell-indexed */
        prev = (double *)malloc((size_t)n * sizeof(double));
        for (size_t c = 0; c < M; ++c)
            if (index_map[c] >= 0) prev[index_map[c]] = p->warm_start[c];
        double *Ap = (double *)malloc((size_t)n * sizeof(double));
        oracle_spmv(n, rowptr, col, val, prev, Ap);
        for (int64_t r = 0; r < n; ++r) b[r] -= Ap[r];
        free(Ap);
    }
    const double t2 = now2_s();
    int iters = 0;
    double reresid = 0;
    const int ok = oracle_pcg(n, rowptr, col, val, b, x, p->residual, p->max_iterations, p->mic_mode, p->mic_omega,
                              p->mic_min_diag_ratio, &iters, &reresid);
    if (prev) { /* :205-208 */
        for (int64_t r = 0; r < n; ++r) x[r] += prev[r];
        free(prev);
    }
    const double t3 = now2_s();
    for (size_t c = 0; c < M; ++c) { /* :210-213 */
        const int64_t r = index_map[c];
        pressure[c] = r >= 0 ? (float)x[r] : 0.0f;
        if (pressure64) pressure64[c] = r >= 0 ? x[r] : 0.0;
    }
    oracle2_velocity_update(g, pressure, area, rho, NULL);
    const double t4 = now2_s();
    res->iterations = iters;
    res->reresid = reresid;
    res->converged = ok;
    res->n_unknowns = n;
    res->t_fractions = t1 - t0;
    res->t_build = t2 - t1;
    res->t_linsolve = t3 - t2;
    res->t_update = t4 - t3;
    res->t_projection = t4 - t0;
    free(index_map);
    free(rowptr);
    free(col);
    free(val);
    free(b);
    free(x);
    for (int dim = 0; dim < 2; ++dim) {
        free(area[dim]);
        free(rho[dim]);
    }
    return ok;
}
