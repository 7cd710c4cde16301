/*
 * shkz_oracle.c -- CPU restatement of Shiokaze's MAC pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function cites the
 * reference lines it restates; the arithmetic keeps the reference's operand order and
 * storage precision (Real = float fields, FP64 linear system) so that fractions, the
 * assembled operator and the right-hand side are bit-identical to the reference's.
 *
 * Threading follows the reference: fractions, assembly and the velocity update are
 * parallel loops (std::thread per loop in the reference, OpenMP here); labelling, the
 * matrix copy, the CG loop and the scatter are serial.
 */
#include "shkz_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int nthreads(const oracle_params *p)
{
#ifdef _OPENMP
    return (p && p->threads > 0) ? p->threads : omp_get_max_threads();
#else
    (void)p;
    return 1;
#endif
}

/* ---------------------------------------------------------------- indexing */
static inline size_t cidx(const oracle_grid *g, int i, int j, int k)
{
    return ((size_t)i * (size_t)g->h + (size_t)j) * (size_t)g->d + (size_t)k;
}
static inline size_t fidx(const oracle_grid *g, int dim, int i, int j, int k)
{
    size_t H = (size_t)g->h + (dim == 1), D = (size_t)g->d + (dim == 2);
    return ((size_t)i * H + (size_t)j) * D + (size_t)k;
}
static inline size_t nidx(const oracle_grid *g, int i, int j, int k)
{
    return ((size_t)i * (size_t)(g->h + 1) + (size_t)j) * (size_t)(g->d + 1) + (size_t)k;
}
static inline size_t face_count(const oracle_grid *g, int dim)
{
    return (size_t)(g->w + (dim == 0)) * (size_t)(g->h + (dim == 1)) * (size_t)(g->d + (dim == 2));
}
static inline double phi_at(const oracle_grid *g, size_t c)
{
    return g->fluid_f64 ? ((const double *)g->fluid)[c] : (double)((const float *)g->fluid)[c];
}
static inline double solid_at(const oracle_grid *g, size_t n)
{
    return g->solid_f64 ? ((const double *)g->solid)[n] : (double)((const float *)g->solid)[n];
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int out_of_bounds(const oracle_grid *g, int i, int j, int k)
{
    return i < 0 || i >= g->w || j < 0 || j >= g->h || k < 0 || k >= g->d; /* shape.h:813-815 */
}

void oracle_default_params(oracle_params *p)
{
    memset(p, 0, sizeof(*p));
    p->residual = 1e-4;
    p->max_iterations = 30000;
    p->mic_mode = ORACLE_MIC_FAITHFUL;
    p->mic_omega = 0.97;
    p->mic_min_diag_ratio = 0.25;
    p->eps_fluid = 1e-2;
    p->eps_solid = 1e-2;
    p->second_order_fluid = 1;
    p->second_order_solid = 1;
}

/* ------------------------------------------------------- utility.h helpers */

/* utility.h:162-170: liquid fraction of a face from the two cell level-set values. */
double oracle_fraction(double phi0, double phi1)
{
    if (phi0 * phi1 >= 0.0) {
        return (phi0 < 0.0 || phi1 < 0.0) ? 1.0 : 0.0;
    }
    double lo = phi0 < phi1 ? phi0 : phi1;
    double gap = fabs(phi1 - phi0);
    if (gap < DBL_MIN) gap = DBL_MIN;
    return -lo / gap;
}

/* utility.h:179-214: area of the negative region of a bilinear-free (piecewise-linear on
 * edges) unit square, walking the corners (0,0),(1,0),(1,1),(0,1), emitting every negative
 * corner and every sign-change crossing, then the shoelace sum. */
double oracle_get_area(const double q[4])
{
    static const int cx[4] = {0, 1, 1, 0};
    static const int cy[4] = {0, 0, 1, 1};
    double v[4];
    for (int n = 0; n < 4; ++n) v[n] = q[cx[n] * 2 + cy[n]];
    double px[8], py[8];
    int m = 0;
    for (int n = 0; n < 4; ++n) {
        int nn = (n + 1) & 3;
        if (v[n] < 0.0) {
            px[m] = cx[n];
            py[m] = cy[n];
            ++m;
        }
        if (v[n] * v[nn] < 0) {
            double y0 = v[n], y1 = v[nn];
            if (y0 - y1) {
                double a = y0 / (y0 - y1);
                px[m] = (1.0 - a) * cx[n] + a * cx[nn];
                py[m] = (1.0 - a) * cy[n] + a * cy[nn];
                ++m;
            }
        }
    }
    double sum = 0.0;
    for (int a = 0; a < m; ++a) {
        int b = (a + 1) % m;
        sum += px[a] * py[b] - py[a] * px[b];
    }
    return 0.5 * sum;
}

/* utility.h:113-120 */
double oracle_box_sdf(const double p[3], const double p0[3], const double p1[3])
{
    double sd = -9999.0;
    for (int dim = 0; dim < 3; ++dim) {
        if (p0[dim] - p[dim] > sd) sd = p0[dim] - p[dim];
        if (p[dim] - p1[dim] > sd) sd = p[dim] - p1[dim];
    }
    return sd;
}

/* ------------------------------------------------------------- fractions */

/* array_utility3.h:112-122: a level set "exists" when it holds a negative value. */
static int levelset_exist_fluid(const oracle_grid *g)
{
    size_t n = (size_t)g->w * g->h * g->d;
    for (size_t c = 0; c < n; ++c)
        if (phi_at(g, c) < 0.0) return 1;
    return 0;
}
static int levelset_exist_solid(const oracle_grid *g)
{
    if (!g->solid) return 0;
    size_t n = (size_t)(g->w + 1) * (g->h + 1) * (g->d + 1);
    for (size_t c = 0; c < n; ++c)
        if (solid_at(g, c) < 0.0) return 1;
    return 0;
}

/* macutility3.cpp:94-165 (area) and :166-194 (rho); macpressuresolver3.cpp:70-80 overrides.
 * The reference evaluates the sparse narrow band and flood-fills the rest; on a dense grid
 * the same values follow from evaluating every face, because get_area and fraction depend
 * only on signs wherever all samples lie outside the band. */
void oracle_face_fractions(const oracle_grid *g, const oracle_params *p, float *area[3], float *rho[3])
{
    const int N[3] = {g->w, g->h, g->d};
    const int has_solid = levelset_exist_solid(g);
    const int has_fluid = levelset_exist_fluid(g);
    const int nt = nthreads(p);
    for (int dim = 0; dim < 3; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1), D = g->d + (dim == 2);
        float *A = area[dim], *R = rho[dim];
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < W; ++i) {
            for (int j = 0; j < H; ++j) {
                for (int k = 0; k < D; ++k) {
                    const int pi[3] = {i, j, k};
                    const size_t f = fidx(g, dim, i, j, k);
                    /* ---- area */
                    double a;
                    if (has_solid) {
                        if (pi[dim] == 0 || pi[dim] == N[dim]) {
                            a = 0.0;
                        } else {
                            double q[4]; /* q[2*x + y] = quadsolid[x][y] */
                            if (dim == 0) {
                                q[0] = solid_at(g, nidx(g, i, j, k));
                                q[2] = solid_at(g, nidx(g, i, j + 1, k));
                                q[3] = solid_at(g, nidx(g, i, j + 1, k + 1));
                                q[1] = solid_at(g, nidx(g, i, j, k + 1));
                            } else if (dim == 1) {
                                q[0] = solid_at(g, nidx(g, i, j, k));
                                q[2] = solid_at(g, nidx(g, i + 1, j, k));
                                q[3] = solid_at(g, nidx(g, i + 1, j, k + 1));
                                q[1] = solid_at(g, nidx(g, i, j, k + 1));
                            } else {
                                q[0] = solid_at(g, nidx(g, i, j, k));
                                q[2] = solid_at(g, nidx(g, i + 1, j, k));
                                q[3] = solid_at(g, nidx(g, i + 1, j + 1, k));
                                q[1] = solid_at(g, nidx(g, i, j + 1, k));
                            }
                            a = 1.0 - oracle_get_area(q);
                        }
                        if (a && a < p->eps_solid) a = p->eps_solid;
                    } else {
                        a = (pi[dim] == 0 || pi[dim] == N[dim]) ? 0.0 : 1.0;
                    }
                    float af = (float)a;
                    if (!p->second_order_solid && af) af = 1.0f;
                    A[f] = af;
                    /* ---- rho */
                    double r;
                    if (has_fluid) {
                        int i0 = clampi(i, 0, g->w - 1), j0 = clampi(j, 0, g->h - 1), k0 = clampi(k, 0, g->d - 1);
                        int i1 = clampi(i - (dim == 0), 0, g->w - 1), j1 = clampi(j - (dim == 1), 0, g->h - 1),
                            k1 = clampi(k - (dim == 2), 0, g->d - 1);
                        r = oracle_fraction(phi_at(g, cidx(g, i0, j0, k0)), phi_at(g, cidx(g, i1, j1, k1)));
                        if (r && r < p->eps_fluid) r = p->eps_fluid;
                    } else {
                        r = 1.0;
                    }
                    float rf = (float)r;
                    if (!p->second_order_fluid && rf) rf = 1.0f;
                    R[f] = rf;
                }
            }
        }
    }
}

/* macpressuresolver3.cpp:85-115.  Curvature is the 7-point Laplacian of the (clamped) level
 * set, summed in the level set's own precision as the reference sums Real values. */
void oracle_surface_tension(const oracle_grid *g, double kappa, float *rho[3])
{
    const size_t M = (size_t)g->w * g->h * g->d;
    float *curv = (float *)malloc(M * sizeof(float));
    const double dx = g->dx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                int im = clampi(i - 1, 0, g->w - 1), ip = clampi(i + 1, 0, g->w - 1);
                int jm = clampi(j - 1, 0, g->h - 1), jp = clampi(j + 1, 0, g->h - 1);
                int km = clampi(k - 1, 0, g->d - 1), kp = clampi(k + 1, 0, g->d - 1);
                double value;
                if (g->fluid_f64) {
                    const double *f = (const double *)g->fluid;
                    double s = +f[cidx(g, im, j, k)] + f[cidx(g, ip, j, k)] + f[cidx(g, i, jm, k)] +
                               f[cidx(g, i, jp, k)] + f[cidx(g, i, j, km)] + f[cidx(g, i, j, kp)];
                    value = (s - 6.0 * f[cidx(g, i, j, k)]) / (dx * dx);
                } else {
                    const float *f = (const float *)g->fluid;
                    float s = +f[cidx(g, im, j, k)] + f[cidx(g, ip, j, k)] + f[cidx(g, i, jm, k)] +
                              f[cidx(g, i, jp, k)] + f[cidx(g, i, j, km)] + f[cidx(g, i, j, kp)];
                    value = (s - 6.0 * f[cidx(g, i, j, k)]) / (dx * dx);
                }
                curv[cidx(g, i, j, k)] = (float)value;
            }
    for (int dim = 0; dim < 3; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1), D = g->d + (dim == 2);
        float *U = g->vel[dim];
        const uint8_t *act = g->active[dim];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j)
                for (int k = 0; k < D; ++k) {
                    size_t f = fidx(g, dim, i, j, k);
                    if (act && !act[f]) continue;
                    double r = rho[dim][f];
                    if (r && r < 1.0) {
                        int i0 = clampi(i, 0, g->w - 1), j0 = clampi(j, 0, g->h - 1), k0 = clampi(k, 0, g->d - 1);
                        int i1 = clampi(i - (dim == 0), 0, g->w - 1), j1 = clampi(j - (dim == 1), 0, g->h - 1),
                            k1 = clampi(k - (dim == 2), 0, g->d - 1);
                        double sgn = phi_at(g, cidx(g, i0, j0, k0)) < 0.0 ? -1.0 : 1.0;
                        double theta = sgn < 0 ? 1.0 - r : r;
                        double face_c = theta * curv[cidx(g, i0, j0, k0)] + (1.0 - theta) * curv[cidx(g, i1, j1, k1)];
                        float inc = (float)(-sgn * g->dt / (dx * r) * kappa * face_c);
                        U[f] += inc;
                    }
                }
    }
    free(curv);
}

/* ------------------------------------------------------------ labelling */

/* Neighbour walk order of macpressuresolver3.cpp:128-130 / :170-173:
 * +x,-x,+y,-y,+z,-z; the face of a +direction neighbour is indexed at the neighbour cell,
 * the face of a -direction neighbour at the cell itself. */
static const int NB_DI[6] = {1, -1, 0, 0, 0, 0};
static const int NB_DJ[6] = {0, 0, 1, -1, 0, 0};
static const int NB_DK[6] = {0, 0, 0, 0, 1, -1};
static const int NB_DIM[6] = {0, 0, 1, 1, 2, 2};
static const int NB_SGN[6] = {1, -1, 1, -1, 1, -1};

static inline size_t nb_face(const oracle_grid *g, int nq, int i, int j, int k)
{
    if (NB_SGN[nq] > 0) return fidx(g, NB_DIM[nq], i + NB_DI[nq], j + NB_DJ[nq], k + NB_DK[nq]);
    return fidx(g, NB_DIM[nq], i, j, k);
}

/* macpressuresolver3.cpp:121-156.  A cell is an unknown when it is liquid and shares an open
 * face (area>0 and rho>0) with an in-bounds liquid neighbour.  Numbers are handed out in a
 * serial sweep (z fastest here; the reference walks its 16^3 tiles -- only summation order
 * in dot products depends on this). */
int64_t oracle_label(const oracle_grid *g, float *area[3], float *rho[3], int64_t *index_map)
{
    int64_t index = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                size_t c = cidx(g, i, j, k);
                int inside = 0;
                if (phi_at(g, c) < 0.0) {
                    for (int nq = 0; nq < 6; ++nq) {
                        int qi = i + NB_DI[nq], qj = j + NB_DJ[nq], qk = k + NB_DK[nq];
                        if (out_of_bounds(g, qi, qj, qk)) continue;
                        if (phi_at(g, cidx(g, qi, qj, qk)) < 0.0) {
                            size_t f = nb_face(g, nq, i, j, k);
                            int dim = NB_DIM[nq];
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

/* ------------------------------------------------------------- assembly */

int64_t oracle_assemble_count(const oracle_grid *g, const int64_t *index_map, float *area[3], float *rho[3])
{
    int64_t nnz = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                size_t c = cidx(g, i, j, k);
                if (index_map[c] < 0) continue;
                nnz += 1;
                for (int nq = 0; nq < 6; ++nq) {
                    int qi = i + NB_DI[nq], qj = j + NB_DJ[nq], qk = k + NB_DK[nq];
                    if (out_of_bounds(g, qi, qj, qk)) continue;
                    size_t f = nb_face(g, nq, i, j, k);
                    int dim = NB_DIM[nq];
                    if (area[dim][f] && rho[dim][f] && phi_at(g, cidx(g, qi, qj, qk)) < 0.0) nnz += 1;
                }
            }
    return nnz;
}

/* macpressuresolver3.cpp:159-199 (RCMatrix::add_to_element keeps each row sorted by column,
 * RCMatrix.cpp:249-272).  For an unknown c and each in-bounds neighbour q in the order
 * +x,-x,+y,-y,+z,-z:  if area>0 and rho>0, w = dt*area/(dx*dx*rho) goes to the diagonal and,
 * when q is liquid, -w to A[c,q]; for any area>0 the rhs gains -sgn*area*u/dx. */
void oracle_assemble(const oracle_grid *g, const int64_t *index_map, float *area[3], float *rho[3],
                     int64_t *rowptr, int64_t *col, double *val, double *b)
{
    const double dt = g->dt, dx = g->dx;
    /* row pointers: serial prefix over unknowns in index order */
    int64_t n = 0, nnz = 0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                size_t c = cidx(g, i, j, k);
                if (index_map[c] < 0) continue;
                int cnt = 1;
                for (int nq = 0; nq < 6; ++nq) {
                    int qi = i + NB_DI[nq], qj = j + NB_DJ[nq], qk = k + NB_DK[nq];
                    if (out_of_bounds(g, qi, qj, qk)) continue;
                    size_t f = nb_face(g, nq, i, j, k);
                    int dim = NB_DIM[nq];
                    if (area[dim][f] && rho[dim][f] && phi_at(g, cidx(g, qi, qj, qk)) < 0.0) ++cnt;
                }
                rowptr[n++] = nnz;
                nnz += cnt;
            }
    rowptr[n] = nnz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                size_t c = cidx(g, i, j, k);
                int64_t row = index_map[c];
                if (row < 0) continue;
                int64_t cols[7];
                double vals[7];
                int cnt = 0;
                double rhs = 0.0;
                double diagonal = 0.0;
                if (phi_at(g, c) < 0.0) {
                    for (int nq = 0; nq < 6; ++nq) {
                        int dim = NB_DIM[nq];
                        int qi = i + NB_DI[nq], qj = j + NB_DJ[nq], qk = k + NB_DK[nq];
                        if (out_of_bounds(g, qi, qj, qk)) continue;
                        size_t f = nb_face(g, nq, i, j, k);
                        double a = area[dim][f];
                        if (a) {
                            double r = rho[dim][f];
                            if (r) {
                                double value = dt * a / (dx * dx * r);
                                size_t q = cidx(g, qi, qj, qk);
                                if (phi_at(g, q) < 0.0) {
                                    cols[cnt] = index_map[q];
                                    vals[cnt] = -value;
                                    ++cnt;
                                }
                                diagonal += value;
                            }
                            rhs += -NB_SGN[nq] * a * (double)g->vel[dim][f] / dx;
                        }
                    }
                    cols[cnt] = row;
                    vals[cnt] = diagonal;
                    ++cnt;
                }
                /* insertion sort by column (at most 7 entries) */
                for (int a = 1; a < cnt; ++a) {
                    int64_t cc = cols[a];
                    double vv = vals[a];
                    int bb = a - 1;
                    while (bb >= 0 && cols[bb] > cc) {
                        cols[bb + 1] = cols[bb];
                        vals[bb + 1] = vals[bb];
                        --bb;
                    }
                    cols[bb + 1] = cc;
                    vals[bb + 1] = vv;
                }
                int64_t base = rowptr[row];
                for (int a = 0; a < cnt; ++a) {
                    col[base + a] = cols[a];
                    val[base + a] = vals[a];
                }
                b[row] = rhs;
            }
}

/* macpressuresolver3.cpp:207-211 */
double oracle_volume_correction(double current_volume, double target_volume, double dt, double gain, double *y_prev)
{
    if (!(gain && target_volume)) return 0.0;
    double x = (current_volume - target_volume) / target_volume;
    double y = *y_prev + x * dt;
    *y_prev = y;
    double kp = gain * 2.3 / (25.0 * 0.01);
    double ki = kp * kp / 16.0;
    return -(kp * x + ki * y) / (x + 1.0);
}

/* ------------------------------------------------------------ solver */

/* sparse_matrix.h:265-275: row sums in ascending column order. */
void oracle_spmv(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) {
        double s = 0;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) s += val[p] * x[col[p]];
        y[i] = s;
    }
}

static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double abs_max(int64_t n, const double *a)
{
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = fabs(a[i]);
        if (v > m) m = v;
    }
    return n ? m : 0.0;
}
static void add_scaled(int64_t n, double alpha, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) y[i] += alpha * x[i];
}

/* Incomplete Cholesky factor L (A ~ L L^T) with the pattern of A's lower triangle.  Column k of
 * L holds the entries below its diagonal, stored in row k's upper part (A is symmetric), in
 * ascending row order; invdiag[k] = 1 / L_kk. */
typedef struct {
    int64_t n;
    int64_t *colstart; /* n+1 */
    int64_t *rowindex;
    double *value;
    double *invdiag;
} mic_factor;

static void mic_free(mic_factor *f)
{
    free(f->colstart);
    free(f->rowindex);
    free(f->value);
    free(f->invdiag);
}

/* Position of row r in column c of the factor (rows ascending), -1 if the pattern lacks it. */
static int64_t mic_find(const mic_factor *F, int64_t c, int64_t r)
{
    for (int64_t q = F->colstart[c]; q < F->colstart[c + 1]; ++q) {
        if (F->rowindex[q] == r) return q;
        if (F->rowindex[q] > r) break;
    }
    return -1;
}

/* Modified incomplete Cholesky with zero fill-in, the factor pcg_solver.h:89-189 builds (omega =
 * 0.97, pivots below min_diag_ratio of the original diagonal replaced by the original diagonal).
 * Right-looking: column k takes its pivot from the running diagonal d, is scaled by 1/L_kk, and
 * then updates the trailing matrix with the outer product of its entries c: for every pair of its
 * rows j < l the product c_j c_l is subtracted from entry (l, j) where the pattern holds one, and
 * otherwise (a dropped fill-in) omega times it is taken from both d_j and d_l, which keeps the row
 * sums of L L^T those of A; each d_j loses c_j^2.  Rows whose original diagonal is 0 are left out
 * (invdiag 0). */
static void mic_factor_build(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
                             double omega, double min_ratio, mic_factor *F)
{
    F->n = n;
    F->colstart = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    F->invdiag = (double *)calloc((size_t)(n ? n : 1), sizeof(double));
    double *a0 = (double *)calloc((size_t)(n ? n : 1), sizeof(double)); /* original diagonal */
    double *d = (double *)calloc((size_t)(n ? n : 1), sizeof(double));  /* running diagonal */
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) total += col[p] > i;
    F->rowindex = (int64_t *)malloc((size_t)(total ? total : 1) * sizeof(int64_t));
    F->value = (double *)malloc((size_t)(total ? total : 1) * sizeof(double));
    int64_t at = 0;
    for (int64_t i = 0; i < n; ++i) {
        F->colstart[i] = at;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            if (col[p] == i) a0[i] = d[i] = val[p];
            if (col[p] <= i) continue;
            /* insertion keeps the column's rows ascending (CSR columns may come in any order) */
            int64_t q = at++;
            while (q > F->colstart[i] && F->rowindex[q - 1] > col[p]) {
                F->rowindex[q] = F->rowindex[q - 1];
                F->value[q] = F->value[q - 1];
                --q;
            }
            F->rowindex[q] = col[p];
            F->value[q] = val[p];
        }
    }
    F->colstart[n] = at;
    for (int64_t k = 0; k < n; ++k) {
        if (a0[k] == 0) continue;
        const double pivot = d[k] < min_ratio * a0[k] ? a0[k] : d[k];
        const double inv = 1.0 / sqrt(pivot);
        F->invdiag[k] = inv;
        const int64_t b0 = F->colstart[k], b1 = F->colstart[k + 1];
        for (int64_t p = b0; p < b1; ++p) F->value[p] *= inv;
        for (int64_t p = b0; p < b1; ++p) {
            const int64_t j = F->rowindex[p];
            const double cj = F->value[p];
            d[j] -= cj * cj;
            for (int64_t q = p + 1; q < b1; ++q) {
                const int64_t l = F->rowindex[q];
                const double prod = cj * F->value[q];
                const int64_t e = mic_find(F, j, l);
                if (e >= 0) {
                    F->value[e] -= prod;
                } else {
                    d[j] -= omega * prod;
                    d[l] -= omega * prod;
                }
            }
        }
    }
    free(a0);
    free(d);
}

/* z = (L L^T)^-1 x, the preconditioner pcg_solver.h:196-223 applies: forward substitution with L
 * column by column (each solved entry is scattered into the rows below it), then backward
 * substitution with L^T, whose row c is L's column c. */
static void mic_apply(const mic_factor *F, const double *x, double *z)
{
    const int64_t n = F->n;
    memcpy(z, x, (size_t)n * sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        z[c] *= F->invdiag[c];
        const double zc = z[c];
        for (int64_t p = F->colstart[c]; p < F->colstart[c + 1]; ++p) z[F->rowindex[p]] -= F->value[p] * zc;
    }
    for (int64_t c = n - 1; c >= 0; --c) {
        double acc = z[c];
        for (int64_t p = F->colstart[c]; p < F->colstart[c + 1]; ++p) acc -= F->value[p] * z[F->rowindex[p]];
        z[c] = acc * F->invdiag[c];
    }
}

/* PCGSolver<double>::solve, pcg_solver.h:246-295.  x starts at zero, the tolerance is
 * residual*||b||_inf and the loop exits on ||r||_inf <= tol.  In the reference the
 * preconditioner output is overwritten by its input (pcg_solver.h:383): ORACLE_MIC_FAITHFUL
 * pays for the factor and both triangular solves and then uses z = r, exactly as the
 * reference does; ORACLE_MIC_APPLIED uses the factor; ORACLE_MIC_NONE skips it. */
static int pcg_run(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *b, double *x,
                   double residual, int max_iterations, int mic_mode, double omega, double min_diag_ratio,
                   int *iterations, double *reresid, double snap_residual, double *x_snap, int *snap_iterations);

int oracle_pcg(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *b, double *x,
               double residual, int max_iterations, int mic_mode, double omega, double min_diag_ratio,
               int *iterations, double *reresid)
{
    return pcg_run(n, rowptr, col, val, b, x, residual, max_iterations, mic_mode, omega, min_diag_ratio, iterations,
                   reresid, 0.0, NULL, NULL);
}

/* oracle_pcg that also records the iterate at the first iteration whose ||r||_inf meets
 * snap_residual * ||b||_inf (the answer the same solver returns at that looser tolerance) in
 * x_snap and that iteration in *snap_iterations (-1 if never met). */
int oracle_pcg_snap(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *b,
                    double *x, double residual, int max_iterations, int mic_mode, double omega,
                    double min_diag_ratio, int *iterations, double *reresid, double snap_residual, double *x_snap,
                    int *snap_iterations)
{
    return pcg_run(n, rowptr, col, val, b, x, residual, max_iterations, mic_mode, omega, min_diag_ratio, iterations,
                   reresid, snap_residual, x_snap, snap_iterations);
}

static int pcg_run(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *b, double *x,
                   double residual, int max_iterations, int mic_mode, double omega, double min_diag_ratio,
                   int *iterations, double *reresid, double snap_residual, double *x_snap, int *snap_iterations)
{
    if (snap_iterations) *snap_iterations = -1;
    if (residual < 1e-30) residual = 1e-30;
    memset(x, 0, (size_t)n * sizeof(double));
    if (n == 0) {
        *iterations = 0;
        *reresid = 0;
        return 1;
    }
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *s = (double *)malloc((size_t)n * sizeof(double));
    double *tmp = (double *)malloc((size_t)n * sizeof(double));
    memcpy(r, b, (size_t)n * sizeof(double));
    double res = abs_max(n, r);
    const double res0 = res;
    int ok = 0;
    mic_factor F;
    memset(&F, 0, sizeof(F));
    if (res == 0) {
        *iterations = 0;
        *reresid = 0;
        ok = 1;
        goto done;
    }
    const double tol = residual * res;
    if (mic_mode != ORACLE_MIC_NONE) mic_factor_build(n, rowptr, col, val, omega, min_diag_ratio, &F);
#define PRECOND(in, out)                                                                                               \
    do {                                                                                                               \
        if (mic_mode == ORACLE_MIC_APPLIED) mic_apply(&F, in, out);                                                    \
        else {                                                                                                         \
            if (mic_mode == ORACLE_MIC_FAITHFUL) mic_apply(&F, in, tmp);                                               \
            memcpy(out, in, (size_t)n * sizeof(double));                                                               \
        }                                                                                                              \
    } while (0)
    PRECOND(r, z);
    double rho = dot(n, z, r);
    if (rho == 0 || rho != rho) { /* pcg_solver.h:263-271 leave residual_out unnormalised */
        *iterations = 0;
        *reresid = res;
        ok = 0;
        goto done;
    }
    memcpy(s, z, (size_t)n * sizeof(double));
    int it;
    for (it = 0; it < max_iterations; ++it) {
        oracle_spmv(n, rowptr, col, val, s, z);
        double alpha = rho / dot(n, s, z);
        add_scaled(n, alpha, s, x);
        add_scaled(n, -alpha, z, r);
        res = abs_max(n, r);
        if (x_snap && *snap_iterations < 0 && res <= snap_residual * res0) {
            memcpy(x_snap, x, (size_t)n * sizeof(double));
            *snap_iterations = it + 1;
        }
        if (res <= tol) {
            *iterations = it + 1;
            *reresid = res / res0;
            ok = 1;
            goto done;
        }
        PRECOND(r, z);
        double rho_new = dot(n, z, r);
        double beta = rho_new / rho;
        add_scaled(n, beta, s, z);
        double *t = s;
        s = z;
        z = t;
        rho = rho_new;
    }
#undef PRECOND
    *iterations = it;
    *reresid = res / res0;
    ok = 0;
done:
    if (F.colstart) mic_free(&F);
    free(r);
    free(z);
    free(s);
    free(tmp);
    return ok;
}

/* ------------------------------------------------------ velocity update */

/* macpressuresolver3.cpp:252-268 on dense faces.  Open faces (area>0 and rho>0) on the domain
 * boundary become 0; interior ones lose dt*(P(c)-P(c-e))/(rho*dx), the difference taken in
 * float as the reference subtracts two Real values; every other face becomes 0 and is switched
 * off, except a boundary face whose inside cell holds liquid, which stays active at 0
 * (:263-265).  active_out (optional) receives the activity after the update: 1 for open faces
 * and those boundary faces, 0 otherwise and for faces that were inactive. */
void oracle_velocity_update(const oracle_grid *g, const float *P, float *area[3], float *rho[3],
                            uint8_t *active_out[3])
{
    const int N[3] = {g->w, g->h, g->d};
    const double dt = g->dt, dx = g->dx;
    for (int dim = 0; dim < 3; ++dim) {
        const int W = g->w + (dim == 0), H = g->h + (dim == 1), D = g->d + (dim == 2);
        float *U = g->vel[dim];
        const uint8_t *act = g->active[dim];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j)
                for (int k = 0; k < D; ++k) {
                    const int pi[3] = {i, j, k};
                    size_t f = fidx(g, dim, i, j, k);
                    uint8_t *aout = active_out ? active_out[dim] : NULL;
                    if (act && !act[f]) {
                        if (aout) aout[f] = 0;
                        continue;
                    }
                    double r = rho[dim][f];
                    uint8_t keep = 1;
                    if (area[dim][f] && r) {
                        if (pi[dim] == 0 || pi[dim] == N[dim]) U[f] = 0.0f;
                        else {
                            float diff = P[cidx(g, i, j, k)] - P[cidx(g, i - (dim == 0), j - (dim == 1), k - (dim == 2))];
                            float delta = (float)(dt * diff / (r * dx));
                            U[f] -= delta;
                        }
                    } else {
                        U[f] = 0.0f;
                        if (pi[dim] == 0) keep = phi_at(g, cidx(g, i, j, k)) < 0.0;
                        else if (pi[dim] == N[dim]) keep = phi_at(g, cidx(g, i - (dim == 0), j - (dim == 1), k - (dim == 2))) < 0.0;
                        else keep = 0;
                    }
                    if (aout) aout[f] = keep;
                }
    }
}

double oracle_divergence(const oracle_grid *g, const int64_t *index_map, float *area[3], double *div)
{
    double mx = 0.0;
    for (int i = 0; i < g->w; ++i)
        for (int j = 0; j < g->h; ++j)
            for (int k = 0; k < g->d; ++k) {
                size_t c = cidx(g, i, j, k);
                if (index_map[c] < 0) {
                    if (div) div[c] = 0.0;
                    continue;
                }
                double s = 0.0;
                for (int nq = 0; nq < 6; ++nq) {
                    int dim = NB_DIM[nq];
                    int qi = i + NB_DI[nq], qj = j + NB_DJ[nq], qk = k + NB_DK[nq];
                    if (out_of_bounds(g, qi, qj, qk)) continue;
                    size_t f = nb_face(g, nq, i, j, k);
                    double a = area[dim][f];
                    if (a) s += NB_SGN[nq] * a * (double)g->vel[dim][f] / g->dx;
                }
                if (div) div[c] = s;
                if (fabs(s) > mx) mx = fabs(s);
            }
    return mx;
}

/* ----------------------------------------------------------- project() */

int oracle_project(const oracle_grid *g, const oracle_params *p, float *pressure, double *pressure64,
                   oracle_result *res)
{
    memset(res, 0, sizeof(*res));
    const double t0 = now_s();
    const size_t M = (size_t)g->w * g->h * g->d;
    float *area[3], *rho[3];
    for (int dim = 0; dim < 3; ++dim) {
        area[dim] = (float *)malloc(face_count(g, dim) * sizeof(float));
        rho[dim] = (float *)malloc(face_count(g, dim) * sizeof(float));
    }
    oracle_face_fractions(g, p, area, rho);
    if (p->surface_tension) oracle_surface_tension(g, p->surface_tension, rho);
    const double t1 = now_s();

    int64_t *index_map = (int64_t *)malloc(M * sizeof(int64_t));
    const int64_t n = oracle_label(g, area, rho, index_map);
    const int64_t nnz = oracle_assemble_count(g, index_map, area, rho);
    int64_t *rowptr = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t *col = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    double *val = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    double *b = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *x = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    oracle_assemble(g, index_map, area, rho, rowptr, col, val, b);
    if (p->rhs_correct)
        for (int64_t r = 0; r < n; ++r) b[r] += p->rhs_correct;
    double *prev = NULL;
    if (p->warm_start && n) { /* macpressuresolver3.cpp:221-230, cell-indexed */
        prev = (double *)malloc((size_t)n * sizeof(double));
        for (size_t c = 0; c < M; ++c)
            if (index_map[c] >= 0) prev[index_map[c]] = p->warm_start[c];
        double *Ap = (double *)malloc((size_t)n * sizeof(double));
        oracle_spmv(n, rowptr, col, val, prev, Ap);
        for (int64_t r = 0; r < n; ++r) b[r] -= Ap[r];
        free(Ap);
    }
    /* pcg.cpp:47-52 copies the matrix before solving; the copy is part of the reference's cost */
    int64_t *col2 = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    double *val2 = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    const double t2 = now_s();
    memcpy(col2, col, (size_t)nnz * sizeof(int64_t));
    memcpy(val2, val, (size_t)nnz * sizeof(double));
    int iters = 0;
    double reresid = 0;
    int ok = oracle_pcg(n, rowptr, col2, val2, b, x, p->residual, p->max_iterations, p->mic_mode, p->mic_omega,
                        p->mic_min_diag_ratio, &iters, &reresid);
    if (prev) { /* macpressuresolver3.cpp:239-242 */
        for (int64_t r = 0; r < n; ++r) x[r] += prev[r];
        free(prev);
    }
    const double t3 = now_s();
    /* macpressuresolver3.cpp:245-248: scatter into the float pressure array */
    for (size_t c = 0; c < M; ++c) {
        int64_t r = index_map[c];
        pressure[c] = r >= 0 ? (float)x[r] : 0.0f;
        if (pressure64) pressure64[c] = r >= 0 ? x[r] : 0.0;
    }
    oracle_velocity_update(g, pressure, area, rho, NULL);
    const double t4 = now_s();
    res->iterations = iters;
    res->reresid = reresid;
    res->converged = ok;
    res->n_unknowns = n;
    res->t_fractions = t1 - t0;
    res->t_build = t2 - t1;
    res->t_linsolve = t3 - t2;
    res->t_update = t4 - t3;
    res->t_projection = t4 - t0;
    free(col2);
    free(val2);
    free(index_map);
    free(rowptr);
    free(col);
    free(val);
    free(b);
    free(x);
    for (int dim = 0; dim < 3; ++dim) {
        free(area[dim]);
        free(rho[dim]);
    }
    return ok;
}
