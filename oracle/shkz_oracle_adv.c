/*
 * shkz_oracle_adv.c -- CPU restatement of Shiokaze's advection (macadvection3), the step that
 * moves the velocity and the level set between projections:
 *   macadvection3::advect_vector / advect_u            src/advection/macadvection3.cpp:47-53,140-177
 *   macadvection3::advect_semiLagrangian_u             :68-137
 *   macadvection3::advect_scalar / advect_cell         :38-45,239-288
 *   macadvection3::advect_semiLagrangian_cell          :179-237
 *   defaults MacCormack on, WENO off, TrimNarrowBand 1  :290-294
 *   the WENO option (config key "WENO")                :59,85-88,200-203
 *   WENO3::interpolate (order 6) on array3<Real>       include/shiokaze/math/WENO3.h:105-135
 *   WENO::interp6                                      include/shiokaze/math/WENO.h:55-88
 *   macarray3::convert_to_full (face and cell forms)   include/shiokaze/array/macarray3.h:305-372
 *   array_interpolator3::interpolate_coef/interpolate  include/shiokaze/array/array_interpolator3.h:40-97
 *   vec arithmetic                                     include/shiokaze/math/vec.h:133-136,208-230,401-403
 * on dense arrays (z fastest, layouts of shkz_oracle.h): face velocities with per-face activity,
 * cell scalars with optional per-cell activity.  Inactive entries read 0 for velocities (the
 * background value) and their own stored value for scalars (the caller stores the background or
 * fill value there, which is what the reference reads); they are never written.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never linked by the product.
 *
 * Types follow the reference: fields are Real (float); a full velocity at a face is a float
 * vector, and the departure point of the value pass is computed in float (dt*u/dx on vec3r)
 * while the min/max pass recomputes it in double (vec3d); interpolation accumulates in float;
 * MacCormack's correction is 0.5 times a float difference.  No contraction (-ffp-contract=off).
 */
#include "shkz_oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

static inline double amin(double a, double b) { return b < a ? b : a; }  /* std::min */
static inline double amax(double a, double b) { return a < b ? b : a; }  /* std::max */
static inline int iclamp(int v, int n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }

typedef struct {
    int s[3];
} Shape;

static inline Shape face_shape3(const oracle_vfield *f, int a)
{
    Shape S = {{f->w + (a == 0), f->h + (a == 1), f->d + (a == 2)}};
    return S;
}

static inline size_t sidx(const Shape *S, int i, int j, int k) { return ((size_t)i * S->s[1] + j) * S->s[2] + k; }

/* face value with the background 0 on inactive faces */
static inline float fval(const float *U, const uint8_t *A, size_t q) { return A[q] ? U[q] : 0.0f; }

/* array_interpolator3::interpolate<Real> (only_actives = false) over a float array with a value
 * accessor: inactive faces read 0 when A is given */
static float interp_f(const float *U, const uint8_t *A, const Shape *S, const double p[3])
{
    int idx[24];
    double coef[8];
    oracle_interp_coef(S->s, p, idx, coef);
    float value = 0.0f;
    for (int n = 0; n < 8; ++n)
        if (coef[n]) {
            const size_t q = sidx(S, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]);
            const float v = A ? fval(U, A, q) : U[q];
            value = (float)((double)value + (double)v * coef[n]);
        }
    return value;
}

/* the same over a level set of either precision (double accumulates in double) */
static double interp_ls(const void *phi, int f64, const Shape *S, const double p[3])
{
    int idx[24];
    double coef[8];
    oracle_interp_coef(S->s, p, idx, coef);
    if (f64) {
        double value = 0.0;
        for (int n = 0; n < 8; ++n)
            if (coef[n]) value += ((const double *)phi)[sidx(S, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2])] * coef[n];
        return value;
    }
    float value = 0.0f;
    for (int n = 0; n < 8; ++n)
        if (coef[n])
            value = (float)((double)value +
                            (double)((const float *)phi)[sidx(S, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2])] * coef[n]);
    return value;
}

static inline double ls_at(const void *phi, int f64, size_t c)
{
    return f64 ? ((const double *)phi)[c] : (double)((const float *)phi)[c];
}

/* macarray3::convert_to_full into a face array (macarray3.h:350-372): the face's own component,
 * and each other component the mean of the 4 values around the face (clamped into that
 * component's array), accumulated in double, divided by 4, stored as float. */
static void full_at_face(const oracle_vfield *f, float *const U[3], const int dim, int i, int j, int k, float out[3])
{
    const int pc[3] = {i, j, k};
    for (int ud = 0; ud < 3; ++ud) {
        if (ud == dim) {
            const Shape S = face_shape3(f, dim);
            out[ud] = fval(U[dim], f->active[dim], sidx(&S, i, j, k));
            continue;
        }
        const Shape S = face_shape3(f, ud);
        double u = 0.0;
        for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj) {
                int q[3];
                for (int a = 0; a < 3; ++a)
                    q[a] = iclamp(pc[a] - (a == dim) + ii * (a == dim) + jj * (a == ud), S.s[a]);
                u += fval(U[ud], f->active[ud], sidx(&S, q[0], q[1], q[2]));
            }
        u /= 4.0;
        out[ud] = (float)u;
    }
}

/* Six-point WENO interpolation (Macdonald & Ruuth, J. Sci. Comput. 35 (2008) 219-240, the method
 * WENO::interp6 implements): f holds samples at the nodes -2..3, x lies in [0, 1].  Three cubics,
 * each through four consecutive nodes r-2..r+1 (Newton form), are blended with weights
 * gamma_r / (eps + beta_r^2) normalised to sum 1, where gamma_r are the linear weights that make
 * the blend the quintic through all six nodes and beta_r the smoothness indicators, quadratic
 * forms in the samples (180 * beta_r tabulated below term by term: coefficient, sample a, sample
 * b).  Each product, sum and quotient is formed in the reference's order, so the result is the
 * same double. */
typedef struct {
    double c;
    int a, b;
} WenoTerm;
static const WenoTerm weno_beta[3][10] = {
    {{814., 3, 3}, {4326., 2, 2}, {2976., 1, 1}, {244., 0, 0}, {-3579., 2, 3},
     {-6927., 2, 1}, {1854., 2, 0}, {2634., 3, 1}, {-683., 3, 0}, {-1659., 1, 0}},
    {{1986., 3, 3}, {1986., 2, 2}, {244., 1, 1}, {244., 4, 4}, {1074., 2, 4},
     {-3777., 2, 3}, {-1269., 2, 1}, {1074., 3, 1}, {-1269., 4, 3}, {-293., 4, 1}},
    {{814., 2, 2}, {4326., 3, 3}, {2976., 4, 4}, {244., 5, 5}, {-683., 2, 5},
     {2634., 2, 4}, {-3579., 2, 3}, {-6927., 3, 4}, {1854., 3, 5}, {-1659., 4, 5}},
};

static inline double weno_term(const WenoTerm *t, const double f[6])
{
    return t->a == t->b ? t->c * (f[t->a] * f[t->a]) : (t->c * f[t->a]) * f[t->b];
}

double oracle_weno6(double x, const double f[6])
{
    const double gamma[3] = {(2.0 - x) * (3.0 - x) / 20.0, (3.0 - x) * (x + 2.0) / 10.0,
                             (x + 2.0) * (x + 1.0) / 20.0};
    double alpha[3], poly[3], total = 0.0;
    for (int r = 0; r < 3; ++r) {
        double beta = weno_term(&weno_beta[r][0], f);
        for (int t = 1; t < 10; ++t) beta += weno_term(&weno_beta[r][t], f);
        beta /= 180.0;
        /* the cubic through nodes r-2 .. r+1 */
        const double *g = f + r;
        const double e0 = x - (double)(r - 2), e1 = x - (double)(r - 1), e2 = x - (double)r;
        const double d1 = g[1] - g[0];
        const double d2 = g[2] - 2. * g[1] + g[0];
        const double d3 = g[3] - 3. * g[2] + 3. * g[1] - g[0];
        poly[r] = g[0] + d1 * e0 + d2 * e0 * e1 / 2.0 + d3 * e0 * e1 * e2 / 6.0;
        alpha[r] = gamma[r] / (DBL_EPSILON + beta * beta);
        total += alpha[r];
    }
    double out = 0.0;
    for (int r = 0; r < 3; ++r) {
        const double t = alpha[r] / total * poly[r];
        out = r ? out + t : t;
    }
    return out;
}

/* WENO3::interpolate(array, p) with order 6 on a float array of shape S: p is clamped into the
 * array, the base index i = min(x, n-2) (truncated), and the 6x6x6 samples at i-2..i+3 (clamped
 * per axis) are reduced along x, then y, then z.  Inactive entries read 0 when A is given. */
static double weno3_f(const float *U, const uint8_t *A, const Shape *S, const double p[3])
{
    double c[3], fr[3];
    int b[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = amax(0.0, amin(S->s[a] - 1.0, p[a]));
        b[a] = (int)amin(c[a], S->s[a] - 2.0);
        fr[a] = c[a] - b[a];
    }
    double vz[6];
    for (int kk = 0; kk < 6; ++kk) {
        double vy[6];
        for (int jj = 0; jj < 6; ++jj) {
            double vx[6];
            for (int ii = 0; ii < 6; ++ii) {
                const size_t q = sidx(S, iclamp(b[0] + ii - 2, S->s[0]), iclamp(b[1] + jj - 2, S->s[1]),
                                      iclamp(b[2] + kk - 2, S->s[2]));
                vx[ii] = A ? fval(U, A, q) : U[q];
            }
            vy[jj] = oracle_weno6(fr[0], vx);
        }
        vz[kk] = oracle_weno6(fr[1], vy);
    }
    return oracle_weno6(fr[2], vz);
}

double oracle_weno3(const float *U, const int s[3], const double p[3])
{
    const Shape S = {{s[0], s[1], s[2]}};
    return weno3_f(U, NULL, &S, p);
}

/* the value a semi-Lagrangian pass samples at p: trilinear (interpolate<Real>) or WENO */
static inline float sample_f(const float *U, const uint8_t *A, const Shape *S, const double p[3], int weno)
{
    return weno ? (float)weno3_f(U, A, S, p) : interp_f(U, A, S, p);
}

typedef struct {
    float lo, hi;
    uint8_t nb;
} MinMax;

/* advect_semiLagrangian_u for every active face: in -> out (float), minmax optional.
 * The full velocity comes from `in` itself (the field being advected, :72-73). */
static void sl_faces(const oracle_vfield *f, float *const in[3], float *const out[3], MinMax *const mm[3],
                     const void *fluid, int fluid_f64, double dt, int trim, int weno)
{
    const double dx = f->dx;
    const Shape Sc = {{f->w, f->h, f->d}};
    for (int dim = 0; dim < 3; ++dim) {
        const Shape S = face_shape3(f, dim);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < S.s[0]; ++i)
            for (int j = 0; j < S.s[1]; ++j)
                for (int k = 0; k < S.s[2]; ++k) {
                    const size_t e = sidx(&S, i, j, k);
                    if (!f->active[dim][e]) continue;
                    float u[3];
                    full_at_face(f, in, dim, i, j, k, u);
                    const int empty = !u[0] && !u[1] && !u[2];
                    const int pc[3] = {i, j, k};
                    if (!empty) {
                        double p[3];
                        for (int a = 0; a < 3; ++a) {
                            const float t = (float)((double)u[a] * dt);  /* dt*u on vec3r */
                            const float sft = (float)((double)t / dx);   /* /m_dx on vec3r */
                            p[a] = (double)pc[a] - (double)sft;
                        }
                        out[dim][e] = sample_f(in[dim], f->active[dim], &S, p, weno);
                    } else {
                        out[dim][e] = fval(in[dim], f->active[dim], e);
                    }
                    if (!mm) continue;
                    MinMax r;
                    double fp[3];
                    if (!empty) {
                        double p[3];
                        for (int a = 0; a < 3; ++a) {
                            const double sft = (dt * (double)u[a]) / dx;  /* dt*u/m_dx on vec3d */
                            p[a] = (double)pc[a] - sft;
                            fp[a] = (pc[a] + 0.5 * (dim != a)) - sft;
                        }
                        int idx[24];
                        double coef[8];
                        oracle_interp_coef(S.s, p, idx, coef);
                        double lo = DBL_MAX, hi = DBL_MIN;  /* numeric_limits<double>::max(), ::min() */
                        for (int n = 0; n < 8; ++n) {
                            const double v = fval(in[dim], f->active[dim], sidx(&S, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]));
                            lo = amin(lo, v);
                            hi = amax(hi, v);
                        }
                        r.lo = (float)lo;
                        r.hi = (float)hi;
                    } else {
                        for (int a = 0; a < 3; ++a) fp[a] = pc[a] + 0.5 * (dim != a);
                        r.lo = r.hi = fval(in[dim], f->active[dim], e);
                    }
                    const double q[3] = {fp[0] - 0.5, fp[1] - 0.5, fp[2] - 0.5};
                    r.nb = interp_ls(fluid, fluid_f64, &Sc, q) > -dx * (double)trim;
                    mm[dim][e] = r;
                }
    }
}

/* macadvection3::advect_vector(u, u, fluid, dt) (advect_u, :140-177) in place on f->vel. */
void oracle_advect_velocity(const oracle_vfield *f, const void *fluid, int fluid_f64, double dt, int maccormack,
                            int trim, int weno)
{
    float *v0[3], *v1[3], *vin[3];
    MinMax *mm[3];
    for (int a = 0; a < 3; ++a) {
        const Shape S = face_shape3(f, a);
        const size_t n = (size_t)S.s[0] * S.s[1] * S.s[2];
        vin[a] = (float *)malloc(n * sizeof(float));
        memcpy(vin[a], f->vel[a], n * sizeof(float));
        v0[a] = (float *)malloc(n * sizeof(float));
        v1[a] = (float *)malloc(n * sizeof(float));
        memcpy(v0[a], vin[a], n * sizeof(float));
        memcpy(v1[a], vin[a], n * sizeof(float));
        mm[a] = (MinMax *)calloc(n, sizeof(MinMax));
    }
    if (!maccormack) {
        sl_faces(f, vin, f->vel, NULL, fluid, fluid_f64, dt, trim, weno);
    } else {
        sl_faces(f, vin, v0, mm, fluid, fluid_f64, dt, trim, weno);
        sl_faces(f, v0, v1, NULL, fluid, fluid_f64, -dt, trim, weno);
        for (int dim = 0; dim < 3; ++dim) {
            const Shape S = face_shape3(f, dim);
            const size_t n = (size_t)S.s[0] * S.s[1] * S.s[2];
            for (size_t e = 0; e < n; ++e) {
                if (!f->active[dim][e]) continue;
                if (mm[dim][e].nb) {
                    f->vel[dim][e] = v0[dim][e];
                    continue;
                }
                const double lo = mm[dim][e].lo, hi = mm[dim][e].hi;
                const double vel0 = v0[dim][e];
                const float diff = vin[dim][e] - v1[dim][e];
                const double corr = 0.5 * (double)diff;
                if (vel0 + corr < lo) f->vel[dim][e] = (float)lo;
                else if (vel0 + corr > hi) f->vel[dim][e] = (float)hi;
                else f->vel[dim][e] = (float)(vel0 + corr);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        free(vin[a]);
        free(v0[a]);
        free(v1[a]);
        free(mm[a]);
    }
}

/* macarray3::convert_to_full into a cell array (macarray3.h:305-343): a cell's full velocity is
 * valid when both faces of every axis are active (mean in double, stored as float); otherwise
 * the cell reads the zero vector. */
static void full_at_cell(const oracle_vfield *f, int i, int j, int k, float out[3])
{
    const int pc[3] = {i, j, k};
    int valid = 0;
    float v[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
        const Shape S = face_shape3(f, a);
        const size_t q0 = sidx(&S, i, j, k);
        const size_t q1 = sidx(&S, pc[0] + (a == 0), pc[1] + (a == 1), pc[2] + (a == 2));
        int wsum = 0;
        double value = 0.0;
        if (f->active[a][q0]) {
            value += f->vel[a][q0];
            wsum++;
        }
        if (f->active[a][q1]) {
            value += f->vel[a][q1];
            wsum++;
        }
        if (wsum == 2) {
            v[a] = (float)(0.5 * value);
            valid++;
        }
    }
    for (int a = 0; a < 3; ++a) out[a] = valid == 3 ? v[a] : 0.0f;
}

/* advect_semiLagrangian_cell for every active cell */
static void sl_cells(const oracle_vfield *f, const float *in, float *out, const uint8_t *act, MinMax *mm,
                     const void *fluid, int fluid_f64, double dt, int trim, int weno)
{
    const double dx = f->dx;
    const Shape S = {{f->w, f->h, f->d}};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->w; ++i)
        for (int j = 0; j < f->h; ++j)
            for (int k = 0; k < f->d; ++k) {
                const size_t c = sidx(&S, i, j, k);
                if (act && !act[c]) continue;
                float uf[3];
                full_at_cell(f, i, j, k, uf);
                const double u[3] = {uf[0], uf[1], uf[2]};
                const int empty = !u[0] && !u[1] && !u[2];
                double p[3];
                const int pc[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) p[a] = (double)pc[a] - (dt * u[a]) / dx;  /* vec3d arithmetic */
                out[c] = empty ? in[c] : sample_f(in, NULL, &S, p, weno);
                if (!mm) continue;
                MinMax r;
                if (!empty) {
                    int idx[24];
                    double coef[8];
                    oracle_interp_coef(S.s, p, idx, coef);
                    double lo = FLT_MAX, hi = FLT_MIN;  /* numeric_limits<Real>::max(), ::min() */
                    for (int n = 0; n < 8; ++n) {
                        const double v = in[sidx(&S, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2])];
                        lo = amin(lo, v);
                        hi = amax(hi, v);
                    }
                    r.lo = (float)lo;
                    r.hi = (float)hi;
                    r.nb = interp_ls(fluid, fluid_f64, &S, p) > -dx * (double)trim;
                } else {
                    r.lo = r.hi = in[c];
                    r.nb = ls_at(fluid, fluid_f64, c) > -dx * (double)trim;
                }
                mm[c] = r;
            }
}

/* macadvection3::advect_scalar(q, velocity, fluid, dt) (advect_cell, :239-288) in place on q.
 * fluid NULL: the narrow band is measured on a copy of q itself, as the level-set tracker calls
 * it (maclevelsetsurfacetracker3.cpp:50-51).  The forward pass reads inactive cells of q at their
 * stored value (the caller's background or fill value); the intermediate field of the backward
 * pass is a fresh array of q's type (:246), whose inactive cells read its background value
 * `background` (no fill information), as in the reference. */
void oracle_advect_scalar(const oracle_vfield *f, float *q, const uint8_t *active_q, const void *fluid, int fluid_f64,
                          double dt, int maccormack, int trim, float background, int weno)
{
    const size_t n = (size_t)f->w * f->h * f->d;
    float *qin = (float *)malloc(n * sizeof(float));
    memcpy(qin, q, n * sizeof(float));
    if (!fluid) {
        fluid = qin;
        fluid_f64 = 0;
    }
    if (!maccormack) {
        sl_cells(f, qin, q, active_q, NULL, fluid, fluid_f64, dt, trim, weno);
        free(qin);
        return;
    }
    float *q0 = (float *)malloc(n * sizeof(float)), *q1 = (float *)malloc(n * sizeof(float));
    MinMax *mm = (MinMax *)calloc(n, sizeof(MinMax));
    for (size_t c = 0; c < n; ++c) q0[c] = (active_q && !active_q[c]) ? background : qin[c];
    memcpy(q1, q0, n * sizeof(float));
    sl_cells(f, qin, q0, active_q, mm, fluid, fluid_f64, dt, trim, weno);
    sl_cells(f, q0, q1, active_q, NULL, fluid, fluid_f64, -dt, trim, weno);
    for (size_t c = 0; c < n; ++c) {
        if (active_q && !active_q[c]) continue;
        if (mm[c].nb) {
            q[c] = q0[c];
            continue;
        }
        const double lo = mm[c].lo, hi = mm[c].hi, v0 = q0[c];
        const float diff = qin[c] - q1[c];
        const double corr = 0.5 * (double)diff;
        if (v0 + corr < lo) q[c] = (float)lo;
        else if (v0 + corr > hi) q[c] = (float)hi;
        else q[c] = (float)(v0 + corr);
    }
    free(qin);
    free(q0);
    free(q1);
    free(mm);
}
