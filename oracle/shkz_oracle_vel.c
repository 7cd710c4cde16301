/*
 * shkz_oracle_vel.c -- CPU restatement of the step that follows a projection in Shiokaze's time
 * steppers: velocity extrapolation and the solid constraint,
 *   macutility3::extrapolate_and_constrain_velocity   src/utility/macutility3.cpp:89-93
 *   macarray_extrapolator3::extrapolate               include/shiokaze/array/macarray_extrapolator3.h
 *   array_extrapolator3::extrapolate_if               include/shiokaze/array/array_extrapolator3.h:47-80
 *     over array3::dilate                             include/shiokaze/array/array3.h:1661-1668
 *     and dilate3::dilate                             src/array/dilate3.h:43-101
 *   macutility3::constrain_velocity                   src/utility/macutility3.cpp:61-88
 *   array_interpolator3::interpolate_coef/interpolate include/shiokaze/array/array_interpolator3.h:40-105
 *   array_derivative3::derivative_interpolate_coef/derivative
 *                                                     include/shiokaze/array/array_derivative3.h:40-117
 *   macarray_interpolator3::interpolate               include/shiokaze/array/macarray_interpolator3.h:43-50
 *   vec arithmetic (dot product, norm2)               include/shiokaze/math/vec.h:208-268
 * on dense face arrays with per-face activity (z fastest, layouts of shkz_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never linked by the product.
 *
 * Arithmetic follows the reference's types: velocities are Real (float); the solid level set is
 * Real too (float), or double for a double-precision level set; interpolation accumulates into
 * the array's element type after a double product (value += array(idx) * coef); the derivative
 * likewise; the normal, the velocity sample and the projection onto the tangent plane are
 * double (vec3d), and the result is stored back as float.  No contraction (-ffp-contract=off).
 */
#include "shkz_oracle.h"

#include <stdlib.h>
#include <string.h>

/* std::min / std::max as <algorithm> defines them (b < a ? b : a, a < b ? b : a) */
static inline double fmin_(double a, double b) { return b < a ? b : a; }
static inline double fmax_(double a, double b) { return a < b ? b : a; }

static inline size_t vf_fidx(const oracle_vfield *f, int dim, int i, int j, int k)
{
    const size_t H = (size_t)f->h + (dim == 1), D = (size_t)f->d + (dim == 2);
    return ((size_t)i * H + j) * D + k;
}

static inline void face_shape(const oracle_vfield *f, int dim, int s[3])
{
    s[0] = f->w + (dim == 0);
    s[1] = f->h + (dim == 1);
    s[2] = f->d + (dim == 2);
}

/* array_interpolator3.h:40-62 */
void oracle_interp_coef(const int shape[3], const double p[3], int idx[24], double coef[8])
{
    const double x = fmax_(0.0, fmin_(shape[0] - 1., p[0]));
    const double y = fmax_(0.0, fmin_(shape[1] - 1., p[1]));
    const double z = fmax_(0.0, fmin_(shape[2] - 1., p[2]));
    const int i = (int)fmin_(x, shape[0] - 2.), j = (int)fmin_(y, shape[1] - 2.), k = (int)fmin_(z, shape[2] - 2.);
    const int c[8][3] = {{i, j, k},         {i + 1, j, k},     {i, j + 1, k},     {i + 1, j + 1, k},
                         {i, j, k + 1},     {i + 1, j, k + 1}, {i, j + 1, k + 1}, {i + 1, j + 1, k + 1}};
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a) idx[3 * n + a] = c[n][a];
    const double zk1 = (k + 1) - z, zk = z - k, xi1 = (i + 1) - x, xi = x - i, yj1 = (j + 1) - y, yj = y - j;
    coef[0] = zk1 * xi1 * yj1;
    coef[1] = zk1 * xi * yj1;
    coef[2] = zk1 * xi1 * yj;
    coef[3] = zk1 * xi * yj;
    coef[4] = zk * xi1 * yj1;
    coef[5] = zk * xi * yj1;
    coef[6] = zk * xi1 * yj;
    coef[7] = zk * xi * yj;
}

/* array_derivative3.h:40-85: coef[8*dim + n] */
void oracle_deriv_coef(const int shape[3], const double p[3], int idx[24], double coef[24])
{
    double unused[8];
    oracle_interp_coef(shape, p, idx, unused);
    const double x = fmax_(0.0, fmin_(shape[0] - 1., p[0]));
    const double y = fmax_(0.0, fmin_(shape[1] - 1., p[1]));
    const double z = fmax_(0.0, fmin_(shape[2] - 1., p[2]));
    const int i = idx[0], j = idx[1], k = idx[2];
    const double zk1 = (k + 1) - z, zk = z - k, xi1 = (i + 1) - x, xi = x - i, yj1 = (j + 1) - y, yj = y - j;
    double *cx = coef, *cy = coef + 8, *cz = coef + 16;
    cx[0] = -zk1 * yj1;
    cx[1] = zk1 * yj1;
    cx[2] = -zk1 * yj;
    cx[3] = zk1 * yj;
    cx[4] = -zk * yj1;
    cx[5] = zk * yj1;
    cx[6] = -zk * yj;
    cx[7] = zk * yj;
    cy[0] = -zk1 * xi1;
    cy[1] = -zk1 * xi;
    cy[2] = zk1 * xi1;
    cy[3] = zk1 * xi;
    cy[4] = -zk * xi1;
    cy[5] = -zk * xi;
    cy[6] = zk * xi1;
    cy[7] = zk * xi;
    cz[0] = -xi1 * yj1;
    cz[1] = -xi * yj1;
    cz[2] = -xi1 * yj;
    cz[3] = -xi * yj;
    cz[4] = xi1 * yj1;
    cz[5] = xi * yj1;
    cz[6] = xi1 * yj;
    cz[7] = xi * yj;
}

static inline double solid_value(const oracle_vfield *f, int i, int j, int k)
{
    const size_t n = ((size_t)i * (f->h + 1) + j) * (f->d + 1) + k;
    return f->solid_f64 ? ((const double *)f->solid)[n] : (double)((const float *)f->solid)[n];
}

/* array_interpolator3.h:76-97 (only_actives = false) on the nodal solid: the accumulator has the
 * array's element type, so a float level set rounds after every term. */
static double interp_solid(const oracle_vfield *f, const double p[3])
{
    const int shape[3] = {f->w + 1, f->h + 1, f->d + 1};
    int idx[24];
    double coef[8];
    oracle_interp_coef(shape, p, idx, coef);
    if (f->solid_f64) {
        double value = 0.0;
        for (int n = 0; n < 8; ++n)
            if (coef[n]) value += solid_value(f, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]) * coef[n];
        return value;
    }
    float value = 0.0f;
    for (int n = 0; n < 8; ++n)
        if (coef[n]) value = (float)((double)value + solid_value(f, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]) * coef[n]);
    return value;
}

/* array_derivative3.h:94-103: result[dim] += coef[dim][n] * array(idx[n]), no zero test */
static void deriv_solid(const oracle_vfield *f, const double p[3], double out[3])
{
    const int shape[3] = {f->w + 1, f->h + 1, f->d + 1};
    int idx[24];
    double coef[24];
    oracle_deriv_coef(shape, p, idx, coef);
    for (int a = 0; a < 3; ++a) {
        if (f->solid_f64) {
            double r = 0.0;
            for (int n = 0; n < 8; ++n) r += coef[8 * a + n] * solid_value(f, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]);
            out[a] = r;
        } else {
            float r = 0.0f;
            for (int n = 0; n < 8; ++n)
                r = (float)((double)r + coef[8 * a + n] * solid_value(f, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]));
            out[a] = r;
        }
    }
}

/* macarray_interpolator3.h:43-50: component a sampled at p - 0.5 (off its own axis) in index
 * space of the face array; inactive faces read the background value 0 (array3::operator()). */
static double interp_face(const oracle_vfield *f, const float *U, const uint8_t *act, int a, const double p[3])
{
    int shape[3];
    face_shape(f, a, shape);
    const double pos[3] = {p[0] - 0.5 * (a != 0), p[1] - 0.5 * (a != 1), p[2] - 0.5 * (a != 2)};
    int idx[24];
    double coef[8];
    oracle_interp_coef(shape, pos, idx, coef);
    float value = 0.0f;
    for (int n = 0; n < 8; ++n)
        if (coef[n]) {
            const size_t q = vf_fidx(f, a, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]);
            const float v = act[q] ? U[q] : 0.0f;
            value = (float)((double)value + v * coef[n]);
        }
    return value;
}

/* `width` passes of array_extrapolator3::extrapolate on each component.  A pass collects the
 * inactive in-bounds 6-neighbours of the active faces (dilate3.h:50-70), gives each the mean of
 * its active neighbours as they were before the pass, summed in float in the order +x, -x, +y,
 * -y, -z, +z and divided by their count (array_extrapolator3.h:54-72), and activates them all
 * at the end of the pass (dilate3.h:90-97).  Returns the number of faces activated. */
int64_t oracle_extrapolate(const oracle_vfield *f, int width)
{
    static const int Q[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}};
    int64_t total = 0;
    for (int dim = 0; dim < 3; ++dim) {
        int s[3];
        face_shape(f, dim, s);
        const size_t n = (size_t)s[0] * s[1] * s[2];
        float *U = f->vel[dim];
        uint8_t *A = f->active[dim];
        float *U0 = (float *)malloc(n * sizeof(float));
        uint8_t *A0 = (uint8_t *)malloc(n);
        for (int pass = 0; pass < width; ++pass) {
            memcpy(U0, U, n * sizeof(float));
            memcpy(A0, A, n);
            for (int i = 0; i < s[0]; ++i)
                for (int j = 0; j < s[1]; ++j)
                    for (int k = 0; k < s[2]; ++k) {
                        const size_t e = vf_fidx(f, dim, i, j, k);
                        if (A0[e]) continue;
                        float sum = 0.0f;
                        int weight = 0;
                        for (int q = 0; q < 6; ++q) {
                            const int a = i + Q[q][0], b = j + Q[q][1], c = k + Q[q][2];
                            if (a < 0 || b < 0 || c < 0 || a >= s[0] || b >= s[1] || c >= s[2]) continue;
                            const size_t g = vf_fidx(f, dim, a, b, c);
                            if (A0[g]) {
                                sum += U0[g];
                                weight++;
                            }
                        }
                        if (weight) {
                            U[e] = sum / (float)weight;
                            A[e] = 1;
                            total++;
                        }
                    }
        }
        free(U0);
        free(A0);
    }
    return total;
}

/* macutility3.cpp:61-88.  Nothing changes unless the solid level set has a negative value
 * (array_utility3.h:112-122).  Otherwise every active face whose centre samples the solid
 * below 0 loses its velocity's component into the solid, the velocity being sampled from a copy
 * taken before any face changes; then active faces on the domain boundary pointing out of the
 * domain become 0.  Returns 1 when the solid exists. */
int oracle_constrain_velocity(const oracle_vfield *f)
{
    if (!f->solid) return 0;
    int exist = 0;
    const size_t nn = (size_t)(f->w + 1) * (f->h + 1) * (f->d + 1);
    for (size_t q = 0; q < nn && !exist; ++q)
        exist = (f->solid_f64 ? ((const double *)f->solid)[q] : (double)((const float *)f->solid)[q]) < 0.0;
    if (!exist) return 0;
    float *S[3];
    uint8_t *SA[3];
    for (int a = 0; a < 3; ++a) {
        int s[3];
        face_shape(f, a, s);
        const size_t n = (size_t)s[0] * s[1] * s[2];
        S[a] = (float *)malloc(n * sizeof(float));
        SA[a] = (uint8_t *)malloc(n);
        memcpy(S[a], f->vel[a], n * sizeof(float));
        memcpy(SA[a], f->active[a], n);
    }
    const int N[3] = {f->w, f->h, f->d};
    for (int dim = 0; dim < 3; ++dim) {
        int s[3];
        face_shape(f, dim, s);
        float *U = f->vel[dim];
        for (int i = 0; i < s[0]; ++i)
            for (int j = 0; j < s[1]; ++j)
                for (int k = 0; k < s[2]; ++k) {
                    const size_t e = vf_fidx(f, dim, i, j, k);
                    if (!f->active[dim][e]) continue;
                    const int pi[3] = {i, j, k};
                    const double p[3] = {i + 0.5 * (dim != 0), j + 0.5 * (dim != 1), k + 0.5 * (dim != 2)};
                    if (interp_solid(f, p) < 0.0) {
                        double der[3], nrm[3];
                        deriv_solid(f, p, der);
                        for (int a = 0; a < 3; ++a) nrm[a] = der[a] / f->dx;
                        double n2 = 0.0;
                        for (int a = 0; a < 3; ++a) n2 += nrm[a] * nrm[a];
                        if (n2) {
                            double u[3];
                            for (int a = 0; a < 3; ++a) u[a] = interp_face(f, S[a], SA[a], a, p);
                            double un = 0.0;
                            for (int a = 0; a < 3; ++a) un += u[a] * nrm[a];
                            if (un < 0.0) U[e] = (float)(u[dim] - nrm[dim] * un);
                        }
                    }
                    if (pi[dim] == 0 && U[e] < 0.0f) U[e] = 0.0f;
                    if (pi[dim] == N[dim] && U[e] > 0.0f) U[e] = 0.0f;
                }
    }
    for (int a = 0; a < 3; ++a) {
        free(S[a]);
        free(SA[a]);
    }
    return 1;
}
