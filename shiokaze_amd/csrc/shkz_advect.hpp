// shkz_advect.hpp -- advection on the device: macadvection3 (src/advection/macadvection3.cpp)
// (MacCormack or semi-Lagrangian, trilinear or WENO sampling, TrimNarrowBand), for the face
// velocity (advect_vector -> advect_u, :47-53,140-177) and for cell scalars such as the liquid
// level set (advect_scalar -> advect_cell, :38-45,239-288).
//
//   k_adv_face_fwd   forward semi-Lagrangian pass of advect_semiLagrangian_u (:68-137) with the
//                    MacCormack min/max and narrow-band record; its result holds 0 at inactive
//                    faces, so the backward pass samples it without the activity bytes (half of
//                    its loads)
//   k_adv_face_bwd   backward pass from the forward result, fused with MacCormack's final
//                    combination (:152-168), written in place over the input
//   k_adv_cell_fwd,
//   k_adv_cell_bwd   the same for cells (advect_semiLagrangian_cell, :179-237; :255-270)
//   k_adv_*_copy     semi-Lagrangian only: the forward result copied over the active entries
//
// The full velocity at a face (macarray3::convert_to_full into a face array, macarray3.h:350-372)
// and at a cell (:305-343) is computed where it is needed instead of being stored.  A face
// velocity field samples inactive faces as 0, its background value.  Cell scalars sample their
// stored value in the forward pass; in the backward pass inactive cells read `background`, the
// value a fresh array of the scalar's type holds (:246).
//
// Arithmetic follows the reference's types operation by operation (no contraction): the value
// pass of the face version computes the departure point from a float vector (vec3r), its min/max
// pass from a double one (vec3d); interpolation accumulates in the array's type; MacCormack's
// correction is 0.5 times a float difference.
//
// The WENO option (config key "WENO", :59) samples the value passes with WENO3::interpolate
// (include/shiokaze/math/WENO3.h:105-135, order 6) in double and stores the float; the MacCormack
// bounds and the narrow-band test stay trilinear (:96-118,207-231).  The kernels take it as a
// template flag, so the default trilinear kernels keep their register budget.
#pragma once
#include <float.h>

#include "shkz_extrap.hpp"
#include "shkz_geometry.hpp"

namespace shkz {

struct AdvScratch {
    float *v0[3];   // forward result (faces) or v0[0] (cells)
    float *lo[3];   // MacCormack bounds
    float *hi[3];
    uint8_t *nb[3]; // inside the narrow band: keep the semi-Lagrangian value
};

// The value of an entry that reads 0 where inactive.  The value is loaded together with the
// activity byte (a select, not a branch), so the two loads share one memory latency.
__device__ __forceinline__ float face_val(const float *U, const uint8_t *A, size_t q) {
    const float u = U[q];
    return A[q] ? u : 0.0f;
}

__device__ __forceinline__ int clamp_i(int v, int n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }

// Flat index of (i, j, k) in an array of shape s, in 32-bit arithmetic: the advection entry points
// reject grids with (w+1)(h+1)(d+1) >= 2^32.
__device__ __forceinline__ size_t flat3(int i, int j, int k, const int s[3]) {
    return ((unsigned)i * (unsigned)s[1] + (unsigned)j) * (unsigned)s[2] + (unsigned)k;
}

// macarray3::convert_to_full at face (dim; i,j,k) of the fields U (activity F.act).  MASKED: U may
// hold anything at inactive faces (read through the activity bytes); otherwise U holds 0 there.
template <bool MASKED>
__device__ __forceinline__ void full_face(const VelField &F, const float *const U[3], int dim, int i, int j, int k,
                                          float out[3]) {
#pragma clang fp contract(off)
    const int pc[3] = {i, j, k};
#pragma unroll
    for (int ud = 0; ud < 3; ++ud) {
        int s[3];
        vf_shape(F, ud, s);
        if (ud == dim) {
            const size_t q = flat3(i, j, k, s);
            out[ud] = MASKED ? face_val(U[ud], F.act[ud], q) : U[ud][q];
            continue;
        }
        double u = 0.0;
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                int q[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) q[a] = clamp_i(pc[a] - (a == dim) + ii * (a == dim) + jj * (a == ud), s[a]);
                const size_t n = flat3(q[0], q[1], q[2], s);
                u += (double)(MASKED ? face_val(U[ud], F.act[ud], n) : U[ud][n]);
            }
        u /= 4.0;
        out[ud] = (float)u;
    }
}

// macarray3::convert_to_full at cell (i,j,k): valid when both faces of every axis are active.
__device__ __forceinline__ void full_cell(const VelField &F, int i, int j, int k, double out[3]) {
#pragma clang fp contract(off)
    int valid = 0;
    float v[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int s[3];
        vf_shape(F, a, s);
        const size_t q0 = flat3(i, j, k, s);
        const size_t q1 = flat3(i + (a == 0), j + (a == 1), k + (a == 2), s);
        // values and activity bytes load together
        const float u0 = F.vel[a][q0], u1 = F.vel[a][q1];
        const bool a0 = F.act[a][q0] != 0, a1 = F.act[a][q1] != 0;
        int wsum = 0;
        double value = 0.0;
        if (a0) {
            value += (double)u0;
            wsum++;
        }
        if (a1) {
            value += (double)u1;
            wsum++;
        }
        if (wsum == 2) {
            v[a] = (float)(0.5 * value);
            valid++;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = valid == 3 ? (double)v[a] : 0.0;
}

// The 8 corner values of c (inactive entries 0 when A is given), and the interpolation and the
// bounds computed from them: the value pass and the min/max pass share their stencil unless the
// two departure points (float and double arithmetic) fall in different cells.
__device__ __forceinline__ void load8(const float *U, const uint8_t *A, const int s[3], const Corner &c, float v[8]) {
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const size_t q = corner_index(c, n, s);
        v[n] = A ? face_val(U, A, q) : U[q];
    }
}
__device__ __forceinline__ float interp8(const Corner &c, const float v[8]) {
#pragma clang fp contract(off)
    double w[8];
    interp_weights(c, w);
    float val = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n)
        if (w[n] != 0.0) val = acc<float>(val, (double)v[n] * w[n]);
    return val;
}
// interpolate<Real> over a float array of shape s; inactive entries read 0 when A is given (the
// eight corners are loaded at once)
__device__ __forceinline__ float interp_arr(const float *U, const uint8_t *A, const int s[3], const double p[3]) {
    const Corner c = corner_of(s, p);
    float v[8];
    load8(U, A, s, c, v);
    return interp8(c, v);
}
__device__ __forceinline__ void bounds8(const float v[8], double &lo, double &hi) {
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double x = (double)v[n];
        lo = x < lo ? x : lo;  // std::min
        hi = hi < x ? x : hi;  // std::max
    }
}
__device__ __forceinline__ bool same_cell(const Corner &a, const Corner &b) {
    return a.i == b.i && a.j == b.j && a.k == b.k;
}

// min and max of the 8 corner values at p (interpolate_coef's stencil), from the given initial
// bounds (the reference's numeric_limits initialisers)
__device__ __forceinline__ void corner_bounds(const float *U, const uint8_t *A, const int s[3], const double p[3],
                                              double &lo, double &hi) {
    const Corner c = corner_of(s, p);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const size_t q = corner_index(c, n, s);
        const double v = A ? (double)face_val(U, A, q) : (double)U[q];
        lo = v < lo ? v : lo;  // std::min
        hi = hi < v ? v : hi;  // std::max
    }
}

// the liquid level set at p (cell index space): float accumulation for a float level set
__device__ __forceinline__ double interp_level(const void *phi, int f64, const int s[3], const double p[3]) {
#pragma clang fp contract(off)
    const Corner c = corner_of(s, p);
    double w[8];
    interp_weights(c, w);
    // all eight corners are loaded at once (advection asks for at least 2 cells per axis, so they
    // lie in the array); zero weights still skip the accumulation
    if (f64) {
        double v[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) v[n] = ((const double *)phi)[corner_index(c, n, s)];
        double val = 0.0;
#pragma unroll
        for (int n = 0; n < 8; ++n)
            if (w[n] != 0.0) val += v[n] * w[n];
        return val;
    }
    float v[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) v[n] = ((const float *)phi)[corner_index(c, n, s)];
    float val = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n)
        if (w[n] != 0.0) val = acc<float>(val, (double)v[n] * w[n]);
    return (double)val;
}

// Six-point WENO interpolation (Macdonald & Ruuth, J. Sci. Comput. 35 (2008) 219-240; the
// reference's WENO::interp6, include/shiokaze/math/WENO.h:55-88).  f holds the samples at the
// nodes -2..3 and x in [0, 1] the position.  The three cubics through four consecutive nodes
// (Newton form) are blended with weights gamma_r / (eps + beta_r^2) normalised to sum 1: gamma_r
// the linear weights (which depend on x only and are passed in, computed once per axis), beta_r
// the smoothness indicators, quadratic forms in the samples whose terms (coefficient, sample a,
// sample b; 180 beta_r) are tabulated in the reference's order.  Every operation is the
// reference's, in its order, so the double is the same.
namespace weno {
struct Term {
    double c;
    int a, b;
};
constexpr Term BETA[3][10] = {
    {{814., 3, 3}, {4326., 2, 2}, {2976., 1, 1}, {244., 0, 0}, {-3579., 2, 3},
     {-6927., 2, 1}, {1854., 2, 0}, {2634., 3, 1}, {-683., 3, 0}, {-1659., 1, 0}},
    {{1986., 3, 3}, {1986., 2, 2}, {244., 1, 1}, {244., 4, 4}, {1074., 2, 4},
     {-3777., 2, 3}, {-1269., 2, 1}, {1074., 3, 1}, {-1269., 4, 3}, {-293., 4, 1}},
    {{814., 2, 2}, {4326., 3, 3}, {2976., 4, 4}, {244., 5, 5}, {-683., 2, 5},
     {2634., 2, 4}, {-3579., 2, 3}, {-6927., 3, 4}, {1854., 3, 5}, {-1659., 4, 5}},
};
}  // namespace weno

__device__ __forceinline__ void weno_gamma(double x, double g[3]) {
#pragma clang fp contract(off)
    g[0] = (2.0 - x) * (3.0 - x) / 20.0;
    g[1] = (3.0 - x) * (x + 2.0) / 10.0;
    g[2] = (x + 2.0) * (x + 1.0) / 20.0;
}

__device__ __forceinline__ double weno6(const double g[3], double x, const double f[6]) {
#pragma clang fp contract(off)
    double alpha[3], poly[3], total = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double beta = 0.0;
#pragma unroll
        for (int t = 0; t < 10; ++t) {
            const weno::Term T = weno::BETA[r][t];
            const double v = T.a == T.b ? T.c * (f[T.a] * f[T.a]) : (T.c * f[T.a]) * f[T.b];
            beta = t ? beta + v : v;
        }
        beta = beta / 180.0;
        const double e0 = x - (double)(r - 2), e1 = x - (double)(r - 1), e2 = x - (double)r;
        const double d1 = f[r + 1] - f[r];
        const double d2 = f[r + 2] - 2. * f[r + 1] + f[r];
        const double d3 = f[r + 3] - 3. * f[r + 2] + 3. * f[r + 1] - f[r];
        poly[r] = f[r] + d1 * e0 + d2 * e0 * e1 / 2.0 + d3 * e0 * e1 * e2 / 6.0;
        alpha[r] = g[r] / (DBL_EPSILON + beta * beta);
        total += alpha[r];
    }
    const double t0 = alpha[0] / total * poly[0], t1 = alpha[1] / total * poly[1], t2 = alpha[2] / total * poly[2];
    return t0 + t1 + t2;
}

__device__ __forceinline__ double dmin_std(double a, double b) { return b < a ? b : a; }  // std::min
__device__ __forceinline__ double dmax_std(double a, double b) { return a < b ? b : a; }  // std::max

// WENO3::interpolate(array, p) with order 6 over a float array of shape s: p clamped into the
// array, base index min(x, n-2) truncated, the 6x6x6 samples at base-2 .. base+3 (clamped per
// axis) reduced along x, then y, then z.  Inactive entries read 0 when A is given.  The y and z
// partials enter 6-deep shift registers, so the two outer loops stay rolled (one copy of the
// x-row code) and index nothing dynamically.
__device__ double weno3_arr(const float *U, const uint8_t *A, const int s[3], const double p[3]) {
#pragma clang fp contract(off)
    int b[3];
    double fr[3], g[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double c = dmax_std(0.0, dmin_std(s[a] - 1.0, p[a]));
        b[a] = (int)dmin_std(c, s[a] - 2.0);
        fr[a] = c - b[a];
        weno_gamma(fr[a], g[a]);
    }
    double vz[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int kk = 0; kk < 6; ++kk) {
        const int k = clamp_i(b[2] + kk - 2, s[2]);
        double vy[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int jj = 0; jj < 6; ++jj) {
            const int j = clamp_i(b[1] + jj - 2, s[1]);
            double vx[6];
#pragma unroll
            for (int ii = 0; ii < 6; ++ii) {
                const size_t q = flat3(clamp_i(b[0] + ii - 2, s[0]), j, k, s);
                vx[ii] = A ? (double)face_val(U, A, q) : (double)U[q];
            }
            const double r = weno6(g[0], fr[0], vx);
#pragma unroll
            for (int n = 0; n < 5; ++n) vy[n] = vy[n + 1];
            vy[5] = r;
        }
        const double r = weno6(g[1], fr[1], vy);
#pragma unroll
        for (int n = 0; n < 5; ++n) vz[n] = vz[n + 1];
        vz[5] = r;
    }
    return weno6(g[2], fr[2], vz);
}

// The value a semi-Lagrangian pass samples at p: trilinear (interpolate<Real>) or WENO.
template <int WENO>
__device__ __forceinline__ float sample_arr(const float *U, const uint8_t *A, const int s[3], const double p[3]) {
    if (WENO) return (float)weno3_arr(U, A, s, p);
    return interp_arr(U, A, s, p);
}

// Forward pass over every active face: S.v0 = the semi-Lagrangian value; with `minmax`, the
// MacCormack bounds and the narrow-band flag.  `in` is the field being advected (its own full
// velocity moves it, :72-73).
template <int WENO>
__global__ void k_adv_face_fwd(VelField F, const float *in0, const float *in1, const float *in2, AdvScratch S,
                               const void *fluid, int fluid64, double dt, int trim, int minmax) {
#pragma clang fp contract(off)
    int a, i, j, k;
    size_t f;
    if (!vf_decode(F, a, f, i, j, k)) return;
    if (!F.act[a][f]) {
        S.v0[a][f] = 0.0f;  // what the activity mask reads there
        return;
    }
    const float *const in[3] = {in0, in1, in2};
    float u[3];
    full_face<true>(F, in, a, i, j, k, u);
    const bool empty = u[0] == 0.0f && u[1] == 0.0f && u[2] == 0.0f;
    int s[3];
    vf_shape(F, a, s);
    const int pc[3] = {i, j, k};
    float value;
    float v8[8];
    Corner cv;
    if (!empty) {
        double p[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float t = (float)((double)u[b] * dt);   // dt*u on a float vector
            const float sft = (float)vf_div_dx(F, (double)t);  // /dx on a float vector
            p[b] = (double)pc[b] - (double)sft;
        }
        if (WENO) {
            value = (float)weno3_arr(in[a], F.act[a], s, p);
        } else {
            cv = corner_of(s, p);
            load8(in[a], F.act[a], s, cv, v8);
            value = interp8(cv, v8);
        }
    } else {
        value = face_val(in[a], F.act[a], f);
    }
    S.v0[a][f] = value;
    if (!minmax) return;
    double fp[3];
    float lo, hi;
    if (!empty) {
        double p[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double sft = vf_div_dx(F, dt * (double)u[b]);  // dt*u/dx on a double vector
            p[b] = (double)pc[b] - sft;
            fp[b] = (pc[b] + 0.5 * (a != b)) - sft;
        }
        double dlo = DBL_MAX, dhi = DBL_MIN;
        const Corner cm = corner_of(s, p);
        if (WENO || !same_cell(cm, cv)) load8(in[a], F.act[a], s, cm, v8);
        bounds8(v8, dlo, dhi);
        lo = (float)dlo;
        hi = (float)dhi;
    } else {
#pragma unroll
        for (int b = 0; b < 3; ++b) fp[b] = pc[b] + 0.5 * (a != b);
        lo = hi = face_val(in[a], F.act[a], f);
    }
    const int sc[3] = {F.w, F.h, F.d};
    const double q[3] = {fp[0] - 0.5, fp[1] - 0.5, fp[2] - 0.5};
    S.lo[a][f] = lo;
    S.hi[a][f] = hi;
    S.nb[a][f] = interp_level(fluid, fluid64, sc, q) > -F.dx * (double)trim ? 1 : 0;
}

// Backward pass from S.v0 (-dt) fused with MacCormack's final step, written over F.vel (the
// input field, read only at the thread's own face).
template <int WENO>
__global__ void k_adv_face_bwd(VelField F, AdvScratch S, double dt_back) {
#pragma clang fp contract(off)
    int a, i, j, k;
    size_t f;
    if (!vf_decode(F, a, f, i, j, k)) return;
    if (!F.act[a][f]) return;
    const float v0 = S.v0[a][f];
    if (S.nb[a][f]) {
        F.vel[a][f] = v0;
        return;
    }
    const float *const in[3] = {S.v0[0], S.v0[1], S.v0[2]};
    float u[3];
    full_face<false>(F, in, a, i, j, k, u);
    int s[3];
    vf_shape(F, a, s);
    const int pc[3] = {i, j, k};
    float v1;
    if (u[0] != 0.0f || u[1] != 0.0f || u[2] != 0.0f) {
        double p[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float t = (float)((double)u[b] * dt_back);
            const float sft = (float)vf_div_dx(F, (double)t);
            p[b] = (double)pc[b] - (double)sft;
        }
        v1 = sample_arr<WENO>(in[a], nullptr, s, p);
    } else {
        v1 = v0;
    }
    const double lo = S.lo[a][f], hi = S.hi[a][f];
    const float diff = F.vel[a][f] - v1;
    const double r = (double)v0 + 0.5 * (double)diff;
    F.vel[a][f] = r < lo ? (float)lo : (r > hi ? (float)hi : (float)r);
}

// semi-Lagrangian only: the forward result over the active faces
__global__ void k_adv_face_copy(VelField F, AdvScratch S) {
    int a, i, j, k;
    size_t f;
    if (!vf_decode(F, a, f, i, j, k)) return;
    if (F.act[a][f]) F.vel[a][f] = S.v0[a][f];
}

// Forward pass over the cells of q (advected by F): S.v0[0] = the semi-Lagrangian value, or
// `background` on inactive cells (what the backward pass reads there); with `minmax`, bounds
// and the narrow-band flag against `fluid` (cells).
template <int WENO>
__global__ void k_adv_cell_fwd(VelField F, const float *q, const uint8_t *aq, AdvScratch S, const void *fluid,
                               int fluid64, double dt, int trim, float background, int minmax) {
#pragma clang fp contract(off)
    const size_t c = (size_t)xcd_brick(blockIdx.x, gridDim.x) * VF_BLOCK + threadIdx.x;
    const size_t M = (size_t)F.w * F.h * F.d;
    if (c >= M) return;
    if (aq && !aq[c]) {
        S.v0[0][c] = background;
        return;
    }
    int i, j, k;
    unflatten((unsigned)c, (unsigned)F.h, (unsigned)F.d, i, j, k);
    double u[3];
    full_cell(F, i, j, k, u);
    const bool empty = u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0;
    const int s[3] = {F.w, F.h, F.d};
    const int pc[3] = {i, j, k};
    double p[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) p[b] = (double)pc[b] - vf_div_dx(F, dt * u[b]);
    float v8[8];
    const Corner cv = corner_of(s, p);
    if (!empty) {
        if (WENO) {
            S.v0[0][c] = (float)weno3_arr(q, nullptr, s, p);
        } else {
            load8(q, nullptr, s, cv, v8);
            S.v0[0][c] = interp8(cv, v8);
        }
    } else {
        S.v0[0][c] = q[c];
    }
    if (!minmax) return;
    if (!empty) {
        double lo = (double)FLT_MAX, hi = (double)FLT_MIN;
        if (WENO) load8(q, nullptr, s, cv, v8);
        bounds8(v8, lo, hi);
        S.lo[0][c] = (float)lo;
        S.hi[0][c] = (float)hi;
        S.nb[0][c] = interp_level(fluid, fluid64, s, p) > -F.dx * (double)trim ? 1 : 0;
    } else {
        S.lo[0][c] = S.hi[0][c] = q[c];
        const double fc = fluid64 ? ((const double *)fluid)[c] : (double)((const float *)fluid)[c];
        S.nb[0][c] = fc > -F.dx * (double)trim ? 1 : 0;
    }
}

// Backward pass from S.v0[0] fused with the final MacCormack step, written over q.
template <int WENO>
__global__ void k_adv_cell_bwd(VelField F, float *q, const uint8_t *aq, AdvScratch S, double dt_back) {
#pragma clang fp contract(off)
    const size_t c = (size_t)xcd_brick(blockIdx.x, gridDim.x) * VF_BLOCK + threadIdx.x;
    const size_t M = (size_t)F.w * F.h * F.d;
    if (c >= M) return;
    if (aq && !aq[c]) return;
    const float v0 = S.v0[0][c];
    if (S.nb[0][c]) {
        q[c] = v0;
        return;
    }
    int i, j, k;
    unflatten((unsigned)c, (unsigned)F.h, (unsigned)F.d, i, j, k);
    double u[3];
    full_cell(F, i, j, k, u);
    const int s[3] = {F.w, F.h, F.d};
    const int pc[3] = {i, j, k};
    float v1;
    if (u[0] != 0.0 || u[1] != 0.0 || u[2] != 0.0) {
        double p[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) p[b] = (double)pc[b] - vf_div_dx(F, dt_back * u[b]);
        v1 = sample_arr<WENO>(S.v0[0], nullptr, s, p);
    } else {
        v1 = v0;
    }
    const double lo = S.lo[0][c], hi = S.hi[0][c];
    const float diff = q[c] - v1;
    const double r = (double)v0 + 0.5 * (double)diff;
    q[c] = r < lo ? (float)lo : (r > hi ? (float)hi : (float)r);
}

__global__ void k_adv_cell_copy(size_t M, float *q, const uint8_t *aq, AdvScratch S) {
    const size_t c = (size_t)blockIdx.x * VF_BLOCK + threadIdx.x;
    if (c < M && (!aq || aq[c])) q[c] = S.v0[0][c];
}

}  // namespace shkz
