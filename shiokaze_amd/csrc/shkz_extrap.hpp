// shkz_extrap.hpp -- the step after a projection in Shiokaze's time steppers, on the device:
// velocity extrapolation and the solid constraint, macutility3::extrapolate_and_constrain_velocity
// (src/utility/macutility3.cpp:89-93), which macliquid3::extend_both (macliquid3.cpp:309-319) and
// macsmoke3 (macsmoke3.cpp:294) call on the velocity the projection left.
//
//   k_pack_activity,
//   k_extrap_pass,
//   k_unpack_activity  array_extrapolator3::extrapolate (array_extrapolator3.h:47-80, over
//                   dilate3.h:43-101) on all three face arrays, on bit-packed activity
//   k_constrain,
//   k_constrain_apply  macutility3::constrain_velocity (macutility3.cpp:61-88) with the
//                   interpolation of array_interpolator3.h:40-97, array_derivative3.h:40-103 and
//                   macarray_interpolator3.h:43-50; the first kernel writes constrained values to
//                   a side buffer, the second applies them and the boundary rule, so the velocity
//                   needs no full copy (the reference samples a copy, velocity_save)
//
// Layout: the dense face arrays of the projection (u (w+1,h,d), v (w,h+1,d), w (w,h,d+1), z
// fastest) with one activity byte per face.  One launch covers the three arrays (blockIdx.y
// selects the array).  The extrapolation passes ping-pong two bit-packed copies of the activity
// and write the values of the faces they activate in place (no thread of a pass reads them).
//
// Arithmetic matches the reference's types operation by operation (no contraction): float sums
// of float velocities, interpolation into the array's element type after a double product, vec3d
// normals and dot products, the result rounded to float.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shkz_device.hpp"

namespace shkz {

struct VelField {
    int w, h, d;
    int solid64;          // solid level set in double (else float, the reference's Real)
    double dx;
    const void *solid;    // nodal (w+1,h+1,d+1) or nullptr
    float *vel[3];
    uint8_t *act[3];
    size_t n[3];          // faces per array
    // dx a power of two (1/256 for a unit box of 256 cells): x / dx then equals x * inv_dx bit for
    // bit (both round the same exact value once; the inverse is a power of two), and the kernels
    // multiply instead of dividing (an FP64 division is about eleven instructions here)
    int dx_pow2;
    double inv_dx;
};

// x / F.dx, exactly as the division rounds.
__device__ __forceinline__ double vf_div_dx(const VelField &F, double x) { return F.dx_pow2 ? x * F.inv_dx : x / F.dx; }

__device__ __forceinline__ void vf_shape(const VelField &F, int a, int s[3]) {
    s[0] = F.w + (a == 0);
    s[1] = F.h + (a == 1);
    s[2] = F.d + (a == 2);
}

// Face kernels launch a (ceil(max n / 256), 3) grid: blockIdx.y is the face array, so the array
// index is uniform across a wavefront and every per-array field (extent, pointers) is a scalar
// load.  (A flat index over the three arrays made it vary inside the wavefront that straddles two
// arrays, and hipcc then lowered a dynamically indexed kernel-argument read with v_readfirstlane,
// as if it were uniform: that wavefront's lanes read the first lane's array extent.)
// Threads per block of every face, cell and activity-word kernel here.  The kernels index with
// this constant, not blockDim.x: the workgroup size is a 16-bit field of the hidden kernel
// arguments, which hipcc reads with a vector load, one more memory round trip before a wave can
// form its index.
constexpr int VF_BLOCK = 256;
inline dim3 vf_grid(const size_t n[3]) {
    const size_t m = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
    return dim3((unsigned)(((m + 255) / 256 + 7) / 8 * 8), 3);
}

// (array, local index, i, j, k) of this thread; false past the end of its array
__device__ __forceinline__ bool vf_decode(const VelField &F, int &a, size_t &f, int &i, int &j, int &k) {
    a = (int)blockIdx.y;
    f = (size_t)xcd_brick(blockIdx.x, gridDim.x) * VF_BLOCK + threadIdx.x;  // XCD-contiguous slabs
    if (f >= F.n[a]) return false;
    int s[3];
    vf_shape(F, a, s);
    // 32-bit decode (each face array holds fewer than 2^32 faces, checked at entry)
    const unsigned u = (unsigned)f, t = u / (unsigned)s[2], ii = t / (unsigned)s[1];
    k = (int)(u - t * (unsigned)s[2]);
    j = (int)(t - ii * (unsigned)s[1]);
    i = (int)ii;
    return true;
}

// The extrapolation works on bit-packed activity: each row (i, j) of a face array is a run of
// 32-bit words along k, bit b of word kw = face k = 32 kw + b.  A pass is then one thread per word:
// it forms the six neighbour masks from the word, its two neighbours along k and the words of the
// rows +-1 in j and i, and activates the inactive faces with an active neighbour.
struct BitField {
    uint32_t *bits[3];
    int W[3];  // words per row
};

__device__ __forceinline__ int vf_rows(const VelField &F, int a) { return (F.w + (a == 0)) * (F.h + (a == 1)); }

// activity bytes -> bits: one lane per face and word bit, a wavefront ballot forms two words; each
// wavefront packs PACK_GROUPS consecutive groups of 64 faces (its loads all issued first), so the
// launch has a quarter of the wavefronts of one group each.
constexpr int PACK_GROUPS = 4;
__global__ void k_pack_activity(VelField F, BitField B) {
    const int a = (int)blockIdx.y;
    int s[3];
    vf_shape(F, a, s);
    const int W = B.W[a];
    const size_t nw = (size_t)vf_rows(F, a) * W;
    const unsigned lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * VF_BLOCK + threadIdx.x) >> 6;
    bool on[PACK_GROUPS];
#pragma unroll
    for (int g = 0; g < PACK_GROUPS; ++g) {
        const size_t slot = (wave * PACK_GROUPS + g) * 64 + lane;
        const size_t w = slot >> 5;
        on[g] = false;
        if (w < nw) {
            const unsigned r = (unsigned)w / (unsigned)W, kw = (unsigned)w - r * (unsigned)W;  // 32-bit: fewer than 2^32 words
            const int k = (int)(kw * 32 + (lane & 31));
            if (k < s[2]) on[g] = F.act[a][(size_t)r * s[2] + k] != 0;
        }
    }
#pragma unroll
    for (int g = 0; g < PACK_GROUPS; ++g) {
        const unsigned long long m = __ballot(on[g]);
        const size_t w = ((wave * PACK_GROUPS + g) * 64 + lane) >> 5;
        if (w < nw && (lane & 31) == 0) B.bits[a][w] = (uint32_t)((lane & 32) ? (m >> 32) : m);
    }
}

// One extrapolation pass: Bo = Bi plus every inactive face with an in-bounds neighbour active in
// Bi, which takes the mean of those neighbours, summed in float in the order +x, -x, +y, -y, -z,
// +z (array_extrapolator3.h:57-72).  The faces a pass activates are inactive in Bi, so no thread
// of the pass reads their values.
__global__ void k_extrap_pass(VelField F, BitField Bi, BitField Bo) {
#pragma clang fp contract(off)
    const int a = (int)blockIdx.y;
    int s[3];
    vf_shape(F, a, s);
    const int W = Bi.W[a];
    const size_t w = (size_t)blockIdx.x * VF_BLOCK + threadIdx.x;
    const size_t nw = (size_t)vf_rows(F, a) * W;
    if (w >= nw) return;
    const unsigned r = (unsigned)w / (unsigned)W, kw = (unsigned)w - r * (unsigned)W;  // 32-bit: fewer than 2^32 words
    const int i = (int)(r / (unsigned)s[1]), j = (int)(r - (unsigned)i * s[1]);
    const uint32_t *B = Bi.bits[a];
    const uint32_t A = B[w];
    const uint32_t nxt = kw + 1 < (unsigned)W ? B[w + 1] : 0u, prv = kw > 0 ? B[w - 1] : 0u;
    const uint32_t kp = (A >> 1) | (nxt << 31);  // bit b: face k+1 active
    const uint32_t km = (A << 1) | (prv >> 31);  // bit b: face k-1 active
    const uint32_t jp = j + 1 < s[1] ? B[w + W] : 0u, jm = j > 0 ? B[w - W] : 0u;
    const size_t sxw = (size_t)s[1] * W;
    const uint32_t ip = i + 1 < s[0] ? B[w + sxw] : 0u, im = i > 0 ? B[w - sxw] : 0u;
    const int nbits = s[2] - (int)kw * 32;
    const uint32_t valid = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
    uint32_t cand = (kp | km | jp | jm | ip | im) & ~A & valid;
    Bo.bits[a][w] = A | cand;
    if (!cand) return;
    float *U = F.vel[a];
    const size_t sx = (size_t)s[1] * s[2], sy = (size_t)s[2];
    const size_t f0 = (size_t)r * s[2] + (size_t)kw * 32;
    while (cand) {
        const int b = __builtin_ctz(cand);
        cand &= cand - 1;
        const size_t f = f0 + b;
        float sum = 0.0f;
        int weight = 0;
        if ((ip >> b) & 1u) { sum += U[f + sx]; weight++; }
        if ((im >> b) & 1u) { sum += U[f - sx]; weight++; }
        if ((jp >> b) & 1u) { sum += U[f + sy]; weight++; }
        if ((jm >> b) & 1u) { sum += U[f - sy]; weight++; }
        if ((km >> b) & 1u) { sum += U[f - 1]; weight++; }
        if ((kp >> b) & 1u) { sum += U[f + 1]; weight++; }
        // float / int in the reference; the double quotient rounded once to float is the
        // correctly rounded float quotient
        U[f] = (float)((double)sum / (double)weight);
    }
}

// bits -> activity bytes for the faces the passes activated (final & ~initial)
__global__ void k_unpack_activity(VelField F, BitField B0, BitField Bf) {
    const int a = (int)blockIdx.y;
    int s[3];
    vf_shape(F, a, s);
    const int W = Bf.W[a];
    const size_t w = (size_t)blockIdx.x * VF_BLOCK + threadIdx.x;
    if (w >= (size_t)vf_rows(F, a) * W) return;
    uint32_t fresh = Bf.bits[a][w] & ~B0.bits[a][w];
    if (!fresh) return;
    const unsigned r = (unsigned)w / (unsigned)W, kw = (unsigned)w - r * (unsigned)W;  // 32-bit: fewer than 2^32 words
    const size_t f0 = (size_t)r * s[2] + (size_t)kw * 32;
    while (fresh) {
        const int b = __builtin_ctz(fresh);
        fresh &= fresh - 1;
        F.act[a][f0 + b] = 1;
    }
}

__device__ __forceinline__ double dmin_(double a, double b) { return b < a ? b : a; }  // std::min
__device__ __forceinline__ double dmax_(double a, double b) { return a < b ? b : a; }  // std::max

// Corner base index and clamped position of array_interpolator3::interpolate_coef (:40-62).
struct Corner {
    int i, j, k;
    double x, y, z;
};
__device__ __forceinline__ Corner corner_of(const int s[3], const double p[3]) {
#pragma clang fp contract(off)
    Corner c;
    c.x = dmax_(0.0, dmin_(s[0] - 1., p[0]));
    c.y = dmax_(0.0, dmin_(s[1] - 1., p[1]));
    c.z = dmax_(0.0, dmin_(s[2] - 1., p[2]));
    c.i = (int)dmin_(c.x, s[0] - 2.);
    c.j = (int)dmin_(c.y, s[1] - 2.);
    c.k = (int)dmin_(c.z, s[2] - 2.);
    return c;
}

// Trilinear weights in corner order (i,j,k), (i+1,j,k), (i,j+1,k), (i+1,j+1,k), then k+1.
__device__ __forceinline__ void interp_weights(const Corner &c, double w[8]) {
#pragma clang fp contract(off)
    const double zk1 = (c.k + 1) - c.z, zk = c.z - c.k, xi1 = (c.i + 1) - c.x, xi = c.x - c.i, yj1 = (c.j + 1) - c.y,
                 yj = c.y - c.j;
    w[0] = zk1 * xi1 * yj1;
    w[1] = zk1 * xi * yj1;
    w[2] = zk1 * xi1 * yj;
    w[3] = zk1 * xi * yj;
    w[4] = zk * xi1 * yj1;
    w[5] = zk * xi * yj1;
    w[6] = zk * xi1 * yj;
    w[7] = zk * xi * yj;
}

// Derivative weights of array_derivative3::derivative_interpolate_coef (:40-85), dw[8*dim+n].
__device__ __forceinline__ void deriv_weights(const Corner &c, double dw[24]) {
#pragma clang fp contract(off)
    const double zk1 = (c.k + 1) - c.z, zk = c.z - c.k, xi1 = (c.i + 1) - c.x, xi = c.x - c.i, yj1 = (c.j + 1) - c.y,
                 yj = c.y - c.j;
    dw[0] = -zk1 * yj1;
    dw[1] = zk1 * yj1;
    dw[2] = -zk1 * yj;
    dw[3] = zk1 * yj;
    dw[4] = -zk * yj1;
    dw[5] = zk * yj1;
    dw[6] = -zk * yj;
    dw[7] = zk * yj;
    dw[8] = -zk1 * xi1;
    dw[9] = -zk1 * xi;
    dw[10] = zk1 * xi1;
    dw[11] = zk1 * xi;
    dw[12] = -zk * xi1;
    dw[13] = -zk * xi;
    dw[14] = zk * xi1;
    dw[15] = zk * xi;
    dw[16] = -xi1 * yj1;
    dw[17] = -xi * yj1;
    dw[18] = -xi1 * yj;
    dw[19] = -xi * yj;
    dw[20] = xi1 * yj1;
    dw[21] = xi * yj1;
    dw[22] = xi1 * yj;
    dw[23] = xi * yj;
}

// Flat index of corner n of c in an array of shape s.  32-bit arithmetic: the entry points that
// sample face, cell and node arrays reject grids with (w+1)(h+1)(d+1) >= 2^32.
__device__ __forceinline__ size_t corner_index(const Corner &c, int n, const int s[3]) {
    const unsigned i = c.i + (n & 1), j = c.j + ((n >> 1) & 1), k = c.k + ((n >> 2) & 1);
    return (i * (unsigned)s[1] + j) * (unsigned)s[2] + k;
}

// Accumulation in the array's element type T after a double product (value += array(idx)*coef).
template <class T>
__device__ __forceinline__ T acc(T v, double term) {
#pragma clang fp contract(off)
    return (T)((double)v + term);
}

// The solid sample and its derivative at p (nodal array), in the level set's type T.  Returns the
// interpolated value where the point is inside (negative); 0 where it is not.
template <class T>
__device__ __forceinline__ double solid_sample(const VelField &F, const double p[3], double der[3], bool &inside) {
#pragma clang fp contract(off)
    const int s[3] = {F.w + 1, F.h + 1, F.d + 1};
    const Corner c = corner_of(s, p);
    const T *S = (const T *)F.solid;
    T v[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) v[n] = S[corner_index(c, n, s)];
    // with no negative corner the interpolation is a sum of non-negative products (a NaN corner
    // makes it NaN): never below 0 whatever the rounding, so the point is outside and the weights
    // are not needed (most faces lie away from the solid)
    bool any_neg = false;
#pragma unroll
    for (int n = 0; n < 8; ++n) any_neg = any_neg || v[n] < T(0);
    if (!any_neg) {
        inside = false;
        return 0.0;
    }
    double w[8];
    interp_weights(c, w);
    T val = T(0);
#pragma unroll
    for (int n = 0; n < 8; ++n)
        if (w[n] != 0.0) val = acc<T>(val, (double)v[n] * w[n]);
    inside = (double)val < 0.0;
    if (inside) {
        double dw[24];
        deriv_weights(c, dw);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T r = T(0);
#pragma unroll
            for (int n = 0; n < 8; ++n) r = acc<T>(r, dw[8 * a + n] * (double)v[n]);
            der[a] = (double)r;
        }
    }
    return (double)val;
}

// Component b of the saved velocity at p (macarray_interpolator3.h:43-50): inactive faces read 0.
__device__ __forceinline__ double vel_sample(const VelField &F, const float *const Sv[3], int b, const double p[3]) {
#pragma clang fp contract(off)
    int s[3];
    vf_shape(F, b, s);
    const double pos[3] = {p[0] - 0.5 * (b != 0), p[1] - 0.5 * (b != 1), p[2] - 0.5 * (b != 2)};
    const Corner c = corner_of(s, pos);
    double w[8];
    interp_weights(c, w);
    // all eight corners are loaded at once (they lie in the array: the constraint asks for at
    // least 2 cells per axis); zero weights still skip the accumulation
    float u[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const size_t q = corner_index(c, n, s);
        const float sv = Sv[b][q];
        u[n] = F.act[b][q] ? sv : 0.0f;
    }
    float val = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n)
        if (w[n] != 0.0) val = acc<float>(val, (double)u[n] * w[n]);
    return (double)val;
}

// Activity code of a face whose constrained value waits in the side buffer (k_constrain_apply).
constexpr uint8_t ACT_CONSTRAINED = 255;

// macutility3::constrain_velocity, first half.  For every active face: the activity codes of the
// passes become 1; when `constrain` is set and the solid level set has a negative value (*exist),
// a face whose centre samples the solid below 0 and whose velocity points into the solid gets
// the velocity minus its normal component, written to X (not to the velocity, which neighbouring
// threads still sample, as the reference samples its velocity_save copy) and marked
// ACT_CONSTRAINED.  The velocity arrays are only read here.
__global__ void k_constrain(VelField F, float *X0, float *X1, float *X2, const int *exist, int constrain) {
#pragma clang fp contract(off)
    int a, i, j, k;
    size_t f;
    if (!vf_decode(F, a, f, i, j, k)) return;
    const bool with_solid = constrain && F.solid && *exist;  // loaded with the activity byte
    const int t = F.act[a][f];
    if (!t) return;
    if (with_solid) {
        const float *const Sv[3] = {F.vel[0], F.vel[1], F.vel[2]};
        const double p[3] = {i + 0.5 * (a != 0), j + 0.5 * (a != 1), k + 0.5 * (a != 2)};
        double der[3] = {0.0, 0.0, 0.0};
        bool inside;
        if (F.solid64) solid_sample<double>(F, p, der, inside);
        else solid_sample<float>(F, p, der, inside);
        if (inside) {
            double nrm[3];
#pragma unroll
            for (int b = 0; b < 3; ++b) nrm[b] = vf_div_dx(F, der[b]);
            double n2 = 0.0;
#pragma unroll
            for (int b = 0; b < 3; ++b) n2 += nrm[b] * nrm[b];
            if (n2 != 0.0) {
                double u[3];
#pragma unroll
                for (int b = 0; b < 3; ++b) u[b] = vel_sample(F, Sv, b, p);
                double un = 0.0;
#pragma unroll
                for (int b = 0; b < 3; ++b) un += u[b] * nrm[b];
                if (un < 0.0) {
                    float *const X[3] = {X0, X1, X2};
                    X[a][f] = (float)(u[a] - nrm[a] * un);
                    F.act[a][f] = ACT_CONSTRAINED;  // nonzero: neighbours still read the face as active
                    return;
                }
            }
        }
    }
    if (t > 1) F.act[a][f] = 1;
}

// Second half: constrained faces take their value from X, then (with the constraint on and a
// solid present) active faces on the domain boundary pointing out of the domain become 0.
__global__ void k_constrain_apply(VelField F, const float *X0, const float *X1, const float *X2, const int *exist,
                                  int constrain) {
    int a, i, j, k;
    size_t f;
    if (!vf_decode(F, a, f, i, j, k)) return;
    const int t = F.act[a][f];
    const int pi[3] = {i, j, k}, N[3] = {F.w, F.h, F.d};
    const bool boundary = pi[a] == 0 || pi[a] == N[a];
    if (t != ACT_CONSTRAINED && !(t && boundary)) return;
    if (!(constrain && F.solid && *exist)) return;
    const float *const X[3] = {X0, X1, X2};
    float out = t == ACT_CONSTRAINED ? X[a][f] : F.vel[a][f];
    if (pi[a] == 0 && out < 0.0f) out = 0.0f;
    if (pi[a] == N[a] && out > 0.0f) out = 0.0f;
    F.vel[a][f] = out;
    if (t == ACT_CONSTRAINED) F.act[a][f] = 1;
}

}  // namespace shkz
