// shkz_synth.hip -- synthetic scenes generated on the device (inputs of serving workloads).
//
// The shapes come from the host generator's parameters (shkz_scene_params_internal), and the
// level sets are evaluated with the host's arithmetic in the host's order (no contraction), so
// they equal shkz_scene_fill's.  Velocities use a counter-based generator instead of the host's
// sequential mt19937_64 stream: face f of array a gets Box-Muller of two 53-bit uniforms from
// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014) of a
// per-array key and 2f, 2f+1.
#include "../../include/shkz_mi355x.h"
#include "shkz_scene_params.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

__device__ double fluid_sdf(const ShkzSceneParams &S, const double p[3]) {
#pragma clang fp contract(off)
    if (S.all_fluid) return -1.0;
    const double q[3] = {p[0] - S.shift[0], p[1] - S.shift[1], p[2] - S.shift[2]};
    double v = 1e9;
    if (S.pool >= 0.0) v = fmin(v, q[1] - S.pool);
    for (int s = 0; s < S.n_liquid; ++s) {
        const double dx = q[0] - S.liquid[s][0], dy = q[1] - S.liquid[s][1], dz = q[2] - S.liquid[s][2];
        v = fmin(v, sqrt(dx * dx + dy * dy + dz * dz) - S.liquid[s][3]);
    }
    return v;
}

__device__ double solid_sdf(const ShkzSceneParams &S, const double p[3]) {
#pragma clang fp contract(off)
    double v = 1e9;
    for (int s = 0; s < S.n_solid; ++s) {
        const double dx = p[0] - S.solid[s][0], dy = p[1] - S.solid[s][1], dz = p[2] - S.solid[s][2];
        v = fmin(v, sqrt(dx * dx + dy * dy + dz * dz) - S.solid[s][3]);
    }
    if (S.has_box) {
        double sd = -9999.0;
        for (int dim = 0; dim < 3; ++dim) {
            sd = fmax(sd, S.box0[dim] - p[dim]);
            sd = fmax(sd, p[dim] - S.box1[dim]);
        }
        v = fmin(v, sd);
    }
    return v;
}

__global__ void k_fill_fluid(ShkzSceneParams S, int W, int H, int D, double dx, int with_solid, int f64, void *fluid) {
#pragma clang fp contract(off)
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)W * H * D) return;
    const unsigned t = (unsigned)c / (unsigned)D;  // 32-bit decode (boxes below 2^32 cells)
    const int k = (int)((unsigned)c - t * (unsigned)D), j = (int)(t % (unsigned)H), i = (int)(t / (unsigned)H);
    const double p[3] = {(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx};
    double val = fluid_sdf(S, p);
    if (with_solid) val = fmax(val, -(solid_sdf(S, p) + dx));
    if (f64) ((double *)fluid)[c] = val;
    else ((float *)fluid)[c] = (float)val;
}

__global__ void k_fill_solid(ShkzSceneParams S, int W, int H, int D, double dx, int with_solid, float *solid) {
#pragma clang fp contract(off)
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W1 = W + 1, H1 = H + 1, D1 = D + 1;
    if (n >= (size_t)W1 * H1 * D1) return;
    const unsigned t = (unsigned)n / (unsigned)D1;
    const int k = (int)((unsigned)n - t * (unsigned)D1), j = (int)(t % (unsigned)H1), i = (int)(t / (unsigned)H1);
    const double p[3] = {i * dx, j * dx, k * dx};
    solid[n] = with_solid ? (float)solid_sdf(S, p) : 1.0f;
}

__global__ void k_fill_velocity(uint64_t key, size_t n, int mode, float add, float *a) {
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    float v = 0.0f;
    if (mode != 1) {
        const double u1 = (double)((splitmix64(key ^ (2 * (uint64_t)f)) >> 11) + 1) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(splitmix64(key ^ (2 * (uint64_t)f + 1)) >> 11) * (1.0 / 9007199254740992.0);
        v = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
    a[f] = v + add;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int shkz_scene_fill_device(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid,
                                      float *u, float *v, float *w, void *stream) {
    ShkzSceneParams S;
    const int rc = shkz_scene_params_internal(desc, &S);
    if (rc != SHKZ_OK) return rc;
    const int W = desc->w, H = desc->h, D = desc->d;
    if ((uint64_t)(W + 1) * (H + 1) * (D + 1) >= ((uint64_t)1 << 32)) return SHKZ_ERR_INVALID;  // 32-bit indices
    const double dx = std::min(1.0 / W, std::min(1.0 / H, 1.0 / D));
    const int with_solid = (S.n_solid > 0 || S.has_box) ? 1 : 0;
    if (has_solid) *has_solid = with_solid;
    hipStream_t st = (hipStream_t)stream;
    const size_t M = (size_t)W * H * D, Nn = (size_t)(W + 1) * (H + 1) * (D + 1);
    if (fluid)
        k_fill_fluid<<<blocks(M), 256, 0, st>>>(S, W, H, D, dx, with_solid && !S.all_fluid, desc->fluid_dtype == SHKZ_F64 ? 1 : 0,
                                                fluid);
    if (solid) k_fill_solid<<<blocks(Nn), 256, 0, st>>>(S, W, H, D, dx, with_solid, solid);
    const uint64_t seed = desc->velocity_seed ? desc->velocity_seed : desc->seed;
    float *vel[3] = {u, v, w};
    for (int dim = 0; dim < 3; ++dim) {
        if (!vel[dim]) continue;
        const size_t n = (size_t)(W + (dim == 0)) * (H + (dim == 1)) * (D + (dim == 2));
        const uint64_t key = seed * 0x632be59bd9b4e019ULL + 0x5851f42d4c957f2dULL * (uint64_t)(dim + 1);
        const float add = dim == 1 ? (float)(desc->gravity * desc->dt) : 0.0f;
        k_fill_velocity<<<blocks(n), 256, 0, st>>>(key, n, desc->velocity_mode, add, vel[dim]);
    }
    return hipGetLastError() == hipSuccess ? SHKZ_OK : SHKZ_ERR_DEVICE;
}
