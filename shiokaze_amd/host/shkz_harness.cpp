// shkz_harness.cpp -- standalone C++ host harness of the C ABI.
//
//   shkz_harness [n] [steps] [mixed|fp64|cg]
//     Reproduces the call contract of macliquid3::idle (src/liquid/macliquid3.cpp:321-367): every
//     step the caller owns the level sets and the velocity, adds gravity, calls
//     project(dt, velocity, solid, fluid) and keeps the result; here the liquid is translated
//     rigidly between steps instead of advected.
//
//   shkz_harness writeback <n> <seed> <out-file>
//     Runs the Shiokaze module's staging and write-back (shkz_stage.hpp, the code
//     macpressuresolver3_mi355x runs) on a sparse velocity: one face in five is inactive in the
//     caller's array and reads its background value 0.  Writes, in order, to <out-file>: int32 n;
//     fluid f32 (n^3), solid f32 ((n+1)^3), the staged u, v, w f32 (the projection's input), the
//     caller's activity u, v, w (u8), the velocity after the write-back u, v, w f32, its activity
//     u, v, w u8, the set of cells m_pressure holds (u8, n^3) and their values f32 (n^3).
#include "../../include/shkz_mi355x.h"
#include "shkz_stage.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// The caller's sparse activity: one face in five inactive (a hash of array and index).
static bool face_active(int dim, size_t f) {
    uint64_t z = (uint64_t)f * 0x9E3779B97F4A7C15ull + (uint64_t)(dim + 1) * 0xBF58476D1CE4E5B9ull;
    z ^= z >> 31;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 29;
    return z % 5 != 0;
}

template <class T>
static bool put(FILE *f, const std::vector<T> &a) {
    return std::fwrite(a.data(), sizeof(T), a.size(), f) == a.size();
}

static int writeback(int n, uint64_t seed, const char *path) {
    shkz_ctx *ctx = nullptr;
    if (shkz_create(0, 0, &ctx) != SHKZ_OK) {
        std::fprintf(stderr, "shkz_create: %s\n", shkz_last_error());
        return 1;
    }
    const size_t M = (size_t)n * n * n, F = (size_t)(n + 1) * n * n, Nn = (size_t)(n + 1) * (n + 1) * (n + 1);
    std::vector<float> fluid(M), solid(Nn), vel0[3], stage[3], p(M);
    std::vector<uint8_t> act_in[3], act_out[3], unknown(M);
    for (int a = 0; a < 3; ++a) vel0[a].resize(F);
    shkz_scene_desc sd = {};
    sd.kind = SHKZ_SCENE_DROP_SOLIDS;
    sd.w = sd.h = sd.d = n;
    sd.seed = seed;
    sd.dt = 1.0 / 120.0;
    sd.gravity = -9.8;
    sd.fluid_dtype = SHKZ_F32;
    int32_t has_solid = 0;
    if (shkz_scene_fill(&sd, fluid.data(), solid.data(), &has_solid, vel0[0].data(), vel0[1].data(), vel0[2].data()) !=
        SHKZ_OK)
        return 1;
    // the caller's sparse velocity and what the module stages from it (operator(): 0 off the actives)
    for (int a = 0; a < 3; ++a) {
        act_in[a].resize(F);
        act_out[a].resize(F);
        for (size_t f = 0; f < F; ++f) {
            act_in[a][f] = face_active(a, f) ? 1 : 0;
            if (!act_in[a][f]) vel0[a][f] = 0.0f;
        }
        const int fw = n + (a == 0), fh = n + (a == 1), fd = n + (a == 2);
        const std::vector<float> &src = vel0[a];
        shkz_host::densify(stage[a], fw, fh, fd,
                           [&](int i, int j, int k) { return src[((size_t)i * fh + j) * fd + k]; });
    }
    shkz_grid g = {};
    g.w = g.h = g.d = n;
    g.dt = sd.dt;
    g.fluid = fluid.data();
    g.fluid_dtype = SHKZ_F32;
    g.solid = has_solid ? solid.data() : nullptr;
    g.solid_dtype = SHKZ_F32;
    std::vector<float> work[3] = {stage[0], stage[1], stage[2]};
    g.u = work[0].data();
    g.v = work[1].data();
    g.w_vel = work[2].data();
    g.pressure = p.data();
    g.active_out_u = act_out[0].data();
    g.active_out_v = act_out[1].data();
    g.active_out_w = act_out[2].data();
    g.unknown_out = unknown.data();
    g.scene_id = 7;
    shkz_params prm;
    shkz_default_params(&prm);
    prm.residual = 1e-8;
    shkz_result r;
    const int rc = shkz_project(ctx, &g, &prm, &r);
    if (rc != SHKZ_OK || !r.converged) {
        std::fprintf(stderr, "shkz_project failed (%d): %s\n", rc, shkz_last_error());
        shkz_destroy(ctx);
        return 1;
    }
    // the module's write-back onto the caller's arrays
    std::vector<float> vel_after[3];
    std::vector<uint8_t> act_after[3];
    for (int a = 0; a < 3; ++a) {
        vel_after[a] = vel0[a];
        act_after[a] = act_in[a];
        for (size_t f = 0; f < F; ++f) {
            switch (shkz_host::face_action(act_in[a][f] != 0, act_out[a][f])) {
            case shkz_host::FACE_SET: vel_after[a][f] = work[a][f]; break;
            case shkz_host::FACE_OFF:
                vel_after[a][f] = 0.0f;  // an inactive face reads its background value
                act_after[a][f] = 0;
                break;
            default: break;
            }
        }
    }
    std::vector<float> pset(M, 0.0f);
    for (size_t c = 0; c < M; ++c)
        if (unknown[c]) pset[c] = p[c];
    FILE *out = std::fopen(path, "wb");
    if (!out) return 1;
    bool ok = std::fwrite(&n, sizeof(int32_t), 1, out) == 1 && put(out, fluid) && put(out, solid);
    for (int a = 0; a < 3; ++a) ok = ok && put(out, stage[a]);
    for (int a = 0; a < 3; ++a) ok = ok && put(out, act_in[a]);
    for (int a = 0; a < 3; ++a) ok = ok && put(out, vel_after[a]);
    for (int a = 0; a < 3; ++a) ok = ok && put(out, act_after[a]);
    ok = ok && put(out, unknown) && put(out, pset);
    std::fclose(out);
    std::printf("writeback: n=%d unknowns %lld iterations %d reresid_inf %.3e\n", n, (long long)r.n_unknowns,
                r.iterations, r.reresid_inf);
    shkz_destroy(ctx);
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "writeback") {
        if (argc < 5) {
            std::fprintf(stderr, "usage: shkz_harness writeback <n> <seed> <out-file>\n");
            return 2;
        }
        return writeback(std::atoi(argv[2]), std::strtoull(argv[3], nullptr, 10), argv[4]);
    }
    const int n = argc > 1 ? std::atoi(argv[1]) : 64;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
    const std::string mode = argc > 3 ? argv[3] : "mixed";
    shkz_ctx *ctx = nullptr;
    if (shkz_create(0, 0, &ctx) != SHKZ_OK) {
        std::fprintf(stderr, "shkz_create: %s\n", shkz_last_error());
        return 1;
    }
    const size_t M = (size_t)n * n * n, F = (size_t)(n + 1) * n * n, Nn = (size_t)(n + 1) * (n + 1) * (n + 1);
    std::vector<double> fluid(M);
    std::vector<float> solid(Nn), u(F), v(F), w(F), p(M);
    shkz_params prm;
    shkz_default_params(&prm);
    prm.residual = 1e-6;
    prm.warm_start = 1;
    if (mode == "fp64") prm.precision = SHKZ_PREC_FP64;
    if (mode == "cg") {
        prm.solver = SHKZ_SOLVER_CG;
        prm.precision = SHKZ_PREC_FP64;
    }
    int rc = 0;
    for (int s = 0; s < steps && rc == 0; ++s) {
        shkz_scene_desc sd = {};
        sd.kind = SHKZ_SCENE_DROP_SOLIDS;
        sd.w = sd.h = sd.d = n;
        sd.seed = 1;
        sd.dt = 1.0 / 120.0;
        sd.gravity = -9.8;
        sd.shift[1] = -0.01 * s;  // the drop falls
        sd.fluid_dtype = SHKZ_F64;
        sd.velocity_seed = 1000 + s;
        int32_t has_solid = 0;
        if (shkz_scene_fill(&sd, fluid.data(), solid.data(), &has_solid, u.data(), v.data(), w.data()) != SHKZ_OK) return 1;
        shkz_grid g = {};
        g.w = g.h = g.d = n;
        g.dt = sd.dt;
        g.fluid = fluid.data();
        g.fluid_dtype = SHKZ_F64;
        g.solid = has_solid ? solid.data() : nullptr;
        g.solid_dtype = SHKZ_F32;
        g.u = u.data();
        g.v = v.data();
        g.w_vel = w.data();
        g.pressure = p.data();
        g.scene_id = 1;
        shkz_result r;
        const auto t0 = std::chrono::steady_clock::now();
        rc = shkz_project(ctx, &g, &prm, &r);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != SHKZ_OK) {
            std::fprintf(stderr, "step %d: shkz_project failed (%d): %s\n", s, rc, shkz_last_error());
            break;
        }
        std::printf("step %d: unknowns %lld bricks %lld levels %d iterations %d (outer %d) reresid_inf %.3e "
                    "converged %d  %.2f ms (host arrays)\n",
                    s, (long long)r.n_unknowns, (long long)r.n_bricks, r.n_levels, r.iterations, r.outer_iterations,
                    r.reresid_inf, r.converged, ms);
    }
    shkz_destroy(ctx);
    return rc == 0 ? 0 : 1;
}
