// shkz_harness.cpp -- standalone C++ host harness of the C ABI.  Reproduces the call contract of
// macliquid3::idle (src/liquid/macliquid3.cpp:321-367): every step the caller owns the level sets
// and the velocity, adds gravity, calls project(dt, velocity, solid, fluid) and keeps the result;
// here the liquid is translated rigidly between steps instead of advected.
//   shkz_harness [n] [steps] [mixed|fp64|cg]
#include "../../include/shkz_mi355x.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char **argv) {
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
