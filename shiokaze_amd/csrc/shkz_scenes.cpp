// shkz_scenes.cpp -- seeded synthetic projection workloads (host only).
//
// Sampling conventions follow macutility3::assign_initial_variables
// (src/utility/macutility3.cpp:336-374): the liquid level set is sampled at cell centres,
// the solid level set at nodes, and where a solid exists the liquid is pushed out of it as
// max(phi_fluid, -(phi_solid + dx)) (:358-364).  Shapes are unions (min) of signed distances;
// the box uses the reference's utility::box (include/shiokaze/utility/utility.h:113-120).
#include "../../include/shkz_mi355x.h"
#include "shkz_scene_params.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

// std::mt19937_64 is fully specified by the C++ standard; the normal sampler is written here
// (Box-Muller) because std::normal_distribution differs between standard libraries.
struct SceneRng {
    std::mt19937_64 g;
    bool has_spare = false;
    double spare = 0.0;
    explicit SceneRng(uint64_t seed) : g(seed) {}
    double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u1 = 1.0 - uniform();  // (0,1]
        double u2 = uniform();
        double rad = std::sqrt(-2.0 * std::log(u1));
        double ang = 6.283185307179586476925286766559 * u2;
        spare = rad * std::sin(ang);
        has_spare = true;
        return rad * std::cos(ang);
    }
};

struct Sphere {
    double c[3];
    double r;
};

struct Scene {
    std::vector<Sphere> liquid_spheres;
    double pool = -1.0;  // pool level (liquid below y < pool), <0 = none
    std::vector<Sphere> solid_spheres;
    bool solid_box = false;
    double box0[3] = {0, 0, 0}, box1[3] = {0, 0, 0};
    double shift[3] = {0, 0, 0};
    bool all_fluid = false;  // every cell liquid (phi = -1), the liquid not pushed out of solids

    bool has_solid() const { return !solid_spheres.empty() || solid_box; }

    double fluid(const double p[3]) const {
        if (all_fluid) return -1.0;
        const double q[3] = {p[0] - shift[0], p[1] - shift[1], p[2] - shift[2]};
        double v = 1e9;
        if (pool >= 0.0) v = std::min(v, q[1] - pool);
        for (const Sphere &s : liquid_spheres) {
            double dx = q[0] - s.c[0], dy = q[1] - s.c[1], dz = q[2] - s.c[2];
            v = std::min(v, std::sqrt(dx * dx + dy * dy + dz * dz) - s.r);
        }
        return v;
    }
    double solid(const double p[3]) const {
        double v = 1e9;
        for (const Sphere &s : solid_spheres) {
            double dx = p[0] - s.c[0], dy = p[1] - s.c[1], dz = p[2] - s.c[2];
            v = std::min(v, std::sqrt(dx * dx + dy * dy + dz * dz) - s.r);
        }
        if (solid_box) {
            double sd = -9999.0;
            for (int dim = 0; dim < 3; ++dim) {
                sd = std::max(sd, box0[dim] - p[dim]);
                sd = std::max(sd, p[dim] - box1[dim]);
            }
            v = std::min(v, sd);
        }
        return v;
    }
};

// Draws the scene's shapes from rng (the velocity stream continues from where this leaves it).
static int build_scene(const shkz_scene_desc *desc, SceneRng &rng, Scene &sc) {
    for (int a = 0; a < 3; ++a) sc.shift[a] = desc->shift[a];
    switch (desc->kind) {
    case SHKZ_SCENE_DROP:
        sc.liquid_spheres.push_back({{0.5, 0.7, 0.5}, 0.15});
        sc.pool = 0.3;
        break;
    case SHKZ_SCENE_DROP_SOLIDS:
        sc.liquid_spheres.push_back({{0.5, 0.7, 0.5}, 0.15});
        sc.pool = 0.3;
        for (int s = 0; s < 3; ++s) {
            Sphere sp;
            for (int a = 0; a < 3; ++a) sp.c[a] = rng.uniform(0.2, 0.8);
            sp.r = 0.1;
            sc.solid_spheres.push_back(sp);
        }
        break;
    case SHKZ_SCENE_SPHERES_BOX: {
        sc.pool = 0.25;
        for (int s = 0; s < 8; ++s) {
            Sphere sp;
            sp.r = rng.uniform(0.05, 0.15);
            sp.c[0] = rng.uniform(0.15, 0.85);
            sp.c[1] = rng.uniform(0.3, 0.85);
            sp.c[2] = rng.uniform(0.15, 0.85);
            sc.liquid_spheres.push_back(sp);
        }
        // one axis-aligned box holding 10% of the unit cube
        double sx = rng.uniform(0.35, 0.65), sy = rng.uniform(0.35, 0.65);
        double sz = std::min(0.9, 0.1 / (sx * sy));
        double s3[3] = {sx, sy, sz};
        for (int a = 0; a < 3; ++a) {
            sc.box0[a] = rng.uniform(0.05, std::max(0.05, 0.95 - s3[a]));
            sc.box1[a] = sc.box0[a] + s3[a];
        }
        sc.solid_box = true;
        break;
    }
    case SHKZ_SCENE_RANDOM_DROP: {
        sc.pool = rng.uniform(0.2, 0.35);
        Sphere sp;
        sp.r = rng.uniform(0.08, 0.18);
        sp.c[0] = rng.uniform(0.3, 0.7);
        sp.c[2] = rng.uniform(0.3, 0.7);
        double lo = sc.pool + sp.r + 0.05, hi = 0.95 - sp.r;
        sp.c[1] = rng.uniform(lo, std::max(lo, hi));
        sc.liquid_spheres.push_back(sp);
        break;
    }
    case SHKZ_SCENE_SMOKE:
        sc.all_fluid = true;
        for (int s = 0; s < 2; ++s) {
            Sphere sp;
            for (int a = 0; a < 3; ++a) sp.c[a] = rng.uniform(0.25, 0.75);
            sp.r = rng.uniform(0.08, 0.15);
            sc.solid_spheres.push_back(sp);
        }
        break;
    default:
        return SHKZ_ERR_INVALID;
    }
    return SHKZ_OK;
}

}  // namespace

extern "C" int shkz_scene_params_internal(const shkz_scene_desc *desc, ShkzSceneParams *out) {
    if (!desc || !out || desc->w <= 0 || desc->h <= 0 || desc->d <= 0) return SHKZ_ERR_INVALID;
    SceneRng rng(desc->seed);
    Scene sc;
    if (build_scene(desc, rng, sc) != SHKZ_OK) return SHKZ_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    if (sc.liquid_spheres.size() > 8 || sc.solid_spheres.size() > 4) return SHKZ_ERR_INVALID;
    out->n_liquid = (int32_t)sc.liquid_spheres.size();
    out->n_solid = (int32_t)sc.solid_spheres.size();
    out->has_box = sc.solid_box ? 1 : 0;
    out->all_fluid = sc.all_fluid ? 1 : 0;
    out->pool = sc.pool;
    for (int q = 0; q < out->n_liquid; ++q) {
        for (int a = 0; a < 3; ++a) out->liquid[q][a] = sc.liquid_spheres[q].c[a];
        out->liquid[q][3] = sc.liquid_spheres[q].r;
    }
    for (int q = 0; q < out->n_solid; ++q) {
        for (int a = 0; a < 3; ++a) out->solid[q][a] = sc.solid_spheres[q].c[a];
        out->solid[q][3] = sc.solid_spheres[q].r;
    }
    for (int a = 0; a < 3; ++a) {
        out->box0[a] = sc.box0[a];
        out->box1[a] = sc.box1[a];
        out->shift[a] = sc.shift[a];
    }
    return SHKZ_OK;
}

extern "C" int shkz_scene_fill(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid, float *u,
                               float *v, float *w) {
    if (!desc || desc->w <= 0 || desc->h <= 0 || desc->d <= 0) return SHKZ_ERR_INVALID;
    const int W = desc->w, H = desc->h, D = desc->d;
    const double dx = std::min(1.0 / W, std::min(1.0 / H, 1.0 / D));
    SceneRng rng(desc->seed);
    Scene sc;
    if (build_scene(desc, rng, sc) != SHKZ_OK) return SHKZ_ERR_INVALID;
    const bool with_solid = sc.has_solid();
    if (has_solid) *has_solid = with_solid ? 1 : 0;

    // liquid level set at cell centres
    if (fluid) {
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j)
                for (int k = 0; k < D; ++k) {
                    double p[3] = {(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx};
                    double val = sc.fluid(p);
                    if (with_solid && !sc.all_fluid) val = std::max(val, -(sc.solid(p) + dx));
                    size_t c = ((size_t)i * H + j) * D + k;
                    if (desc->fluid_dtype == SHKZ_F64) ((double *)fluid)[c] = val;
                    else ((float *)fluid)[c] = (float)val;
                }
    }
    // solid level set at nodes
    if (solid) {
        for (int i = 0; i <= W; ++i)
            for (int j = 0; j <= H; ++j)
                for (int k = 0; k <= D; ++k) {
                    double p[3] = {i * dx, j * dx, k * dx};
                    size_t n = ((size_t)i * (H + 1) + j) * (D + 1) + k;
                    solid[n] = with_solid ? (float)sc.solid(p) : 1.0f;
                }
    }
    // face velocities: i.i.d. N(0,1) over u, then v, then w (index order), gravity*dt on v
    SceneRng vrng(desc->velocity_seed ? desc->velocity_seed : 0);
    SceneRng &vr = desc->velocity_seed ? vrng : rng;
    float *vel[3] = {u, v, w};
    for (int dim = 0; dim < 3; ++dim) {
        size_t n = (size_t)(W + (dim == 0)) * (H + (dim == 1)) * (D + (dim == 2));
        float *a = vel[dim];
        if (!a) continue;
        if (desc->velocity_mode == 1) {
            std::memset(a, 0, n * sizeof(float));
        } else {
            for (size_t f = 0; f < n; ++f) a[f] = (float)vr.normal();
        }
        if (dim == 1 && desc->gravity != 0.0) {
            const float g = (float)(desc->gravity * desc->dt);
            for (size_t f = 0; f < n; ++f) a[f] += g;
        }
    }
    return SHKZ_OK;
}

// 2D scenes: the same seeded shapes cut at z = 0 (circles, pool, the box's x-y rectangle) on a
// (w, h) grid with dx = min(1/w, 1/h): level sets at cell centres, solid at nodes, N(0,1)
// velocities over u then v (index order), gravity*dt added to v.
extern "C" int shkz_scene_fill2d(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid, float *u,
                                 float *v) {
    if (!desc || desc->w <= 0 || desc->h <= 0) return SHKZ_ERR_INVALID;
    const int W = desc->w, H = desc->h;
    const double dx = std::min(1.0 / W, 1.0 / H);
    SceneRng rng(desc->seed);
    Scene sc;
    if (build_scene(desc, rng, sc) != SHKZ_OK) return SHKZ_ERR_INVALID;
    // the 3D shapes evaluated in the plane through their centres (z of each sphere, box at its middle)
    auto fluid2 = [&](double x, double y) {
        if (sc.all_fluid) return -1.0;
        const double q[2] = {x - sc.shift[0], y - sc.shift[1]};
        double val = 1e9;
        if (sc.pool >= 0.0) val = std::min(val, q[1] - sc.pool);
        for (const Sphere &s : sc.liquid_spheres) {
            const double ddx = q[0] - s.c[0], ddy = q[1] - s.c[1];
            val = std::min(val, std::sqrt(ddx * ddx + ddy * ddy) - s.r);
        }
        return val;
    };
    auto solid2 = [&](double x, double y) {
        double val = 1e9;
        for (const Sphere &s : sc.solid_spheres) {
            const double ddx = x - s.c[0], ddy = y - s.c[1];
            val = std::min(val, std::sqrt(ddx * ddx + ddy * ddy) - s.r);
        }
        if (sc.solid_box) {
            double sd = -9999.0;
            const double p[2] = {x, y};
            for (int dim = 0; dim < 2; ++dim) {
                sd = std::max(sd, sc.box0[dim] - p[dim]);
                sd = std::max(sd, p[dim] - sc.box1[dim]);
            }
            val = std::min(val, sd);
        }
        return val;
    };
    const bool with_solid = sc.has_solid();
    if (has_solid) *has_solid = with_solid ? 1 : 0;
    if (fluid)
        for (int i = 0; i < W; ++i)
            for (int j = 0; j < H; ++j) {
                const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
                double val = fluid2(x, y);
                if (with_solid && !sc.all_fluid) val = std::max(val, -(solid2(x, y) + dx));
                const size_t c = (size_t)i * H + j;
                if (desc->fluid_dtype == SHKZ_F64) ((double *)fluid)[c] = val;
                else ((float *)fluid)[c] = (float)val;
            }
    if (solid)
        for (int i = 0; i <= W; ++i)
            for (int j = 0; j <= H; ++j) solid[(size_t)i * (H + 1) + j] = with_solid ? (float)solid2(i * dx, j * dx) : 1.0f;
    SceneRng vrng(desc->velocity_seed ? desc->velocity_seed : 0);
    SceneRng &vr = desc->velocity_seed ? vrng : rng;
    float *vel[2] = {u, v};
    for (int dim = 0; dim < 2; ++dim) {
        const size_t n = (size_t)(W + (dim == 0)) * (H + (dim == 1));
        float *a = vel[dim];
        if (!a) continue;
        if (desc->velocity_mode == 1) std::memset(a, 0, n * sizeof(float));
        else
            for (size_t f = 0; f < n; ++f) a[f] = (float)vr.normal();
        if (dim == 1 && desc->gravity != 0.0) {
            const float g = (float)(desc->gravity * desc->dt);
            for (size_t f = 0; f < n; ++f) a[f] += g;
        }
    }
    return SHKZ_OK;
}
