// macpressuresolver2_mi355x.cpp -- Shiokaze module: the MI355X pressure projection as a drop-in
// for macpressuresolver2 (src/projection/macpressuresolver2.cpp).  Select it with
//     Projection=macpressuresolver2_mi355x
// It implements macproject2_interface (include/shiokaze/projection/macproject2_interface.h:38-99)
// on top of the C ABI in include/shkz_mi355x.h.  It reads the configuration keys of
// macpressuresolver2 (:253-260), the pcg solver (src/linsolver/pcg.cpp:39-44) and macutility2.
#include <shiokaze/projection/macproject2_interface.h>
#include <shiokaze/array/array2.h>
#include <shiokaze/array/macarray2.h>
#include <shiokaze/array/array_utility2.h>
#include <shiokaze/utility/utility.h>
#include <shiokaze/core/console.h>
#include <shiokaze/core/timer.h>

#include "../../include/shkz_mi355x.h"
#include "shkz_stage.hpp"

#include <cstdint>
#include <string>
#include <vector>

SHKZ_USING_NAMESPACE

class macpressuresolver2_mi355x : public macproject2_interface {
protected:
    LONG_NAME("MAC Pressure Solver 2D (MI355X)")

    virtual void set_target_volume(double current_volume, double target_volume) override {
        m_current_volume = current_volume;
        m_target_volume = target_volume;
    }

    virtual void project(double dt, macarray2<Real> &velocity, const array2<Real> &solid, const array2<Real> &fluid,
                         double surface_tension, const std::vector<signed_rigidbody2_interface *> *rigidbodies) override {
        const int w = m_shape.w, h = m_shape.h;
        const size_t M = (size_t)w * h;
        // ---- densify (operator() yields the background / fill values where the sparse grid is
        // inactive, array2.h), as the reference's routines read them
        shkz_host::densify(m_phi, w, h, 1, [&](int i, int j, int) { return fluid(i, j); });
        const bool has_solid = array_utility2::levelset_exist(solid) && solid.shape() == m_shape.nodal();
        if (has_solid) shkz_host::densify(m_solid, w + 1, h + 1, 1, [&](int i, int j, int) { return solid(i, j); });
        for (int dim = 0; dim < 2; ++dim) {
            const shape2 fs = m_shape.face(dim);
            const array2<Real> &u = velocity[dim];
            shkz_host::densify(m_vel0[dim], fs.w, fs.h, 1, [&](int i, int j, int) { return u(i, j); });
            shkz_host::densify(m_act[dim], fs.w, fs.h, 1, [&](int i, int j, int) { return u.active(i, j) ? 1 : 0; });
            m_vel[dim] = m_vel0[dim];
        }
        m_p.resize(M);
        m_unknown.resize(M);
        shkz_grid2d g = {};
        g.w = w;
        g.h = h;
        g.dx = m_dx;
        g.dt = dt;
        g.fluid = m_phi.data();
        g.fluid_dtype = SHKZ_F32;
        g.solid = has_solid ? m_solid.data() : nullptr;
        g.solid_dtype = SHKZ_F32;
        g.u = m_vel[0].data();
        g.v = m_vel[1].data();
        g.active_u = m_act[0].data();
        g.active_v = m_act[1].data();
        g.pressure = m_p.data();
        g.unknown_out = m_unknown.data();
        g.scene_id = scene_id();
        shkz_params p = m_params;
        p.surface_tension = surface_tension;
        p.current_volume = m_current_volume;
        p.target_volume = m_target_volume;
        shkz_result r = {};
        const int rc = shkz_project2d(m_ctx, &g, &p, &r);
        if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) {
            console::dump("shkz_project2d failed (%d): %s\n", rc, shkz_last_error());
            throw std::runtime_error(shkz_last_error());
        }
        // ---- write back: m_pressure on every unknown, zeros included (macpressuresolver2.cpp sets
        // each of them), velocities with the 2D module's activity rule
        m_pressure.clear();
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < h; ++j) {
                const size_t c = (size_t)i * h + j;
                if (m_unknown[c]) m_pressure.set(i, j, m_p[c]);
            }
        const bool has_fluid = array_utility2::levelset_exist(fluid);
        for (int dim = 0; dim < 2; ++dim) {
            const shape2 fs = m_shape.face(dim);
            velocity[dim].parallel_actives([&](int i, int j, auto &it, int tn) {
                const size_t f = (size_t)i * fs.h + j;
                if (face_open(dim, i, j, has_solid, has_fluid)) {
                    it.set(m_vel[dim][f]);
                } else if (boundary_outflow(dim, i, j, m_vel0[dim][f])) {
                    it.set(0.0);  // macpressuresolver2.cpp:224-225
                } else {
                    it.set_off();  // :226
                }
            });
        }
    }

    // area > 0 and rho > 0 as macutility2 computes them on these level sets
    bool face_open(int dim, int i, int j, bool has_solid, bool has_fluid) const {
        const int pi[2] = {i, j};
        const int N[2] = {(int)m_shape.w, (int)m_shape.h};
        const bool boundary = pi[dim] == 0 || pi[dim] == N[dim];
        if (boundary) return false;
        if (has_solid) {
            const double s0 = m_solid[(size_t)i * (m_shape.h + 1) + j];
            const double s1 = m_solid[(size_t)(i + (dim != 0)) * (m_shape.h + 1) + (j + (dim != 1))];
            if (1.0 - utility::fraction(s0, s1) == 0.0) return false;
        }
        if (!has_fluid) return true;
        const vec2i a = m_shape.clamp(i, j), b = m_shape.clamp(i - (dim == 0), j - (dim == 1));
        return m_phi[(size_t)a[0] * m_shape.h + a[1]] < 0.0 || m_phi[(size_t)b[0] * m_shape.h + b[1]] < 0.0;
    }
    // a closed boundary face with liquid inside whose velocity pointed out of the domain
    bool boundary_outflow(int dim, int i, int j, float u0) const {
        const int pi[2] = {i, j};
        const int N[2] = {(int)m_shape.w, (int)m_shape.h};
        if (pi[dim] == 0) return m_phi[(size_t)i * m_shape.h + j] < 0.0 && u0 < 0.0f;
        if (pi[dim] == N[dim]) {
            const int a = i - (dim == 0), b = j - (dim == 1);
            return m_phi[(size_t)a * m_shape.h + b] < 0.0 && u0 > 0.0f;
        }
        return false;
    }

    virtual void configure(configuration &config) override {
        shkz_default_params(&m_params);
        bool so_fluid = true, so_solid = true, warm = false;
        double gain = m_params.gain, residual = m_params.residual, eps_fluid = m_params.eps_fluid,
               eps_solid = m_params.eps_solid;
        unsigned maxit = (unsigned)m_params.max_iterations;
        std::string precision = "mixed", solver = "mgpcg";
        config.get_bool("SecondOrderAccurateFluid", so_fluid, "Whether to enforce second order accuracy for free surfaces");
        config.get_bool("SecondOrderAccurateSolid", so_solid, "Whether to enforce second order accuracy for solid surfaces");
        config.get_double("Gain", gain, "Rate for volume correction");
        config.get_bool("WarmStart", warm, "Start from the solution of previous pressure");
        config.get_double("Residual", residual, "Tolerable residual");
        config.get_unsigned("MaxIterations", maxit, "Maximal iteration count");
        config.get_double("EpsFluid", eps_fluid, "Minimal bound for fluid fraction");
        config.get_double("EpsSolid", eps_solid, "Minimal bound for solid fraction");
        config.get_string("Precision", precision, "mixed (FP32 multigrid in FP64 defect correction) or fp64");
        config.get_string("Solver", solver, "mgpcg or cg (the reference pcg iteration)");
        config.get_integer("Device", m_device, "HIP device");
        m_params.second_order_fluid = so_fluid;
        m_params.second_order_solid = so_solid;
        m_params.gain = gain;
        m_params.warm_start = warm;
        m_params.residual = residual;
        m_params.max_iterations = (int32_t)maxit;
        m_params.eps_fluid = eps_fluid;
        m_params.eps_solid = eps_solid;
        m_params.precision = precision == "fp64" ? SHKZ_PREC_FP64 : SHKZ_PREC_MIXED;
        m_params.solver = solver == "cg" ? SHKZ_SOLVER_CG : SHKZ_SOLVER_MGPCG;
    }
    virtual void initialize(const shape2 &shape, double dx) override {
        m_shape = shape;
        m_dx = dx;
    }
    virtual void post_initialize() override {
        m_pressure.initialize(m_shape);
        m_target_volume = m_current_volume = 0.0;
        if (!m_ctx && shkz_create(m_device, 0, &m_ctx) != SHKZ_OK) throw std::runtime_error(shkz_last_error());
        // m_y_prev = 0: dropping the scene's resident state resets the library's volume integrator
        // (SHKZ_ERR_NOT_FOUND before the first projection is expected)
        (void)shkz_release_scene(m_ctx, scene_id());
    }
    uint64_t scene_id() const { return (uint64_t) reinterpret_cast<uintptr_t>(this); }
    virtual const array2<Real> *get_pressure() const override { return &m_pressure; }

public:
    virtual ~macpressuresolver2_mi355x() {
        if (m_ctx) shkz_destroy(m_ctx);
    }

protected:
    shape2 m_shape;
    double m_dx{0.0};
    array2<Real> m_pressure{this};
    shkz_params m_params{};
    shkz_ctx *m_ctx{nullptr};
    int m_device{0};
    double m_target_volume{0.0}, m_current_volume{0.0};
    std::vector<float> m_phi, m_solid, m_vel[2], m_vel0[2], m_p;
    std::vector<uint8_t> m_act[2], m_unknown;
};

extern "C" module *create_instance() { return new macpressuresolver2_mi355x(); }

extern "C" const char *license() { return "MIT"; }
