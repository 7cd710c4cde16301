// macpressuresolver3_mi355x.cpp -- Shiokaze module: the MI355X pressure projection as a drop-in
// for macpressuresolver3 (src/projection/macpressuresolver3.cpp).  Select it with
//     Projection=macpressuresolver3_mi355x
// It implements macproject3_interface (include/shiokaze/projection/macproject3_interface.h:38-102)
// on top of the C ABI in include/shkz_mi355x.h.  It is built inside a Shiokaze tree, next to the
// other projection modules, linked against libshkz_mi355x.so; it reads the same configuration keys as macpressuresolver3 (:274-280), the pcg solver
// (src/linsolver/pcg.cpp:39-44) and macutility3 (src/utility/macutility3.cpp:408-412).
#include <shiokaze/projection/macproject3_interface.h>
#include <shiokaze/array/array3.h>
#include <shiokaze/array/macarray3.h>
#include <shiokaze/array/array_utility3.h>
#include <shiokaze/core/console.h>
#include <shiokaze/core/timer.h>

#include "../../include/shkz_mi355x.h"
#include "shkz_stage.hpp"

#include <cstdint>
#include <string>
#include <vector>

SHKZ_USING_NAMESPACE

class macpressuresolver3_mi355x : public macproject3_interface {
protected:
    LONG_NAME("MAC Pressure Solver 3D (MI355X)")

    virtual void set_target_volume(double current_volume, double target_volume) override {
        m_current_volume = current_volume;
        m_target_volume = target_volume;
    }

    virtual void project(double dt, macarray3<Real> &velocity, const array3<Real> &solid, const array3<Real> &fluid,
                         double surface_tension, const std::vector<signed_rigidbody3_interface *> *rigidbodies) override {
        scoped_timer timer(this);
        timer.tick();
        console::dump(">>> Pressure Projection (MI355X) started...\n");
        const int w = m_shape.w, h = m_shape.h, d = m_shape.d;
        const size_t M = (size_t)w * h * d;
        // ---- densify on every host thread: operator() yields the background / fill value where
        // the sparse grid is inactive (array3.h:796-800), which is what the reference's routines
        // read too.  Every face is staged (inactive ones read their background value, as the
        // reference's assembly reads them, macpressuresolver3.cpp:187-194); the write-back below
        // visits only the faces the caller holds active.
        shkz_host::densify(m_phi, w, h, d, [&](int i, int j, int k) { return fluid(i, j, k); });
        const bool has_solid = array_utility3::levelset_exist(solid) && solid.shape() == m_shape.nodal();
        if (has_solid)
            shkz_host::densify(m_solid, w + 1, h + 1, d + 1, [&](int i, int j, int k) { return solid(i, j, k); });
        for (int dim = 0; dim < 3; ++dim) {
            const shape3 fs = m_shape.face(dim);
            const array3<Real> &u = velocity[dim];
            shkz_host::densify(m_vel[dim], fs.w, fs.h, fs.d, [&](int i, int j, int k) { return u(i, j, k); });
            m_act_out[dim].resize(fs.count());
        }
        m_p.resize(M);
        m_unknown.resize(M);
        // ---- project on the device
        shkz_grid g = {};
        g.w = w;
        g.h = h;
        g.d = d;
        g.dx = m_dx;
        g.dt = dt;
        g.fluid = m_phi.data();
        g.fluid_dtype = SHKZ_F32;
        g.solid = has_solid ? m_solid.data() : nullptr;
        g.solid_dtype = SHKZ_F32;
        g.u = m_vel[0].data();
        g.v = m_vel[1].data();
        g.w_vel = m_vel[2].data();
        g.pressure = m_p.data();
        g.scene_id = scene_id();
        g.active_out_u = m_act_out[0].data();
        g.active_out_v = m_act_out[1].data();
        g.active_out_w = m_act_out[2].data();
        g.unknown_out = m_unknown.data();
        shkz_params p = m_params;
        p.surface_tension = surface_tension;
        p.current_volume = m_current_volume;
        p.target_volume = m_target_volume;
        shkz_result r = {};
        const int rc = shkz_project(m_ctx, &g, &p, &r);
        if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) {
            console::dump("shkz_project failed (%d): %s\n", rc, shkz_last_error());
            throw std::runtime_error(shkz_last_error());
        }
        console::write(get_argument_name() + "_number_projection_iteration", r.iterations);
        console::dump("Done. Took %d iterations, Reresid=%e.\n", r.iterations, r.reresid);
        // ---- write back: m_pressure on every unknown, zeros included (:245-248), serially as the
        // reference's const_serial_actives; velocities on the faces the caller holds active (:252-266)
        m_pressure.clear();
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < h; ++j)
                for (int k = 0; k < d; ++k) {
                    const size_t c = ((size_t)i * h + j) * d + k;
                    if (m_unknown[c]) m_pressure.set(i, j, k, m_p[c]);
                }
        for (int dim = 0; dim < 3; ++dim) {
            const shape3 fs = m_shape.face(dim);
            const std::vector<float> &vel = m_vel[dim];
            const std::vector<uint8_t> &act = m_act_out[dim];
            velocity[dim].parallel_actives([&](int i, int j, int k, auto &it, int tn) {
                const size_t f = ((size_t)i * fs.h + j) * fs.d + k;
                if (shkz_host::face_action(true, act[f]) == shkz_host::FACE_SET)
                    it.set(vel[f]);
                else
                    it.set_off();  // macpressuresolver3.cpp:266
            });
        }
        console::dump("<<< Projection done. Took %s.\n", timer.stock("projection").c_str());
    }

    // The library keys this module's resident state (warm-start pressure, volume integrator) by
    // the module's address.
    uint64_t scene_id() const { return (uint64_t) reinterpret_cast<uintptr_t>(this); }

    virtual void configure(configuration &config) override {
        shkz_default_params(&m_params);
        bool so_fluid = true, so_solid = true, warm = false;
        double gain = m_params.gain, residual = m_params.residual, eps_fluid = m_params.eps_fluid,
               eps_solid = m_params.eps_solid;
        unsigned maxit = (unsigned)m_params.max_iterations;
        std::string precision = "mixed", solver = "mgpcg";
        config.get_bool("SecondOrderAccurateFluid", so_fluid, "Whether to enforce second order accuracy");
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
    virtual void initialize(const shape3 &shape, double dx) override {
        m_shape = shape;
        m_dx = dx;
    }
    virtual void post_initialize() override {
        m_pressure.initialize(m_shape);
        m_target_volume = m_current_volume = 0.0;
        if (!m_ctx && shkz_create(m_device, 0, &m_ctx) != SHKZ_OK) throw std::runtime_error(shkz_last_error());
        // m_y_prev = 0 (macpressuresolver3.cpp:290): dropping the scene's resident state resets the
        // library's volume integrator (its warm-start pressure goes with it; SHKZ_ERR_NOT_FOUND
        // before the first projection is expected)
        (void)shkz_release_scene(m_ctx, scene_id());
    }
    virtual const array3<Real> *get_pressure() const override { return &m_pressure; }

public:
    virtual ~macpressuresolver3_mi355x() {
        if (m_ctx) shkz_destroy(m_ctx);
    }

protected:
    shape3 m_shape;
    double m_dx{0.0};
    array3<Real> m_pressure{this};
    shkz_params m_params{};
    shkz_ctx *m_ctx{nullptr};
    int m_device{0};
    double m_target_volume{0.0}, m_current_volume{0.0};
    std::vector<float> m_phi, m_solid, m_vel[3], m_p;
    std::vector<uint8_t> m_act_out[3], m_unknown;
};

extern "C" module *create_instance() { return new macpressuresolver3_mi355x(); }

extern "C" const char *license() { return "MIT"; }
