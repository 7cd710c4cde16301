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
#include <shiokaze/utility/utility.h>
#include <shiokaze/core/console.h>
#include <shiokaze/core/timer.h>

#include "../../include/shkz_mi355x.h"

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
        // ---- densify: operator() yields background / fill values where the sparse grid is
        // inactive (array3.h:796-800), which is what the reference's routines read too.
        m_phi.resize(M);
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < h; ++j)
                for (int k = 0; k < d; ++k) m_phi[((size_t)i * h + j) * d + k] = fluid(i, j, k);
        const bool has_solid = array_utility3::levelset_exist(solid) && solid.shape() == m_shape.nodal();
        if (has_solid) {
            m_solid.resize((size_t)(w + 1) * (h + 1) * (d + 1));
            for (int i = 0; i <= w; ++i)
                for (int j = 0; j <= h; ++j)
                    for (int k = 0; k <= d; ++k) m_solid[((size_t)i * (h + 1) + j) * (d + 1) + k] = solid(i, j, k);
        }
        for (int dim = 0; dim < 3; ++dim) {
            const shape3 fs = m_shape.face(dim);
            m_vel[dim].resize(fs.count());
            m_act[dim].resize(fs.count());
            for (int i = 0; i < (int)fs.w; ++i)
                for (int j = 0; j < (int)fs.h; ++j)
                    for (int k = 0; k < (int)fs.d; ++k) {
                        const size_t f = ((size_t)i * fs.h + j) * fs.d + k;
                        m_vel[dim][f] = velocity[dim](i, j, k);
                        m_act[dim][f] = velocity[dim].active(i, j, k) ? 1 : 0;
                    }
        }
        m_p.assign(M, 0.0f);
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
        g.active_u = m_act[0].data();
        g.active_v = m_act[1].data();
        g.active_w = m_act[2].data();
        g.pressure = m_p.data();
        g.scene_id = reinterpret_cast<uintptr_t>(this);
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
        // ---- write back: m_pressure on the unknowns, velocities on the faces the reference visits
        m_pressure.clear();
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < h; ++j)
                for (int k = 0; k < d; ++k) {
                    const float v = m_p[((size_t)i * h + j) * d + k];
                    if (v) m_pressure.set(i, j, k, v);
                }
        const bool has_fluid = array_utility3::levelset_exist(fluid);
        for (int dim = 0; dim < 3; ++dim) {
            const shape3 fs = m_shape.face(dim);
            velocity[dim].parallel_actives([&](int i, int j, int k, auto &it, int tn) {
                const size_t f = ((size_t)i * fs.h + j) * fs.d + k;
                if (face_open(dim, i, j, k, has_solid, has_fluid, solid) || boundary_liquid(dim, i, j, k))
                    it.set(m_vel[dim][f]);
                else
                    it.set_off();  // macpressuresolver3.cpp:266
            });
        }
        console::dump("<<< Projection done. Took %s.\n", timer.stock("projection").c_str());
    }

    // area > 0 and rho > 0 (macpressuresolver3.cpp:255), evaluated as the reference's
    // macutility3 routines would on these level sets.
    bool face_open(int dim, int i, int j, int k, bool has_solid, bool has_fluid, const array3<Real> &solid) const {
        const int pi[3] = {i, j, k};
        const unsigned N[3] = {m_shape.w, m_shape.h, m_shape.d};
        const bool boundary = pi[dim] == 0 || pi[dim] == (int)N[dim];
        bool area = !boundary;
        if (has_solid && !boundary) {
            double q[2][2];
            if (dim == 0) {
                q[0][0] = solid(i, j, k); q[1][0] = solid(i, j + 1, k); q[1][1] = solid(i, j + 1, k + 1); q[0][1] = solid(i, j, k + 1);
            } else if (dim == 1) {
                q[0][0] = solid(i, j, k); q[1][0] = solid(i + 1, j, k); q[1][1] = solid(i + 1, j, k + 1); q[0][1] = solid(i, j, k + 1);
            } else {
                q[0][0] = solid(i, j, k); q[1][0] = solid(i + 1, j, k); q[1][1] = solid(i + 1, j + 1, k); q[0][1] = solid(i, j + 1, k);
            }
            area = (1.0 - utility::get_area(q)) != 0.0;
        } else if (has_solid && boundary) {
            area = false;
        }
        if (!area) return false;
        if (!has_fluid) return true;
        const vec3i a = m_shape.clamp(i, j, k), b = m_shape.clamp(i - (dim == 0), j - (dim == 1), k - (dim == 2));
        return m_phi[((size_t)a[0] * m_shape.h + a[1]) * m_shape.d + a[2]] < 0.0 ||
               m_phi[((size_t)b[0] * m_shape.h + b[1]) * m_shape.d + b[2]] < 0.0;
    }
    // a boundary face with liquid on its inner side stays active at 0 (macpressuresolver3.cpp:264-265)
    bool boundary_liquid(int dim, int i, int j, int k) const {
        const int pi[3] = {i, j, k};
        const unsigned N[3] = {m_shape.w, m_shape.h, m_shape.d};
        if (pi[dim] == 0) return m_phi[((size_t)i * m_shape.h + j) * m_shape.d + k] < 0.0;
        if (pi[dim] == (int)N[dim]) {
            const int a = i - (dim == 0), b = j - (dim == 1), c = k - (dim == 2);
            return m_phi[((size_t)a * m_shape.h + b) * m_shape.d + c] < 0.0;
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
    std::vector<uint8_t> m_act[3];
};

extern "C" module *create_instance() { return new macpressuresolver3_mi355x(); }

extern "C" const char *license() { return "MIT"; }
