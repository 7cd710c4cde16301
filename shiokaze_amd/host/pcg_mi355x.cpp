// pcg_mi355x.cpp -- Shiokaze module: the LinSolver drop-in.  Select it with
//     LinSolver=pcg_mi355x
// It implements RCMatrix_solver_interface<size_t,double> (include/shiokaze/linsolver/
// RCMatrix_solver.h:36-77) like the pcg module (src/linsolver/pcg.cpp:32-73) and reads the same
// configuration keys (:39-44; ModifiedIC and MinDiagRatio are accepted: the reference computes
// its MIC(0) preconditioner and discards it, pcg_solver.h:379-384, so its iteration is plain CG,
// which this module runs on the GPU).  Extra keys: Preconditioner=none|jacobi, Device=<n>.
#include <shiokaze/linsolver/RCMatrix_solver.h>
#include <shiokaze/core/console.h>

#include "../../include/shkz_mi355x.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

SHKZ_USING_NAMESPACE

template <class N, class T>
class pcg_mi355x : public RCMatrix_solver_interface<N, T> {
protected:
    LONG_NAME("Conjugate Gradient Solver (MI355X)")
    ARGUMENT_NAME("PCG")

    virtual void configure(configuration &config) override {
        shkz_linsolve_default_params(&m_params);
        double residual = m_params.residual, mic = 0.97, min_ratio = 0.25;
        unsigned maxit = (unsigned)m_params.max_iterations;
        std::string precond = "none";
        config.get_double("Residual", residual, "Tolerable residual");
        config.get_double("ModifiedIC", mic, "Modified incomplete cholesky");
        config.get_double("MinDiagRatio", min_ratio, "Minimal diagonal ratio");
        config.get_unsigned("MaxIterations", maxit, "Maximal iteration count");
        config.get_string("Preconditioner", precond, "none (the reference iteration) or jacobi");
        config.get_integer("Device", m_device, "HIP device");
        m_params.residual = residual;
        m_params.max_iterations = (int32_t)maxit;
        m_params.preconditioner = precond == "jacobi" ? SHKZ_LS_JACOBI : SHKZ_LS_NONE;
    }

    virtual typename RCMatrix_solver_interface<N, T>::Result solve(const RCMatrix_interface<N, T> *A,
                                                                    const RCMatrix_vector_interface<N, T> *b,
                                                                    RCMatrix_vector_interface<N, T> *x) const override {
        if (!m_solver && shkz_linsolver_create(m_device, &m_solver) != SHKZ_OK)
            throw std::runtime_error(shkz_last_error());
        const N n = A->rows();
        std::vector<int64_t> rowptr(n + 1, 0), col;
        std::vector<double> val;
        for (N row = 0; row < n; ++row) {
            A->const_for_each(row, [&](N column, T value) {
                col.push_back((int64_t)column);
                val.push_back((double)value);
            });
            rowptr[row + 1] = (int64_t)col.size();
        }
        std::vector<T> rhs;
        b->convert_to(rhs);
        std::vector<double> bd(rhs.begin(), rhs.end()), xd(n, 0.0);
        shkz_linsolve_result r = {};
        const int rc = shkz_linsolve_csr(m_solver, (int64_t)n, rowptr.data(), col.data(), val.data(), bd.data(),
                                         xd.data(), &m_params, &r);
        if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) throw std::runtime_error(shkz_last_error());
        x->convert_from(std::vector<T>(xd.begin(), xd.end()));
        return {(N)r.iterations, (T)r.reresid};
    }

public:
    virtual ~pcg_mi355x() {
        if (m_solver) shkz_linsolver_destroy(m_solver);
    }

protected:
    shkz_linsolve_params m_params{};
    int m_device{0};
    mutable shkz_linsolver *m_solver{nullptr};
};

extern "C" module *create_instance() { return new pcg_mi355x<INDEX_TYPE, FLOAT_TYPE>(); }

extern "C" const char *license() { return "MIT"; }
