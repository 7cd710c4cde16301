// ref_driver.cpp -- exposes a few header-only pieces of the reference itself through a C ABI,
// compiled straight from the reference's own headers where they lie (see oracle/Makefile).
// Used only by tests in a container that holds the reference, to pin the oracle:
//   utility::fraction / get_area / box   include/shiokaze/utility/utility.h:113-214
//   SparseMatrix::add_to_element,
//   FixedSparseMatrix::construct_from_matrix, multiply
//                                        local/include/pcgsolver/sparse_matrix.h:75-90,217-275
//   array_interpolator3::interpolate_coef
//                                        include/shiokaze/array/array_interpolator3.h:40-62
//   array_derivative3::derivative_interpolate_coef
//                                        include/shiokaze/array/array_derivative3.h:40-85
//   WENO::interp6                        include/shiokaze/math/WENO.h:55-88
// The reference's solver header (pcg_solver.h) needs GSL's cblas headers, which this image
// lacks, so it is not built here.
#include <shiokaze/utility/utility.h>
#include <pcgsolver/sparse_matrix.h>
#include <shiokaze/array/array_interpolator3.h>
#include <shiokaze/array/array_derivative3.h>
#include <shiokaze/math/WENO.h>
#include <cstdint>
#include <vector>

SHKZ_USING_NAMESPACE

extern "C" double ref_fraction(double a, double b) { return utility::fraction(a, b); }

extern "C" double ref_get_area(const double q[4]) {
    double iso[2][2] = {{q[0], q[1]}, {q[2], q[3]}};
    return utility::get_area(iso);
}

extern "C" double ref_box(const double p[3], const double p0[3], const double p1[3]) {
    return utility::box(vec3d(p[0], p[1], p[2]), vec3d(p0[0], p0[1], p0[2]), vec3d(p1[0], p1[1], p1[2]));
}

// Builds SparseMatrix<double> from triplets in the given order (add_to_element), freezes it into
// a FixedSparseMatrix and multiplies x.  Also returns the frozen CSR (rowstart n+1, col/val nnz).
extern "C" int64_t ref_spmv(int64_t n, int64_t nt, const int64_t *ti, const int64_t *tj, const double *tv,
                            const double *x, double *y, int64_t *rowstart, int64_t *colindex, double *value) {
    SparseMatrix<double> m((unsigned)n);
    for (int64_t t = 0; t < nt; ++t) m.add_to_element((unsigned)ti[t], (unsigned)tj[t], tv[t]);
    FixedSparseMatrix<double> f;
    f.construct_from_matrix(m);
    std::vector<double> xv(x, x + n), yv;
    multiply(f, xv, yv);
    for (int64_t i = 0; i < n; ++i) y[i] = yv[i];
    if (rowstart) {
        for (int64_t i = 0; i <= n; ++i) rowstart[i] = f.rowstart[i];
        for (size_t p = 0; p < f.colindex.size(); ++p) {
            colindex[p] = f.colindex[p];
            value[p] = f.value[p];
        }
    }
    return (int64_t)f.colindex.size();
}

// Corner indices (idx[3n+a]) and trilinear weights at p for an array of the given shape.
extern "C" void ref_interp_coef(const int shape[3], const double p[3], int idx[24], double coef[8]) {
    vec3i ind[8];
    array_interpolator3::interpolate_coef(shape3(shape[0], shape[1], shape[2]), vec3d(p[0], p[1], p[2]), ind, coef);
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a) idx[3 * n + a] = ind[n][a];
}

// Corner indices and the derivative weights coef[8*dim+n] at p.
extern "C" void ref_deriv_coef(const int shape[3], const double p[3], int idx[24], double coef[24]) {
    vec3i ind[8];
    double c[DIM3][8];
    array_derivative3::derivative_interpolate_coef(shape3(shape[0], shape[1], shape[2]), vec3d(p[0], p[1], p[2]),
                                                   ind, c);
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a) {
            idx[3 * n + a] = ind[n][a];
            coef[8 * a + n] = c[a][n];
        }
}

// Six-point WENO interpolation of samples at nodes -2..3 at x in [0, 1].
extern "C" double ref_weno6(double x, const double v[6]) { return WENO::interp6(x, v); }
