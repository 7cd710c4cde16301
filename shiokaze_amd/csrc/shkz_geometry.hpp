// shkz_geometry.hpp -- face fractions of the variational ghost-fluid operator, evaluated on the
// device directly from the level sets (matrix-free: nothing per face is stored).
//
//   area  = open (non-solid) fraction of a face, macutility3::compute_area_fraction
//           (src/utility/macutility3.cpp:94-165) with utility::get_area
//           (include/shiokaze/utility/utility.h:179-214);
//   rho   = liquid fraction of a face, macutility3::compute_fluid_fraction (:166-194) with
//           utility::fraction (utility.h:162-170);
//   both are rounded to float as the reference stores them (Real = float), and the
//   first-order switches of macpressuresolver3.cpp:70-80 are applied.
// Floating-point contraction is off here so that the values match the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shkz {

struct GridDesc {
    int w, h, d;
    int phi64, solid64;
    double dx, dt;
    const void *phi;     // cells
    const void *solid;   // nodes or nullptr
    const int *exist;    // device: [0] liquid level set has a negative value, [1] solid has one
    double eps_fluid, eps_solid;
    int so_fluid, so_solid;
    float *vel[3];
    const uint8_t *active[3];
    uint8_t *active_out[3]; // optional: activity after the update (nullptr = not wanted)
    int flat;            // a 2D grid (w, h) held as d = 1: solid nodes (w+1, h+1), no z faces
    int liquid;          // whether the liquid level set has a negative value: 0 = read exist[0],
                         // 1 = no, 2 = yes (known on the host once the labelling has run)
};

// Dense indices in 32-bit arithmetic (in-range, non-negative coordinates): every dense box holds
// fewer than 2^32 entries (checked at entry), and a 64-bit multiply is several instructions.
__device__ __forceinline__ size_t cell_index(const GridDesc &G, int i, int j, int k) {
    return ((unsigned)i * (unsigned)G.h + (unsigned)j) * (unsigned)G.d + (unsigned)k;
}
__device__ __forceinline__ size_t face_index(const GridDesc &G, int dim, int i, int j, int k) {
    const unsigned H = G.h + (dim == 1), D = G.d + (dim == 2);
    return ((unsigned)i * H + (unsigned)j) * D + (unsigned)k;
}
__device__ __forceinline__ size_t node_index(const GridDesc &G, int i, int j, int k) {
    return ((unsigned)i * (unsigned)(G.h + 1) + (unsigned)j) * (unsigned)(G.d + 1) + (unsigned)k;
}
__device__ __forceinline__ double phi_at(const GridDesc &G, int i, int j, int k) {
    const size_t c = cell_index(G, i, j, k);
    return G.phi64 ? ((const double *)G.phi)[c] : (double)((const float *)G.phi)[c];
}
__device__ __forceinline__ double solid_at(const GridDesc &G, int i, int j, int k) {
    const size_t n = G.flat ? (size_t)((unsigned)i * (unsigned)(G.h + 1) + (unsigned)j) : node_index(G, i, j, k);
    return G.solid64 ? ((const double *)G.solid)[n] : (double)((const float *)G.solid)[n];
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// N samples of a float-or-double array at in-range indices, with one branch on the element type
// for the group: the loads are issued together instead of one type test (and one wait) apiece.
template <int N>
__device__ __forceinline__ void load_samples(const void *base, bool is64, const size_t (&idx)[N], double (&out)[N]) {
    if (is64) {
        const double *a = (const double *)base;
#pragma unroll
        for (int q = 0; q < N; ++q) out[q] = a[idx[q]];
    } else {
        const float *a = (const float *)base;
#pragma unroll
        for (int q = 0; q < N; ++q) out[q] = (double)a[idx[q]];
    }
}

// The grid's "has a negative value" flags (liquid, solid), read once per thread: the face
// functions would otherwise reload them for every face, behind every store that might alias them.
// The solid flag is always set: with no negative solid sample, every face's marching-squares
// area is exactly 1 (no vertex, no crossing: a zero sum), the value of the reference's no-solid
// branch, so the solid branch runs whenever a solid array is given and no scan of it is needed.
struct Exist {
    bool liquid, solid;
};
__device__ __forceinline__ Exist exist_of(const GridDesc &G) {
    return Exist{G.liquid ? G.liquid == 2 : G.exist[0] != 0, true};
}

// (i, j, k) of flat index e of a z-fastest (ni, nj, nk) box, in 32-bit arithmetic: every dense box
// here holds fewer than 2^32 entries (checked at entry), and 64-bit division has no hardware
// support (a long software sequence per thread).
__device__ __forceinline__ void unflatten(unsigned e, unsigned nj, unsigned nk, int &i, int &j, int &k) {
    const unsigned t = e / nk;
    k = (int)(e - t * nk);
    const unsigned a = t / nj;
    i = (int)a;
    j = (int)(t - a * nj);
}

// Liquid fraction between two level-set samples (utility.h:162-170).
__device__ __forceinline__ double liquid_fraction(double phi0, double phi1) {
#pragma clang fp contract(off)
    if (phi0 * phi1 >= 0.0) return (phi0 < 0.0 || phi1 < 0.0) ? 1.0 : 0.0;
    const double lo = phi0 < phi1 ? phi0 : phi1;
    double gap = fabs(phi1 - phi0);
    if (gap < 2.2250738585072014e-308) gap = 2.2250738585072014e-308;
    return -lo / gap;
}

// Negative-region area of a unit square from corner samples v[0..3] at (0,0),(1,0),(1,1),(0,1):
// polygon of negative corners and edge crossings, shoelace sum (utility.h:179-214).
__device__ __forceinline__ double negative_area(const double v[4]) {
#pragma clang fp contract(off)
    // no negative corner: no vertex and no crossing (a product of two non-negative samples is not
    // negative), the sum is 0; all four negative: the unit square, whose shoelace sum is exactly 2
    const bool n0 = v[0] < 0.0, n1 = v[1] < 0.0, n2 = v[2] < 0.0, n3 = v[3] < 0.0;
    if (!(n0 || n1 || n2 || n3)) return 0.0;
    if (n0 && n1 && n2 && n3) return 1.0;
    const int cx[4] = {0, 1, 1, 0};
    const int cy[4] = {0, 0, 1, 1};
    double px[8], py[8];
    int m = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const int nn = (n + 1) & 3;
        if (v[n] < 0.0) {
            px[m] = cx[n];
            py[m] = cy[n];
            ++m;
        }
        if (v[n] * v[nn] < 0) {
            const double y0 = v[n], y1 = v[nn];
            if (y0 - y1) {
                const double a = y0 / (y0 - y1);
                px[m] = (1.0 - a) * cx[n] + a * cx[nn];
                py[m] = (1.0 - a) * cy[n] + a * cy[nn];
                ++m;
            }
        }
    }
    double sum = 0.0;
    for (int a = 0; a < m; ++a) {
        const int b = (a + 1) % m;
        sum += px[a] * py[b] - py[a] * px[b];
    }
    return 0.5 * sum;
}

// Open fraction of an interior 3D face from its four solid corner samples (corner order of
// macutility3.cpp:123-138 mapped to (0,0),(1,0),(1,1),(0,1)), float as stored by the reference.
__device__ __forceinline__ float area_from_corners(const GridDesc &G, const double v[4]) {
#pragma clang fp contract(off)
    double a = 1.0 - negative_area(v);
    if (a && a < G.eps_solid) a = G.eps_solid;
    float af = (float)a;
    if (!G.so_solid && af) af = 1.0f;
    return af;
}

// Liquid fraction of a face from the level-set samples on its two sides, float as stored by the
// reference (face_rho with the samples given).
__device__ __forceinline__ float rho_from_samples(const GridDesc &G, const Exist &E, double phi0, double phi1) {
#pragma clang fp contract(off)
    double r;
    if (E.liquid) {
        r = liquid_fraction(phi0, phi1);
        if (r && r < G.eps_fluid) r = G.eps_fluid;
    } else {
        r = 1.0;
    }
    float rf = (float)r;
    if (!G.so_fluid && rf) rf = 1.0f;
    return rf;
}

// Open fraction of face (dim; i,j,k), float as stored by the reference.
__device__ __forceinline__ float face_area(const GridDesc &G, const Exist &E, int dim, int i, int j, int k) {
#pragma clang fp contract(off)
    const int N[3] = {G.w, G.h, G.d};
    const int pi[3] = {i, j, k};
    const bool boundary = (pi[dim] == 0 || pi[dim] == N[dim]);
    double a;
    if (G.solid && E.solid) {
        if (boundary) {
            a = 0.0;
        } else if (G.flat) {
            // 2D face = an edge between two nodes: 1 - fraction of the solid samples at its ends
            // (macutility2.cpp:104)
            a = 1.0 - liquid_fraction(solid_at(G, i, j, 0), solid_at(G, i + (dim != 0), j + (dim != 1), 0));
        } else {
            // corner order of macutility3.cpp:123-138 mapped to (0,0),(1,0),(1,1),(0,1)
            // the face's second and third corner step along a and b (the two axes other than dim)
            const int ai = dim != 0, aj = dim == 0, bj = dim == 2, bk = dim != 2;
            const size_t idx[4] = {node_index(G, i, j, k), node_index(G, i + ai, j + aj, k),
                                   node_index(G, i + ai, j + aj + bj, k + bk), node_index(G, i, j + bj, k + bk)};
            double v[4];
            load_samples<4>(G.solid, G.solid64, idx, v);
            return area_from_corners(G, v);
        }
        if (a && a < G.eps_solid) a = G.eps_solid;
    } else {
        a = boundary ? 0.0 : 1.0;
    }
    float af = (float)a;
    if (!G.so_solid && af) af = 1.0f;
    return af;
}

// Liquid fraction of face (dim; i,j,k) between cell (i,j,k) and (i,j,k)-e_dim, clamped into the
// grid, float as stored by the reference.
__device__ __forceinline__ float face_rho(const GridDesc &G, const Exist &E, int dim, int i, int j, int k) {
    if (!E.liquid) return rho_from_samples(G, E, 0.0, 0.0);
    const int i0 = clampi(i, 0, G.w - 1), j0 = clampi(j, 0, G.h - 1), k0 = clampi(k, 0, G.d - 1);
    const int i1 = clampi(i - (dim == 0), 0, G.w - 1), j1 = clampi(j - (dim == 1), 0, G.h - 1),
              k1 = clampi(k - (dim == 2), 0, G.d - 1);
    const size_t idx[2] = {cell_index(G, i0, j0, k0), cell_index(G, i1, j1, k1)};
    double ph[2];
    load_samples<2>(G.phi, G.phi64, idx, ph);
    return rho_from_samples(G, E, ph[0], ph[1]);
}

__device__ __forceinline__ float face_area(const GridDesc &G, int dim, int i, int j, int k) {
    return face_area(G, exist_of(G), dim, i, j, k);
}
__device__ __forceinline__ float face_rho(const GridDesc &G, int dim, int i, int j, int k) {
    return face_rho(G, exist_of(G), dim, i, j, k);
}

}  // namespace shkz
