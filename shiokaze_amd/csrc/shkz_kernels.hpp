// shkz_kernels.hpp -- HIP kernels of the MI355X pressure projection (gfx950).
//
// Stages, in the order of macpressuresolver3::project (src/projection/macpressuresolver3.cpp):
//   k_any_negative     levelset_exist() of both level sets (array_utility3.h:112-122)
//   k_curvature,
//   k_surface_tension  surface-tension impulse (:85-115)
//   k_flags            unknown labelling (:121-156) + which 8^3 bricks hold unknowns
//   k_operator         matrix-free assembly (:159-199): conductances, Dirichlet sums, rhs
//   k_coarsen, k_idiag multigrid operators (no counterpart: they stand in for MIC(0))
//   k_defect4 ... k_update_p, V-cycle kernels
//                      Krylov solve (stands in for PCGSolver::solve, pcg_solver.h:246-295)
//   k_scatter_pressure m_pressure (:245-248)
//   k_velocity_update  velocity update (:252-268)
//
// Build and output kernels work on one grid's dense arrays.  Solver kernels work on a batch: the
// stored bricks of every grid of the batch, concatenated, one 512-thread block per brick (thread =
// cell) with an XCD-aware block-to-brick remap.  Stencil neighbours in the same brick are read at
// +-1/8/64, the others through the brick's neighbour table.  Reductions are per grid and
// deterministic (per-brick partials, combined in brick order).
#pragma once
#include "shkz_device.hpp"
#include "shkz_geometry.hpp"

// No floating-point contraction in any solver kernel: every a*b+c rounds twice, exactly as
// written.  The same sweep, residual and prolongation are computed by different kernels depending
// on a level's size (one launch per sweep, or several sweeps in LDS), and the size of a level
// depends on the batch a grid is in; with the operations fixed by the source, every form gives the
// same bits, so a grid's result does not depend on its batch.
#pragma clang fp contract(off)

namespace shkz {

// ------------------------------------------------------------------------------ helpers
constexpr int UNIFORM_COUNT_BIT = 1 << 20;  // brick counts: the brick is BM_UNIFORM (counts <= 512)

struct Nbrs {
    long base;
    int n[6];
};

__device__ __forceinline__ Nbrs nbrs_of(const BrickMeta &m, int b) {
    Nbrs N;
    N.base = (long)b * BRICK_CELLS;
#pragma unroll
    for (int d = 0; d < 6; ++d) N.n[d] = m.nbr[d];
    return N;
}

__device__ __forceinline__ Nbrs load_nbrs(const LevelDesc &L, int b) { return nbrs_of(L.meta[b], b); }

// Storage index of the six face neighbours of local cell t, -1 where no brick is stored.
__device__ __forceinline__ void nb_index(const Nbrs &N, int t, long idx[6]) {
    const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
    idx[NB_XM] = li > 0 ? N.base + t - 64 : (N.n[NB_XM] >= 0 ? (long)N.n[NB_XM] * BRICK_CELLS + t + 448 : -1);
    idx[NB_XP] = li < 7 ? N.base + t + 64 : (N.n[NB_XP] >= 0 ? (long)N.n[NB_XP] * BRICK_CELLS + t - 448 : -1);
    idx[NB_YM] = lj > 0 ? N.base + t - 8 : (N.n[NB_YM] >= 0 ? (long)N.n[NB_YM] * BRICK_CELLS + t + 56 : -1);
    idx[NB_YP] = lj < 7 ? N.base + t + 8 : (N.n[NB_YP] >= 0 ? (long)N.n[NB_YP] * BRICK_CELLS + t - 56 : -1);
    idx[NB_ZM] = lk > 0 ? N.base + t - 1 : (N.n[NB_ZM] >= 0 ? (long)N.n[NB_ZM] * BRICK_CELLS + t + 7 : -1);
    idx[NB_ZP] = lk < 7 ? N.base + t + 1 : (N.n[NB_ZP] >= 0 ? (long)N.n[NB_ZP] * BRICK_CELLS + t - 7 : -1);
}

template <class T>
__device__ __forceinline__ T ld(const T *a, long i) {
    return i >= 0 ? a[i] : T(0);
}

// (A p)_c for an unknown c; also returns the diagonal.
template <class T>
__device__ __forceinline__ T apply_row(const OpDesc<T> &A, const T *__restrict__ p, long c, const long idx[6], T dirc,
                                       T pc, T &diag) {
    const T gxm = A.gx[c], gym = A.gy[c], gzm = A.gz[c];
    const T gxp = ld(A.gx, idx[NB_XP]), gyp = ld(A.gy, idx[NB_YP]), gzp = ld(A.gz, idx[NB_ZP]);
    const T off = gxm * ld(p, idx[NB_XM]) + gxp * ld(p, idx[NB_XP]) + gym * ld(p, idx[NB_YM]) +
                  gyp * ld(p, idx[NB_YP]) + gzm * ld(p, idx[NB_ZM]) + gzp * ld(p, idx[NB_ZP]);
    diag = dirc + ((gxm + gxp) + (gym + gyp) + (gzm + gzp));
    return diag * pc - off;
}

// Sum of g_f * v_q over the six faces of cell c.
template <class T>
__device__ __forceinline__ T offdiag_sum(const OpDesc<T> &A, const T *__restrict__ v, long c, const long idx[6]) {
    const T gxp = ld(A.gx, idx[NB_XP]), gyp = ld(A.gy, idx[NB_YP]), gzp = ld(A.gz, idx[NB_ZP]);
    return A.gx[c] * ld(v, idx[NB_XM]) + gxp * ld(v, idx[NB_XP]) + A.gy[c] * ld(v, idx[NB_YM]) +
           gyp * ld(v, idx[NB_YP]) + A.gz[c] * ld(v, idx[NB_ZM]) + gzp * ld(v, idx[NB_ZP]);
}

// apply_row and offdiag_sum on a BM_UNIFORM brick: every coupling is g and the Dirichlet sum 0
// (the same operations in the same order, so the same values).
template <class T>
__device__ __forceinline__ T offdiag_uniform(T g, const T *__restrict__ v, const long idx[6]) {
    return g * ld(v, idx[NB_XM]) + g * ld(v, idx[NB_XP]) + g * ld(v, idx[NB_YM]) + g * ld(v, idx[NB_YP]) +
           g * ld(v, idx[NB_ZM]) + g * ld(v, idx[NB_ZP]);
}
template <class T>
__device__ __forceinline__ T apply_row_uniform(T g, const T *__restrict__ p, long c, const long idx[6], T pc) {
    (void)c;
    const T off = offdiag_uniform(g, p, idx);
    const T diag = T(0) + ((g + g) + (g + g) + (g + g));
    return diag * pc - off;
}

__device__ __forceinline__ int brick_or_idle(const LevelDesc &L) {
    const int b = xcd_brick(blockIdx.x, gridDim.x);
    return b < L.nbricks ? b : -1;
}

__device__ __forceinline__ bool all_done(const Gate &Q) { return Q.ctl && Q.ctl->all_done; }

// The brick's grid is done (in a batch of one, all_done says so already: no load).
__device__ __forceinline__ bool grid_done(const Gate &Q, const LevelDesc &L, const BrickMeta &m) {
    return L.ngrids > 1 && Q.S && Q.S[m.grid].done;
}

// Smoothing-side V-cycle work on a brick is skipped for a grid that is done or whose last level
// this is (its coarse solve handles it).
__device__ __forceinline__ bool skip_smooth(const Gate &Q, const LevelDesc &L, const BrickMeta &m) {
    return (m.flags & BM_COARSEST) || grid_done(Q, L, m);
}

// Value of level vector v at cell (i,j,k) of the grid described by G (0 outside or unstored).
template <class T>
__device__ __forceinline__ T level_value(const LevelDesc &L, const GridLevel &G, const T *__restrict__ v, int i, int j,
                                         int k) {
    if (i < 0 || j < 0 || k < 0 || i >= G.ni || j >= G.nj || k >= G.nk) return T(0);
    const int b = L.bmap[G.bmap_off + ((i >> 3) * G.nby + (j >> 3)) * G.nbz + (k >> 3)];
    if (b < 0) return T(0);
    return v[(long)b * BRICK_CELLS + (((i & 7) * 8 + (j & 7)) * 8 + (k & 7))];
}

// ------------------------------------------------------------------------------ build (one grid)
__global__ void k_reset_ints(int *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// flag |= 1 when any of the n values is negative.  A 16-byte aligned array is read 16 bytes per
// load, four loads in flight per thread.
__global__ void k_any_negative(const void *a, int is64, size_t n, int *flag) {
    int found = 0;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
    if (((uintptr_t)a & 15) == 0) {
        const size_t per = is64 ? 2 : 4;  // values per 16 bytes
        const size_t nv = n / per;
        const float4 *v4 = (const float4 *)a;
        auto neg = [&](const float4 x) -> int {
            if (is64) {
                const double2 d = *reinterpret_cast<const double2 *>(&x);
                return (d.x < 0.0) | (d.y < 0.0);
            }
            return (x.x < 0.0f) | (x.y < 0.0f) | (x.z < 0.0f) | (x.w < 0.0f);
        };
        size_t i = tid;
        for (; i + 3 * nt < nv; i += 4 * nt) {
            float4 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = v4[i + u * nt];
#pragma unroll
            for (int u = 0; u < 4; ++u) found |= neg(x[u]);
        }
        for (; i < nv; i += nt) found |= neg(v4[i]);
        for (size_t j = nv * per + tid; j < n; j += nt)
            found |= is64 ? (((const double *)a)[j] < 0.0) : (((const float *)a)[j] < 0.0f);
    } else {
        for (size_t i = tid; i < n; i += nt) {
            const double v = is64 ? ((const double *)a)[i] : (double)((const float *)a)[i];
            found |= (v < 0.0);
        }
    }
    found = __syncthreads_or(found);
    // the flag only ever goes from 0 to 1: a plain store, skipped once it is seen set (thousands of
    // blocks finding a negative value would otherwise queue read-modify-writes on one address)
    if (threadIdx.x == 0 && found && __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 7-point Laplacian of the level set with clamped neighbours (macpressuresolver3.cpp:92-100);
// float level sets are summed in float as the reference sums Real values.
__global__ void k_curvature(GridDesc G, float *curv) {
    const size_t M = (size_t)G.w * G.h * G.d;
    const size_t c = (size_t)xcd_brick(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M) return;
    int i, j, k;
    unflatten((unsigned)c, (unsigned)G.h, (unsigned)G.d, i, j, k);
    const int im = clampi(i - 1, 0, G.w - 1), ip = clampi(i + 1, 0, G.w - 1);
    const int jm = clampi(j - 1, 0, G.h - 1), jp = clampi(j + 1, 0, G.h - 1);
    const int km = clampi(k - 1, 0, G.d - 1), kp = clampi(k + 1, 0, G.d - 1);
    double value;
    if (G.flat) {
        // 5-point Laplacian of the 2D module (macpressuresolver2.cpp:82-89), summed in the level
        // set's precision
        if (G.phi64) {
            const double s = +phi_at(G, im, j, k) + phi_at(G, ip, j, k) + phi_at(G, i, jm, k) + phi_at(G, i, jp, k);
            value = (s - 4.0 * phi_at(G, i, j, k)) / (G.dx * G.dx);
        } else {
            const float *f = (const float *)G.phi;
            const float s = +f[cell_index(G, im, j, k)] + f[cell_index(G, ip, j, k)] + f[cell_index(G, i, jm, k)] +
                            f[cell_index(G, i, jp, k)];
            value = (s - 4.0 * f[c]) / (G.dx * G.dx);
        }
    } else if (G.phi64) {
        const double s = +phi_at(G, im, j, k) + phi_at(G, ip, j, k) + phi_at(G, i, jm, k) + phi_at(G, i, jp, k) +
                         phi_at(G, i, j, km) + phi_at(G, i, j, kp);
        value = (s - 6.0 * phi_at(G, i, j, k)) / (G.dx * G.dx);
    } else {
        const float *f = (const float *)G.phi;
        const float s = +f[cell_index(G, im, j, k)] + f[cell_index(G, ip, j, k)] + f[cell_index(G, i, jm, k)] +
                        f[cell_index(G, i, jp, k)] + f[cell_index(G, i, j, km)] + f[cell_index(G, i, j, kp)];
        value = (s - 6.0 * f[c]) / (G.dx * G.dx);
    }
    curv[c] = (float)value;
}

// Surface-tension impulse on faces with 0 < rho < 1 (macpressuresolver3.cpp:102-113).
__global__ void k_surface_tension(GridDesc G, int dim, const float *curv, double kappa) {
    const int W = G.w + (dim == 0), H = G.h + (dim == 1), D = G.d + (dim == 2);
    const size_t n = (size_t)W * H * D;
    const size_t f = (size_t)xcd_brick(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (f >= n) return;
    if (G.active[dim] && !G.active[dim][f]) return;
    int i, j, k;
    unflatten((unsigned)f, (unsigned)H, (unsigned)D, i, j, k);
    const double rho = face_rho(G, exist_of(G), dim, i, j, k);
    if (rho && rho < 1.0) {
        const int i0 = clampi(i, 0, G.w - 1), j0 = clampi(j, 0, G.h - 1), k0 = clampi(k, 0, G.d - 1);
        const int i1 = clampi(i - (dim == 0), 0, G.w - 1), j1 = clampi(j - (dim == 1), 0, G.h - 1),
                  k1 = clampi(k - (dim == 2), 0, G.d - 1);
        const double sgn = phi_at(G, i0, j0, k0) < 0.0 ? -1.0 : 1.0;
        const double theta = sgn < 0 ? 1.0 - rho : rho;
        const double face_c = theta * curv[cell_index(G, i0, j0, k0)] + (1.0 - theta) * curv[cell_index(G, i1, j1, k1)];
        G.vel[dim][f] += (float)(-sgn * G.dt / (G.dx * rho) * kappa * face_c);
    }
}

// Unknown labelling (macpressuresolver3.cpp:121-156): a liquid cell is an unknown when an
// in-bounds liquid neighbour shares a face with nonzero area (rho is 1 between two liquid
// cells).  One block per brick position of the dense grid; writes cell flags (bit0 liquid, bit1
// unknown) and per brick bit0 = holds an unknown, bit1 = holds a liquid cell (so the host learns
// levelset_exist() of the liquid, array_utility3.h:112-122, on the way).  The face fractions here
// are only taken beside a liquid cell, so the liquid exists.
__global__ __launch_bounds__(BLOCK) void k_flags(GridDesc G, int nbx, int nby, int nbz, uint8_t *cflag, uint8_t *bflag) {
    const int bid = xcd_brick(blockIdx.x, gridDim.x);  // each XCD takes a contiguous slab
    if (bid >= nbx * nby * nbz) return;
    const int bz = bid % nbz, by = (bid / nbz) % nby, bx = bid / (nbz * nby);
    const int t = threadIdx.x;
    const int i = bx * BRICK + (t >> 6), j = by * BRICK + ((t >> 3) & 7), k = bz * BRICK + (t & 7);
    int unknown = 0, any_liquid = 0;
    const Exist E{true, true};
    if (i < G.w && j < G.h && k < G.d) {
        const double pc = phi_at(G, i, j, k);
        const bool liquid = pc < 0.0;
        any_liquid = liquid;
        if (liquid) {
            const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
            for (int nq = 0; nq < 6 && !unknown; ++nq) {
                const int qi = i + di[nq], qj = j + dj[nq], qk = k + dk[nq];
                if (qi < 0 || qi >= G.w || qj < 0 || qj >= G.h || qk < 0 || qk >= G.d) continue;
                const double pq = phi_at(G, qi, qj, qk);
                if (!(pq < 0.0)) continue;
                const int dim = nq >> 1;
                const int fi = (nq & 1) ? i : qi, fj = (nq & 1) ? j : qj, fk = (nq & 1) ? k : qk;
                // face_rho of this interior face from the two samples in hand: the one at the
                // face's index cell first (the neighbour for a high face, the cell for a low one)
                if (face_area(G, E, dim, fi, fj, fk) != 0.0f &&
                    ((nq & 1) ? rho_from_samples(G, E, pc, pq) : rho_from_samples(G, E, pq, pc)) != 0.0f)
                    unknown = 1;
            }
        }
        cflag[cell_index(G, i, j, k)] = (uint8_t)((liquid ? 1 : 0) | (unknown ? 2 : 0));
    }
    // both bits with one barrier: per-wavefront ballots combined in LDS
    __shared__ int wbits[BLOCK / 64];
    const bool wu = __ballot(unknown) != 0, wl = __ballot(any_liquid) != 0;
    if ((t & 63) == 0) wbits[t >> 6] = (wu ? 1 : 0) | (wl ? 2 : 0);
    __syncthreads();
    if (t == 0) {
        int v = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) v |= wbits[w];
        bflag[bid] = (uint8_t)v;
    }
}

// Matrix-free assembly on the stored bricks of one grid at level 0 (macpressuresolver3.cpp:159-199).
// For an unknown c and each in-bounds neighbour q (+x,-x,+y,-y,+z,-z): with area>0 and rho>0
// the conductance w = dt*area/(dx*dx*rho) joins the diagonal, and couples c to q when q is
// liquid (then q is an unknown too); with area>0 the rhs gains -sgn*area*u/dx.  Stored: the
// low-face couplings gx/gy/gz and dir = the sum of conductances to non-unknown neighbours
// (so diag = dir + sum of the six couplings); dir = -1 marks a non-unknown.
//
// The same block also finishes the brick for the solver: the inverse diagonal the smoothers use
// (k_idiag's arithmetic on the stored FP32 (id32) or FP64 (id64) values, each coupling of a high
// face being the neighbour's low one, formed here from the same samples), the brick's unknown count
// (the count k_idiag forms on the coarse levels), and whether its operator is one
// constant: every cell an unknown, dir = 0, all six couplings and the inverse diagonal equal to the
// brick's first cell's (then BM_UNIFORM and its UniCoef record).
struct OperatorFinish {
    float *id32;      // FP32 inverse diagonal (the mixed V-cycle) or nullptr
    double *id64;     // FP64 inverse diagonal (FP64 V-cycle) or nullptr
    BrickMeta *meta;  // level-0 brick records (BM_UNIFORM is or-ed in)
    UniCoef *uni;
    int *counts;      // unknowns per brick (| UNIFORM_COUNT_BIT)
};
__global__ __launch_bounds__(BLOCK) void k_operator(GridDesc G, LevelDesc L, GridLevel GL, const uint8_t *__restrict__ cflag,
                                                    double *dir64, double *gx64, double *gy64, double *gz64,
                                                    float *dir32, float *gx32, float *gy32, float *gz32, double *b,
                                                    double *x, const double *__restrict__ warm, double rhs_correct,
                                                    OperatorFinish FIN) {
#pragma clang fp contract(off)
    const int local = xcd_brick(blockIdx.x, gridDim.x);
    if (local >= GL.brick_end - GL.brick_start) return;
    const int br = GL.brick_start + local;
    const int t = threadIdx.x;
    const int bx = L.bxyz[3 * br], by = L.bxyz[3 * br + 1], bz = L.bxyz[3 * br + 2];
    const int i = bx * BRICK + (t >> 6), j = by * BRICK + ((t >> 3) & 7), k = bz * BRICK + (t & 7);
    const long c = (long)br * BRICK_CELLS + t;
    const bool inside = i < G.w && j < G.h && k < G.d;
    const bool unknown = inside && (cflag[cell_index(G, i, j, k)] & 2);
    double dir = -1.0, g[3] = {0.0, 0.0, 0.0}, rhs = 0.0, x0 = 0.0;
    double gp[3] = {0.0, 0.0, 0.0};  // couplings of the high faces (the neighbours' low ones)
    if (unknown && !G.flat) {
        // 3D: the cell's 8 solid corners, 7 level-set samples and 6 face velocities are loaded at
        // once; each face's area and liquid fraction then use the face functions' arithmetic
        const Exist E = exist_of(G);
        const double dt = G.dt, dx = G.dx;
        const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
        const int sgn[6] = {1, -1, 1, -1, 1, -1};
        double sc[8];  // solid at node (i + a, j + b, k + c), index (a * 2 + b) * 2 + c
        if (G.solid) {
            size_t nidx[8];
#pragma unroll
            for (int n = 0; n < 8; ++n) nidx[n] = node_index(G, i + (n >> 2), j + ((n >> 1) & 1), k + (n & 1));
            load_samples<8>(G.solid, G.solid64, nidx, sc);
        }
        // level set at the six neighbours (clamped into the grid: out-of-grid ones go unused) and
        // at the cell (last)
        size_t pidx[7];
        float uf[6];  // velocity at the six faces
#pragma unroll
        for (int nq = 0; nq < 6; ++nq) {
            pidx[nq] = cell_index(G, clampi(i + di[nq], 0, G.w - 1), clampi(j + dj[nq], 0, G.h - 1),
                                  clampi(k + dk[nq], 0, G.d - 1));
            const int dim = nq >> 1, o = (nq & 1) ? 0 : 1;
            uf[nq] = G.vel[dim][face_index(G, dim, i + (dim == 0) * o, j + (dim == 1) * o, k + (dim == 2) * o)];
        }
        pidx[6] = cell_index(G, i, j, k);
        double ph[7];
        load_samples<7>(G.phi, G.phi64, pidx, ph);
        const double pc = ph[6];
        double dsum = 0.0;
#pragma unroll
        for (int nq = 0; nq < 6; ++nq) {
            const int qi = i + di[nq], qj = j + dj[nq], qk = k + dk[nq];
            if (qi < 0 || qi >= G.w || qj < 0 || qj >= G.h || qk < 0 || qk >= G.d) continue;
            const int dim = nq >> 1, o = (nq & 1) ? 0 : 1;  // the face lies at offset o along dim
            float area = 1.0f;  // an interior face without solid (face_area's no-solid branch)
            if (G.solid) {
                // face_area's corner order for the face of axis dim at offset o
                double v[4];
                auto n = [&](int a, int b, int c) { return sc[(a * 2 + b) * 2 + c]; };
                if (dim == 0) {
                    v[0] = n(o, 0, 0); v[1] = n(o, 1, 0); v[2] = n(o, 1, 1); v[3] = n(o, 0, 1);
                } else if (dim == 1) {
                    v[0] = n(0, o, 0); v[1] = n(1, o, 0); v[2] = n(1, o, 1); v[3] = n(0, o, 1);
                } else {
                    v[0] = n(0, 0, o); v[1] = n(1, 0, o); v[2] = n(1, 1, o); v[3] = n(0, 1, o);
                }
                area = area_from_corners(G, v);
            }
            if (area) {
                // face_rho: the sample at the face's index cell first, then the one below it
                const double rho = o ? rho_from_samples(G, E, ph[nq], pc) : rho_from_samples(G, E, pc, ph[nq]);
                if (rho) {
                    const double value = dt * (double)area / (dx * dx * rho);
                    if (ph[nq] < 0.0) {
                        if (nq & 1) g[dim] = value;
                        else gp[dim] = value;
                    } else {
                        dsum += value;
                    }
                }
                rhs += -sgn[nq] * (double)area * (double)uf[nq] / dx;
            }
        }
        dir = dsum;
        rhs += rhs_correct;
        if (warm) x0 = warm[cell_index(G, i, j, k)];
    } else if (unknown) {
        const Exist E = exist_of(G);
        const double dt = G.dt, dx = G.dx;
        const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
        const int sgn[6] = {1, -1, 1, -1, 1, -1};
        double dsum = 0.0;
        for (int nq = 0; nq < 6; ++nq) {
            const int qi = i + di[nq], qj = j + dj[nq], qk = k + dk[nq];
            if (qi < 0 || qi >= G.w || qj < 0 || qj >= G.h || qk < 0 || qk >= G.d) continue;
            const int dim = nq >> 1;
            const int fi = (nq & 1) ? i : qi, fj = (nq & 1) ? j : qj, fk = (nq & 1) ? k : qk;
            const double area = face_area(G, E, dim, fi, fj, fk);
            if (area) {
                const double rho = face_rho(G, E, dim, fi, fj, fk);
                if (rho) {
                    const double value = dt * area / (dx * dx * rho);
                    if (phi_at(G, qi, qj, qk) < 0.0) {
                        if (nq & 1) g[dim] = value;
                        else gp[dim] = value;
                    } else {
                        dsum += value;
                    }
                }
                if (!G.flat) {
                    const float u = G.vel[dim][face_index(G, dim, fi, fj, fk)];
                    rhs += -sgn[nq] * area * (double)u / dx;
                }
            }
            if (G.flat) {
                // the 2D module adds the term for every in-bounds neighbour (macpressuresolver2.cpp:176)
                const float u = G.vel[dim][face_index(G, dim, fi, fj, fk)];
                rhs += -sgn[nq] * area * (double)u / dx;
            }
        }
        dir = dsum;
        rhs += rhs_correct;
        if (warm) x0 = warm[cell_index(G, i, j, k)];
    }
    dir64[c] = dir;
    gx64[c] = g[0];
    gy64[c] = g[1];
    gz64[c] = g[2];
    b[c] = unknown ? rhs : 0.0;
    x[c] = x0;
    if (dir32) {
        dir32[c] = (float)dir;
        gx32[c] = (float)g[0];
        gy32[c] = (float)g[1];
        gz32[c] = (float)g[2];
    }
    // k_idiag: diag = dir + the six couplings, in the stored type
    float i32 = 0.0f;
    double i64 = 0.0;
    if (dir >= 0) {
        if (FIN.id32) {
            const float d = (float)dir + (((float)g[0] + (float)gp[0]) + ((float)g[1] + (float)gp[1]) +
                                          ((float)g[2] + (float)gp[2]));
            if (d > 0) i32 = 1.0f / d;
        }
        if (FIN.id64) {
            const double d = dir + ((g[0] + gp[0]) + (g[1] + gp[1]) + (g[2] + gp[2]));
            if (d > 0) i64 = 1.0 / d;
        }
    }
    if (FIN.id32) FIN.id32[c] = i32;
    if (FIN.id64) FIN.id64[c] = i64;
    // unknown count and the uniformity vote: the brick's first cell gives the reference values
    __shared__ double gref, iref64;
    __shared__ float iref32;
    if (t == 0) {
        gref = g[0];
        iref64 = i64;
        iref32 = i32;
    }
    const int n = __syncthreads_count(dir >= 0);
    const double Gu = gref;
    const bool ok = Gu > 0.0 && dir == 0.0 && g[0] == Gu && g[1] == Gu && g[2] == Gu && gp[0] == Gu && gp[1] == Gu &&
                    gp[2] == Gu && (!FIN.id64 || i64 == iref64) && (!FIN.id32 || i32 == iref32);
    const bool uniform = __syncthreads_and(ok);
    if (t == 0) {
        if (uniform) {
            FIN.meta[br].flags = L.meta[br].flags | BM_UNIFORM;
            UniCoef u;
            u.g64 = Gu;
            u.id64 = iref64;
            u.g32 = (float)Gu;
            u.id32 = iref32;
            u.pad[0] = u.pad[1] = 0.0f;
            FIN.uni[br] = u;
        }
        FIN.counts[br] = n | (uniform ? UNIFORM_COUNT_BIT : 0);
    }
}

// ------------------------------------------------------------------------------ hierarchy (batch)
// Coarse operator from the piecewise-constant aggregation P0 of the 2x2x2 children: the couplings
// are 1/2 of the Galerkin product (g_H = 1/2 sum of the four fine couplings across the coarse
// face, the usual rediscretisation factor for a 2h grid), the Dirichlet (air) part is kept whole
// (dir_H = sum of the children's dir), which keeps the free-surface condition as strong on the
// coarse grids as on the fine one.  Every child of a coarse cell and every fine face it needs lie
// in one fine brick.
template <class Ti, class To>
__global__ __launch_bounds__(BLOCK) void k_coarsen(LevelDesc Lf, OpDesc<Ti> Af, LevelDesc Lc, To *dir, To *gx, To *gy,
                                                   To *gz) {
    const int br = brick_or_idle(Lc);
    if (br < 0) return;
    const int t = threadIdx.x;
    const GridLevel F = Lf.gl[Lc.meta[br].grid];
    const int I = Lc.bxyz[3 * br] * BRICK + (t >> 6), J = Lc.bxyz[3 * br + 1] * BRICK + ((t >> 3) & 7),
              K = Lc.bxyz[3 * br + 2] * BRICK + (t & 7);
    const long c = (long)br * BRICK_CELLS + t;
    double sdir = 0.0, sg[3] = {0.0, 0.0, 0.0};
    bool any = false;
    const int fi0 = 2 * I, fj0 = 2 * J, fk0 = 2 * K;
    if (fi0 < F.ni && fj0 < F.nj && fk0 < F.nk) {
        const int fb = Lf.bmap[F.bmap_off + ((fi0 >> 3) * F.nby + (fj0 >> 3)) * F.nbz + (fk0 >> 3)];
        if (fb >= 0) {
            for (int a = 0; a < 2; ++a)
                for (int bb = 0; bb < 2; ++bb)
                    for (int cc = 0; cc < 2; ++cc) {
                        const int li = (fi0 + a) & 7, lj = (fj0 + bb) & 7, lk = (fk0 + cc) & 7;
                        const long f = (long)fb * BRICK_CELLS + (li * 8 + lj) * 8 + lk;
                        const double d = (double)Af.dir[f];
                        if (d >= 0.0) {
                            any = true;
                            sdir += d;
                        }
                        if (a == 0) sg[0] += (double)Af.gx[f];
                        if (bb == 0) sg[1] += (double)Af.gy[f];
                        if (cc == 0) sg[2] += (double)Af.gz[f];
                    }
        }
    }
    dir[c] = any ? (To)sdir : (To)(-1);
    gx[c] = any ? (To)(0.5 * sg[0]) : (To)0;
    gy[c] = any ? (To)(0.5 * sg[1]) : (To)0;
    gz[c] = any ? (To)(0.5 * sg[2]) : (To)0;
}

// Brick (x, y, z) of grid g on a level, -1 outside the grid or where none is stored.
__device__ __forceinline__ int brick_at(const LevelDesc &L, int g, int x, int y, int z) {
    const GridLevel &gl = L.gl[g];
    if (x < 0 || y < 0 || z < 0 || x >= gl.nbx || y >= gl.nby || z >= gl.nbz) return -1;
    return L.bmap[gl.bmap_off + (x * gl.nby + y) * gl.nbz + z];
}

// Transfer and neighbourhood tables of one level L from the base tables of L and its neighbours
// in the hierarchy (Lf one level finer, Lc one coarser; a table is skipped where its pointer is
// null).  One thread per (brick, slot), 99 slots: 64 of child64 (the finer bricks 2B-1 .. 2B+2
// per axis), 8 of parent8 (the coarser bricks (B+1)/2 - 1 + {0,1} per axis; left at -1 for a grid
// that ends on this level) and 27 of nb27 (B-1 .. B+1 per axis, in k_smooth_k's order).
constexpr int LINK_SLOTS = 64 + 8 + 27;
__global__ __launch_bounds__(256) void k_links(LevelDesc L, LevelDesc Lf, LevelDesc Lc, int *child64, int *parent8,
                                               int *nb27) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int br = t / LINK_SLOTS, q = t - br * LINK_SLOTS;
    if (br >= L.nbricks) return;
    const int g = L.meta[br].grid;
    const int X = L.bxyz[3 * br], Y = L.bxyz[3 * br + 1], Z = L.bxyz[3 * br + 2];
    if (q < 64) {
        if (child64)
            child64[64 * (size_t)br + q] = brick_at(Lf, g, 2 * X - 1 + (q >> 4), 2 * Y - 1 + ((q >> 2) & 3), 2 * Z - 1 + (q & 3));
    } else if (q < 72) {
        if (parent8) {
            const int c = q - 64;
            parent8[8 * (size_t)br + c] =
                L.gl[g].coarsest ? -1
                                 : brick_at(Lc, g, (X + 1) / 2 - 1 + (c >> 2), (Y + 1) / 2 - 1 + ((c >> 1) & 1),
                                            (Z + 1) / 2 - 1 + (c & 1));
        }
    } else if (nb27) {
        const int c = q - 72;
        nb27[27 * (size_t)br + c] = brick_at(L, g, X + c / 9 - 1, Y + (c / 3) % 3 - 1, Z + c % 3 - 1);
    }
}

// The smoothers use a stored inverse diagonal: damped Jacobi is
//   z' = (1 - omega) z + omega * idiag * (rhs + sum_f g_f z_q),
// and idiag = 0 marks a non-unknown (or a cell whose diagonal vanished).
// With counts, the same block also counts the brick's unknowns (dir >= 0), plus UNIFORM_COUNT_BIT
// on BM_UNIFORM bricks.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_idiag(LevelDesc L, OpDesc<T> A, T *__restrict__ idiag, int *counts) {
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    const T dirc = A.dir[c];
    T out = 0;
    if (dirc >= 0) {
        const Nbrs N = load_nbrs(L, br);
        long idx[6];
        nb_index(N, t, idx);
        const T diag = dirc + ((A.gx[c] + ld(A.gx, idx[NB_XP])) + (A.gy[c] + ld(A.gy, idx[NB_YP])) +
                               (A.gz[c] + ld(A.gz, idx[NB_ZP])));
        if (diag > 0) out = T(1) / diag;
    }
    idiag[c] = out;
    if (counts) {
        const int n = __syncthreads_count(dirc >= 0);
        if (t == 0) counts[br] = n | ((L.meta[br].flags & BM_UNIFORM) ? UNIFORM_COUNT_BIT : 0);
    }
}

// ------------------------------------------------------------------------------ Krylov (batch)
// Grid of brick br (no load in a batch of one).
__device__ __forceinline__ int grid_of(const LevelDesc &L, int br) { return L.ngrids > 1 ? L.meta[br].grid : 0; }

// q = A p and p.q partials (FIN_PQ forms alpha = (r.z) / (p.q)).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_apply_dot(LevelDesc L, OpDesc<T> A, const T *__restrict__ p,
                                                     T *__restrict__ q, const SolveScalars *S, const BatchCtl *ctl,
                                                     RedBuf R) {
    __shared__ double lds[32];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    // single grid: p is loaded before the gate resolves; batch: finished grids load nothing
    if (L.ngrids > 1 && (ctl->all_done || S[m.grid].done)) return;
    const T pc = p[c];
    if (ctl->all_done) return;
    T out = 0;
    double part = 0.0;
    if (m.flags & BM_UNIFORM) {
        const UniT<T> u = uni_load<T>(L, br);
        long idx[6];
        nb_index(nbrs_of(m, br), t, idx);
        out = apply_row_uniform(u.g, p, c, idx, pc);
        part = (double)pc * (double)out;
    } else {
        // every load is issued before the unknown test (the values of other cells are discarded)
        const Nbrs N = nbrs_of(m, br);
        long idx[6];
        nb_index(N, t, idx);
        const T dirc = A.dir[c];
        T diag;
        const T v = apply_row(A, p, c, idx, dirc, pc, diag);
        out = dirc >= 0 ? v : T(0);
        part = dirc >= 0 ? (double)pc * (double)v : 0.0;
    }
    q[c] = out;
    part = block_sum(part, lds);
    store_partials(R, br, part, 0.0);
}

// x += alpha p, r -= alpha q with ||r||_inf and r.r partials (FIN_UPDATE runs the convergence
// test of pcg_solver.h:278-285).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_cg_update(LevelDesc L, T *__restrict__ x, T *__restrict__ r,
                                                     const T *__restrict__ p, const T *__restrict__ q,
                                                     const SolveScalars *S, const BatchCtl *ctl, RedBuf R) {
    __shared__ double lds[32];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const long c = (long)br * BRICK_CELLS + threadIdx.x;
    // A single grid issues the vectors' loads before the gate's loads resolve (they are wasted
    // only on an iteration past convergence); in a batch, bricks of finished grids load nothing.
    if (L.ngrids > 1 && (ctl->all_done || S[grid_of(L, br)].done)) return;
    const T xv = x[c], pv = p[c], r0 = r[c], qv = q[c];
    const SolveScalars &Sg = S[grid_of(L, br)];
    if (ctl->all_done || Sg.done) return;
    const T alpha = (T)Sg.alpha;
    x[c] = xv + alpha * pv;
    const T rv = r0 - alpha * qv;
    r[c] = rv;
    double mx = fabs((double)rv), sq = (double)rv * (double)rv;
    block_reduce2<true>(mx, sq, lds);
    store_partials(R, br, mx, sq);
}

// r.z partials for the next direction (FIN_RZ forms beta).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_dot_rz(LevelDesc L, const T *__restrict__ r, const T *__restrict__ z,
                                                  const SolveScalars *S, const BatchCtl *ctl, RedBuf R) {
    __shared__ double lds[32];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    if (ctl->all_done || S[grid_of(L, br)].done) return;
    const long c = (long)br * BRICK_CELLS + threadIdx.x;
    const double part = block_sum((double)r[c] * (double)z[c], lds);
    store_partials(R, br, part, 0.0);
}

// p = z + beta p (first: p = z).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_update_p(LevelDesc L, const T *__restrict__ z, T *__restrict__ p,
                                                    const SolveScalars *S, const BatchCtl *ctl, int first) {
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const long c = (long)br * BRICK_CELLS + threadIdx.x;
    // single grid: loads before the gate resolves; batch: finished grids load nothing
    if (L.ngrids > 1 && (ctl->all_done || S[grid_of(L, br)].done)) return;
    const T zv = z[c], pv = first ? T(0) : p[c];
    const SolveScalars &Sg = S[grid_of(L, br)];
    if (ctl->all_done || Sg.done) return;
    p[c] = first ? zv : zv + (T)Sg.beta * pv;
}

enum FinOp {
    FIN_PQ = 0,           // alpha = rz / pq
    FIN_UPDATE = 1,       // rinf, rl2, iteration count, convergence test
    FIN_UPDATE_CG = 2,    // FIN_UPDATE plus beta = r.r / rz (unpreconditioned CG)
    FIN_RZ_FIRST = 3,     // rz of the first direction
    FIN_RZ = 4,           // beta = rz_new / rz
    FIN_DEFECT_FIRST = 5, // ||r0|| and the absolute tolerance  (param = relative residual)
    FIN_DEFECT = 6,       // ||r|| only
    FIN_INNER_INIT = 7,   // inner solve: ||r0||, tol = max(param, tol_outer/||r_outer||) * ||r0||,
                          // maxit = the grid's iterations left (0 when its defect correction is done)
    FIN_OUTER = 8         // defect correction: ||r||, iterations += the inner solve's; done when the
                          // defect converged, the iterations ran out or the defect fell by < 10 %
};

// Fresh per-grid scalars (norm, iteration limit; everything else 0) for n grids.
// Iterations without a new minimum of the residual after which an inner FP32 solve of the defect
// correction stops (and the FP64 defect is recomputed).
constexpr int INNER_STALL_LIMIT = 12;

__global__ void k_init_scalars(SolveScalars *S, int n, int norm, int maxit) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    SolveScalars v{};
    v.norm = norm;
    v.maxit = maxit;
    S[g] = v;
}

// Resets a control block for a new Krylov solve over ngrids grids (a kernel, so that it orders
// with the work around it and needs no host buffer).
__global__ void k_init_ctl(BatchCtl *ctl, int ngrids) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->ngrids = ngrids;
        ctl->ndone = 0;
        ctl->all_done = 0;
        ctl->pad = 0;
    }
}

__device__ __forceinline__ void mark_done(SolveScalars &S, BatchCtl *ctl) {
    S.done = 1;
    if (ctl) {
        const int n = atomicAdd(&ctl->ndone, 1) + 1;
        if (n >= ctl->ngrids) ctl->all_done = 1;
    }
}

// One block per grid combines the partials of the grid's bricks in brick order and updates the
// grid's scalars.  Grids that finished are left alone (their kernels stored no partials).
// Every load is issued before the done test resolves: a single grid's brick range is the level's
// (no dependent load of its GridLevel), and the grid's scalars are copied in the same round trip,
// so the update after the reduction reads no memory (a chain of four dependent round trips became
// one).  The summation order is the brick order as before.
__global__ __launch_bounds__(1024) void k_finalize(const double *__restrict__ partials, LevelDesc L0,
                                                   SolveScalars *Sarr, BatchCtl *ctl, int op, double param, int iparam,
                                                   const SolveScalars *outer) {
    __shared__ double lds[32];
    const int g = blockIdx.x;
    SolveScalars &S = Sarr[g];
    const SolveScalars s0 = S;
    const bool stop = s0.done || (ctl && ctl->all_done);
    int bs = 0, be = L0.nbricks;
    if (L0.ngrids > 1) {
        const GridLevel G = L0.gl[g];
        bs = G.brick_start;
        be = G.brick_end;
    }
    const bool always = (op == FIN_DEFECT_FIRST || op == FIN_DEFECT || op == FIN_INNER_INIT);
    const bool a_is_max = (op == FIN_UPDATE || op == FIN_UPDATE_CG || op == FIN_DEFECT_FIRST || op == FIN_DEFECT ||
                           op == FIN_INNER_INIT || op == FIN_OUTER);
    double a = 0.0, b = 0.0;
    const double2 *pp = reinterpret_cast<const double2 *>(partials);
    // up to 16 loads in flight per thread (16K bricks per pass); past the grid's end a thread adds
    // zeros, which change neither the sums nor the maxima (of norms).  A finished grid's partials
    // are stale but in bounds, and its result is dropped below.
    constexpr int U = 16;
    for (int i = bs + threadIdx.x; i < be; i += U * blockDim.x) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = i + u * (int)blockDim.x;
            v[u] = j < be ? pp[j] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a = a_is_max ? fmax(a, v[u].x) : a + v[u].x;
            b += v[u].y;
        }
    }
    if (!always && stop) return;
    if (a_is_max) block_reduce2<true>(a, b, lds);
    else block_reduce2<false>(a, b, lds);
    if (threadIdx.x != 0) return;
    switch (op) {
    case FIN_PQ: {
        S.pq = a;
        const double al = s0.rz / a;
        S.alpha = al;
        if (!(al == al) || al - al != 0.0) {
            S.status = 1;
            mark_done(S, ctl);
        }
        break;
    }
    case FIN_UPDATE:
    case FIN_UPDATE_CG: {
        const double rl2 = sqrt(b);
        S.rinf = a;
        S.rl2 = rl2;
        const int it = s0.iter + 1;
        S.iter = it;
        const double cur = s0.norm == 0 ? a : rl2;
        if (op == FIN_UPDATE_CG) {
            S.beta = b / s0.rz;
            S.rz = b;
        }
        // an inner solve whose FP32 residual stops reaching new minima hands back to the FP64
        // defect correction (a consistent singular system, for one, keeps a null-space component
        // of the FP32 rounding that no iteration removes)
        bool stalled = false;
        if (s0.stall_limit > 0) {
            if (cur < s0.rmin) {
                S.rmin = cur;
                S.stall = 0;
            } else {
                const int st = s0.stall + 1;
                S.stall = st;
                stalled = st >= s0.stall_limit;
            }
        }
        if (!(b == b)) {
            S.status = 1;
            mark_done(S, ctl);
        } else if (cur <= s0.tol) {
            S.converged = 1;
            mark_done(S, ctl);
        } else if (it >= s0.maxit || stalled) {
            mark_done(S, ctl);
        }
        break;
    }
    case FIN_RZ_FIRST:
    case FIN_RZ: {
        S.beta = op == FIN_RZ_FIRST ? 0.0 : a / s0.rz;
        S.rz = a;
        if (!(a == a)) {
            S.status = 1;
            mark_done(S, ctl);
        } else if (op == FIN_RZ_FIRST && a == 0.0) {
            mark_done(S, ctl);
        }
        break;
    }
    case FIN_DEFECT_FIRST:
    case FIN_DEFECT:
    case FIN_INNER_INIT: {
        const double rl2 = sqrt(b);
        S.rinf = a;
        S.rl2 = rl2;
        if (op != FIN_DEFECT) {
            S.r0inf = a;
            S.r0l2 = rl2;
            const double r0 = s0.norm == 0 ? a : rl2;
            double rel = param;
            int maxit = s0.maxit;
            if (op == FIN_INNER_INIT) {
                const SolveScalars &O = outer[g];
                const double ocur = O.norm == 0 ? O.rinf : O.rl2;
                rel = ocur > 0.0 ? fmax(param, O.tol / ocur) : param;
                maxit = (O.done || ocur <= O.tol || O.status) ? 0 : max(1, O.maxit - O.iter);
                S.maxit = maxit;
                S.rmin = r0;
                S.stall = 0;
                S.stall_limit = INNER_STALL_LIMIT;
            }
            S.tol = rel * r0;
            S.iter = 0;
            S.done = 0;
            const bool bad = !(b == b);
            S.status = bad ? 1 : 0;
            S.converged = (r0 == 0.0) ? 1 : 0;
            if (r0 == 0.0 || bad || maxit <= 0) mark_done(S, ctl);
        }
        break;
    }
    case FIN_OUTER: {
        const SolveScalars &I = outer[g];
        const double prev = s0.norm == 0 ? s0.rinf : s0.rl2;
        const double rl2 = sqrt(b);
        S.rinf = a;
        S.rl2 = rl2;
        const int it = s0.iter + I.iter;
        S.iter = it;
        const double cur = s0.norm == 0 ? a : rl2;
        if (!(b == b) || I.status) {
            S.status = 1;
            mark_done(S, ctl);
        } else if (cur <= s0.tol) {
            S.converged = 1;
            mark_done(S, ctl);
        } else if (it >= s0.maxit || cur > 0.9 * prev) {
            mark_done(S, ctl);
        }
        break;
    }
    default:
        break;
    }
}

// ------------------------------------------------------------------------------ V-cycle (batch)
// z = omega * idiag * rhs (one damped-Jacobi sweep from a zero guess).
// FINEST: 1 on the finest level (a distinct symbol for profiling), 0 on coarse levels.
template <class T, int FINEST>
__global__ __launch_bounds__(BLOCK) void k_smooth_first(LevelDesc L, const T *__restrict__ idiag,
                                                        const T *__restrict__ rhs, T *__restrict__ z, T omega,
                                                        Gate Q) {
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const long c = (long)br * BRICK_CELLS + threadIdx.x;
    const T id = (m.flags & BM_UNIFORM) ? uni_load<T>(L, br).id : idiag[c];
    z[c] = omega * id * rhs[c];
}

// Two damped-Jacobi sweeps from a zero guess in one pass: the first sweep's values at the
// neighbours, omega * idiag_q * rhs_q, are formed on the fly.
template <class T, int FINEST>
__global__ __launch_bounds__(BLOCK) void k_smooth2(LevelDesc L, OpDesc<T> A, const T *__restrict__ idiag,
                                                   const T *__restrict__ rhs, T *__restrict__ z, T omega,
                                                   Gate Q) {
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    if (m.flags & BM_UNIFORM) {
        const UniT<T> u = uni_load<T>(L, br);
        const Nbrs N = nbrs_of(m, br);
        long idx[6];
        nb_index(N, t, idx);
        // neighbours in the brick share its inverse diagonal; the others are read
        auto z1 = [&](long i) {
            return i >= 0 ? omega * ((i >= N.base && i < N.base + BRICK_CELLS) ? u.id : idiag[i]) * rhs[i] : T(0);
        };
        const T off = u.g * z1(idx[NB_XM]) + u.g * z1(idx[NB_XP]) + u.g * z1(idx[NB_YM]) + u.g * z1(idx[NB_YP]) +
                      u.g * z1(idx[NB_ZM]) + u.g * z1(idx[NB_ZP]);
        const T r = rhs[c];
        z[c] = (T(1) - omega) * (omega * u.id * r) + omega * u.id * (r + off);
        return;
    }
    const T id = idiag[c];
    T out = 0;
    if (id > 0) {
        const Nbrs N = nbrs_of(m, br);
        long idx[6];
        nb_index(N, t, idx);
        auto z1 = [&](long i) { return i >= 0 ? omega * idiag[i] * rhs[i] : T(0); };
        const T gxp = ld(A.gx, idx[NB_XP]), gyp = ld(A.gy, idx[NB_YP]), gzp = ld(A.gz, idx[NB_ZP]);
        const T off = A.gx[c] * z1(idx[NB_XM]) + gxp * z1(idx[NB_XP]) + A.gy[c] * z1(idx[NB_YM]) +
                      gyp * z1(idx[NB_YP]) + A.gz[c] * z1(idx[NB_ZM]) + gzp * z1(idx[NB_ZP]);
        const T r = rhs[c];
        out = (T(1) - omega) * (omega * id * r) + omega * id * (r + off);
    }
    z[c] = out;
}

// zout = (1 - omega) zin + omega idiag (rhs + offdiag(zin)).  DOT: also rhs.zout partials (the
// finest level's last post-smoothing sweep yields r.z of the preconditioned CG for free).
template <class T, int DOT>
__device__ __forceinline__ void jacobi_brick(const LevelDesc &L, const OpDesc<T> &A, const T *__restrict__ idiag,
                                             const T *__restrict__ rhs, const T *__restrict__ zin,
                                             T *__restrict__ zout, T omega, const RedBuf &R, int br, const BrickMeta &m,
                                             double *lds);

// Persistent form: the grid holds at most PERSIST_BLOCKS workgroups, each looping over virtual
// block ids vb = blockIdx.x + k * gridDim.x (gridDim.x a multiple of 8, so vb keeps its XCD).
constexpr int PERSIST_BLOCKS = 1024;  // 4 resident 512-thread workgroups on each of 256 CUs

template <class T, int FINEST, int DOT>
__global__ __launch_bounds__(BLOCK) void k_jacobi(LevelDesc L, OpDesc<T> A, const T *__restrict__ idiag,
                                                  const T *__restrict__ rhs, const T *__restrict__ zin,
                                                  T *__restrict__ zout, T omega, Gate Q, RedBuf R, int nvirt) {
    __shared__ double lds[32];
    if (DOT) {  // one brick per workgroup (its reduction costs two barriers per brick)
        const int br = brick_or_idle(L);
        if (br < 0) return;
        const BrickMeta m = L.meta[br];
        if (all_done(Q) || skip_smooth(Q, L, m)) return;
        jacobi_brick<T, DOT>(L, A, idiag, rhs, zin, zout, omega, R, br, m, lds);
        return;
    }
    if (all_done(Q)) return;
    for (int vb = blockIdx.x; vb < nvirt; vb += gridDim.x) {
        const int br = xcd_brick(vb, nvirt);
        if (br >= L.nbricks) continue;
        const BrickMeta m = L.meta[br];
        if (skip_smooth(Q, L, m)) continue;
        jacobi_brick<T, DOT>(L, A, idiag, rhs, zin, zout, omega, R, br, m, lds);
    }
}

template <class T, int DOT>
__device__ __forceinline__ void jacobi_brick(const LevelDesc &L, const OpDesc<T> &A, const T *__restrict__ idiag,
                                             const T *__restrict__ rhs, const T *__restrict__ zin,
                                             T *__restrict__ zout, T omega, const RedBuf &R, int br, const BrickMeta &m,
                                             double *lds) {
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    T out = 0;
    double part = 0.0;
    if (m.flags & BM_UNIFORM) {
        const UniT<T> u = uni_load<T>(L, br);
        long idx[6];
        nb_index(nbrs_of(m, br), t, idx);
        const T r = rhs[c];
        out = (T(1) - omega) * zin[c] + omega * u.id * (r + offdiag_uniform(u.g, zin, idx));
        if (DOT) part = (double)r * (double)out;
    } else {
        // every load is issued before the unknown test (the values of other cells are discarded)
        const Nbrs N = nbrs_of(m, br);
        long idx[6];
        nb_index(N, t, idx);
        const T id = idiag[c];
        const T r = rhs[c];
        const T v = (T(1) - omega) * zin[c] + omega * id * (r + offdiag_sum(A, zin, c, idx));
        out = id > 0 ? v : T(0);
        if (DOT) part = id > 0 ? (double)r * (double)v : 0.0;
    }
    zout[c] = out;
    if (DOT) {
        part = block_sum(part, lds);
        store_partials(R, br, part, 0.0);
    }
}

// ---- FP32 sweeps on cell rows: four consecutive cells (one float4 along z) per thread, 128
// threads per brick.  Lane q of a brick holds the cells (q >> 4, (q >> 1) & 7, 4h .. 4h + 3) with
// h = q & 1; the two lanes of a row are neighbours in the wavefront and swap their edge values, so
// only the row's two end cells read a value across the brick's z faces.  A quarter of the load
// instructions of the one-cell form, and 16-byte accesses.
constexpr int ROW4_THREADS = BRICK_CELLS / 4;

__device__ __forceinline__ float4 ld4(const float *__restrict__ a, long i) {
    return *reinterpret_cast<const float4 *>(a + i);
}
__device__ __forceinline__ float4 ld4z(const float *__restrict__ a, long i) {
    return i >= 0 ? ld4(a, i) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ void unpack4(const float4 v, float o[4]) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}
// Four consecutive values (16-byte loads: one float4 or two double2), zeros where i < 0.
__device__ __forceinline__ void load4(const float *__restrict__ a, long i, float o[4]) { unpack4(ld4(a, i), o); }
__device__ __forceinline__ void load4(const double *__restrict__ a, long i, double o[4]) {
    const double2 u = *reinterpret_cast<const double2 *>(a + i), w = *reinterpret_cast<const double2 *>(a + i + 2);
    o[0] = u.x;
    o[1] = u.y;
    o[2] = w.x;
    o[3] = w.y;
}
template <class T>
__device__ __forceinline__ void load4z(const T *__restrict__ a, long i, T o[4]) {
    if (i >= 0) {
        load4(a, i, o);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = T(0);
    }
}
__device__ __forceinline__ void store4(float *__restrict__ a, long i, const float v[4]) {
    *reinterpret_cast<float4 *>(a + i) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void store4(double *__restrict__ a, long i, const double v[4]) {
    *reinterpret_cast<double2 *>(a + i) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(a + i + 2) = make_double2(v[2], v[3]);
}

// The six neighbour rows of lane q's cells: x and y neighbours as whole rows (index of their first
// cell, -1 where no brick is stored), z neighbours as the row shifted by one cell.
struct Row4Nbrs {
    long c, ixm, ixp, iym, iyp, iz;  // iz: the end cell across the z face (k = -1 for h = 0, 8 for h = 1)
    int h;
};
__device__ __forceinline__ Row4Nbrs row4_nbrs(const BrickMeta &m, int br, int q) {
    Row4Nbrs N;
    const int li = q >> 4, lj = (q >> 1) & 7, h = q & 1;
    const int o = li * 64 + lj * 8 + 4 * h;
    N.h = h;
    N.c = (long)br * BRICK_CELLS + o;
    N.ixm = li > 0 ? N.c - 64 : (m.nbr[NB_XM] >= 0 ? (long)m.nbr[NB_XM] * BRICK_CELLS + o + 448 : -1);
    N.ixp = li < 7 ? N.c + 64 : (m.nbr[NB_XP] >= 0 ? (long)m.nbr[NB_XP] * BRICK_CELLS + o - 448 : -1);
    N.iym = lj > 0 ? N.c - 8 : (m.nbr[NB_YM] >= 0 ? (long)m.nbr[NB_YM] * BRICK_CELLS + o + 56 : -1);
    N.iyp = lj < 7 ? N.c + 8 : (m.nbr[NB_YP] >= 0 ? (long)m.nbr[NB_YP] * BRICK_CELLS + o - 56 : -1);
    const int nz = h ? m.nbr[NB_ZP] : m.nbr[NB_ZM];
    N.iz = nz >= 0 ? (long)nz * BRICK_CELLS + li * 64 + lj * 8 + (h ? 0 : 7) : -1;
    return N;
}

// The values of v at the six neighbours of the lane's four cells (0 where no brick is stored).
template <class T>
struct Row4Vals {
    T c[4], xm[4], xp[4], ym[4], yp[4], zm[4], zp[4];
};
template <class T>
__device__ __forceinline__ Row4Vals<T> row4_vals(const T *__restrict__ v, const Row4Nbrs &N) {
    Row4Vals<T> o;
    load4(v, N.c, o.c);
    load4z(v, N.ixm, o.xm);
    load4z(v, N.ixp, o.xp);
    load4z(v, N.iym, o.ym);
    load4z(v, N.iyp, o.yp);
    const T end = N.iz >= 0 ? v[N.iz] : T(0);
    // lane swap: h = 0 receives the cell k = 4 of its partner, h = 1 the cell k = 3
    const T swap = __shfl_xor(N.h ? o.c[0] : o.c[3], 1, 64);
    o.zm[0] = N.h ? swap : end;
    o.zp[3] = N.h ? end : swap;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        o.zm[j + 1] = o.c[j];
        o.zp[j] = o.c[j + 1];
    }
    return o;
}

// The couplings of the lane's cells: low faces from the cell itself, high faces from the next cell
// along each axis (0 where no brick is stored).
template <class T>
struct Row4Coef {
    T gxm[4], gym[4], gzm[4], gxp[4], gyp[4], gzp[4];
};
template <class T>
__device__ __forceinline__ Row4Coef<T> row4_coef(const OpDesc<T> &A, const Row4Nbrs &N) {
    Row4Coef<T> o;
    load4(A.gx, N.c, o.gxm);
    load4(A.gy, N.c, o.gym);
    load4(A.gz, N.c, o.gzm);
    load4z(A.gx, N.ixp, o.gxp);
    load4z(A.gy, N.iyp, o.gyp);
    const T end = (N.h && N.iz >= 0) ? A.gz[N.iz] : T(0);  // the coupling across k = 8
    const T swap = __shfl_xor(o.gzm[0], 1, 64);            // h = 0: the partner's k = 4
#pragma unroll
    for (int j = 0; j < 3; ++j) o.gzp[j] = o.gzm[j + 1];
    o.gzp[3] = N.h ? end : swap;
    return o;
}

// jacobi_brick on the lane's four cells (same per-cell arithmetic); DOT adds rhs.zout of the
// unknowns, in cell order, to part.
template <int DOT>
__device__ __forceinline__ void jacobi_row4(const LevelDesc &L, const OpDesc<float> &A, const float *__restrict__ idiag,
                                            const float *__restrict__ rhs, const float *__restrict__ zin,
                                            float *__restrict__ zout, float omega, int br, const BrickMeta &m, int q,
                                            double &part) {
    const Row4Nbrs N = row4_nbrs(m, br, q);
    const Row4Vals<float> z = row4_vals(zin, N);
    float r[4], out[4];
    unpack4(ld4(rhs, N.c), r);
    if (m.flags & BM_UNIFORM) {
        const UniT<float> u = uni_load<float>(L, br);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float off = u.g * z.xm[j] + u.g * z.xp[j] + u.g * z.ym[j] + u.g * z.yp[j] + u.g * z.zm[j] +
                              u.g * z.zp[j];
            out[j] = (1.0f - omega) * z.c[j] + omega * u.id * (r[j] + off);
            if (DOT) part += (double)r[j] * (double)out[j];
        }
    } else {
        const Row4Coef<float> g = row4_coef(A, N);
        float id[4];
        unpack4(ld4(idiag, N.c), id);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float off = g.gxm[j] * z.xm[j] + g.gxp[j] * z.xp[j] + g.gym[j] * z.ym[j] + g.gyp[j] * z.yp[j] +
                              g.gzm[j] * z.zm[j] + g.gzp[j] * z.zp[j];
            const float v = (1.0f - omega) * z.c[j] + omega * id[j] * (r[j] + off);
            out[j] = id[j] > 0 ? v : 0.0f;
            if (DOT) part += id[j] > 0 ? (double)r[j] * (double)v : 0.0;
        }
    }
    *reinterpret_cast<float4 *>(zout + N.c) = make_float4(out[0], out[1], out[2], out[3]);
}

// FP32 sweep: workgroups take blockDim.x / 128 consecutive bricks per virtual block (virtual blocks
// vb = blockIdx.x + k * gridDim.x < nvirt, nvirt a multiple of 8, XCD-preserving); each wavefront
// works alone on half a brick (no barrier).
template <int FINEST>
__global__ __launch_bounds__(BLOCK) void k_jacobi4(LevelDesc L, OpDesc<float> A, const float *__restrict__ idiag,
                                                   const float *__restrict__ rhs, const float *__restrict__ zin,
                                                   float *__restrict__ zout, float omega, Gate Q, int nvirt) {
    if (all_done(Q)) return;
    const int bpb = blockDim.x / ROW4_THREADS, slot = threadIdx.x / ROW4_THREADS, q = threadIdx.x % ROW4_THREADS;
    for (int vb = blockIdx.x; vb < nvirt; vb += gridDim.x) {
        const int br = bpb * xcd_brick(vb, nvirt) + slot;
        if (br >= L.nbricks) continue;
        const BrickMeta m = L.meta[br];
        if (skip_smooth(Q, L, m)) continue;
        double part = 0.0;
        jacobi_row4<0>(L, A, idiag, rhs, zin, zout, omega, br, m, q, part);
    }
}

// The finest level's last post-smoothing sweep with the rhs.z partial of each brick: one brick
// per 128-thread workgroup.
__global__ __launch_bounds__(ROW4_THREADS) void k_jacobi4_dot(LevelDesc L, OpDesc<float> A,
                                                              const float *__restrict__ idiag,
                                                              const float *__restrict__ rhs,
                                                              const float *__restrict__ zin, float *__restrict__ zout,
                                                              float omega, Gate Q, RedBuf R) {
    __shared__ double lds[ROW4_THREADS / 64];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    double part = 0.0;
    jacobi_row4<1>(L, A, idiag, rhs, zin, zout, omega, br, m, threadIdx.x, part);
    part = wave_sum(part);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < ROW4_THREADS / 64; ++w) sum += lds[w];
        R.partials[2 * br + 0] = sum;
        R.partials[2 * br + 1] = 0.0;
    }
}

// ---- Brick-local smoother (finest level, FP32): z += M_b (rhs - A z), where M_b applies
// LOCAL_SGS_SWEEPS symmetric red-black Gauss-Seidel sweeps from zero to the brick's own operator
// (its couplings to cells in other bricks dropped, the diagonal whole), i.e. the half-sweeps red,
// black, red, ..., red.  Across bricks it is block Jacobi; inside a brick, Gauss-Seidel.  M_b is
// symmetric, so the V-cycle built on it is a symmetric preconditioner.  One pass over the level's
// arrays per application, like one Jacobi sweep, smoothing about as much as three.  The coarse
// levels use it too (three sweeps per application): there the pre-smoothing from zero runs inside
// the restriction that forms the level's right-hand side (k_restrict_pre).
constexpr int LOCAL_SGS_SWEEPS = 2;         // finest level: the default (2 or 3, a tuned control)
constexpr int LOCAL_SGS_SWEEPS_COARSE = 3;  // coarse levels (one application per smoothing step)

// The couplings of the lane's four cells to cells of the same brick (0 across the brick's faces):
// from the loaded Row4Coef of the level (whose high couplings may come from neighbour bricks).
__device__ __forceinline__ Row4Coef<float> row4_local(const Row4Coef<float> &g, int q) {
    const int li = q >> 4, lj = (q >> 1) & 7, h = q & 1;
    Row4Coef<float> o;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        o.gxm[j] = li > 0 ? g.gxm[j] : 0.0f;
        o.gxp[j] = li < 7 ? g.gxp[j] : 0.0f;
        o.gym[j] = lj > 0 ? g.gym[j] : 0.0f;
        o.gyp[j] = lj < 7 ? g.gyp[j] : 0.0f;
        const int lk = 4 * h + j;
        o.gzm[j] = lk > 0 ? g.gzm[j] : 0.0f;
        o.gzp[j] = lk < 7 ? g.gzp[j] : 0.0f;
    }
    return o;
}
__device__ __forceinline__ Row4Coef<float> row4_local_uniform(float g, int q) {
    Row4Coef<float> u;
#pragma unroll
    for (int j = 0; j < 4; ++j) u.gxm[j] = u.gxp[j] = u.gym[j] = u.gyp[j] = u.gzm[j] = u.gzp[j] = g;
    return row4_local(u, q);
}

// One cell of a lane in the brick-local sweep: what its update needs besides the neighbour values
// (its x couplings and LDS slots, y couplings, three z weights, inverse diagonal and right-hand side).
struct SgsCell {
    float ax_m, ax_p, ay_m, ay_p;  // couplings to the x and y neighbours (0 across the brick's faces)
    float za, zb, zs;              // weights of the z terms: the two own cells of the other group and
                                   // the partner lane's swapped value (see local_sgs4)
    float id, r;
    int xm, xp, st;                // LDS slots: x neighbours (clamped in-bounds where the coupling is
                                   // 0) and the cell's own slot
};

// e = M_b r on the brick of a 128-thread workgroup (lane q: cells 4q .. 4q+3 of the brick, the
// storage order).  lds: 512 floats.  Every lane of the workgroup must call it (barriers, DPP).
//
// Cell j of the lane has colour col0 ^ (j & 1) (col0 = (li + lj) & 1), so each half-sweep updates
// two cells of every lane: group P (j = col0, col0 + 2) in the even half-sweeps (the first one, from
// zero, included), group Q (j = 1 - col0, 3 - col0) in the odd ones.  Held by group, the update has
// no divergent branch (adjacent lanes have opposite col0, so a branch on the colour would run both
// sides in every wavefront), and the neighbours of an updated cell are fixed registers:
//  - y: the lane two below / above in the same 16-lane row has the opposite col0, so the same cell
//    there belongs to the other group at the same position: one DPP row shift per value (0 outside
//    the row: lj = 0 / 7, outside the brick);
//  - z: the lane's own cells of the other group, and at the row's ends the partner lane's
//    (q ^ 1, same col0) by a DPP lane swap; the weights za, zb, zs place the two couplings;
//  - x: lanes q -+ 16 (other wavefront at li = 3 / 4), through LDS.
// A coupling across the brick's faces is 0 and its value is read from an in-bounds slot that holds
// a sweep value, so the product is a zero.  The sums use fused multiply-adds in a fixed order.
constexpr int DPP_ROW_SHR2 = 0x112, DPP_ROW_SHL2 = 0x102, DPP_SWAP_PAIRS = 0xB1;  // quad_perm [1,0,3,2]
template <int CTRL>
__device__ __forceinline__ float sgs_dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float sgs_update(const SgsCell &c, const float *lds, float ym, float yp, float z0, float z1,
                                            float zsw) {
    float acc = c.r;
    acc = __builtin_fmaf(c.ax_m, lds[c.xm], acc);
    acc = __builtin_fmaf(c.ax_p, lds[c.xp], acc);
    acc = __builtin_fmaf(c.ay_m, ym, acc);
    acc = __builtin_fmaf(c.ay_p, yp, acc);
    acc = __builtin_fmaf(c.za, z0, acc);
    acc = __builtin_fmaf(c.zb, z1, acc);
    acc = __builtin_fmaf(c.zs, zsw, acc);
    return c.id * acc;
}

template <int K>
__device__ __forceinline__ void local_sgs4(const Row4Coef<float> &c, const float id[4], const float r[4], float e[4],
                                           float *lds, int q) {
    const int li = q >> 4, lj = (q >> 1) & 7;
    const int base = 4 * q;
    const bool c1 = ((li + lj) & 1) != 0;  // col0
    // The lane's choice between its cells is made on loaded values (one select per field), never on
    // an array index: an index chosen per lane would keep the arrays in memory (LDS, addressed
    // through the dispatch packet), on the critical path of every launch.
    float gxm[4], gxp[4], gym[4], gyp[4], gzm[4], gzp[4], idv[4], rv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        gxm[j] = c.gxm[j];
        gxp[j] = c.gxp[j];
        gym[j] = c.gym[j];
        gyp[j] = c.gyp[j];
        gzm[j] = c.gzm[j];
        gzp[j] = c.gzp[j];
        idv[j] = id[j];
        rv[j] = r[j];
    }
    // cell ja when col0 = 0, jb when col0 = 1, with the z weights of each case
    auto cell_of = [&](int ja, float za_a, float zb_a, float zs_a, int jb, float za_b, float zb_b, float zs_b) {
        SgsCell o;
        o.ax_m = c1 ? gxm[jb] : gxm[ja];
        o.ax_p = c1 ? gxp[jb] : gxp[ja];
        o.ay_m = c1 ? gym[jb] : gym[ja];
        o.ay_p = c1 ? gyp[jb] : gyp[ja];
        o.za = c1 ? za_b : za_a;
        o.zb = c1 ? zb_b : zb_a;
        o.zs = c1 ? zs_b : zs_a;
        o.id = c1 ? idv[jb] : idv[ja];
        o.r = c1 ? rv[jb] : rv[ja];
        const int j = c1 ? jb : ja;
        o.st = base + j;
        o.xm = li > 0 ? base + j - 64 : base + j + 64;
        o.xp = li < 7 ? base + j + 64 : base + j - 64;
        return o;
    };
    // group P: cells 0, 2 (col0 = 0) or 1, 3 (col0 = 1); group Q: the others.  For col0 = 0:
    //   P0 = cell 0: z- partner's cell 3 (= its Q1), z+ cell 1 (Q0)
    //   P1 = cell 2: z- cell 1 (Q0), z+ cell 3 (Q1)
    //   Q0 = cell 1: z- cell 0 (P0), z+ cell 2 (P1)
    //   Q1 = cell 3: z- cell 2 (P1), z+ partner's cell 0 (= its P0)
    // and for col0 = 1:
    //   P0 = cell 1: z- cell 0 (Q0), z+ cell 2 (Q1)
    //   P1 = cell 3: z- cell 2 (Q1), z+ partner's cell 0 (= its Q0)
    //   Q0 = cell 0: z- partner's cell 3 (= its P1), z+ cell 1 (P0)
    //   Q1 = cell 2: z- cell 1 (P0), z+ cell 3 (P1)
    // P0 and P1 take the swap of Q1 and Q0, Q0 and Q1 the swap of P1 and P0.
    const SgsCell P0 = cell_of(0, gzp[0], 0.0f, gzm[0], 1, gzm[1], gzp[1], 0.0f);
    const SgsCell P1 = cell_of(2, gzm[2], gzp[2], 0.0f, 3, 0.0f, gzm[3], gzp[3]);
    const SgsCell Q0 = cell_of(1, gzm[1], gzp[1], 0.0f, 0, gzp[0], 0.0f, gzm[0]);
    const SgsCell Q1 = cell_of(3, 0.0f, gzm[3], gzp[3], 2, gzm[2], gzp[2], 0.0f);
    // the first half-sweep, from zero: no neighbour values yet
    float p0 = P0.id * P0.r, p1 = P1.id * P1.r, q0 = 0.0f, q1 = 0.0f;
    lds[P0.st] = p0;
    lds[P1.st] = p1;
    __syncthreads();
#pragma unroll
    for (int s = 1; s <= 2 * K; ++s) {
        if (s & 1) {
            const float ym0 = sgs_dpp<DPP_ROW_SHR2>(p0), yp0 = sgs_dpp<DPP_ROW_SHL2>(p0);
            const float ym1 = sgs_dpp<DPP_ROW_SHR2>(p1), yp1 = sgs_dpp<DPP_ROW_SHL2>(p1);
            const float sw0 = sgs_dpp<DPP_SWAP_PAIRS>(p0), sw1 = sgs_dpp<DPP_SWAP_PAIRS>(p1);
            q0 = sgs_update(Q0, lds, ym0, yp0, p0, p1, sw1);
            q1 = sgs_update(Q1, lds, ym1, yp1, p0, p1, sw0);
            lds[Q0.st] = q0;
            lds[Q1.st] = q1;
        } else {
            const float ym0 = sgs_dpp<DPP_ROW_SHR2>(q0), yp0 = sgs_dpp<DPP_ROW_SHL2>(q0);
            const float ym1 = sgs_dpp<DPP_ROW_SHR2>(q1), yp1 = sgs_dpp<DPP_ROW_SHL2>(q1);
            const float sw0 = sgs_dpp<DPP_SWAP_PAIRS>(q0), sw1 = sgs_dpp<DPP_SWAP_PAIRS>(q1);
            p0 = sgs_update(P0, lds, ym0, yp0, q0, q1, sw1);
            p1 = sgs_update(P1, lds, ym1, yp1, q0, q1, sw0);
            lds[P0.st] = p0;
            lds[P1.st] = p1;
        }
        __syncthreads();
    }
    e[0] = c1 ? q0 : p0;
    e[1] = c1 ? p0 : q0;
    e[2] = c1 ? q1 : p1;
    e[3] = c1 ? p1 : q1;
}

// Pre-smoothing from zero on the finest level: z = M_b rhs (one brick per 128-thread workgroup), K
// sweeps.
template <int K>
__global__ __launch_bounds__(ROW4_THREADS) void k_pre_local4(LevelDesc L, OpDesc<float> A,
                                                             const float *__restrict__ idiag,
                                                             const float *__restrict__ rhs, float *__restrict__ z,
                                                             Gate Q) {
    __shared__ float lds[BRICK_CELLS];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const int q = threadIdx.x;
    const Row4Nbrs N = row4_nbrs(m, br, q);
    float r[4], id[4], e[4];
    unpack4(ld4(rhs, N.c), r);
    Row4Coef<float> c;
    if (m.flags & BM_UNIFORM) {
        const UniT<float> u = uni_load<float>(L, br);
        c = row4_local_uniform(u.g, q);
#pragma unroll
        for (int j = 0; j < 4; ++j) id[j] = u.id;
    } else {
        c = row4_local(row4_coef(A, N), q);
        unpack4(ld4(idiag, N.c), id);
    }
    local_sgs4<K>(c, id, r, e, lds, q);
    *reinterpret_cast<float4 *>(z + N.c) = make_float4(e[0], e[1], e[2], e[3]);
}

// Post-smoothing (and further pre-smoothing) with the brick-local smoother: zout = zin + M_b (rhs - A zin);
// DOT: also the rhs.zout partials of the unknowns (the last post-smoothing pass yields r.z of the
// preconditioned CG).  One brick per 128-thread workgroup.  FINEST: 1 on the finest level, 0 on
// the coarse ones (the same code; distinct symbols for profiling).
template <int DOT, int K = LOCAL_SGS_SWEEPS, int FINEST = 1>
__global__ __launch_bounds__(ROW4_THREADS) void k_post_local4(LevelDesc L, OpDesc<float> A,
                                                              const float *__restrict__ idiag,
                                                              const float *__restrict__ rhs,
                                                              const float *__restrict__ zin, float *__restrict__ zout,
                                                              Gate Q, RedBuf R) {
    __shared__ float lds[BRICK_CELLS];
    __shared__ double red[2];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const int q = threadIdx.x;
    const Row4Nbrs N = row4_nbrs(m, br, q);
    const Row4Vals<float> v = row4_vals(zin, N);
    float r[4], id[4], e[4], res[4];
    unpack4(ld4(rhs, N.c), r);
    Row4Coef<float> c;
    bool unk[4];
    if (m.flags & BM_UNIFORM) {
        const UniT<float> u = uni_load<float>(L, br);
        const float diag = 0.0f + ((u.g + u.g) + (u.g + u.g) + (u.g + u.g));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float off = u.g * v.xm[j] + u.g * v.xp[j] + u.g * v.ym[j] + u.g * v.yp[j] + u.g * v.zm[j] +
                              u.g * v.zp[j];
            res[j] = r[j] - (diag * v.c[j] - off);
            id[j] = u.id;
            unk[j] = true;
        }
        c = row4_local_uniform(u.g, q);
    } else {
        const Row4Coef<float> g = row4_coef(A, N);
        float dir[4];
        load4(A.dir, N.c, dir);
        unpack4(ld4(idiag, N.c), id);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float off = g.gxm[j] * v.xm[j] + g.gxp[j] * v.xp[j] + g.gym[j] * v.ym[j] + g.gyp[j] * v.yp[j] +
                              g.gzm[j] * v.zm[j] + g.gzp[j] * v.zp[j];
            const float diag = dir[j] + ((g.gxm[j] + g.gxp[j]) + (g.gym[j] + g.gyp[j]) + (g.gzm[j] + g.gzp[j]));
            unk[j] = dir[j] >= 0;
            res[j] = unk[j] ? r[j] - (diag * v.c[j] - off) : 0.0f;
        }
        c = row4_local(g, q);
    }
    local_sgs4<K>(c, id, res, e, lds, q);
    float out[4];
    double part = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        out[j] = unk[j] ? v.c[j] + e[j] : 0.0f;
        if (DOT) part += unk[j] ? (double)r[j] * (double)out[j] : 0.0;
    }
    *reinterpret_cast<float4 *>(zout + N.c) = make_float4(out[0], out[1], out[2], out[3]);
    if (DOT) {
        part = wave_sum(part);
        if ((q & 63) == 0) red[q >> 6] = part;
        __syncthreads();
        if (q == 0) {
            R.partials[2 * br + 0] = red[0] + red[1];
            R.partials[2 * br + 1] = 0.0;
        }
    }
}

// (A v) on the lane's four cells, with the unknown mask: apply_row / apply_row_uniform per cell.
template <class T>
__device__ __forceinline__ void apply_row4(const LevelDesc &L, const OpDesc<T> &A, int br, const BrickMeta &m,
                                           const Row4Nbrs &N, const Row4Vals<T> &v, T av[4], bool unk[4]) {
    if (m.flags & BM_UNIFORM) {
        const UniT<T> u = uni_load<T>(L, br);
        const T diag = T(0) + ((u.g + u.g) + (u.g + u.g) + (u.g + u.g));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const T off = u.g * v.xm[j] + u.g * v.xp[j] + u.g * v.ym[j] + u.g * v.yp[j] + u.g * v.zm[j] + u.g * v.zp[j];
            av[j] = diag * v.c[j] - off;
            unk[j] = true;
        }
    } else {
        const Row4Coef<T> g = row4_coef(A, N);
        T dir[4];
        load4(A.dir, N.c, dir);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const T off = g.gxm[j] * v.xm[j] + g.gxp[j] * v.xp[j] + g.gym[j] * v.ym[j] + g.gyp[j] * v.yp[j] +
                          g.gzm[j] * v.zm[j] + g.gzp[j] * v.zp[j];
            const T diag = dir[j] + ((g.gxm[j] + g.gxp[j]) + (g.gym[j] + g.gyp[j]) + (g.gzm[j] + g.gzp[j]));
            av[j] = diag * v.c[j] - off;
            unk[j] = dir[j] >= 0;
        }
    }
}

// Brick and lane of a workgroup that takes BPB consecutive bricks (block ids remapped XCD-aware;
// gridDim.x a multiple of 8); br >= L.nbricks for an idle slot.  Row-4 launches give the finest
// level BLOCK threads (four bricks) and coarse levels ROW4_THREADS (one brick): the kernels' FINEST
// flag fixes BPB at compile time (blockDim.x is a 16-bit hidden kernel argument that hipcc reads
// with a vector load, one more memory round trip at the start of every wave).
template <int FINEST>
__device__ __forceinline__ void row4_slot(const LevelDesc &L, int &br, int &q) {
    constexpr int BPB = FINEST ? BLOCK / ROW4_THREADS : 1;
    br = BPB * xcd_brick(blockIdx.x, gridDim.x) + (int)(threadIdx.x / ROW4_THREADS);
    q = threadIdx.x % ROW4_THREADS;
}

// Sum / max over the two wavefronts of a 128-thread workgroup (result valid in every thread).
template <bool AMAX>
__device__ __forceinline__ void row4_reduce2(double &a, double &b, double *lds) {
    a = AMAX ? wave_max(a) : wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[w] = a;
        lds[2 + w] = b;
    }
    __syncthreads();
    a = AMAX ? fmax(lds[0], lds[1]) : lds[0] + lds[1];
    b = lds[2] + lds[3];
}

// res = rhs - A z (k_residual on rows).
template <int FINEST>
__global__ __launch_bounds__(BLOCK) void k_residual4(LevelDesc L, OpDesc<float> A, const float *__restrict__ rhs,
                                                     const float *__restrict__ z, float *__restrict__ res, Gate Q) {
    int br, q;
    row4_slot<FINEST>(L, br, q);
    if (br >= L.nbricks) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const Row4Nbrs N = row4_nbrs(m, br, q);
    const Row4Vals<float> v = row4_vals(z, N);
    float r[4], av[4], out[4];
    bool unk[4];
    unpack4(ld4(rhs, N.c), r);
    apply_row4(L, A, br, m, N, v, av, unk);
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = unk[j] ? r[j] - av[j] : 0.0f;
    *reinterpret_cast<float4 *>(res + N.c) = make_float4(out[0], out[1], out[2], out[3]);
}

// z = omega * idiag * rhs (k_smooth_first on rows).
template <int FINEST>
__global__ __launch_bounds__(BLOCK) void k_smooth_first4(LevelDesc L, const float *__restrict__ idiag,
                                                         const float *__restrict__ rhs, float *__restrict__ z,
                                                         float omega, Gate Q) {
    int br, q;
    row4_slot<FINEST>(L, br, q);
    if (br >= L.nbricks) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const long c = (long)br * BRICK_CELLS + 4 * q;
    float r[4], id[4], out[4];
    unpack4(ld4(rhs, c), r);
    if (m.flags & BM_UNIFORM) {
        const float u = uni_load<float>(L, br).id;
#pragma unroll
        for (int j = 0; j < 4; ++j) id[j] = u;
    } else {
        unpack4(ld4(idiag, c), id);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = omega * id[j] * r[j];
    *reinterpret_cast<float4 *>(z + c) = make_float4(out[0], out[1], out[2], out[3]);
}

// ---- The FP32 Krylov iteration of the mixed-precision solve.  Its vectors r, p, q, z are stored in
// FP32, but the iterate is the FP64 pressure itself (x64 += alpha p in FP64) and q = A p is formed
// from the FP64 operator in FP64 arithmetic, rounded once when stored.  Then the recursive
// residual follows b - A x64 to FP32 storage accuracy of r and q: neither the rounding of an FP32
// iterate nor an FP32 operator apply (whose cancellation amplifies the rounding by about the
// condition number, 1e4 and more at 256^3) enters it, so one pass reaches the final tolerance.

// (A v) in FP64 on the lane's four cells from FP32 values v and the FP64 operator (k_apply_dot's
// arithmetic in FP64); unk: the unknown mask.
__device__ __forceinline__ void apply_row4_64(const LevelDesc &L, const OpDesc<double> &A, int br, const BrickMeta &m,
                                              const Row4Nbrs &N, const Row4Vals<float> &v, double av[4],
                                              bool unk[4]) {
    if (m.flags & BM_UNIFORM) {
        const double g = L.uni[br].g64;
        const double diag = 0.0 + ((g + g) + (g + g) + (g + g));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double off = g * (double)v.xm[j] + g * (double)v.xp[j] + g * (double)v.ym[j] + g * (double)v.yp[j] +
                               g * (double)v.zm[j] + g * (double)v.zp[j];
            av[j] = diag * (double)v.c[j] - off;
            unk[j] = true;
        }
    } else {
        const Row4Coef<double> g = row4_coef(A, N);
        double dir[4];
        load4(A.dir, N.c, dir);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double off = g.gxm[j] * (double)v.xm[j] + g.gxp[j] * (double)v.xp[j] + g.gym[j] * (double)v.ym[j] +
                               g.gyp[j] * (double)v.yp[j] + g.gzm[j] * (double)v.zm[j] + g.gzp[j] * (double)v.zp[j];
            const double diag = dir[j] + ((g.gxm[j] + g.gxp[j]) + (g.gym[j] + g.gyp[j]) + (g.gzm[j] + g.gzp[j]));
            av[j] = diag * (double)v.c[j] - off;
            unk[j] = dir[j] >= 0;
        }
    }
}

// q = A p (FP64 operator and arithmetic, q stored in FP32) with the p.q partials (FP64 q): one
// brick per 128-thread workgroup.
__global__ __launch_bounds__(ROW4_THREADS) void k_apply_dot4(LevelDesc L, OpDesc<double> A, const float *__restrict__ p,
                                                             float *__restrict__ qv, const SolveScalars *S,
                                                             const BatchCtl *ctl, RedBuf R) {
    __shared__ double lds[4];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    // a batch skips finished grids before loading anything; a single grid loads first
    if (L.ngrids > 1 && (ctl->all_done || S[m.grid].done)) return;
    const Row4Nbrs N = row4_nbrs(m, br, threadIdx.x);
    const Row4Vals<float> v = row4_vals(p, N);
    if (ctl->all_done) return;
    double av[4];
    float out[4];
    bool unk[4];
    apply_row4_64(L, A, br, m, N, v, av, unk);
    double part = 0.0, unused = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        out[j] = unk[j] ? (float)av[j] : 0.0f;
        part += unk[j] ? (double)v.c[j] * av[j] : 0.0;
    }
    *reinterpret_cast<float4 *>(qv + N.c) = make_float4(out[0], out[1], out[2], out[3]);
    row4_reduce2<false>(part, unused, lds);
    store_partials(R, br, part, 0.0);
}

// k_update_p4 (p = z + beta p, FP32) fused into k_apply_dot4: the new direction of the lane's cells
// and of their neighbours is formed on the fly from z and the previous direction (the same FP32
// arithmetic), written to pnew (a different buffer than pold: other workgroups still read pold),
// and q = A pnew with the pnew.q partials.
__global__ __launch_bounds__(ROW4_THREADS) void k_apply_dot4p(LevelDesc L, OpDesc<double> A,
                                                              const float *__restrict__ z,
                                                              const float *__restrict__ pold,
                                                              float *__restrict__ pnew, float *__restrict__ qv,
                                                              const SolveScalars *S, const BatchCtl *ctl, RedBuf R) {
    __shared__ double lds[4];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    const int g = L.ngrids > 1 ? m.grid : 0;
    if (L.ngrids > 1 && (ctl->all_done || S[g].done)) return;
    const Row4Nbrs N = row4_nbrs(m, br, threadIdx.x);
    const Row4Vals<float> zv = row4_vals(z, N);
    const Row4Vals<float> po = row4_vals(pold, N);
    if (ctl->all_done) return;
    const float beta = (float)S[g].beta;
    Row4Vals<float> v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v.c[j] = zv.c[j] + beta * po.c[j];
        v.xm[j] = zv.xm[j] + beta * po.xm[j];
        v.xp[j] = zv.xp[j] + beta * po.xp[j];
        v.ym[j] = zv.ym[j] + beta * po.ym[j];
        v.yp[j] = zv.yp[j] + beta * po.yp[j];
        v.zm[j] = zv.zm[j] + beta * po.zm[j];
        v.zp[j] = zv.zp[j] + beta * po.zp[j];
    }
    double av[4];
    float out[4];
    bool unk[4];
    apply_row4_64(L, A, br, m, N, v, av, unk);
    double part = 0.0, unused = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        out[j] = unk[j] ? (float)av[j] : 0.0f;
        part += unk[j] ? (double)v.c[j] * av[j] : 0.0;
    }
    *reinterpret_cast<float4 *>(pnew + N.c) = make_float4(v.c[0], v.c[1], v.c[2], v.c[3]);
    *reinterpret_cast<float4 *>(qv + N.c) = make_float4(out[0], out[1], out[2], out[3]);
    row4_reduce2<false>(part, unused, lds);
    store_partials(R, br, part, 0.0);
}

// x64 += alpha p (FP64), r -= alpha q (FP32) with ||r||_inf and r.r partials; ZFIRST: also the first
// pre-smoothing step of the next V-cycle from zero on the finest level, on the bricks it covers
// (not those of grids ending on this level), with the same arithmetic: 1 = a damped Jacobi sweep
// (k_smooth_first4<1>), z = omega * idiag * r; 2 = the brick-local smoother (k_pre_local4<K>),
// whose brick is the workgroup's.  A grid that this update finds converged gets a z it never uses.
template <int ZFIRST, int K = LOCAL_SGS_SWEEPS>
__global__ __launch_bounds__(ROW4_THREADS) void k_cg_update4(LevelDesc L, double *__restrict__ x,
                                                             float *__restrict__ r, const float *__restrict__ p,
                                                             const float *__restrict__ qv, const SolveScalars *S,
                                                             const BatchCtl *ctl, RedBuf R,
                                                             const float *__restrict__ idiag, float *__restrict__ z,
                                                             float omega, OpDesc<float> A32) {
    __shared__ double lds[4];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const long c = (long)br * BRICK_CELLS + 4 * threadIdx.x;
    const BrickMeta m = L.meta[br];
    const int g = L.ngrids > 1 ? m.grid : 0;
    if (L.ngrids > 1 && (ctl->all_done || S[g].done)) return;
    double xv[4];
    float pv[4], rv[4], qq[4], id[4];
    load4(x, c, xv);
    unpack4(ld4(p, c), pv);
    unpack4(ld4(r, c), rv);
    unpack4(ld4(qv, c), qq);
    Row4Coef<float> cf;  // ZFIRST == 2: the brick's own couplings
    if (ZFIRST) {
        if (m.flags & BM_UNIFORM) {
            const UniT<float> u = uni_load<float>(L, br);
#pragma unroll
            for (int j = 0; j < 4; ++j) id[j] = u.id;
            if (ZFIRST == 2) cf = row4_local_uniform(u.g, threadIdx.x);
        } else {
            unpack4(ld4(idiag, c), id);
            if (ZFIRST == 2) cf = row4_local(row4_coef(A32, row4_nbrs(m, br, threadIdx.x)), threadIdx.x);
        }
    }
    const SolveScalars &Sg = S[g];
    if (ctl->all_done || Sg.done) return;
    const double alpha = Sg.alpha;
    const float alpha32 = (float)alpha;
    double mx = 0.0, sq = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        xv[j] = xv[j] + alpha * (double)pv[j];
        rv[j] = rv[j] - alpha32 * qq[j];
        mx = fmax(mx, fabs((double)rv[j]));
        sq += (double)rv[j] * (double)rv[j];
    }
    store4(x, c, xv);
    *reinterpret_cast<float4 *>(r + c) = make_float4(rv[0], rv[1], rv[2], rv[3]);
    if (ZFIRST == 1 && !(m.flags & BM_COARSEST)) {
        float out[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) out[j] = omega * id[j] * rv[j];
        *reinterpret_cast<float4 *>(z + c) = make_float4(out[0], out[1], out[2], out[3]);
    }
    if (ZFIRST == 2 && !(m.flags & BM_COARSEST)) {
        // the brick-local smoother's pre-smoothing from zero (k_pre_local4) on the updated residual
        __shared__ float zl[BRICK_CELLS];
        float e[4];
        local_sgs4<K>(cf, id, rv, e, zl, threadIdx.x);
        *reinterpret_cast<float4 *>(z + c) = make_float4(e[0], e[1], e[2], e[3]);
    }
    row4_reduce2<true>(mx, sq, lds);
    store_partials(R, br, mx, sq);
}

// p = z + beta p (first: p = z), k_update_p on rows.
__global__ __launch_bounds__(BLOCK) void k_update_p4(LevelDesc L, const float *__restrict__ z, float *__restrict__ p,
                                                     const SolveScalars *S, const BatchCtl *ctl, int first) {
    int br, q;
    row4_slot<1>(L, br, q);  // the finest level only (row-4 launch of BLOCK threads)
    if (br >= L.nbricks) return;
    const long c = (long)br * BRICK_CELLS + 4 * q;
    if (L.ngrids > 1 && (ctl->all_done || S[grid_of(L, br)].done)) return;
    float zv[4], pv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    unpack4(ld4(z, c), zv);
    if (!first) unpack4(ld4(p, c), pv);
    const SolveScalars &Sg = S[grid_of(L, br)];
    if (ctl->all_done || Sg.done) return;
    const float beta = (float)Sg.beta;
#pragma unroll
    for (int j = 0; j < 4; ++j) pv[j] = first ? zv[j] : zv[j] + beta * pv[j];
    *reinterpret_cast<float4 *>(p + c) = make_float4(pv[0], pv[1], pv[2], pv[3]);
}

// r = b - A x in FP64 with per-brick ||r||_inf and ||r||^2 partials (finalized by FIN_DEFECT*):
// one brick per 128-thread workgroup, four cells per thread (two 16-byte loads per array).
__global__ __launch_bounds__(ROW4_THREADS) void k_defect4(LevelDesc L, OpDesc<double> A, const double *__restrict__ b,
                                                          const double *__restrict__ x, double *__restrict__ r,
                                                          RedBuf R) {
    __shared__ double lds[4];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    const Row4Nbrs N = row4_nbrs(m, br, threadIdx.x);
    const Row4Vals<double> v = row4_vals(x, N);
    double bv[4], av[4], out[4];
    bool unk[4];
    load4(b, N.c, bv);
    apply_row4(L, A, br, m, N, v, av, unk);
    double mx = 0.0, sq = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        out[j] = unk[j] ? bv[j] - av[j] : 0.0;
        mx = fmax(mx, fabs(out[j]));
        sq += out[j] * out[j];
    }
    store4(r, N.c, out);
    row4_reduce2<true>(mx, sq, lds);
    store_partials(R, br, mx, sq);
}

// Inner (FP32) solve set-up of the defect correction: r32 = (float) r64 with ||r32|| partials
// (finalized by FIN_INNER_INIT); four cells per thread.  The inner solve adds its correction to the
// FP64 iterate directly (k_cg_update4).
__global__ __launch_bounds__(ROW4_THREADS) void k_inner_init4(LevelDesc L, const double *__restrict__ r64,
                                                              float *__restrict__ r32, RedBuf R) {
    __shared__ double lds[4];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const long c = (long)br * BRICK_CELLS + 4 * threadIdx.x;
    double rv[4];
    float v[4];
    load4(r64, c, rv);
    double mx = 0.0, sq = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = (float)rv[j];
        mx = fmax(mx, fabs((double)v[j]));
        sq += (double)v[j] * (double)v[j];
    }
    store4(r32, c, v);
    row4_reduce2<true>(mx, sq, lds);
    store_partials(R, br, mx, sq);
}

// res = rhs - A z
template <class T, int FINEST>
__global__ __launch_bounds__(BLOCK) void k_residual(LevelDesc L, OpDesc<T> A, const T *__restrict__ rhs,
                                                    const T *__restrict__ z, T *__restrict__ res, Gate Q) {
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    T out = 0;
    if (m.flags & BM_UNIFORM) {
        const UniT<T> u = uni_load<T>(L, br);
        long idx[6];
        nb_index(nbrs_of(m, br), t, idx);
        out = rhs[c] - apply_row_uniform(u.g, z, c, idx, z[c]);
    } else {
        // every load is issued before the unknown test (the values of other cells are discarded)
        const Nbrs N = nbrs_of(m, br);
        long idx[6];
        nb_index(N, t, idx);
        const T dirc = A.dir[c];
        T diag;
        const T v = rhs[c] - apply_row(A, z, c, idx, dirc, z[c], diag);
        out = dirc >= 0 ? v : T(0);
    }
    res[c] = out;
}

// Restriction = transpose of trilinear prolongation: coarse cell I gathers fine cells
// 2I-1..2I+2 per axis with weights 1/4, 3/4, 3/4, 1/4.  One block per coarse brick B stages the
// 18^3 fine residuals it needs (fine cells 16B-1 .. 16B+16 per axis, in the fine bricks 2B-1 ..
// 2B+2 that the brick's child table lists) in LDS.
// M_b r for one cell per thread of a 512-thread workgroup holding one brick (cell t, storage
// order): K symmetric red-black Gauss-Seidel sweeps from zero on the brick's own operator (c: the
// cell's couplings to cells of the brick, 0 across its faces), as local_sgs4.  lds: 512 floats.
template <int K>
__device__ __forceinline__ float local_sgs1(const float c[6], float id, float r, float *lds, int t) {
    const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
    const int colour = (li + lj + lk) & 1;
    float e = colour == 0 ? id * r : 0.0f;
    if (colour == 0) lds[t] = e;
    __syncthreads();
#pragma unroll
    for (int s = 1; s <= 2 * K; ++s) {
        if (colour == (s & 1)) {
            const float xm = li > 0 ? lds[t - 64] : 0.0f, xp = li < 7 ? lds[t + 64] : 0.0f;
            const float ym = lj > 0 ? lds[t - 8] : 0.0f, yp = lj < 7 ? lds[t + 8] : 0.0f;
            const float zm = lk > 0 ? lds[t - 1] : 0.0f, zp = lk < 7 ? lds[t + 1] : 0.0f;
            const float nb = c[0] * xm + c[1] * xp + c[2] * ym + c[3] * yp + c[4] * zm + c[5] * zp;
            e = id * (r + nb);
            lds[t] = e;
        }
        __syncthreads();
    }
    return e;
}

// k_restrict (FP32) that also runs the coarse level's pre-smoothing from zero with the brick-local
// smoother: zc = M_b rhs_c on the coarse brick (skipped on bricks of grids that end on the coarse
// level: their coarse solve starts from the right-hand side).
template <int FINEST>
__global__ __launch_bounds__(BLOCK) void k_restrict_pre(LevelDesc Lc, OpDesc<float> Ac, const float *__restrict__ idc,
                                                        const float *__restrict__ res, float *__restrict__ rhs_c,
                                                        float *__restrict__ zc, Gate Q) {
    constexpr int E = 18;
    __shared__ float tile[E * E * E];
    __shared__ int bid[64];
    const int br = brick_or_idle(Lc);
    if (br < 0) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    const float dirc = Ac.dir[c];
    const int cid = t < 64 ? Lc.child64[64 * br + t] : -1;
    const BrickMeta m = Lc.meta[br];
    if (all_done(Q) || grid_done(Q, Lc, m)) return;
    if (t < 64) bid[t] = cid;
    __syncthreads();
    constexpr int NQ = (E * E * E + BLOCK - 1) / BLOCK;
    float v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + q * BLOCK;
        v[q] = 0.0f;
        if (e < E * E * E) {
            const int c3 = e % E, b = (e / E) % E, a = e / (E * E);
            const int slot = (((a + 7) >> 3) * 4 + ((b + 7) >> 3)) * 4 + ((c3 + 7) >> 3);
            const int id = bid[slot];
            if (id >= 0) v[q] = res[(long)id * BRICK_CELLS + ((((a + 7) & 7) * 8 + ((b + 7) & 7)) * 8 + ((c3 + 7) & 7))];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + q * BLOCK;
        if (e < E * E * E) tile[e] = v[q];
    }
    __syncthreads();
    const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
    float out = 0.0f;
    if (dirc >= 0) {
        const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        const float wk[4] = {Lc.flat ? 0.0f : w[0], Lc.flat ? 1.0f : w[1], Lc.flat ? 0.0f : w[2],
                             Lc.flat ? 0.0f : w[3]};
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float acc_a = 0.0f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float *row = &tile[((2 * li + a) * E + (2 * lj + b)) * E + 2 * lk];
                const float acc_b = wk[0] * row[0] + wk[1] * row[1] + wk[2] * row[2] + wk[3] * row[3];
                acc_a += w[b] * acc_b;
            }
            acc += w[a] * acc_a;
        }
        out = acc;
    }
    rhs_c[c] = out;
    if (m.flags & BM_COARSEST) return;  // the whole block: the flag is the brick's
    __syncthreads();                     // the tile becomes the smoother's scratch
    float cf[6];
    cf[0] = li > 0 ? Ac.gx[c] : 0.0f;
    cf[1] = li < 7 ? Ac.gx[c + 64] : 0.0f;
    cf[2] = lj > 0 ? Ac.gy[c] : 0.0f;
    cf[3] = lj < 7 ? Ac.gy[c + 8] : 0.0f;
    cf[4] = lk > 0 ? Ac.gz[c] : 0.0f;
    cf[5] = lk < 7 ? Ac.gz[c + 1] : 0.0f;
    const float e = local_sgs1<LOCAL_SGS_SWEEPS_COARSE>(cf, idc[c], out, tile, t);
    zc[c] = e;
}

template <class T, int FINEST>
__global__ __launch_bounds__(BLOCK) void k_restrict(LevelDesc Lc, OpDesc<T> Ac, const T *__restrict__ res,
                                                    T *__restrict__ rhs_c, Gate Q) {
    constexpr int E = 18;
    __shared__ T tile[E * E * E];
    __shared__ int bid[64];
    const int br = brick_or_idle(Lc);
    if (br < 0) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    // loads that depend on the brick only, issued before the gate's loads resolve
    const T dirc = Ac.dir[c];
    const int cid = t < 64 ? Lc.child64[64 * br + t] : -1;
    const BrickMeta m = Lc.meta[br];
    if (all_done(Q) || grid_done(Q, Lc, m)) return;
    if (t < 64) bid[t] = cid;
    __syncthreads();
    // all NQ loads of a thread are issued before the first LDS store (one memory latency, not NQ)
    constexpr int NQ = (E * E * E + BLOCK - 1) / BLOCK;
    T v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + q * BLOCK;
        v[q] = T(0);
        if (e < E * E * E) {
            const int c3 = e % E, b = (e / E) % E, a = e / (E * E);
            // tile offset a is fine cell 16B - 1 + a: fine brick 2B - 1 + ((a + 7) >> 3), local (a + 7) & 7
            const int slot = (((a + 7) >> 3) * 4 + ((b + 7) >> 3)) * 4 + ((c3 + 7) >> 3);
            const int id = bid[slot];
            if (id >= 0) v[q] = res[(long)id * BRICK_CELLS + ((((a + 7) & 7) * 8 + ((b + 7) & 7)) * 8 + ((c3 + 7) & 7))];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + q * BLOCK;
        if (e < E * E * E) tile[e] = v[q];
    }
    __syncthreads();
    T out = 0;
    if (dirc >= 0) {
        const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
        const T w[4] = {T(0.25), T(0.75), T(0.75), T(0.25)};
        // a flat (2D) level has one cell layer: the transfer is bilinear, weight 1 on that layer
        const T wk[4] = {Lc.flat ? T(0) : w[0], Lc.flat ? T(1) : w[1], Lc.flat ? T(0) : w[2], Lc.flat ? T(0) : w[3]};
        T acc = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            T acc_a = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const T *row = &tile[((2 * li + a) * E + (2 * lj + b)) * E + 2 * lk];
                const T acc_b = wk[0] * row[0] + wk[1] * row[1] + wk[2] * row[2] + wk[3] * row[3];
                acc_a += w[b] * acc_b;
            }
            acc += w[a] * acc_a;
        }
        out = acc;
    }
    rhs_c[c] = out;
}

// z_f += trilinear interpolation of the coarse correction (fine unknowns only).  One block per
// fine brick B stages the 6^3 coarse values it needs (coarse cells 4B-1 .. 4B+4 per axis, in the
// coarse bricks its parent table lists) in LDS.  Coarse cells outside the coarse grid read 0: they
// lie in unstored bricks or are non-unknowns, whose correction every V-cycle kernel keeps at 0.
template <class T, int FINEST>
__global__ __launch_bounds__(BLOCK) void k_prolong(LevelDesc Lf, OpDesc<T> Af, const T *__restrict__ zc,
                                                   T *__restrict__ zf, Gate Q) {
    constexpr int E = 6;
    __shared__ T tile[E * E * E];
    __shared__ int bid[8];
    const int br = brick_or_idle(Lf);
    if (br < 0) return;
    const int t = threadIdx.x;
    const long c = (long)br * BRICK_CELLS + t;
    // loads that depend on the brick only, issued before the gate's loads resolve
    const T zf0 = zf[c];
    const int pid = t < 8 ? Lf.parent8[8 * br + t] : -1;
    // with B = 2M + p, tile offset a is coarse cell 8M + 4p - 1 + a: coarse brick
    // M - 1 + p + slot with slot = ((4p + 7 + a) >> 3) - p, local (4p + 7 + a) & 7
    const int px = Lf.bxyz[3 * br] & 1, py = Lf.bxyz[3 * br + 1] & 1, pz = Lf.bxyz[3 * br + 2] & 1;
    const BrickMeta m = Lf.meta[br];
    if (all_done(Q) || skip_smooth(Q, Lf, m)) return;
    const bool upd = (m.flags & BM_UNIFORM) || Af.dir[c] >= 0;
    if (t < 8) bid[t] = pid;
    __syncthreads();
    if (t < E * E * E) {
        const int c = t % E, b = (t / E) % E, a = t / (E * E);
        const int ia = 4 * px + 7 + a, ib = 4 * py + 7 + b, ic = 4 * pz + 7 + c;
        const int slot = (((ia >> 3) - px) * 2 + ((ib >> 3) - py)) * 2 + ((ic >> 3) - pz);
        const int id = bid[slot];
        tile[t] = id >= 0 ? zc[(long)id * BRICK_CELLS + (((ia & 7) * 8 + (ib & 7)) * 8 + (ic & 7))] : T(0);
    }
    __syncthreads();
    if (!upd) return;
    const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
    const int a0 = (li >> 1) + 1, b0 = (lj >> 1) + 1, c0 = (lk >> 1) + 1;
    const int a1 = (li & 1) ? a0 + 1 : a0 - 1, b1 = (lj & 1) ? b0 + 1 : b0 - 1, c1 = (lk & 1) ? c0 + 1 : c0 - 1;
    auto at = [&](int a, int b, int cc) { return tile[(a * E + b) * E + cc]; };
    const T n = T(0.75), f = T(0.25);
    // a flat (2D) level has one cell layer: weight 1 on it, none on the empty neighbour layer
    const T nk = Lf.flat ? T(1) : n, fk = Lf.flat ? T(0) : f;
    const T v = n * (n * (nk * at(a0, b0, c0) + fk * at(a0, b0, c1)) + f * (nk * at(a0, b1, c0) + fk * at(a0, b1, c1))) +
                f * (n * (nk * at(a1, b0, c0) + fk * at(a1, b0, c1)) + f * (nk * at(a1, b1, c0) + fk * at(a1, b1, c1)));
    zf[c] = zf0 + v;
}

// k_prolong on cell rows (FP32): 512-thread workgroups take four fine bricks, each staging its own
// 6^3 coarse tile; lane q adds the correction to its four cells with k_prolong's arithmetic (cells
// that are not unknowns are written back unchanged).
template <int FINEST>
__global__ __launch_bounds__(BLOCK) void k_prolong4(LevelDesc Lf, OpDesc<float> Af, const float *__restrict__ zc,
                                                    float *__restrict__ zf, Gate Q) {
    constexpr int E = 6, NT = E * E * E;
    __shared__ float tile[BLOCK / ROW4_THREADS][NT];
    __shared__ int bid[BLOCK / ROW4_THREADS][8];
    if (all_done(Q)) return;
    int br, q;
    row4_slot<FINEST>(Lf, br, q);
    const int slot = threadIdx.x / ROW4_THREADS;
    // every thread reaches the barriers; a slot past the end or of a skipped brick only idles
    bool work = br < Lf.nbricks;
    BrickMeta m{};
    if (work) {
        m = Lf.meta[br];
        work = !skip_smooth(Q, Lf, m);
    }
    const long c = (long)br * BRICK_CELLS + 4 * q;
    float z0[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int px = 0, py = 0, pz = 0;
    if (work) {
        unpack4(ld4(zf, c), z0);
        px = Lf.bxyz[3 * br] & 1;
        py = Lf.bxyz[3 * br + 1] & 1;
        pz = Lf.bxyz[3 * br + 2] & 1;
        if (q < 8) bid[slot][q] = Lf.parent8[8 * br + q];
    }
    __syncthreads();
    // with B = 2M + p, tile offset a is coarse cell 8M + 4p - 1 + a: coarse brick
    // M - 1 + p + slot with slot = ((4p + 7 + a) >> 3) - p, local (4p + 7 + a) & 7
    if (work) {
        float v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = q + u * ROW4_THREADS;
            v[u] = 0.0f;
            if (e < NT) {
                const int cc = e % E, b = (e / E) % E, a = e / (E * E);
                const int ia = 4 * px + 7 + a, ib = 4 * py + 7 + b, ic = 4 * pz + 7 + cc;
                const int s8 = (((ia >> 3) - px) * 2 + ((ib >> 3) - py)) * 2 + ((ic >> 3) - pz);
                const int id = bid[slot][s8];
                if (id >= 0) v[u] = zc[(long)id * BRICK_CELLS + (((ia & 7) * 8 + (ib & 7)) * 8 + (ic & 7))];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = q + u * ROW4_THREADS;
            if (e < NT) tile[slot][e] = v[u];
        }
    }
    __syncthreads();
    if (!work) return;
    float dir[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool uni = m.flags & BM_UNIFORM;
    if (!uni) unpack4(ld4(Af.dir, c), dir);
    const int li = q >> 4, lj = (q >> 1) & 7, h = q & 1;
    const int a0 = (li >> 1) + 1, b0 = (lj >> 1) + 1;
    const int a1 = (li & 1) ? a0 + 1 : a0 - 1, b1 = (lj & 1) ? b0 + 1 : b0 - 1;
    auto at = [&](int a, int b, int cc) { return tile[slot][(a * E + b) * E + cc]; };
    const float n = 0.75f, f = 0.25f;
    // a flat (2D) level has one cell layer: weight 1 on it, none on the empty neighbour layer
    const float nk = Lf.flat ? 1.0f : n, fk = Lf.flat ? 0.0f : f;
    float out[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int lk = 4 * h + j;
        const int c0 = (lk >> 1) + 1, c1 = (lk & 1) ? c0 + 1 : c0 - 1;
        const float v = n * (n * (nk * at(a0, b0, c0) + fk * at(a0, b0, c1)) + f * (nk * at(a0, b1, c0) + fk * at(a0, b1, c1))) +
                        f * (n * (nk * at(a1, b0, c0) + fk * at(a1, b0, c1)) + f * (nk * at(a1, b1, c0) + fk * at(a1, b1, c1)));
        out[j] = (uni || dir[j] >= 0) ? z0[j] + v : z0[j];
    }
    *reinterpret_cast<float4 *>(zf + c) = make_float4(out[0], out[1], out[2], out[3]);
}

// K damped-Jacobi sweeps in one launch (coarse levels, where every launch is latency-bound).  The
// block of brick B stages the cube of cells within distance K of B (bricks B-1 .. B+1 per axis,
// found through the grid's brick map) and runs all K sweeps on it in LDS; values near the cube's
// faces go stale by one cell per sweep, which after K sweeps has not reached B, so B's cells are
// exactly K sweeps of k_jacobi (same arithmetic).  FROM_ZERO: the first sweep starts from z = 0
// (pre-smoothing), else from zin.  1024 threads; the couplings sit in LDS, idiag and rhs in
// registers.
template <class T, int K>
struct SmoothK {
    static constexpr int R = BRICK + 2 * K;          // cube side
    static constexpr int NR = R * R * R;
    static constexpr int THREADS = 1024;
    static constexpr int CPT = (NR + THREADS - 1) / THREADS;  // cube cells per thread
};

//
// RES: the cube carries one more ghost layer, so that after the K sweeps the brick's neighbours are
// exact too, and the kernel also writes the residual rhs - A z of the brick (what k_residual
// computes next, in the same arithmetic): one launch fewer per V-cycle on the level.
//
// PRO (post-smoothing, with zin): the level's z first receives the trilinear prolongation of the
// next coarser level's correction zcoarse, as k_prolong adds it (same arithmetic), on every cell
// of the cube: one launch fewer per V-cycle on the level.  The block stages the coarse cells the
// cube needs, in the (up to) 2x2x2 coarse bricks of its parent table.
template <class T, int K>
struct ProTile {
    static constexpr int R = BRICK + 2 * K;
    static constexpr int CT = R / 2 + 2;  // coarse cells per axis
    static constexpr int N = CT * CT * CT;
};

template <class T, int K, int FROM_ZERO, int RES = 0, int PRO = 0>
__global__ __launch_bounds__(1024) void k_smooth_k(LevelDesc L, OpDesc<T> A, const T *__restrict__ idiag,
                                                   const T *__restrict__ rhs, const T *__restrict__ zin,
                                                   T *__restrict__ zout, T omega, Gate Q, T *__restrict__ res,
                                                   const T *__restrict__ zcoarse) {
    static_assert(!(PRO && (FROM_ZERO || RES)), "the prolongation joins post-smoothing only");
    constexpr int GH = K + RES;  // ghost cells per side
    using C = SmoothK<T, GH>;
    constexpr int R = C::R, NR = C::NR, CPT = C::CPT;
    using PT = ProTile<T, GH>;
    __shared__ T zs[2][NR];
    __shared__ T gxs[NR], gys[NR], gzs[NR];
    __shared__ T ctile[PRO ? PT::N : 1];
    __shared__ int nb27[27];
    __shared__ int pb[8];
    const int br = brick_or_idle(L);
    if (br < 0) return;
    const int t = threadIdx.x;
    // loads that depend on the brick only, issued before the gate's loads resolve
    const int nid = t < 27 ? L.nb27[27 * br + t] : -1;
    const int pid = (PRO && t < 8) ? L.parent8[8 * br + t] : -1;
    int px = 0, py = 0, pz = 0;
    if (PRO) {
        px = L.bxyz[3 * br] & 1;
        py = L.bxyz[3 * br + 1] & 1;
        pz = L.bxyz[3 * br + 2] & 1;
    }
    const BrickMeta m = L.meta[br];
    if (all_done(Q) || skip_smooth(Q, L, m)) return;
    if (t < 27) nb27[t] = nid;
    if (PRO && t < 8) pb[t] = pid;
    __syncthreads();
    // Coarse cells are indexed relative to coarse cell 8M - 8 of the brick B = 2M + p (per axis):
    // cube offset a is fine cell 16M + 8p + a - GH, i.e. fi = 16 + 8p + a - GH relative to fine
    // cell 16M - 16; its coarse cells are fi >> 1 (weight 3/4) and the one beside it toward fi's
    // parity side (1/4).  The tile starts at coarse index cmin = ((16 + 8p - GH) >> 1) - 1; coarse
    // index I lies in parent slot (I >> 3) - p, local cell I & 7.
    const int cmx = ((16 + 8 * px - GH) >> 1) - 1, cmy = ((16 + 8 * py - GH) >> 1) - 1,
              cmz = ((16 + 8 * pz - GH) >> 1) - 1;
    if (PRO) {
        constexpr int CT = PT::CT, NQ = (PT::N + C::THREADS - 1) / C::THREADS;
        T cv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + q * C::THREADS;
            cv[q] = T(0);
            if (e < PT::N) {
                const int tc = e % CT, tb = (e / CT) % CT, ta = e / (CT * CT);
                const int Ia = cmx + ta, Ib = cmy + tb, Ic = cmz + tc;
                const int slot = (((Ia >> 3) - px) * 2 + ((Ib >> 3) - py)) * 2 + ((Ic >> 3) - pz);
                const int id = pb[slot];
                if (id >= 0) cv[q] = zcoarse[(long)id * BRICK_CELLS + (((Ia & 7) * 8 + (Ib & 7)) * 8 + (Ic & 7))];
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + q * C::THREADS;
            if (e < PT::N) ctile[e] = cv[q];
        }
        __syncthreads();
    }
    T rid[CPT], rr[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int e = t + q * C::THREADS;
        rid[q] = 0;
        rr[q] = 0;
        if (e >= NR) continue;
        const int a = e / (R * R), b = (e / R) % R, c = e % R;
        // cube offset a is brick-local cell a - GH, in brick (-1, 0 or +1) + local (0..7)
        const int xa = a - GH + BRICK, xb = b - GH + BRICK, xc = c - GH + BRICK;  // 0 .. 23
        const int bi = xa >> 3, bj = xb >> 3, bk = xc >> 3;
        const int id = nb27[(bi * 3 + bj) * 3 + bk];
        T g0 = 0, g1 = 0, g2 = 0, z0 = 0;
        if (id >= 0) {
            const long cell = (long)id * BRICK_CELLS + (((xa & 7) * 8 + (xb & 7)) * 8 + (xc & 7));
            rid[q] = idiag[cell];
            rr[q] = rhs[cell];
            g0 = A.gx[cell];
            g1 = A.gy[cell];
            g2 = A.gz[cell];
            if (!FROM_ZERO) z0 = zin[cell];
            if (PRO && A.dir[cell] >= 0) {
                // k_prolong's correction of this cell
                constexpr int CT = PT::CT;
                const int fa = 16 + 8 * px + a - GH, fb = 16 + 8 * py + b - GH, fc = 16 + 8 * pz + c - GH;
                const int a0 = (fa >> 1) - cmx, b0 = (fb >> 1) - cmy, c0 = (fc >> 1) - cmz;
                const int a1 = (fa & 1) ? a0 + 1 : a0 - 1, b1 = (fb & 1) ? b0 + 1 : b0 - 1,
                          c1 = (fc & 1) ? c0 + 1 : c0 - 1;
                auto at = [&](int i, int j, int k) { return ctile[(i * CT + j) * CT + k]; };
                const T n = T(0.75), f = T(0.25);
                const T nk = L.flat ? T(1) : n, fk = L.flat ? T(0) : f;
                const T v = n * (n * (nk * at(a0, b0, c0) + fk * at(a0, b0, c1)) +
                                 f * (nk * at(a0, b1, c0) + fk * at(a0, b1, c1))) +
                            f * (n * (nk * at(a1, b0, c0) + fk * at(a1, b0, c1)) +
                                 f * (nk * at(a1, b1, c0) + fk * at(a1, b1, c1)));
                z0 = z0 + v;
            }
        }
        gxs[e] = g0;
        gys[e] = g1;
        gzs[e] = g2;
        zs[0][e] = z0;
    }
    __syncthreads();
    // each thread keeps its cells' six couplings and own value in registers: a sweep reads only
    // the six neighbour values from LDS
    T gxm[CPT], gxp[CPT], gym[CPT], gyp[CPT], gzm[CPT], gzp[CPT], zr[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int e = t + q * C::THREADS;
        gxm[q] = gxp[q] = gym[q] = gyp[q] = gzm[q] = gzp[q] = zr[q] = T(0);
        if (e >= NR) continue;
        const int a = e / (R * R), b = (e / R) % R, c = e % R;
        gxm[q] = gxs[e];
        gym[q] = gys[e];
        gzm[q] = gzs[e];
        gxp[q] = a < R - 1 ? gxs[e + R * R] : T(0);
        gyp[q] = b < R - 1 ? gys[e + R] : T(0);
        gzp[q] = c < R - 1 ? gzs[e + 1] : T(0);
        zr[q] = zs[0][e];
    }
#pragma unroll 1
    for (int s = 0; s < K; ++s) {
        const T *zc = zs[s & 1];
        T *zn = zs[(s & 1) ^ 1];
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int e = t + q * C::THREADS;
            if (e >= NR) continue;
            const int a = e / (R * R), b = (e / R) % R, c = e % R;
            T out = 0;
            if (rid[q] > T(0)) {
                const T zxm = a > 0 ? zc[e - R * R] : T(0), zxp = a < R - 1 ? zc[e + R * R] : T(0);
                const T zym = b > 0 ? zc[e - R] : T(0), zyp = b < R - 1 ? zc[e + R] : T(0);
                const T zzm = c > 0 ? zc[e - 1] : T(0), zzp = c < R - 1 ? zc[e + 1] : T(0);
                const T off = gxm[q] * zxm + gxp[q] * zxp + gym[q] * zym + gyp[q] * zyp + gzm[q] * zzm + gzp[q] * zzp;
                out = (T(1) - omega) * zr[q] + omega * rid[q] * (rr[q] + off);
            }
            zn[e] = out;
            zr[q] = out;
        }
        __syncthreads();
    }
    if (t < BRICK_CELLS) {
        const int li = t >> 6, lj = (t >> 3) & 7, lk = t & 7;
        const T *zz = zs[K & 1];
        const int e = ((li + GH) * R + (lj + GH)) * R + (lk + GH);
        const long c = (long)br * BRICK_CELLS + t;
        zout[c] = zz[e];
        if (RES) {
            // apply_row on the cube (the neighbours of the brick are exact after the K sweeps)
            const T dirc = A.dir[c];
            T out = 0;
            if (dirc >= 0) {
                const T gxm = gxs[e], gym = gys[e], gzm = gzs[e];
                const T gxp = gxs[e + R * R], gyp = gys[e + R], gzp = gzs[e + 1];
                const T off = gxm * zz[e - R * R] + gxp * zz[e + R * R] + gym * zz[e - R] + gyp * zz[e + R] +
                              gzm * zz[e - 1] + gzp * zz[e + 1];
                const T diag = dirc + ((gxm + gxp) + (gym + gyp) + (gzm + gzp));
                out = rhs[c] - (diag * zz[e] - off);
            }
            res[c] = out;
        }
    }
}

// Connected components of the coarsest brick of every grid ending on level L (block g: grid g):
// unknowns joined by a nonzero coupling, labelled by min-label propagation in LDS.  code[cell] =
// -1 off the unknowns, -2 in a component with a Dirichlet term (a nonsingular block), else the
// smallest cell index of its component: a pure-Neumann component, singular with the constants of
// the component as its null space (an all-fluid box, a pocket enclosed by solid).  Built once per
// projection; k_coarse_solve makes the right-hand side of such a component consistent.
template <class T>
__global__ __launch_bounds__(BRICK_CELLS) void k_coarse_components(LevelDesc L, OpDesc<T> A, int16_t *code) {
    __shared__ int lab[BRICK_CELLS];
    __shared__ int hasd[BRICK_CELLS];
    const int g = blockIdx.x;
    const GridLevel G = L.gl[g];
    if (!G.coarsest || G.brick_end <= G.brick_start) return;
    const int e = threadIdx.x;
    const long base = (long)G.brick_start * BRICK_CELLS;
    const long c = base + e;
    const int li = e >> 6, lj = (e >> 3) & 7, lk = e & 7;
    const T dirc = A.dir[c];
    const bool unk = dirc >= 0;
    // couplings to the six neighbours (the brick is the whole grid here: nothing outside it)
    const bool cxm = li > 0 && A.gx[c] > T(0), cxp = li < 7 && A.gx[c + 64] > T(0);
    const bool cym = lj > 0 && A.gy[c] > T(0), cyp = lj < 7 && A.gy[c + 8] > T(0);
    const bool czm = lk > 0 && A.gz[c] > T(0), czp = lk < 7 && A.gz[c + 1] > T(0);
    lab[e] = unk ? e : BRICK_CELLS;
    __syncthreads();
    // labels only decrease and always name a cell of the same component, so the propagation (with
    // pointer jumping) reaches the unique fixpoint, the smallest index of each component; at most
    // 512 rounds
    for (int it = 0; it < BRICK_CELLS; ++it) {
        int m = lab[e];
        if (unk) {
            m = min(m, lab[m]);  // pointer jumping: the label of the cell this one points to
            if (cxm) m = min(m, lab[e - 64]);
            if (cxp) m = min(m, lab[e + 64]);
            if (cym) m = min(m, lab[e - 8]);
            if (cyp) m = min(m, lab[e + 8]);
            if (czm) m = min(m, lab[e - 1]);
            if (czp) m = min(m, lab[e + 1]);
        }
        const bool changed = m < lab[e];
        __syncthreads();
        if (changed) lab[e] = m;
        if (!__syncthreads_or(changed)) break;
    }
    // whether each component has a Dirichlet term: dir of an unknown is a sum of conductances
    // (non-negative), so the component's sum of dir is positive exactly when one of its cells has
    // dir > 0 (a sum of non-negative numbers rounds to 0 only when every term is 0); each such cell
    // marks its root (concurrent stores of the same value)
    hasd[e] = 0;
    __syncthreads();
    if (unk && dirc > T(0)) hasd[lab[e]] = 1;
    __syncthreads();
    code[c] = !unk ? (int16_t)-1 : (hasd[lab[e]] ? (int16_t)-2 : (int16_t)lab[e]);
}

// Coarse solve of every grid whose last level L is (one brick): block g handles grid g with
// `sweeps` damped-Jacobi sweeps from zero, in LDS.  DOT (finest level only): also the r.z partial
// of the grid's brick, as k_jacobi<..., DOT> stores it for grids with coarser levels.  CPT cells per
// thread (512 / CPT threads): fewer wavefronts make each sweep's barrier cheaper; every thread keeps
// its cells' couplings and own values in registers.
template <class T, int DOT, int CPT = 1>
__global__ __launch_bounds__(BRICK_CELLS / CPT) void k_coarse_solve(LevelDesc L, OpDesc<T> A,
                                                                    const T *__restrict__ rhs, T *__restrict__ z,
                                                                    int sweeps, T omega, Gate Q, RedBuf R,
                                                                    const int16_t *__restrict__ code) {
    constexpr int NT = BRICK_CELLS / CPT;
    __shared__ T zs[2][BRICK_CELLS];
    __shared__ double lds[32];
    __shared__ int16_t cc[BRICK_CELLS];
    if (all_done(Q)) return;
    const int g = blockIdx.x;
    const GridLevel G = L.gl[g];
    if (!G.coarsest || G.brick_end <= G.brick_start || (Q.S && Q.S[g].done)) return;
    const int t = threadIdx.x;
    const long base = (long)G.brick_start * BRICK_CELLS;
    T gxm[CPT], gym[CPT], gzm[CPT], gxp[CPT], gyp[CPT], gzp[CPT], diag[CPT], inv[CPT], r[CPT], zr[CPT];
    bool unk[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int e = t + u * NT;
        const int li = e >> 6, lj = (e >> 3) & 7, lk = e & 7;
        const long c = base + e;
        const T dirc = A.dir[c];
        gxm[u] = A.gx[c];
        gym[u] = A.gy[c];
        gzm[u] = A.gz[c];
        gxp[u] = li < 7 ? A.gx[c + 64] : T(0);
        gyp[u] = lj < 7 ? A.gy[c + 8] : T(0);
        gzp[u] = lk < 7 ? A.gz[c + 1] : T(0);
        diag[u] = dirc + ((gxm[u] + gxp[u]) + (gym[u] + gyp[u]) + (gzm[u] + gzp[u]));
        unk[u] = dirc >= 0 && diag[u] > 0;
        r[u] = rhs[c];
        inv[u] = unk[u] ? omega / diag[u] : T(0);
        cc[e] = code[c];
    }
    // A pure-Neumann component (k_coarse_components) is singular: its right-hand side is made
    // consistent (its mean over the component removed) so that the sweeps do not grow the
    // component's constant mode, which the Krylov iteration would carry into the pressure.
    bool sing = false;
#pragma unroll
    for (int u = 0; u < CPT; ++u) sing = sing || cc[t + u * NT] >= 0;
    T rs[CPT];  // the right-hand side the sweeps solve (r itself, or made consistent)
#pragma unroll
    for (int u = 0; u < CPT; ++u) rs[u] = r[u];
    if (__syncthreads_or(sing)) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) zs[0][t + u * NT] = r[u];
        __syncthreads();
        // the root of each singular component sums it in cell order and stores its mean
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = t + u * NT;
            if (cc[e] == e) {
                double sum = 0.0;
                int cnt = 0;
                for (int k = e; k < BRICK_CELLS; ++k)
                    if (cc[k] == e) {
                        sum += (double)zs[0][k];
                        ++cnt;
                    }
                zs[1][e] = (T)(sum / cnt);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = t + u * NT;
            if (cc[e] >= 0) rs[u] = r[u] - zs[1][cc[e]];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        zr[u] = unk[u] ? inv[u] * rs[u] : T(0);
        zs[0][t + u * NT] = zr[u];
    }
    __syncthreads();
    int cur = 0;
    for (int s = 1; s < sweeps; ++s) {
        const T *zz = zs[cur];
        T out[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = t + u * NT;
            const int li = e >> 6, lj = (e >> 3) & 7, lk = e & 7;
            out[u] = T(0);
            if (unk[u]) {
                const T off = gxm[u] * (li > 0 ? zz[e - 64] : T(0)) + gxp[u] * (li < 7 ? zz[e + 64] : T(0)) +
                              gym[u] * (lj > 0 ? zz[e - 8] : T(0)) + gyp[u] * (lj < 7 ? zz[e + 8] : T(0)) +
                              gzm[u] * (lk > 0 ? zz[e - 1] : T(0)) + gzp[u] * (lk < 7 ? zz[e + 1] : T(0));
                out[u] = zr[u] + inv[u] * (rs[u] - (diag[u] * zr[u] - off));
            }
        }
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            zs[cur ^ 1][t + u * NT] = out[u];
            zr[u] = out[u];
        }
        __syncthreads();
        cur ^= 1;
    }
    double part = 0.0;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        z[base + t + u * NT] = zr[u];
        part += (double)r[u] * (double)zr[u];
    }
    if (DOT) {
        part = block_sum(part, lds);
        store_partials(R, G.brick_start, part, 0.0);
    }
}

// ------------------------------------------------------------------------------ output (one grid)
// m_pressure: dense float (and optional FP64) pressure, 0 off the unknowns
// (macpressuresolver3.cpp:245-248).
// With unknown_out, also the unknown mask (the labelling's cell flag, bit 1).
__global__ void k_scatter_pressure(GridDesc G, LevelDesc L, GridLevel GL, const double *__restrict__ x, float *P,
                                   double *P64, const uint8_t *__restrict__ cflag, uint8_t *unknown_out) {
    const size_t M = (size_t)G.w * G.h * G.d;
    const size_t c = (size_t)xcd_brick(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M) return;
    int i, j, k;
    unflatten((unsigned)c, (unsigned)G.h, (unsigned)G.d, i, j, k);
    const double v = level_value(L, GL, x, i, j, k);
    if (P) P[c] = (float)v;
    if (P64) P64[c] = v;
    if (unknown_out) unknown_out[c] = (cflag[c] >> 1) & 1;
}

// Dense export of the level-0 system: diag (0 off the unknowns), the low-face couplings and the
// right-hand side, all cell-indexed.
__global__ void k_export_system(GridDesc G, LevelDesc L, GridLevel GL, OpDesc<double> A, const double *__restrict__ b,
                                double *diag, double *gx, double *gy, double *gz, double *rhs) {
    const size_t M = (size_t)G.w * G.h * G.d;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M) return;
    int i, j, k;
    unflatten((unsigned)c, (unsigned)G.h, (unsigned)G.d, i, j, k);
    const bool stored = L.bmap[GL.bmap_off + ((i >> 3) * GL.nby + (j >> 3)) * GL.nbz + (k >> 3)] >= 0;
    const double dirc = stored ? level_value(L, GL, A.dir, i, j, k) : -1.0;
    const bool unk = dirc >= 0.0;
    const double gxm = level_value(L, GL, A.gx, i, j, k), gym = level_value(L, GL, A.gy, i, j, k),
                 gzm = level_value(L, GL, A.gz, i, j, k);
    const double gxp = level_value(L, GL, A.gx, i + 1, j, k), gyp = level_value(L, GL, A.gy, i, j + 1, k),
                 gzp = level_value(L, GL, A.gz, i, j, k + 1);
    diag[c] = unk ? dirc + ((gxm + gxp) + (gym + gyp) + (gzm + gzp)) : 0.0;
    gx[c] = gxm;
    gy[c] = gym;
    gz[c] = gzm;
    rhs[c] = unk ? level_value(L, GL, b, i, j, k) : 0.0;
}

// Velocity update (macpressuresolver3.cpp:252-268) of the three faces indexed at node position
// (i,j,k) in one thread (each level-set sample is read once for the three faces): open faces on
// the domain boundary become 0, interior open faces lose dt*(P1-P0)/(rho*dx) with P stored in
// float and P1-P0 taken in float; all other (active) faces are switched off (0).
__device__ __forceinline__ void update_face(const GridDesc &G, const Exist &E, int dim, int i, int j, int k, float p1,
                                            float p0, float rho_face) {
#pragma clang fp contract(off)
    const size_t f = face_index(G, dim, i, j, k);
    uint8_t *aout = G.active_out[dim];
    if (G.active[dim] && !G.active[dim][f]) {
        if (aout) aout[f] = 0;
        return;
    }
    const int pi[3] = {i, j, k};
    const int N[3] = {G.w, G.h, G.d};
    float *U = G.vel[dim];
    // the area only matters where rho is nonzero: most faces of the domain have air on both sides
    const double rho = rho_face;
    const float area = rho != 0.0 ? face_area(G, E, dim, i, j, k) : 0.0f;
    uint8_t keep = 1;
    if (area != 0.0f && rho != 0.0) {
        if (pi[dim] == 0 || pi[dim] == N[dim]) {
            U[f] = 0.0f;
        } else {
            const float diff = p1 - p0;
            const float delta = (float)(G.dt * (double)diff / (rho * G.dx));
            U[f] = U[f] - delta;
        }
    } else {
        U[f] = 0.0f;
        // a closed boundary face stays active (at 0) where the cell inside holds liquid
        if (pi[dim] == 0) keep = phi_at(G, i, j, k) < 0.0;
        else if (pi[dim] == N[dim]) keep = phi_at(G, i - (dim == 0), j - (dim == 1), k - (dim == 2)) < 0.0;
        else keep = 0;
    }
    if (aout) aout[f] = keep;
}

// P is the grid's dense float pressure (0 off the unknowns), as k_scatter_pressure writes it.
__global__ void k_velocity_update(GridDesc G, const float *__restrict__ P) {
    const int W = G.w + 1, H = G.h + 1, D = G.d + 1;
    const size_t n = (size_t)W * H * D;
    const size_t e = (size_t)xcd_brick(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    int i, j, k;
    unflatten((unsigned)e, (unsigned)H, (unsigned)D, i, j, k);
    const bool in_i = i < G.w, in_j = j < G.h, in_k = k < G.d;
    const float pc = (in_i && in_j && in_k) ? P[cell_index(G, i, j, k)] : 0.0f;
    const Exist E = exist_of(G);
    // face_rho of the three faces from four level-set samples, read once: the (clamped) node cell
    // and the (clamped) cell below it along each axis
    const bool v0 = in_j && in_k, v1 = in_i && in_k, v2 = in_i && in_j && !G.flat;
    float rho[3] = {1.0f, 1.0f, 1.0f};
    // the four samples and the three low-side pressures are loaded together; the clamped indices
    // are in range for every node, and a face that does not exist never uses its values
    const int ci = clampi(i, 0, G.w - 1), cj = clampi(j, 0, G.h - 1), ck = clampi(k, 0, G.d - 1);
    const int ci1 = clampi(i - 1, 0, G.w - 1), cj1 = clampi(j - 1, 0, G.h - 1), ck1 = clampi(k - 1, 0, G.d - 1);
    const size_t idx[4] = {cell_index(G, ci, cj, ck), cell_index(G, ci1, cj, ck), cell_index(G, ci, cj1, ck),
                           cell_index(G, ci, cj, ck1)};
    if (E.liquid) {
        double sm[4];
        load_samples<4>(G.phi, G.phi64, idx, sm);
        rho[0] = rho_from_samples(G, E, sm[0], v0 ? sm[1] : 0.0);
        rho[1] = rho_from_samples(G, E, sm[0], v1 ? sm[2] : 0.0);
        rho[2] = rho_from_samples(G, E, sm[0], v2 ? sm[3] : 0.0);
    } else {
        rho[0] = rho[1] = rho[2] = rho_from_samples(G, E, 0.0, 0.0);
    }
    const float px = P[idx[1]], py = P[idx[2]], pz = P[idx[3]];
    if (v0) update_face(G, E, 0, i, j, k, pc, i > 0 ? px : 0.0f, rho[0]);
    if (v1) update_face(G, E, 1, i, j, k, pc, j > 0 ? py : 0.0f, rho[1]);
    if (v2) update_face(G, E, 2, i, j, k, pc, k > 0 ? pz : 0.0f, rho[2]);
}

}  // namespace shkz
