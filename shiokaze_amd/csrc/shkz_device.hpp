// shkz_device.hpp -- device-side data structures shared by the kernels and the host runtime.
//
// Storage model.  Every grid (and every multigrid level of it) is cut into 8x8x8 bricks.  Only
// bricks that hold at least one unknown are stored, brick-major, 512 cells each, with the cell
// order (li*8 + lj)*8 + lk (z fastest).  A dense brick map (one int per brick position) gives
// the compact id of a stored brick or -1.  Vectors are zero on every stored cell that is not an
// unknown, so the operator needs no separate mask: dir < 0 marks a non-unknown.
//
// The operator of one level is four arrays over the stored cells:
//   dir      sum of the conductances of the open faces towards non-unknown (air) neighbours,
//            -1 on a stored cell that is not an unknown,
//   gx,gy,gz conductance of the cell's low x/y/z face when both cells are unknowns, else 0,
// so that diag_c = dir_c + the six couplings and (A p)_c = diag_c p_c - sum_f g_f p_q with
// p = 0 off the unknowns.  On the
// finest level the conductance is dt*area/(dx*dx*rho), the reference's matrix entry
// (src/projection/macpressuresolver3.cpp:185).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shkz {

constexpr int BRICK = 8;
constexpr int BRICK_CELLS = 512;
constexpr int BLOCK = 512;  // brick kernels: one thread per cell
constexpr int MAX_LEVELS = 32;  // a grid side of 8 * 2^31 cells would need more

// Neighbour slots of a brick in its neighbour table.
enum { NB_XM = 0, NB_XP = 1, NB_YM = 2, NB_YP = 3, NB_ZM = 4, NB_ZP = 5 };

// Per grid and level: the grid's cell and brick extents, where its dense brick map starts in the
// level's combined map, and its range of stored bricks in the level's combined brick list.
struct GridLevel {
    int ni, nj, nk;     // cells
    int nbx, nby, nbz;  // bricks per axis
    int bmap_off;       // offset of this grid's dense brick map
    int brick_start;    // first stored brick of this grid
    int brick_end;      // one past the last
    int coarsest;       // 1: the grid fits one brick here, so this is its last level (coarse solve)
    int pad[2];
};

// Per stored brick: its six neighbour bricks (NB_*, -1 = none), its grid and flags, in one
// 32-byte record (one scalar load per workgroup).
enum {
    BM_COARSEST = 1,  // the brick's grid ends on this level (its coarse solve handles it)
    BM_UNIFORM = 2,   // level 0: the operator is one constant on the brick (UniCoef)
};
struct __align__(32) BrickMeta {
    int nbr[6];
    int grid;
    int flags;
};

// A level-0 brick on which every cell is an unknown without a Dirichlet term and every face of
// every cell (the high faces of its last layers included) has the same conductance g -- the
// interior of the liquid away from surfaces and solids, most of the domain.  The smoothing and
// operator kernels take g and the inverse diagonal from this record instead of streaming the
// four operator arrays and idiag: the same values, so the same results, with 5 fewer arrays read.
struct __align__(32) UniCoef {
    double g64, id64;  // FP64 conductance and inverse diagonal (FP64 smoothers)
    float g32, id32;   // the FP32 copies of the mixed-precision V-cycle
    float pad[2];
};

template <class T>
struct UniT {
    T g, id;
};
__device__ __forceinline__ void uni_pick(const UniCoef &u, UniT<double> &o) {
    o.g = u.g64;
    o.id = u.id64;
}
__device__ __forceinline__ void uni_pick(const UniCoef &u, UniT<float> &o) {
    o.g = u.g32;
    o.id = u.id32;
}

// One multigrid level of a batch of grids: the stored bricks of every grid, concatenated
// (grid-major, Morton order inside a grid).  A single grid is a batch of one.
struct LevelDesc {
    int nbricks;             // stored bricks, all grids
    int ngrids;
    int flat;                // 2D grids (one cell layer): bilinear transfers
    int pad;
    const int *bmap;         // per grid g: bmap[gl[g].bmap_off + (bx*nby + by)*nbz + bz] -> brick or -1
    const int *bxyz;         // brick -> bx, by, bz (in its grid)
    const BrickMeta *meta;   // brick -> neighbours, grid, flags
    const GridLevel *gl;
    const int *child64;      // brick -> the 4x4x4 finer-level bricks 2B-1 .. 2B+2 per axis (-1 = none)
    const int *parent8;      // brick -> the 2x2x2 coarser-level bricks (B+1)/2 - 1 + {0,1} per axis
    const UniCoef *uni;      // level 0: constants of the BM_UNIFORM bricks (nullptr elsewhere)
    const int *nb27;         // coarse levels: brick -> the 3x3x3 bricks B-1 .. B+1 per axis (-1 = none)
};

template <class T>
__device__ __forceinline__ UniT<T> uni_load(const LevelDesc &L, int br) {
    UniT<T> o;
    uni_pick(L.uni[br], o);
    return o;
}

template <class T>
struct OpDesc {
    const T *dir;  // sum of conductances to non-unknown neighbours; -1 marks a non-unknown
    const T *gx;
    const T *gy;
    const T *gz;
};

// Scalars of one grid's Krylov solve, resident on the device so that no iteration waits on the
// host.
struct SolveScalars {
    double rz;        // r.z of the current iteration
    double pq;        // p.Ap
    double alpha, beta;
    double rinf;      // ||r||_inf after the last update
    double rl2;       // ||r||_2 after the last update
    double r0inf;     // ||r0||_inf
    double r0l2;
    double tol;       // absolute tolerance in the chosen norm
    double rmin;      // smallest residual norm reached (stagnation test)
    int32_t iter;     // iterations done
    int32_t maxit;
    int32_t done;     // 1 = stop (converged or out of iterations)
    int32_t converged;
    int32_t norm;     // shkz_norm
    int32_t status;   // 0 ok, 1 = NaN met
    int32_t stall;        // iterations since the residual last reached a new minimum
    int32_t stall_limit;  // > 0: stop after that many (inner solves of the defect correction)
};

// Batch-wide state of one Krylov solve: how many grids are done; every kernel returns at once when
// all are.
struct BatchCtl {
    int32_t ngrids;
    int32_t ndone;
    int32_t all_done;
    int32_t pad;
};

// What a V-cycle launch may skip: everything once all grids are done, and the bricks of a grid
// that is done.  S may be NULL (no per-grid test).
struct Gate {
    const BatchCtl *ctl;
    const SolveScalars *S;
};

// Deterministic two-pass reduction: every brick of a reducing kernel stores two partials (plain
// stores, no atomics); a finalize kernel with one block per grid combines the partials of that
// grid's bricks in brick order and updates the grid's scalars.
struct RedBuf {
    double *partials;  // 2 per brick
};

// Remaps a linear block id so that each of the 8 XCDs (which receive blocks round-robin) gets a
// contiguous range of bricks: neighbouring bricks then share that XCD's L2.
__device__ __forceinline__ int xcd_brick(int block, int nblocks) {
    const int per = (nblocks + 7) >> 3;
    return (block & 7) * per + (block >> 3);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide sum / max of one value per thread (BLOCK threads).  Result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s += lds[i];
    return s;
}
__device__ __forceinline__ double block_max(double v, double *lds) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s = fmax(s, lds[i]);
    return s;
}

// Two block-wide reductions in one LDS round: a by max (AMAX) or sum, b by sum.  Same operations
// in the same order as block_max / block_sum, so the results are identical to theirs.  lds holds
// 32 doubles.
template <bool AMAX>
__device__ __forceinline__ void block_reduce2(double &a, double &b, double *lds) {
    a = AMAX ? wave_max(a) : wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        lds[wid] = a;
        lds[16 + wid] = b;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) {
        sa = AMAX ? fmax(sa, lds[i]) : sa + lds[i];
        sb += lds[16 + i];
    }
    a = sa;
    b = sb;
}

__device__ __forceinline__ void store_partials(const RedBuf &R, int brick, double a, double b) {
    if (threadIdx.x == 0) {
        R.partials[2 * brick + 0] = a;
        R.partials[2 * brick + 1] = b;
    }
}

}  // namespace shkz
