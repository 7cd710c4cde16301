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
constexpr int MAX_LEVELS = 12;

// Neighbour slots of a brick in its neighbour table.
enum { NB_XM = 0, NB_XP = 1, NB_YM = 2, NB_YP = 3, NB_ZM = 4, NB_ZP = 5 };

struct LevelDesc {
    int ni, nj, nk;     // cells
    int nbx, nby, nbz;  // bricks per axis
    int nbricks;        // stored bricks
    int pad;
    const int *bmap;    // (bx*nby + by)*nbz + bz -> compact id or -1
    const int *bxyz;    // compact id -> bx, by, bz
    const int *bnbr;    // compact id -> 6 neighbour ids (NB_*), -1 = none
};

template <class T>
struct OpDesc {
    const T *dir;  // sum of conductances to non-unknown neighbours; -1 marks a non-unknown
    const T *gx;
    const T *gy;
    const T *gz;
};

// Scalars of one Krylov solve, resident on the device so that no iteration waits on the host.
struct SolveScalars {
    double rz;        // r.z of the current iteration
    double pq;        // p.Ap
    double alpha, beta;
    double rinf;      // ||r||_inf after the last update
    double rl2;       // ||r||_2 after the last update
    double r0inf;     // ||r0||_inf
    double r0l2;
    double tol;       // absolute tolerance in the chosen norm
    int32_t iter;     // iterations done
    int32_t maxit;
    int32_t done;     // 1 = stop (converged or out of iterations); kernels return at once
    int32_t converged;
    int32_t norm;     // shkz_norm
    int32_t status;   // 0 ok, 1 = NaN met
    int32_t pad[2];
};

// Deterministic two-pass reduction: every block of a reducing kernel stores two partials (plain
// stores, no atomics); a one-block finalize kernel combines them in block order and updates the
// solve's scalars.
struct RedBuf {
    double *partials;  // 2 per block
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

__device__ __forceinline__ void store_partials(const RedBuf &R, double a, double b) {
    if (threadIdx.x == 0) {
        R.partials[2 * blockIdx.x + 0] = a;
        R.partials[2 * blockIdx.x + 1] = b;
    }
}

}  // namespace shkz
