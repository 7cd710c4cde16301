// shkz_linsolve.hip -- the LinSolver drop-in: conjugate gradients on a general sparse symmetric
// system in CSR form, on the GPU (gfx950).
//
// It stands in for pcg_solver::solve (src/linsolver/pcg.cpp:45-73 -> PCGSolver<double>::solve,
// local/include/pcgsolver/pcg_solver.h:246-295), which serves every module that assembles an
// RCMatrix and asks the LinSolver for x (the stock macpressuresolver3, macstreamfuncsolver3, ...).
// Iteration: x0 = 0, r = b, stop when ||r||_inf <= residual * ||b||_inf, at most max_iterations;
// the reference computes MIC(0) and discards it (:379-384), so its iterates are plain CG's, which
// SHKZ_LS_NONE reproduces; SHKZ_LS_JACOBI applies the inverse diagonal.  The returned relative
// residual is ||r||_inf / ||b||_inf as in :283,293.
//
// Layout: CSR with 32-bit column indices on the device.  SpMV gives each row 8 lanes of a
// 64-wide wavefront (rows of a few to a few dozen entries, the shape of the Poisson systems the
// reference assembles) and reduces them with lane shuffles.  Dot products and norms are
// deterministic two-pass reductions (per-block partials, summed in block order by one block).
// One iteration is captured once into a hipGraph and replayed; the scalars stay on the device.
#include "../../include/shkz_mi355x.h"
#include "shkz_error.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr int LS_BLOCK = 256;
constexpr int LS_LANES = 8;                        // lanes per row in SpMV
constexpr int LS_ROWS = LS_BLOCK / LS_LANES;       // rows per SpMV block

struct LsScalars {
    double rz, alpha, beta, rinf, r0, tol;
    int32_t iter, maxit, done, converged, status, pad;
};

__device__ __forceinline__ double ls_wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double ls_wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block reduction of (a: max or sum, b: sum) into partials[2*block].
template <bool AMAX>
__device__ __forceinline__ void ls_block_partials(double a, double b, double *partials) {
    __shared__ double lds[2][LS_BLOCK / 64];
    a = AMAX ? ls_wave_max(a) : ls_wave_sum(a);
    b = ls_wave_sum(b);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[0][wid] = a;
        lds[1][wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < LS_BLOCK / 64; ++i) {
            sa = AMAX ? fmax(sa, lds[0][i]) : sa + lds[0][i];
            sb += lds[1][i];
        }
        partials[2 * blockIdx.x] = sa;
        partials[2 * blockIdx.x + 1] = sb;
    }
}

// inverse diagonal (0 where the row has no positive diagonal entry)
__global__ void k_ls_idiag(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                           const double *__restrict__ val, double *__restrict__ idiag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
        if (col[e] == i) d += val[e];
    idiag[i] = d > 0.0 ? 1.0 / d : 0.0;
}

// x = 0, r = b, z = M^-1 r; partials: max|r|, r.z
__global__ void k_ls_init(int n, const double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
                          double *__restrict__ z, const double *__restrict__ idiag, double *partials) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0, rz = 0.0;
    if (i < n) {
        const double v = b[i];
        x[i] = 0.0;
        r[i] = v;
        const double zi = idiag ? idiag[i] * v : v;
        z[i] = zi;
        mx = fabs(v);
        rz = v * zi;
    }
    ls_block_partials<true>(mx, rz, partials);
}

// q = A p, partial p.q
__global__ void k_ls_spmv_dot(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ q,
                              const LsScalars *S, double *partials) {
    if (S->done) return;
    const int row = blockIdx.x * LS_ROWS + (threadIdx.x / LS_LANES);
    const int lane = threadIdx.x % LS_LANES;
    double acc = 0.0;
    if (row < n) {
        const int e1 = rowptr[row + 1];
        for (int e = rowptr[row] + lane; e < e1; e += LS_LANES) acc += val[e] * p[col[e]];
    }
#pragma unroll
    for (int off = LS_LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    double part = 0.0;
    if (row < n && lane == 0) {
        q[row] = acc;
        part = p[row] * acc;
    }
    ls_block_partials<false>(part, 0.0, partials);
}

// x += alpha p, r -= alpha q, z = M^-1 r; partials: max|r|, r.z
__global__ void k_ls_update(int n, double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
                            const double *__restrict__ p, const double *__restrict__ q,
                            const double *__restrict__ idiag, const LsScalars *S, double *partials) {
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0, rz = 0.0;
    if (i < n) {
        const double a = S->alpha;
        x[i] += a * p[i];
        const double rv = r[i] - a * q[i];
        r[i] = rv;
        const double zi = idiag ? idiag[i] * rv : rv;
        if (idiag) z[i] = zi;
        mx = fabs(rv);
        rz = rv * zi;
    }
    ls_block_partials<true>(mx, rz, partials);
}

// p = z + beta p (first: p = z)
__global__ void k_ls_direction(int n, const double *__restrict__ z, double *__restrict__ p, const LsScalars *S,
                               int first) {
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = first ? z[i] : z[i] + S->beta * p[i];
}

enum { LS_FIN_INIT = 0, LS_FIN_PQ = 1, LS_FIN_UPDATE = 2 };

// One block combines the partials of nb blocks in block order and updates the scalars, as the
// reference's loop does between its BLAS calls (pcg_solver.h:250-290).
__global__ __launch_bounds__(1024) void k_ls_finalize(const double *__restrict__ partials, int nb, LsScalars *S, int op,
                                                      double residual, int maxit) {
    __shared__ double lds[2][16];
    if (op != LS_FIN_INIT && S->done) return;
    const bool amax = op != LS_FIN_PQ;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a = amax ? fmax(a, partials[2 * i]) : a + partials[2 * i];
        b += partials[2 * i + 1];
    }
    a = amax ? ls_wave_max(a) : ls_wave_sum(a);
    b = ls_wave_sum(b);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[0][wid] = a;
        lds[1][wid] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a = 0.0;
    b = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        a = amax ? fmax(a, lds[0][i]) : a + lds[0][i];
        b += lds[1][i];
    }
    LsScalars &s = *S;
    if (op == LS_FIN_INIT) {
        s.r0 = a;
        s.rinf = a;
        s.rz = b;
        s.iter = 0;
        s.maxit = maxit;
        s.tol = (residual < 1e-30 ? 1e-30 : residual) * a;
        s.status = 0;
        s.converged = 0;
        s.done = 0;
        if (a == 0.0) {  // :253-256
            s.converged = 1;
            s.done = 1;
        } else if (b == 0.0 || !(b == b)) {  // :261-270
            s.status = b == b ? 0 : 1;
            s.done = 1;
        } else if (maxit <= 0) {
            s.done = 1;
        }
    } else if (op == LS_FIN_PQ) {
        s.alpha = s.rz / a;
    } else {
        s.rinf = a;
        s.iter += 1;
        if (!(a == a) || !(b == b)) {
            s.status = 1;
            s.done = 1;
        } else if (a <= s.tol) {
            s.converged = 1;
            s.done = 1;
        } else if (s.iter >= s.maxit) {
            s.done = 1;
        } else {
            s.beta = b / s.rz;
            s.rz = b;
        }
    }
}

inline unsigned blocks_of(long n, int per) { return (unsigned)((n + per - 1) / per); }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct shkz_linsolver {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // device buffers (grown on demand)
    struct Buf {
        void *p = nullptr;
        size_t cap = 0;
    } rowptr, col, val, b, x, r, z, p, q, idiag, partials, scalars;
    LsScalars *host_s = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> exec_key;
    int exec_n = -1, exec_prec = -1;
};

namespace {

int ls_get(shkz_linsolver::Buf &b, size_t bytes, void **out) {
    if (bytes == 0) bytes = 64;
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        const size_t want = bytes + bytes / 8;
        if (hipMalloc(&b.p, want) != hipSuccess) {
            b.p = nullptr;
            return shkz_fail(SHKZ_ERR_NOMEM, "hipMalloc(%zu) failed in the linear solver", want);
        }
        b.cap = want;
    }
    *out = b.p;
    return SHKZ_OK;
}

#define LS_HIP(expr)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return shkz_fail(SHKZ_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)
#define LS_TRY(expr)                                                                                                   \
    do {                                                                                                               \
        int rc_ = (expr);                                                                                              \
        if (rc_ != SHKZ_OK) return rc_;                                                                                \
    } while (0)

// Device arrays of one solve.
struct LsArrays {
    const int *rowptr, *col;
    const double *val, *b;
    double *x, *r, *z, *p, *q, *idiag, *partials;
    LsScalars *S;
};

// The solve on device arrays (b filled); x receives the result.  With `graph`, one iteration is
// captured once into a hipGraph kept by the solver and replayed; without, the same kernels are
// launched directly (same launch shapes, so the same results).  Scalars come back through the
// solver's pinned buffer.
int ls_solve_device(shkz_linsolver *L, hipStream_t s, const LsArrays &D, int n, const shkz_linsolve_params *prm,
                    shkz_linsolve_result *res, bool graph) {
    const int *rp = D.rowptr, *cl = D.col;
    const double *vl = D.val, *b = D.b;
    double *x = D.x, *r = D.r, *z = D.z, *p = D.p, *q = D.q, *id = D.idiag, *part = D.partials;
    LsScalars *S = D.S;
    const bool jacobi = prm->preconditioner == SHKZ_LS_JACOBI;
    const double *M = jacobi ? id : nullptr;
    double *zz = jacobi ? z : r;  // without a preconditioner z is r itself
    const unsigned nb_v = blocks_of(n, LS_BLOCK), nb_m = blocks_of(n, LS_ROWS);
    const int maxit = prm->max_iterations > 0 ? prm->max_iterations : 30000;
    if (jacobi) k_ls_idiag<<<nb_v, LS_BLOCK, 0, s>>>(n, rp, cl, vl, id);
    k_ls_init<<<nb_v, LS_BLOCK, 0, s>>>(n, b, x, r, z, M, part);
    k_ls_finalize<<<1, 1024, 0, s>>>(part, (int)nb_v, S, LS_FIN_INIT, prm->residual, maxit);
    k_ls_direction<<<nb_v, LS_BLOCK, 0, s>>>(n, zz, p, S, 1);
    auto iteration = [&]() {
        k_ls_spmv_dot<<<nb_m, LS_BLOCK, 0, s>>>(n, rp, cl, vl, p, q, S, part);
        k_ls_finalize<<<1, 1024, 0, s>>>(part, (int)nb_m, S, LS_FIN_PQ, 0.0, 0);
        k_ls_update<<<nb_v, LS_BLOCK, 0, s>>>(n, x, r, z, p, q, M, S, part);
        k_ls_finalize<<<1, 1024, 0, s>>>(part, (int)nb_v, S, LS_FIN_UPDATE, 0.0, 0);
        k_ls_direction<<<nb_v, LS_BLOCK, 0, s>>>(n, zz, p, S, 0);
    };
    if (graph) {
        // one captured iteration, reused while the arrays and sizes stay the same
        std::vector<const void *> key = {rp, cl, vl, x, r, z, p, q, id, part, S};
        if (!L->exec || L->exec_key != key || L->exec_n != n || L->exec_prec != (int)jacobi) {
            if (L->exec) (void)hipGraphExecDestroy(L->exec);
            L->exec = nullptr;
            hipGraph_t g = nullptr;
            LS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            iteration();
            LS_HIP(hipStreamEndCapture(s, &g));
            const hipError_t e = hipGraphInstantiate(&L->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                L->exec = nullptr;
                return shkz_fail(SHKZ_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            }
            L->exec_key = key;
            L->exec_n = n;
            L->exec_prec = (int)jacobi;
        }
    }
    int chunk = 8;
    for (;;) {
        LS_HIP(hipMemcpyAsync(L->host_s, S, sizeof(LsScalars), hipMemcpyDeviceToHost, s));
        LS_HIP(hipStreamSynchronize(s));
        const LsScalars h = *L->host_s;
        if (h.done) break;
        // size the next chunk from the observed contraction rate (the replays after convergence
        // return at once)
        int next = chunk;
        if (h.iter > 0 && h.rinf > 0 && h.r0 > 0 && h.rinf < h.r0) {
            const double rate = std::pow(h.rinf / h.r0, 1.0 / h.iter);
            if (rate > 0 && rate < 1) next = (int)std::ceil(std::log(h.tol / h.rinf) / std::log(rate));
        }
        chunk = std::max(1, std::min({next, h.maxit - h.iter, 256}));
        for (int it = 0; it < chunk; ++it) {
            if (graph) LS_HIP(hipGraphLaunch(L->exec, s));
            else iteration();
        }
    }
    LS_HIP(hipGetLastError());
    const LsScalars h = *L->host_s;
    res->iterations = h.iter;
    res->converged = h.converged;
    res->reresid = h.r0 > 0 ? h.rinf / h.r0 : 0.0;
    res->status = h.status ? SHKZ_ERR_NUMERIC : SHKZ_OK;
    return res->status == SHKZ_OK ? SHKZ_OK : shkz_fail(SHKZ_ERR_NUMERIC, "non-finite value met in the linear solve");
}

// Scratch of the allocation-free entry point: a fixed device array (each GPU the library runs on
// holds its own copy until the process ends).  Layout in doubles: r, z, p, q, idiag (each
// SHKZ_LS_NOALLOC_MAX_ROWS), the block partials (two per SpMV block of the largest system), then
// the scalar record.
constexpr size_t NA_ROWS = SHKZ_LS_NOALLOC_MAX_ROWS;
constexpr size_t NA_PARTIALS = 2 * ((NA_ROWS + LS_ROWS - 1) / LS_ROWS);
constexpr size_t NA_SCALARS = (sizeof(LsScalars) + sizeof(double) - 1) / sizeof(double);
constexpr size_t NA_DOUBLES = 5 * NA_ROWS + NA_PARTIALS + NA_SCALARS;
static_assert(NA_ROWS % LS_ROWS == 0 && LS_BLOCK % LS_ROWS == 0, "partials cover both kernels' blocks");
static __device__ __align__(256) double scratch[NA_DOUBLES];

// Per-device address of `scratch` (the array's name is a host-side stand-in, not a device
// address), looked up once per device with that device current.
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
double *g_na_ptr[NA_MAX_DEVICES] = {};

int noalloc_scratch(int device, double **out) {
    if (device < 0 || device >= NA_MAX_DEVICES) return shkz_fail(SHKZ_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return shkz_fail(SHKZ_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                             hipGetErrorString(e));
        g_na_ptr[device] = (double *)ptr;
    }
    *out = g_na_ptr[device];
    return SHKZ_OK;
}

}  // namespace

extern "C" {

void shkz_linsolve_default_params(shkz_linsolve_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->residual = 1e-4;  // pcg.cpp:76-77
    p->max_iterations = 30000;
    p->preconditioner = SHKZ_LS_NONE;
}

int shkz_linsolver_create(int device, shkz_linsolver **out) {
    if (!out) return shkz_fail(SHKZ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    LS_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return shkz_fail(SHKZ_ERR_INVALID, "device %d out of range", device);
    LS_HIP(hipSetDevice(device));
    auto *L = new shkz_linsolver();
    L->device = device;
    if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&L->host_s, sizeof(LsScalars), hipHostMallocDefault) != hipSuccess) {
        if (L->stream) (void)hipStreamDestroy(L->stream);
        delete L;
        return shkz_fail(SHKZ_ERR_DEVICE, "stream or pinned buffer creation failed");
    }
    double *na = nullptr;
    const int rc = noalloc_scratch(device, &na);
    if (rc != SHKZ_OK) {
        (void)hipStreamDestroy(L->stream);
        (void)hipHostFree(L->host_s);
        delete L;
        return rc;
    }
    *out = L;
    return SHKZ_OK;
}

void shkz_linsolver_destroy(shkz_linsolver *L) {
    if (!L) return;
    (void)hipSetDevice(L->device);
    (void)hipStreamSynchronize(L->stream);
    if (L->exec) (void)hipGraphExecDestroy(L->exec);
    for (shkz_linsolver::Buf *b : {&L->rowptr, &L->col, &L->val, &L->b, &L->x, &L->r, &L->z, &L->p, &L->q, &L->idiag,
                                   &L->partials, &L->scalars})
        if (b->p) (void)hipFree(b->p);
    if (L->host_s) (void)hipHostFree(L->host_s);
    (void)hipStreamDestroy(L->stream);
    delete L;
}

int shkz_linsolve_csr(shkz_linsolver *L, int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
                      const double *b, double *x, const shkz_linsolve_params *params, shkz_linsolve_result *result) {
    if (!L || !result || n < 0 || (n > 0 && (!rowptr || !b || !x)))
        return shkz_fail(SHKZ_ERR_INVALID, "NULL argument or negative size");
    std::memset(result, 0, sizeof(*result));
    shkz_linsolve_params def;
    shkz_linsolve_default_params(&def);
    const shkz_linsolve_params *prm = params ? params : &def;
    if (prm->preconditioner != SHKZ_LS_NONE && prm->preconditioner != SHKZ_LS_JACOBI)
        return shkz_fail(SHKZ_ERR_INVALID, "unknown preconditioner %d", prm->preconditioner);
    const int64_t nnz = n > 0 ? rowptr[n] : 0;
    if (n >= ((int64_t)1 << 31) || nnz >= ((int64_t)1 << 31))
        return shkz_fail(SHKZ_ERR_INVALID, "system too large for 32-bit indices (n=%lld, nnz=%lld)", (long long)n,
                         (long long)nnz);
    // validate and narrow the indices on the host (the device trusts them)
    std::vector<int> rp32(n + 1), col32(nnz);
    for (int64_t i = 0; i <= n; ++i) {
        if (rowptr[i] < 0 || rowptr[i] > nnz || (i > 0 && rowptr[i] < rowptr[i - 1]))
            return shkz_fail(SHKZ_ERR_INVALID, "row pointers are not monotone at row %lld", (long long)i);
        rp32[i] = (int)rowptr[i];
    }
    if (nnz > 0 && (!col || !val)) return shkz_fail(SHKZ_ERR_INVALID, "NULL column or value array");
    for (int64_t e = 0; e < nnz; ++e) {
        if (col[e] < 0 || col[e] >= n) return shkz_fail(SHKZ_ERR_INVALID, "column %lld out of range", (long long)col[e]);
        col32[e] = (int)col[e];
    }
    std::lock_guard<std::mutex> lock(L->mu);
    LS_HIP(hipSetDevice(L->device));
    const double t0 = now_ms();
    if (n == 0) {
        result->converged = 1;
        return SHKZ_OK;
    }
    void *pp;
    LS_TRY(ls_get(L->rowptr, (n + 1) * sizeof(int), &pp));
    LS_TRY(ls_get(L->col, nnz * sizeof(int), &pp));
    LS_TRY(ls_get(L->val, nnz * sizeof(double), &pp));
    for (shkz_linsolver::Buf *bb : {&L->b, &L->x, &L->r, &L->z, &L->p, &L->q, &L->idiag})
        LS_TRY(ls_get(*bb, n * sizeof(double), &pp));
    LS_TRY(ls_get(L->partials, 2 * sizeof(double) * std::max(blocks_of(n, LS_BLOCK), blocks_of(n, LS_ROWS)), &pp));
    LS_TRY(ls_get(L->scalars, sizeof(LsScalars), &pp));
    LS_HIP(hipMemcpyAsync(L->rowptr.p, rp32.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, L->stream));
    if (nnz) {
        LS_HIP(hipMemcpyAsync(L->col.p, col32.data(), nnz * sizeof(int), hipMemcpyHostToDevice, L->stream));
        LS_HIP(hipMemcpyAsync(L->val.p, val, nnz * sizeof(double), hipMemcpyHostToDevice, L->stream));
    }
    LS_HIP(hipMemcpyAsync(L->b.p, b, n * sizeof(double), hipMemcpyHostToDevice, L->stream));
    LS_HIP(hipStreamSynchronize(L->stream));  // the host staging vectors go out of scope
    const double t1 = now_ms();
    const LsArrays D{(const int *)L->rowptr.p, (const int *)L->col.p, (const double *)L->val.p, (const double *)L->b.p,
                     (double *)L->x.p, (double *)L->r.p, (double *)L->z.p, (double *)L->p.p, (double *)L->q.p,
                     (double *)L->idiag.p, (double *)L->partials.p, (LsScalars *)L->scalars.p};
    const int rc = ls_solve_device(L, L->stream, D, (int)n, prm, result, true);
    LS_HIP(hipMemcpyAsync(x, L->x.p, n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    LS_HIP(hipStreamSynchronize(L->stream));
    const double t2 = now_ms();
    result->t_upload = t1 - t0;
    result->t_solve = t2 - t1;
    return rc;
}

int shkz_linsolve_csr_device_noalloc(shkz_linsolver *L, int32_t n, const int32_t *rowptr, const int32_t *col,
                                     const double *val, const double *b, double *x,
                                     const shkz_linsolve_params *params, shkz_linsolve_result *result, void *stream) {
    if (!L || !result) return shkz_fail(SHKZ_ERR_INVALID, "NULL solver or result");
    std::memset(result, 0, sizeof(*result));
    if (n < 0 || (n > 0 && (!rowptr || !col || !val || !b || !x)))
        return shkz_fail(SHKZ_ERR_INVALID, "NULL argument or negative size");
    if ((size_t)n > NA_ROWS)
        return shkz_fail(SHKZ_ERR_INVALID, "%d rows exceed the allocation-free solver's limit of %zu rows", (int)n,
                         NA_ROWS);
    shkz_linsolve_params def;
    shkz_linsolve_default_params(&def);
    const shkz_linsolve_params *prm = params ? params : &def;
    if (prm->preconditioner != SHKZ_LS_NONE && prm->preconditioner != SHKZ_LS_JACOBI)
        return shkz_fail(SHKZ_ERR_INVALID, "unknown preconditioner %d", prm->preconditioner);
    std::lock_guard<std::mutex> lock(L->mu);
    LS_HIP(hipSetDevice(L->device));
    const double t0 = now_ms();
    if (n == 0) {
        result->converged = 1;
        return SHKZ_OK;
    }
    double *base = nullptr;
    LS_TRY(noalloc_scratch(L->device, &base));
    const hipStream_t s = (hipStream_t)stream;
    const LsArrays D{rowptr, col, val, b, x, base, base + NA_ROWS, base + 2 * NA_ROWS, base + 3 * NA_ROWS,
                     base + 4 * NA_ROWS, base + 5 * NA_ROWS, (LsScalars *)(base + 5 * NA_ROWS + NA_PARTIALS)};
    const int rc = ls_solve_device(L, s, D, (int)n, prm, result, false);
    result->t_solve = now_ms() - t0;
    return rc;
}

}  // extern "C"
