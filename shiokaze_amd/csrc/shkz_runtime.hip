// shkz_runtime.hip -- host runtime of the MI355X pressure projection and its C ABI.
//
// One shkz_project*() call is one macpressuresolver3::project (src/projection/
// macpressuresolver3.cpp:50-272):
//   fractions + labelling          k_any_negative, k_flags            (:61-156)
//   assembly (matrix-free)         k_operator, k_coarsen              (:159-199)
//   volume correction, warm start  host PI controller, scene state    (:203-230, :239-242)
//   linear solve                   CG / MGPCG / mixed-precision MGPCG (:233-237 -> pcg.cpp:45-73)
//   scatter + velocity update      k_scatter_pressure, k_velocity_update (:245-268)
// The context owns every device buffer (grown on demand, never freed per call), the brick
// hierarchy of the last problem, the per-scene resident state and the kernel timers.
#include "../../include/shkz_mi355x.h"
#include "shkz_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace shkz {

static thread_local std::string g_last_error;

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define SHKZ_HIP(expr)                                                                                                 \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return fail(SHKZ_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));            \
    } while (0)

#define SHKZ_TRY(expr)                                                                                                 \
    do {                                                                                                               \
        int rc_ = (expr);                                                                                              \
        if (rc_ != SHKZ_OK) return rc_;                                                                                \
    } while (0)

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------------ device memory
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
};

struct Arena {
    std::unordered_map<std::string, DevBuf> bufs;
    size_t total = 0;
    size_t budget = 0;

    int get(const std::string &name, size_t bytes, void **out) {
        DevBuf &b = bufs[name];
        if (bytes == 0) bytes = 256;
        if (b.cap < bytes) {
            size_t want = bytes + bytes / 8;
            want = (want + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
            if (total - b.cap + want > budget)
                return fail(SHKZ_ERR_BUDGET, "device budget exceeded: %s needs %zu bytes (held %zu of %zu)", name.c_str(),
                            want, total, budget);
            if (b.ptr) {
                SHKZ_HIP(hipFree(b.ptr));
                total -= b.cap;
                b.ptr = nullptr;
                b.cap = 0;
            }
            hipError_t e = hipMalloc(&b.ptr, want);
            if (e != hipSuccess) {
                b.ptr = nullptr;
                return fail(SHKZ_ERR_NOMEM, "hipMalloc(%zu) for %s failed: %s", want, name.c_str(), hipGetErrorString(e));
            }
            b.cap = want;
            total += want;
        }
        *out = b.ptr;
        return SHKZ_OK;
    }
    void release(const std::string &name) {
        auto it = bufs.find(name);
        if (it == bufs.end()) return;
        if (it->second.ptr) {
            (void)hipFree(it->second.ptr);
            total -= it->second.cap;
        }
        bufs.erase(it);
    }
    void clear() {
        for (auto &kv : bufs)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
        bufs.clear();
        total = 0;
    }
};

template <class T>
static int arena_get(Arena &A, const std::string &name, size_t count, T **out) {
    void *p = nullptr;
    SHKZ_TRY(A.get(name, count * sizeof(T), &p));
    *out = (T *)p;
    return SHKZ_OK;
}

// ------------------------------------------------------------------------ hierarchy
struct LevelHost {
    int ni = 0, nj = 0, nk = 0;
    int nbx = 0, nby = 0, nbz = 0;
    int nbricks = 0;
    std::vector<int> bmap, bxyz, bnbr;
    LevelDesc desc{};
    // operator and V-cycle vectors (element type = level precision)
    void *dir = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr, *idiag = nullptr;
    void *rhs = nullptr, *z = nullptr, *tmp = nullptr, *res = nullptr;
    int64_t unknowns = 0;
};

static inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}
static inline uint64_t morton3(int x, int y, int z) {
    return spread3((uint32_t)x) << 2 | spread3((uint32_t)y) << 1 | spread3((uint32_t)z);
}

// Orders the active bricks along a Morton curve, fills the dense map and the neighbour table.
static void finish_level(LevelHost &L, const std::vector<uint8_t> &active) {
    std::vector<std::pair<uint64_t, int>> keys;
    const int nb = L.nbx * L.nby * L.nbz;
    for (int b = 0; b < nb; ++b)
        if (active[b]) {
            const int bz = b % L.nbz, by = (b / L.nbz) % L.nby, bx = b / (L.nbz * L.nby);
            keys.push_back({morton3(bx, by, bz), b});
        }
    std::sort(keys.begin(), keys.end());
    L.nbricks = (int)keys.size();
    L.bmap.assign(nb, -1);
    L.bxyz.resize(3 * (size_t)L.nbricks);
    for (int id = 0; id < L.nbricks; ++id) {
        const int b = keys[id].second;
        L.bmap[b] = id;
        L.bxyz[3 * id] = b / (L.nbz * L.nby);
        L.bxyz[3 * id + 1] = (b / L.nbz) % L.nby;
        L.bxyz[3 * id + 2] = b % L.nbz;
    }
    L.bnbr.resize(6 * (size_t)L.nbricks);
    const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int id = 0; id < L.nbricks; ++id) {
        for (int d = 0; d < 6; ++d) {
            const int x = L.bxyz[3 * id] + off[d][0], y = L.bxyz[3 * id + 1] + off[d][1], z = L.bxyz[3 * id + 2] + off[d][2];
            int n = -1;
            if (x >= 0 && y >= 0 && z >= 0 && x < L.nbx && y < L.nby && z < L.nbz) n = L.bmap[(x * L.nby + y) * L.nbz + z];
            L.bnbr[6 * id + d] = n;
        }
    }
}

static void level_dims(LevelHost &L, int ni, int nj, int nk) {
    L.ni = ni;
    L.nj = nj;
    L.nk = nk;
    L.nbx = (ni + BRICK - 1) / BRICK;
    L.nby = (nj + BRICK - 1) / BRICK;
    L.nbz = (nk + BRICK - 1) / BRICK;
}

// ------------------------------------------------------------------------ timing
struct TimedLaunch {
    int kernel;   // index into names
    int level;
    double bytes_per_unknown;
    double bytes_fixed;
    hipEvent_t ev0, ev1;
};

struct KernelTotals {
    std::string name;
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

// ------------------------------------------------------------------------ scenes
struct SceneState {
    int w = 0, h = 0, d = 0;
    bool has_pressure = false;
    double y_prev = 0.0;  // volume-correction integrator (macpressuresolver3.cpp:208)
};

}  // namespace shkz

using namespace shkz;

struct shkz_ctx {
    int device = 0;
    Arena arena;
    std::mutex mu;  // one projection at a time per context
    std::unordered_map<uint64_t, SceneState> scenes;
    double y_prev = 0.0;  // stateless calls share the module-level integrator, as m_y_prev does
    // hierarchy of the current problem
    std::vector<LevelHost> levels;
    // pinned scalars read-back
    SolveScalars *host_scalars = nullptr;
    // timing
    bool timing = false;
    std::vector<std::string> kernel_names;
    std::unordered_map<std::string, int> kernel_ids;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<KernelTotals> totals;
    // execution: one work stream (default priority) ordered after the caller's stream
    hipStream_t work = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    bool use_graphs = true;
    bool trace = false;
    // replayable PCG iteration graphs, keyed by everything their kernel arguments depend on
    struct GraphSlot {
        std::string key;
        hipGraphExec_t exec = nullptr;
    } graphs[4];
};

namespace shkz {

// ------------------------------------------------------------------------ launch helper
struct Launcher {
    shkz_ctx *C;
    hipStream_t s;
    bool capturing = false;

    int kernel_id(const char *name) {
        auto it = C->kernel_ids.find(name);
        if (it != C->kernel_ids.end()) return it->second;
        const int id = (int)C->kernel_names.size();
        C->kernel_names.push_back(name);
        C->kernel_ids[name] = id;
        return id;
    }
    hipEvent_t event() {
        if (!C->event_pool.empty()) {
            hipEvent_t e = C->event_pool.back();
            C->event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    template <class K, class... Args>
    void go(const char *name, int level, double bytes_per_unknown, double bytes_fixed, dim3 grid, dim3 block, K kernel,
            Args... args) {
        if (grid.x == 0) return;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (C->timing && !capturing) {
            e0 = event();
            e1 = event();
            if (e0) (void)hipEventRecord(e0, s);
        }
        kernel<<<grid, block, 0, s>>>(args...);
        if (C->timing && !capturing && e0 && e1) {
            (void)hipEventRecord(e1, s);
            C->pending.push_back({kernel_id(name), level, bytes_per_unknown, bytes_fixed, e0, e1});
        }
    }
};

static inline unsigned brick_grid(int nbricks) { return (unsigned)(((nbricks + 7) / 8) * 8); }

// Resolves pending timing events into totals (after a stream synchronisation).
static void resolve_timing(shkz_ctx *C) {
    for (const TimedLaunch &t : C->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t.ev0, t.ev1);
        if ((int)C->totals.size() <= t.kernel) C->totals.resize(t.kernel + 1);
        KernelTotals &K = C->totals[t.kernel];
        K.name = C->kernel_names[t.kernel];
        K.launches += 1;
        K.ms += ms;
        const int64_t unk = (t.level >= 0 && t.level < (int)C->levels.size()) ? C->levels[t.level].unknowns : 0;
        K.bytes += t.bytes_per_unknown * (double)unk + t.bytes_fixed;
        C->event_pool.push_back(t.ev0);
        C->event_pool.push_back(t.ev1);
    }
    C->pending.clear();
}

// ------------------------------------------------------------------------ solve plumbing
struct Problem {
    GridDesc G{};
    int nlev = 0;
    bool mixed = false;
    bool fp64_levels = true;  // V-cycle element type
    // level-0 FP64 arrays
    double *dir64 = nullptr, *gx64 = nullptr, *gy64 = nullptr, *gz64 = nullptr;
    double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr;
    // Krylov vectors of the level-0 solve (element type of the solve)
    void *kr = nullptr, *kp = nullptr, *kq = nullptr, *kx = nullptr, *kz = nullptr, *ktmp = nullptr;
    SolveScalars *S_out = nullptr, *S_in = nullptr;
    RedBuf red{};
    int pre = 2, post = 2, pre_c = 2, post_c = 2, coarse_sweeps = 48;
    int lb = 0;  // first level of the one-workgroup bottom cycle
    double omega = 0.8;
};

template <class T>
static OpDesc<T> op_of(const LevelHost &L) {
    return OpDesc<T>{(const T *)L.dir, (const T *)L.gx, (const T *)L.gy, (const T *)L.gz};
}

// One V-cycle z = M r on the level-0 vectors (rhs0 -> returns the pointer holding z).
// One V-cycle z = M r on the level-0 vectors; returns the pointer holding z.  Levels below P.lb
// run as per-brick kernels, levels P.lb.. in one k_bottom launch.  With dot_partials, the last
// post-smoothing sweep on the finest level also stores r.z partials in P.red (the caller then
// finalizes them instead of running a separate dot kernel).
template <class T>
static T *vcycle(shkz_ctx *C, Launcher &K, Problem &P, const T *rhs0, T *z0, T *tmp0, const SolveScalars *S,
                 bool dot_partials) {
    const int nl = P.nlev;
    std::vector<T *> zcur(nl), zoth(nl);
    const T om = (T)P.omega;
    const double es = (double)sizeof(T);
    for (int l = 0; l < nl; ++l) {
        LevelHost &L = C->levels[l];
        const T *rhs = l == 0 ? rhs0 : (const T *)L.rhs;
        T *z = l == 0 ? z0 : (T *)L.z;
        T *tmp = l == 0 ? tmp0 : (T *)L.tmp;
        const OpDesc<T> A = op_of<T>(L);
        const T *id = (const T *)L.idiag;
        const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
        const bool f = l == 0;
        const int pre = f ? P.pre : P.pre_c;
        if (l == P.lb) {
            BottomDesc<T> B{};
            B.nlev = nl - l;
            B.pre = P.pre_c;
            B.post = P.post_c;
            B.coarse_sweeps = P.coarse_sweeps;
            B.omega = om;
            for (int q = 0; q < B.nlev; ++q) {
                LevelHost &Q = C->levels[l + q];
                BottomLevel<T> &V = B.lv[q];
                V.L = Q.desc;
                V.A = op_of<T>(Q);
                V.idiag = (const T *)Q.idiag;
                V.rhs = q == 0 ? rhs : (const T *)Q.rhs;
                V.z = q == 0 ? z : (T *)Q.z;
                V.tmp = q == 0 ? tmp : (T *)Q.tmp;
                V.res = (T *)Q.res;
            }
            K.go("bottom", -1, 0, 0, dim3(1), dim3(BOTTOM_THREADS), k_bottom<T>, B, S);
            // where the bottom's first level ended: each Jacobi sweep after smooth2 swaps buffers
            const int swaps = B.nlev == 1 ? 0 : (std::max(0, P.pre_c - 2) + P.post_c);
            zcur[l] = (swaps & 1) ? tmp : z;
            zoth[l] = (swaps & 1) ? z : tmp;
            break;
        }
        T *cur = z, *oth = tmp;
        if (pre <= 1) {
            K.go(f ? "smooth_first_L0" : "smooth_first_Lc", l, 3 * es, 0, g, b,
                 f ? k_smooth_first<T, 1> : k_smooth_first<T, 0>, L.desc, id, rhs, z, om, S);
        } else {
            K.go(f ? "smooth2_L0" : "smooth2_Lc", l, 6 * es, 0, g, b, f ? k_smooth2<T, 1> : k_smooth2<T, 0>, L.desc, A,
                 id, rhs, z, om, S);
            for (int s = 2; s < pre; ++s) {
                K.go(f ? "jacobi_L0" : "jacobi_Lc", l, 7 * es, 0, g, b, f ? k_jacobi<T, 1, 0> : k_jacobi<T, 0, 0>,
                     L.desc, A, id, rhs, (const T *)cur, oth, om, S, P.red);
                std::swap(cur, oth);
            }
        }
        K.go(f ? "residual_L0" : "residual_Lc", l, 7 * es, 0, g, b, f ? k_residual<T, 1> : k_residual<T, 0>, L.desc, A,
             rhs, (const T *)cur, (T *)L.res, S);
        LevelHost &Lc = C->levels[l + 1];
        K.go(f ? "restrict_L0" : "restrict_Lc", l, es, 2 * es * (double)Lc.unknowns, dim3(brick_grid(Lc.nbricks)), b,
             f ? k_restrict<T, 1> : k_restrict<T, 0>, L.desc, Lc.desc, op_of<T>(Lc), (const T *)L.res, (T *)Lc.rhs, S);
        zcur[l] = cur;
        zoth[l] = oth;
    }
    for (int l = P.lb - 1; l >= 0; --l) {
        LevelHost &L = C->levels[l];
        LevelHost &Lc = C->levels[l + 1];
        const T *rhs = l == 0 ? rhs0 : (const T *)L.rhs;
        const OpDesc<T> A = op_of<T>(L);
        const T *id = (const T *)L.idiag;
        const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
        const bool f = l == 0;
        const int post = f ? P.post : P.post_c;
        K.go(f ? "prolong_L0" : "prolong_Lc", l, 3 * es, es * (double)Lc.unknowns, g, b,
             f ? k_prolong<T, 1> : k_prolong<T, 0>, L.desc, A, Lc.desc, (const T *)zcur[l + 1], zcur[l], S);
        T *cur = zcur[l], *oth = zoth[l];
        for (int s = 0; s < post; ++s) {
            const bool last_dot = f && dot_partials && s == post - 1;
            if (last_dot)
                K.go("jacobi_dot_L0", l, 7 * es, 0, g, b, k_jacobi<T, 1, 1>, L.desc, A, id, rhs, (const T *)cur, oth, om,
                     S, P.red);
            else
                K.go(f ? "jacobi_L0" : "jacobi_Lc", l, 7 * es, 0, g, b, f ? k_jacobi<T, 1, 0> : k_jacobi<T, 0, 0>,
                     L.desc, A, id, rhs, (const T *)cur, oth, om, S, P.red);
            std::swap(cur, oth);
        }
        zcur[l] = cur;
        zoth[l] = oth;
    }
    return zcur[0];
}

static int read_scalars(shkz_ctx *C, hipStream_t s, const SolveScalars *dS, SolveScalars *out) {
    SHKZ_HIP(hipMemcpyAsync(C->host_scalars, dS, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    SHKZ_HIP(hipGetLastError());
    *out = *C->host_scalars;
    if (C->timing) resolve_timing(C);
    return SHKZ_OK;
}

static void finalize(Launcher &K, const Problem &P, unsigned nblocks, SolveScalars *S, int op, double param = 0.0,
                     int iparam = 0) {
    K.go("finalize", -1, 0, 16.0 * nblocks, dim3(1), dim3(1024), k_finalize, (const double *)P.red.partials,
         (int)nblocks, S, op, param, iparam);
}

template <class T>
static std::string graph_key(shkz_ctx *C, const Problem &P, const void *const *vecs, int nvec, bool precond) {
    std::string k;
    char buf[160];
    std::snprintf(buf, sizeof(buf), "T%zu pc%d pre%d post%d prc%d poc%d lb%d cs%d om%.17g red%p|", sizeof(T),
                  precond ? 1 : 0, P.pre, P.post, P.pre_c, P.post_c, P.lb, P.coarse_sweeps, P.omega,
                  (void *)P.red.partials);
    k += buf;
    for (int i = 0; i < nvec; ++i) {
        std::snprintf(buf, sizeof(buf), "%p,", vecs[i]);
        k += buf;
    }
    const int nl = precond ? P.nlev : 1;
    for (int l = 0; l < nl; ++l) {
        const LevelHost &L = C->levels[l];
        std::snprintf(buf, sizeof(buf), "|%d %d %d %d %p %p %p %p %p %p %p %p %p %p %p", L.ni, L.nj, L.nk, L.nbricks,
                      (const void *)L.desc.bmap, L.dir, L.gx, L.gy, L.gz, L.rhs, L.z, L.tmp, L.res,
                      (const void *)L.desc.bxyz, (const void *)L.desc.bnbr);
        k += buf;
    }
    return k;
}

// Krylov loop on level 0: T = element type; precond = V-cycle or identity (plain CG).
// Expects r (residual), and S with tol/maxit/iter set.  x is updated in place.  One iteration is
// captured once into a hipGraph and replayed (kernels return at once after convergence, so a
// chunk of replays may overshoot by a few no-op iterations).
template <class T>
static int krylov(shkz_ctx *C, Launcher &K, Problem &P, T *x, T *r, T *p, T *q, T *z, T *tmp, SolveScalars *S,
                  bool precond, SolveScalars *hs) {
    LevelHost &L0 = C->levels[0];
    const unsigned nb = brick_grid(L0.nbricks);
    const dim3 g(nb), b(BLOCK);
    const OpDesc<T> A = op_of<T>(L0);
    const double es = (double)sizeof(T);
    // the finest level's last post-smoothing sweep forms r.z when there is one
    const bool fused_dot = precond && P.lb > 0 && P.post > 0;
    // first direction
    const T *zp = r;
    if (precond) zp = vcycle<T>(C, K, P, r, z, tmp, S, fused_dot);
    if (!fused_dot)
        K.go("dot_rz", 0, 2 * es, 0, g, b, k_dot_rz<T>, L0.desc, (const T *)r, zp, (const SolveScalars *)S, P.red);
    finalize(K, P, nb, S, FIN_RZ_FIRST);
    K.go("update_p", 0, 2 * es, 0, g, b, k_update_p<T>, L0.desc, zp, p, (const SolveScalars *)S, 1);

    auto iteration = [&]() {
        K.go("apply_dot", 0, 6 * es, 0, g, b, k_apply_dot<T>, L0.desc, A, (const T *)p, q, (const SolveScalars *)S,
             P.red);
        finalize(K, P, nb, S, FIN_PQ);
        K.go("cg_update", 0, 6 * es, 0, g, b, k_cg_update<T>, L0.desc, x, r, (const T *)p, (const T *)q,
             (const SolveScalars *)S, P.red);
        finalize(K, P, nb, S, precond ? FIN_UPDATE : FIN_UPDATE_CG);
        if (precond) {
            const T *zz = vcycle<T>(C, K, P, r, z, tmp, S, fused_dot);
            if (!fused_dot)
                K.go("dot_rz", 0, 2 * es, 0, g, b, k_dot_rz<T>, L0.desc, (const T *)r, zz, (const SolveScalars *)S,
                     P.red);
            finalize(K, P, nb, S, FIN_RZ);
            K.go("update_p", 0, 3 * es, 0, g, b, k_update_p<T>, L0.desc, zz, p, (const SolveScalars *)S, 0);
        } else {
            K.go("update_p", 0, 3 * es, 0, g, b, k_update_p<T>, L0.desc, (const T *)r, p, (const SolveScalars *)S, 0);
        }
    };
    hipGraphExec_t exec = nullptr;
    if (C->use_graphs && !C->timing) {
        const void *vecs[] = {x, r, p, q, z, tmp, S};
        const std::string key = graph_key<T>(C, P, vecs, 7, precond);
        auto &slot = C->graphs[(sizeof(T) == 8 ? 2 : 0) + (precond ? 1 : 0)];
        if (slot.exec && slot.key == key) {
            exec = slot.exec;
        } else {
            if (slot.exec) {
                SHKZ_HIP(hipGraphExecDestroy(slot.exec));
                slot.exec = nullptr;
                slot.key.clear();
            }
            hipGraph_t graph = nullptr;
            SHKZ_HIP(hipStreamBeginCapture(K.s, hipStreamCaptureModeThreadLocal));
            K.capturing = true;
            iteration();
            K.capturing = false;
            SHKZ_HIP(hipStreamEndCapture(K.s, &graph));
            hipError_t e = hipGraphInstantiate(&slot.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                slot.exec = nullptr;
                return fail(SHKZ_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            }
            slot.key = key;
            exec = slot.exec;
        }
    }
    int chunk = 4;
    for (;;) {
        for (int it = 0; it < chunk; ++it) {
            if (exec) SHKZ_HIP(hipGraphLaunch(exec, K.s));
            else iteration();
        }
        SHKZ_TRY(read_scalars(C, K.s, S, hs));
        if (hs->done) break;
        // size the next chunk from the observed contraction rate
        const double cur = hs->norm == 0 ? hs->rinf : hs->rl2;
        const double r0 = hs->norm == 0 ? hs->r0inf : hs->r0l2;
        int next = chunk;
        if (hs->iter > 0 && cur > 0 && r0 > 0 && cur < r0) {
            const double rate = std::pow(cur / r0, 1.0 / hs->iter);
            if (rate > 0 && rate < 1) next = (int)std::ceil(std::log(hs->tol / cur) / std::log(rate));
        }
        chunk = std::max(1, std::min(next, 64));
        if (hs->iter + chunk > hs->maxit) chunk = std::max(1, hs->maxit - hs->iter);
    }
    return SHKZ_OK;
}

static int upload_scalars(shkz_ctx *C, hipStream_t s, SolveScalars *dS, int norm, int maxit) {
    SolveScalars h{};
    h.norm = norm;
    h.maxit = maxit;
    *C->host_scalars = h;
    SHKZ_HIP(hipMemcpyAsync(dS, C->host_scalars, sizeof(SolveScalars), hipMemcpyHostToDevice, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    return SHKZ_OK;
}

// ------------------------------------------------------------------------ the projection
struct ExportOut {
    double *diag, *gx, *gy, *gz, *rhs;  // device, cell-indexed
};

static int project_stages(shkz_ctx *C, const shkz_grid *grid, const shkz_params *prm, shkz_result *res,
                          hipStream_t s, const ExportOut *exp);

// Runs the projection on the context's work stream and orders the caller's stream after it, so
// that work the caller queues next (and synchronous copies on the null stream) sees the outputs.
static int project_impl(shkz_ctx *C, const shkz_grid *grid, const shkz_params *prm, shkz_result *res,
                        hipStream_t caller, const ExportOut *exp = nullptr) {
    const int rc = project_stages(C, grid, prm, res, caller, exp);
    if (C->work) {
        if (hipEventRecord(C->ev_out, C->work) != hipSuccess || hipStreamWaitEvent(caller, C->ev_out, 0) != hipSuccess) {
            (void)hipStreamSynchronize(C->work);
            return rc != SHKZ_OK ? rc : fail(SHKZ_ERR_DEVICE, "ordering the caller's stream after the projection failed");
        }
        if (rc != SHKZ_OK) (void)hipStreamSynchronize(C->work);
    }
    return rc;
}

static int project_stages(shkz_ctx *C, const shkz_grid *grid, const shkz_params *prm, shkz_result *res,
                          hipStream_t s, const ExportOut *exp) {
    const double t_start = now_ms();
    std::memset(res, 0, sizeof(*res));
    const int w = grid->w, h = grid->h, d = grid->d;
    if (w <= 0 || h <= 0 || d <= 0) return fail(SHKZ_ERR_INVALID, "grid dimensions must be positive (%d,%d,%d)", w, h, d);
    if (!grid->fluid || !grid->u || !grid->v || !grid->w_vel) return fail(SHKZ_ERR_INVALID, "fluid and velocities are required");
    if (!(grid->dt > 0)) return fail(SHKZ_ERR_INVALID, "dt must be positive");
    if (prm->solver != SHKZ_SOLVER_MGPCG && prm->solver != SHKZ_SOLVER_CG) return fail(SHKZ_ERR_INVALID, "unknown solver %d", prm->solver);
    if (prm->precision != SHKZ_PREC_MIXED && prm->precision != SHKZ_PREC_FP64)
        return fail(SHKZ_ERR_INVALID, "unknown precision %d", prm->precision);
    if ((int64_t)w * h * d > (int64_t)1 << 31) return fail(SHKZ_ERR_INVALID, "grid too large for 32-bit brick ids");
    SHKZ_HIP(hipSetDevice(C->device));
    // all work runs on the context's stream, ordered after what the caller queued on its stream
    SHKZ_HIP(hipEventRecord(C->ev_in, s));
    SHKZ_HIP(hipStreamWaitEvent(C->work, C->ev_in, 0));
    s = C->work;
    Launcher K{C, s};
    Arena &AR = C->arena;
    const size_t M = (size_t)w * h * d;
    const double dx = grid->dx > 0 ? grid->dx : std::min(1.0 / w, std::min(1.0 / h, 1.0 / d));

    Problem P;
    P.mixed = (prm->solver == SHKZ_SOLVER_MGPCG && prm->precision == SHKZ_PREC_MIXED);
    P.fp64_levels = !P.mixed;
    P.pre = prm->mg_pre_sweeps > 0 ? prm->mg_pre_sweeps : 4;
    P.post = prm->mg_post_sweeps > 0 ? prm->mg_post_sweeps : 4;
    P.pre_c = prm->mg_coarse_pre_sweeps > 0 ? prm->mg_coarse_pre_sweeps : P.pre;
    P.post_c = prm->mg_coarse_post_sweeps > 0 ? prm->mg_coarse_post_sweeps : P.post;
    P.coarse_sweeps = prm->mg_coarse_sweeps > 0 ? prm->mg_coarse_sweeps : 48;
    P.omega = prm->mg_omega > 0 ? prm->mg_omega : 0.9;
    const int bottom_bricks = prm->mg_bottom_bricks > 0 ? prm->mg_bottom_bricks : 1;

    GridDesc &G = P.G;
    G.w = w;
    G.h = h;
    G.d = d;
    G.phi64 = grid->fluid_dtype == SHKZ_F64;
    G.solid64 = grid->solid_dtype == SHKZ_F64;
    G.dx = dx;
    G.dt = grid->dt;
    G.phi = grid->fluid;
    G.solid = grid->solid;
    G.eps_fluid = prm->eps_fluid;
    G.eps_solid = prm->eps_solid;
    G.so_fluid = prm->second_order_fluid;
    G.so_solid = prm->second_order_solid;
    G.vel[0] = grid->u;
    G.vel[1] = grid->v;
    G.vel[2] = grid->w_vel;
    G.active[0] = grid->active_u;
    G.active[1] = grid->active_v;
    G.active[2] = grid->active_w;

    int *exist = nullptr;
    SHKZ_TRY(arena_get(AR, "exist", 4, &exist));
    G.exist = exist;

    // ---------------------------------------------------------------- fractions + labelling
    K.go("reset", -1, 0, 8, dim3(1), dim3(64), k_reset_ints, exist, 4);
    {
        const unsigned nbk = (unsigned)std::min<size_t>((M + 255) / 256, 4096);
        K.go("any_negative", -1, 0, (double)M * (G.phi64 ? 8 : 4), dim3(nbk), dim3(256), k_any_negative, grid->fluid,
             G.phi64, M, exist + 0);
        if (grid->solid) {
            const size_t Nn = (size_t)(w + 1) * (h + 1) * (d + 1);
            const unsigned nbs = (unsigned)std::min<size_t>((Nn + 255) / 256, 4096);
            K.go("any_negative", -1, 0, (double)Nn * (G.solid64 ? 8 : 4), dim3(nbs), dim3(256), k_any_negative,
                 grid->solid, G.solid64, Nn, exist + 1);
        }
    }
    if (prm->surface_tension != 0.0) {
        float *curv = nullptr;
        SHKZ_TRY(arena_get(AR, "curvature", M, &curv));
        K.go("curvature", -1, 0, (double)M * ((G.phi64 ? 8 : 4) + 4), dim3((unsigned)((M + 255) / 256)), dim3(256),
             k_curvature, G, curv);
        for (int dim = 0; dim < 3; ++dim) {
            const size_t nf = (size_t)(w + (dim == 0)) * (h + (dim == 1)) * (d + (dim == 2));
            K.go("surface_tension", -1, 0, (double)nf * 8, dim3((unsigned)((nf + 255) / 256)), dim3(256),
                 k_surface_tension, G, dim, (const float *)curv, prm->surface_tension);
        }
    }
    LevelHost L0;
    level_dims(L0, w, h, d);
    const int nb0 = L0.nbx * L0.nby * L0.nbz;
    uint8_t *cflag = nullptr, *bflag = nullptr;
    SHKZ_TRY(arena_get(AR, "cflag", M, &cflag));
    SHKZ_TRY(arena_get(AR, "bflag", (size_t)nb0, &bflag));
    K.go("flags", -1, 0, (double)M * ((G.phi64 ? 8 : 4) + 1 + (grid->solid ? 4 : 0)), dim3((unsigned)nb0), dim3(BLOCK),
         k_flags, G, L0.nbx, L0.nby, L0.nbz, cflag, bflag);
    std::vector<uint8_t> hflag(nb0);
    SHKZ_HIP(hipMemcpyAsync(hflag.data(), bflag, nb0, hipMemcpyDeviceToHost, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    SHKZ_HIP(hipGetLastError());
    const double t_frac = now_ms();

    // ---------------------------------------------------------------- hierarchy (host)
    C->levels.clear();
    finish_level(L0, hflag);
    C->levels.push_back(std::move(L0));
    const bool multigrid = prm->solver == SHKZ_SOLVER_MGPCG;
    if (multigrid) {
        while ((int)C->levels.size() < MAX_LEVELS) {
            const LevelHost &F = C->levels.back();
            if (F.nbx <= 1 && F.nby <= 1 && F.nbz <= 1) break;
            LevelHost Lc;
            level_dims(Lc, (F.ni + 1) / 2, (F.nj + 1) / 2, (F.nk + 1) / 2);
            std::vector<uint8_t> act((size_t)Lc.nbx * Lc.nby * Lc.nbz, 0);
            for (int id = 0; id < F.nbricks; ++id) {
                const int bx = F.bxyz[3 * id] / 2, by = F.bxyz[3 * id + 1] / 2, bz = F.bxyz[3 * id + 2] / 2;
                act[((size_t)bx * Lc.nby + by) * Lc.nbz + bz] = 1;
            }
            finish_level(Lc, act);
            C->levels.push_back(std::move(Lc));
        }
    }
    P.nlev = (int)C->levels.size();
    P.lb = P.nlev - 1;
    for (int l = 1; l < P.nlev; ++l)
        if (C->levels[l].nbricks <= bottom_bricks) {
            P.lb = l;
            break;
        }
    if (P.nlev - P.lb > BOTTOM_MAX) P.lb = P.nlev - BOTTOM_MAX;
    if (C->trace) {
        std::fprintf(stderr, "[shkz] levels:");
        for (int l = 0; l < P.nlev; ++l)
            std::fprintf(stderr, " %dx%dx%d/%d%s", C->levels[l].ni, C->levels[l].nj, C->levels[l].nk,
                         C->levels[l].nbricks, l == P.lb ? "(bottom)" : "");
        std::fprintf(stderr, "\n");
    }
    if (P.nlev > 1 && C->levels.back().nbricks > 1) return fail(SHKZ_ERR_INVALID, "coarsest level holds more than one brick");

    // upload the brick tables (one buffer per level)
    int max_blocks = 8;
    for (int l = 0; l < P.nlev; ++l) {
        LevelHost &L = C->levels[l];
        const size_t nmap = L.bmap.size(), nxyz = L.bxyz.size(), nnbr = L.bnbr.size();
        int *meta = nullptr;
        SHKZ_TRY(arena_get(AR, "meta" + std::to_string(l), nmap + nxyz + nnbr + 4, &meta));
        std::vector<int> host(nmap + nxyz + nnbr + 4, -1);
        std::copy(L.bmap.begin(), L.bmap.end(), host.begin());
        std::copy(L.bxyz.begin(), L.bxyz.end(), host.begin() + nmap);
        std::copy(L.bnbr.begin(), L.bnbr.end(), host.begin() + nmap + nxyz);
        SHKZ_HIP(hipMemcpyAsync(meta, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice, s));
        L.desc.ni = L.ni;
        L.desc.nj = L.nj;
        L.desc.nk = L.nk;
        L.desc.nbx = L.nbx;
        L.desc.nby = L.nby;
        L.desc.nbz = L.nbz;
        L.desc.nbricks = L.nbricks;
        L.desc.bmap = meta;
        L.desc.bxyz = meta + nmap;
        L.desc.bnbr = meta + nmap + nxyz;
        max_blocks = std::max(max_blocks, (int)brick_grid(L.nbricks));
    }
    // reduction scratch and scalars
    SHKZ_TRY(arena_get(AR, "partials", (size_t)2 * max_blocks, &P.red.partials));
    SolveScalars *Sbuf = nullptr;
    SHKZ_TRY(arena_get(AR, "scalars", 2, &Sbuf));
    P.S_out = Sbuf;
    P.S_in = Sbuf + 1;

    // ---------------------------------------------------------------- operator + vectors
    LevelHost &F0 = C->levels[0];
    const size_t n0 = (size_t)F0.nbricks * BRICK_CELLS;
    res->n_bricks = F0.nbricks;
    res->n_levels = P.nlev;
    SHKZ_TRY(arena_get(AR, "dir64", n0, &P.dir64));
    SHKZ_TRY(arena_get(AR, "gx64", n0, &P.gx64));
    SHKZ_TRY(arena_get(AR, "gy64", n0, &P.gy64));
    SHKZ_TRY(arena_get(AR, "gz64", n0, &P.gz64));
    SHKZ_TRY(arena_get(AR, "b64", n0, &P.b64));
    SHKZ_TRY(arena_get(AR, "x64", n0, &P.x64));
    SHKZ_TRY(arena_get(AR, "r64", n0, &P.r64));
    float *dir32 = nullptr, *gx32 = nullptr, *gy32 = nullptr, *gz32 = nullptr;
    if (P.mixed) {
        SHKZ_TRY(arena_get(AR, "dir32", n0, &dir32));
        SHKZ_TRY(arena_get(AR, "gx32", n0, &gx32));
        SHKZ_TRY(arena_get(AR, "gy32", n0, &gy32));
        SHKZ_TRY(arena_get(AR, "gz32", n0, &gz32));
    }
    // level-0 operator as seen by the V-cycle / Krylov loop
    if (P.mixed) {
        F0.dir = dir32;
        F0.gx = gx32;
        F0.gy = gy32;
        F0.gz = gz32;
    } else {
        F0.dir = P.dir64;
        F0.gx = P.gx64;
        F0.gy = P.gy64;
        F0.gz = P.gz64;
    }
    const size_t es = P.fp64_levels ? 8 : 4;
    for (int l = 1; l < P.nlev; ++l) {
        LevelHost &L = C->levels[l];
        const size_t n = (size_t)L.nbricks * BRICK_CELLS;
        const std::string sfx = std::to_string(l);
        void *pp = nullptr;
        SHKZ_TRY(AR.get("dir_l" + sfx, n * es, &pp)); L.dir = pp;
        SHKZ_TRY(AR.get("gx_l" + sfx, n * es, &pp)); L.gx = pp;
        SHKZ_TRY(AR.get("gy_l" + sfx, n * es, &pp)); L.gy = pp;
        SHKZ_TRY(AR.get("gz_l" + sfx, n * es, &pp)); L.gz = pp;
        SHKZ_TRY(AR.get("idiag_l" + sfx, n * es, &pp)); L.idiag = pp;
        SHKZ_TRY(AR.get("rhs_l" + sfx, n * es, &pp)); L.rhs = pp;
        SHKZ_TRY(AR.get("z_l" + sfx, n * es, &pp)); L.z = pp;
        SHKZ_TRY(AR.get("tmp_l" + sfx, n * es, &pp)); L.tmp = pp;
        SHKZ_TRY(AR.get("res_l" + sfx, n * es, &pp)); L.res = pp;
    }
    {
        void *pp = nullptr;
        SHKZ_TRY(AR.get("res_l0", n0 * es, &pp));
        F0.res = pp;
        SHKZ_TRY(AR.get("idiag_l0", n0 * es, &pp));
        F0.idiag = pp;
    }
    // Krylov vectors on level 0
    const size_t ks = (prm->solver == SHKZ_SOLVER_MGPCG && P.mixed) ? 4 : 8;
    SHKZ_TRY(AR.get("kp", n0 * ks, &P.kp));
    SHKZ_TRY(AR.get("kq", n0 * ks, &P.kq));
    SHKZ_TRY(AR.get("kz", n0 * ks, &P.kz));
    SHKZ_TRY(AR.get("ktmp", n0 * ks, &P.ktmp));
    if (P.mixed) {
        SHKZ_TRY(AR.get("kr32", n0 * 4, &P.kr));
        SHKZ_TRY(AR.get("kx32", n0 * 4, &P.kx));
    }

    // volume correction (macpressuresolver3.cpp:203-217) and warm start state
    SceneState *scene = nullptr;
    if (grid->scene_id) {
        scene = &C->scenes[grid->scene_id];
        if (scene->w != w || scene->h != h || scene->d != d) {
            scene->w = w;
            scene->h = h;
            scene->d = d;
            scene->has_pressure = false;
        }
    }
    double rhs_correct = 0.0;
    if (prm->gain && prm->target_volume) {
        double &yprev = scene ? scene->y_prev : C->y_prev;
        const double xv = (prm->current_volume - prm->target_volume) / prm->target_volume;
        const double y = yprev + xv * grid->dt;
        yprev = y;
        const double kp = prm->gain * 2.3 / (25.0 * 0.01);
        const double ki = kp * kp / 16.0;
        rhs_correct = -(kp * xv + ki * y) / (xv + 1.0);
    }
    res->rhs_correct = rhs_correct;
    double *scene_p = nullptr;
    if (scene) SHKZ_TRY(arena_get(AR, "scene" + std::to_string(grid->scene_id), M, &scene_p));
    const double *warm = (prm->warm_start && scene && scene->has_pressure) ? scene_p : nullptr;

    const unsigned g0 = brick_grid(F0.nbricks);
    K.go("operator", -1, 0,
         (double)n0 * ((G.phi64 ? 8 : 4) + (grid->solid ? 4 : 0) + 12 + 1 + 6 * 8 + (P.mixed ? 16 : 0)), dim3(g0),
         dim3(BLOCK), k_operator, G, F0.desc, (const uint8_t *)cflag, P.dir64, P.gx64, P.gy64, P.gz64, dir32, gx32,
         gy32, gz32, P.b64, P.x64, warm, rhs_correct);
    for (int l = 1; l < P.nlev; ++l) {
        LevelHost &Lf = C->levels[l - 1];
        LevelHost &Lc = C->levels[l];
        const dim3 g(brick_grid(Lc.nbricks)), b(BLOCK);
        if (P.fp64_levels) {
            K.go("coarsen", l, 0, 0, g, b, k_coarsen<double, double>, Lf.desc, op_of<double>(Lf), Lc.desc,
                 (double *)Lc.dir, (double *)Lc.gx, (double *)Lc.gy, (double *)Lc.gz);
        } else if (l == 1) {
            OpDesc<double> A64{P.dir64, P.gx64, P.gy64, P.gz64};
            K.go("coarsen", l, 0, 0, g, b, k_coarsen<double, float>, Lf.desc, A64, Lc.desc, (float *)Lc.dir,
                 (float *)Lc.gx, (float *)Lc.gy, (float *)Lc.gz);
        } else {
            K.go("coarsen", l, 0, 0, g, b, k_coarsen<float, float>, Lf.desc, op_of<float>(Lf), Lc.desc,
                 (float *)Lc.dir, (float *)Lc.gx, (float *)Lc.gy, (float *)Lc.gz);
        }
    }
    // inverse diagonals for the smoothers
    if (multigrid) {
        for (int l = 0; l < P.nlev; ++l) {
            LevelHost &L = C->levels[l];
            const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
            if (P.fp64_levels)
                K.go("idiag", -1, 0, 0, g, b, k_idiag<double>, L.desc, op_of<double>(L), (double *)L.idiag);
            else
                K.go("idiag", -1, 0, 0, g, b, k_idiag<float>, L.desc, op_of<float>(L), (float *)L.idiag);
        }
    }
    // unknowns per level (result + byte model of the kernel timers)
    {
        int64_t *cnt = nullptr;
        SHKZ_TRY(arena_get(AR, "counts", MAX_LEVELS, &cnt));
        K.go("reset", -1, 0, 0, dim3(1), dim3(64), k_reset_ints, (int *)cnt, 2 * MAX_LEVELS);
        for (int l = 0; l < P.nlev; ++l) {
            LevelHost &L = C->levels[l];
            const bool f64 = (l == 0) || P.fp64_levels;
            const void *dp = (l == 0) ? (const void *)P.dir64 : L.dir;
            const size_t n = (size_t)L.nbricks * BRICK_CELLS;
            if (n == 0) continue;
            K.go("count", -1, 0, (double)n * (f64 ? 8 : 4), dim3((unsigned)std::min<size_t>((n + 255) / 256, 2048)),
                 dim3(256), k_count_unknowns, dp, f64 ? 1 : 0, n, (unsigned long long *)(cnt + l));
        }
        std::vector<int64_t> hc(MAX_LEVELS, 0);
        SHKZ_HIP(hipMemcpyAsync(hc.data(), cnt, MAX_LEVELS * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SHKZ_HIP(hipStreamSynchronize(s));
        SHKZ_HIP(hipGetLastError());
        for (int l = 0; l < P.nlev; ++l) C->levels[l].unknowns = hc[l];
        res->n_unknowns = hc[0];
    }
    const double t_build_issue = now_ms();
    if (exp) {
        K.go("export_system", -1, 0, 0, dim3((unsigned)((M + 255) / 256)), dim3(256), k_export_system, G, F0.desc,
             OpDesc<double>{P.dir64, P.gx64, P.gy64, P.gz64}, (const double *)P.b64, exp->diag, exp->gx, exp->gy,
             exp->gz, exp->rhs);
        SHKZ_HIP(hipStreamSynchronize(s));
        SHKZ_HIP(hipGetLastError());
        if (C->timing) resolve_timing(C);
        return SHKZ_OK;
    }

    // ---------------------------------------------------------------- solve
    SolveScalars hs{};
    const int norm = prm->norm == SHKZ_NORM_L2 ? 1 : 0;
    const int maxit = prm->max_iterations > 0 ? prm->max_iterations : 30000;
    const double residual = prm->residual < 1e-30 ? 1e-30 : prm->residual;
    OpDesc<double> A64{P.dir64, P.gx64, P.gy64, P.gz64};
    const dim3 gb(g0), bb(BLOCK);
    SHKZ_TRY(upload_scalars(C, s, P.S_out, norm, maxit));
    const double t_build = now_ms();
    int total_iters = 0, outer = 0;
    bool converged = false;
    int status = SHKZ_OK;
    if (res->n_unknowns == 0) {
        converged = true;
        outer = 0;
        hs.r0inf = hs.r0l2 = 0.0;
    } else if (!P.mixed) {
        // one FP64 Krylov solve: CG (reference iteration) or MGPCG
        K.go("defect", 0, 7 * 8, 0, gb, bb, k_defect<double>, F0.desc, A64, (const double *)P.b64,
             (const double *)P.x64, P.r64, P.red);
        finalize(K, P, g0, P.S_out, FIN_DEFECT_FIRST, residual);
        SHKZ_TRY(read_scalars(C, s, P.S_out, &hs));
        if (!hs.done) {
            SHKZ_TRY(krylov<double>(C, K, P, P.x64, P.r64, (double *)P.kp, (double *)P.kq, (double *)P.kz,
                                    (double *)P.ktmp, P.S_out, prm->solver == SHKZ_SOLVER_MGPCG, &hs));
        }
        total_iters = hs.iter;
        converged = hs.converged != 0;
        outer = 1;
        if (hs.status) status = SHKZ_ERR_NUMERIC;
    } else {
        // FP64 defect correction around FP32 MGPCG solves
        const double eta = prm->inner_residual > 0 ? prm->inner_residual : 1e-3;
        SolveScalars hi{};
        SHKZ_TRY(upload_scalars(C, s, P.S_in, norm, maxit));
        double prev = -1.0;
        for (outer = 0; outer < 64; ++outer) {
            K.go("defect", 0, 7 * 8, 0, gb, bb, k_defect<double>, F0.desc, A64, (const double *)P.b64,
                 (const double *)P.x64, P.r64, P.red);
            finalize(K, P, g0, P.S_out, outer == 0 ? FIN_DEFECT_FIRST : FIN_DEFECT, residual);
            SHKZ_TRY(read_scalars(C, s, P.S_out, &hs));
            if (hs.status) {
                status = SHKZ_ERR_NUMERIC;
                break;
            }
            const double cur = norm == 0 ? hs.rinf : hs.rl2;
            if (outer == 0 && hs.done) {
                converged = hs.converged != 0;
                break;
            }
            if (cur <= hs.tol) {
                converged = true;
                break;
            }
            if (total_iters >= maxit) break;
            if (prev >= 0 && cur > 0.9 * prev) break;  // FP32 inner solves no longer reduce the defect
            prev = cur;
            K.go("inner_init", 0, 16, 0, gb, bb, k_inner_init, F0.desc, (const double *)P.r64, (float *)P.kr,
                 (float *)P.kx, P.red);
            finalize(K, P, g0, P.S_in, FIN_INNER_INIT, std::max(eta, hs.tol / std::max(cur, 1e-300)),
                     maxit - total_iters);
            SHKZ_TRY(read_scalars(C, s, P.S_in, &hi));
            if (!hi.done) {
                SHKZ_TRY(krylov<float>(C, K, P, (float *)P.kx, (float *)P.kr, (float *)P.kp, (float *)P.kq,
                                       (float *)P.kz, (float *)P.ktmp, P.S_in, true, &hi));
            }
            total_iters += hi.iter;
            if (C->trace)
                std::fprintf(stderr, "[shkz] outer %d: defect inf %.6e (tol %.3e)  inner it %d rinf %.6e r0 %.6e tol %.3e rz %.9e\n",
                             outer, hs.rinf, hs.tol, hi.iter, hi.rinf, hi.r0inf, hi.tol, hi.rz);
            if (hi.status) {
                status = SHKZ_ERR_NUMERIC;
                break;
            }
            K.go("add_correction", 0, 20, 0, gb, bb, k_add_correction, F0.desc, P.x64, (const float *)P.kx);
        }
        outer = std::max(outer, 1);
    }
    // true FP64 residual of the returned pressure
    if (res->n_unknowns > 0) {
        SolveScalars hf{};
        K.go("defect", 0, 7 * 8, 0, gb, bb, k_defect<double>, F0.desc, A64, (const double *)P.b64,
             (const double *)P.x64, P.r64, P.red);
        finalize(K, P, g0, P.S_in, FIN_DEFECT);
        SHKZ_TRY(read_scalars(C, s, P.S_in, &hf));
        const double r0i = hs.r0inf > 0 ? hs.r0inf : 1.0, r0l = hs.r0l2 > 0 ? hs.r0l2 : 1.0;
        res->reresid_inf = hf.rinf / r0i;
        res->reresid_l2 = hf.rl2 / r0l;
        res->reresid = norm == 0 ? res->reresid_inf : res->reresid_l2;
        if (!P.mixed) converged = hs.converged != 0;
    }
    const double t_solve = now_ms();

    // ---------------------------------------------------------------- scatter + update
    if (grid->pressure || grid->pressure64)
        K.go("scatter_pressure", -1, 0, (double)M * ((grid->pressure ? 4 : 0) + (grid->pressure64 ? 8 : 0)),
             dim3((unsigned)((M + 255) / 256)), dim3(256), k_scatter_pressure, G, F0.desc, (const double *)P.x64,
             grid->pressure, grid->pressure64);
    if (scene)
        K.go("scatter_pressure", -1, 0, (double)M * 8, dim3((unsigned)((M + 255) / 256)), dim3(256), k_scatter_pressure,
             G, F0.desc, (const double *)P.x64, (float *)nullptr, scene_p);
    for (int dim = 0; dim < 3; ++dim) {
        const size_t nf = (size_t)(w + (dim == 0)) * (h + (dim == 1)) * (d + (dim == 2));
        K.go("velocity_update", -1, 0, (double)nf * (8 + (G.phi64 ? 16 : 8) + (grid->solid ? 4 : 0)),
             dim3((unsigned)((nf + 255) / 256)), dim3(256), k_velocity_update, G, dim, F0.desc, (const double *)P.x64);
    }
    if (scene) scene->has_pressure = true;
    if (C->timing) resolve_timing(C);
    const double t_end = now_ms();
    res->converged = converged ? 1 : 0;
    res->iterations = total_iters;
    res->outer_iterations = outer;
    res->status = status;
    res->t_solid_fluid_fractions = t_frac - t_start;
    res->t_build_highres_linsystem = t_build - t_frac;
    res->t_linsolve = t_solve - t_build;
    res->t_update_velocity = t_end - t_solve;
    res->t_projection = t_end - t_start;
    (void)t_build_issue;
    if (status != SHKZ_OK) return fail(status, "non-finite value met in the linear solve");
    return SHKZ_OK;
}

}  // namespace shkz

// ============================================================================ C ABI
extern "C" {

int shkz_abi_version(void) { return SHKZ_ABI_VERSION; }

const char *shkz_last_error(void) { return g_last_error.c_str(); }

void shkz_default_params(shkz_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->residual = 1e-4;
    p->norm = SHKZ_NORM_INF;
    p->max_iterations = 30000;
    p->solver = SHKZ_SOLVER_MGPCG;
    p->precision = SHKZ_PREC_MIXED;
    p->eps_fluid = 1e-2;
    p->eps_solid = 1e-2;
    p->second_order_fluid = 1;
    p->second_order_solid = 1;
    p->gain = 1.0;
}

int shkz_create(int device, uint64_t hbm_budget_bytes, shkz_ctx **out) {
    if (!out) return fail(SHKZ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    SHKZ_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SHKZ_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
    SHKZ_HIP(hipSetDevice(device));
    size_t free_b = 0, total_b = 0;
    SHKZ_HIP(hipMemGetInfo(&free_b, &total_b));
    shkz_ctx *C = new shkz_ctx();
    C->device = device;
    C->arena.budget = hbm_budget_bytes ? (size_t)hbm_budget_bytes : (size_t)(0.9 * (double)total_b);
    if (hipHostMalloc((void **)&C->host_scalars, sizeof(SolveScalars), hipHostMallocDefault) != hipSuccess) {
        delete C;
        return fail(SHKZ_ERR_NOMEM, "pinned scalar buffer allocation failed");
    }
    if (hipStreamCreateWithFlags(&C->work, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_out, hipEventDisableTiming) != hipSuccess) {
        (void)hipHostFree(C->host_scalars);
        delete C;
        return fail(SHKZ_ERR_DEVICE, "stream/event creation failed");
    }
    const char *ng = std::getenv("SHKZ_NO_GRAPHS");
    C->use_graphs = !(ng && ng[0] == '1');
    const char *tr = std::getenv("SHKZ_TRACE");
    C->trace = tr && tr[0] == '1';
    *out = C;
    return SHKZ_OK;
}

void shkz_destroy(shkz_ctx *C) {
    if (!C) return;
    (void)hipSetDevice(C->device);
    (void)hipDeviceSynchronize();
    for (auto &t : C->pending) {
        (void)hipEventDestroy(t.ev0);
        (void)hipEventDestroy(t.ev1);
    }
    for (hipEvent_t e : C->event_pool) (void)hipEventDestroy(e);
    for (auto &g : C->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (C->ev_in) (void)hipEventDestroy(C->ev_in);
    if (C->ev_out) (void)hipEventDestroy(C->ev_out);
    if (C->work) (void)hipStreamDestroy(C->work);
    C->arena.clear();
    if (C->host_scalars) (void)hipHostFree(C->host_scalars);
    delete C;
}

int shkz_project_device(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result,
                        void *stream) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    shkz_params def;
    shkz_default_params(&def);
    const shkz_params *p = params ? params : &def;
    std::memset(result, 0, sizeof(*result));
    std::lock_guard<std::mutex> lock(C->mu);
    int rc = project_impl(C, grid, p, result, (hipStream_t)stream);
    result->status = rc;
    return rc;
}

static int project_host(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result,
                        double *const *export_host);

int shkz_project(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result) {
    if (result) std::memset(result, 0, sizeof(*result));
    const int rc = project_host(C, grid, params, result, nullptr);
    if (result) result->status = rc;
    return rc;
}

int shkz_export_system(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, double *diag, double *gx,
                       double *gy, double *gz, double *rhs) {
    if (!diag || !gx || !gy || !gz || !rhs) return fail(SHKZ_ERR_INVALID, "export arrays are required");
    double *outs[5] = {diag, gx, gy, gz, rhs};
    shkz_result r;
    return project_host(C, grid, params, &r, outs);
}

static int project_host(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result,
                        double *const *export_host) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    if (grid->w <= 0 || grid->h <= 0 || grid->d <= 0) return fail(SHKZ_ERR_INVALID, "grid dimensions must be positive");
    if (!grid->fluid || !grid->u || !grid->v || !grid->w_vel) return fail(SHKZ_ERR_INVALID, "fluid and velocities are required");
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    Arena &AR = C->arena;
    const size_t w = grid->w, h = grid->h, d = grid->d;
    const size_t M = w * h * d, Nn = (w + 1) * (h + 1) * (d + 1);
    const size_t nf[3] = {(w + 1) * h * d, w * (h + 1) * d, w * h * (d + 1)};
    shkz_grid dg = *grid;
    void *phi = nullptr, *solid = nullptr;
    const size_t pe = grid->fluid_dtype == SHKZ_F64 ? 8 : 4, se = grid->solid_dtype == SHKZ_F64 ? 8 : 4;
    SHKZ_TRY(AR.get("h_phi", M * pe, &phi));
    SHKZ_HIP(hipMemcpy(phi, grid->fluid, M * pe, hipMemcpyHostToDevice));
    dg.fluid = phi;
    if (grid->solid) {
        SHKZ_TRY(AR.get("h_solid", Nn * se, &solid));
        SHKZ_HIP(hipMemcpy(solid, grid->solid, Nn * se, hipMemcpyHostToDevice));
        dg.solid = solid;
    }
    float *hv[3] = {grid->u, grid->v, grid->w_vel};
    const uint8_t *ha[3] = {grid->active_u, grid->active_v, grid->active_w};
    float *dv[3];
    uint8_t *da[3] = {nullptr, nullptr, nullptr};
    const char *vn[3] = {"h_u", "h_v", "h_w"}, *an[3] = {"h_au", "h_av", "h_aw"};
    for (int dim = 0; dim < 3; ++dim) {
        void *pp = nullptr;
        SHKZ_TRY(AR.get(vn[dim], nf[dim] * 4, &pp));
        dv[dim] = (float *)pp;
        SHKZ_HIP(hipMemcpy(dv[dim], hv[dim], nf[dim] * 4, hipMemcpyHostToDevice));
        if (ha[dim]) {
            SHKZ_TRY(AR.get(an[dim], nf[dim], &pp));
            da[dim] = (uint8_t *)pp;
            SHKZ_HIP(hipMemcpy(da[dim], ha[dim], nf[dim], hipMemcpyHostToDevice));
        }
    }
    dg.u = dv[0];
    dg.v = dv[1];
    dg.w_vel = dv[2];
    dg.active_u = da[0];
    dg.active_v = da[1];
    dg.active_w = da[2];
    float *dp = nullptr;
    double *dp64 = nullptr;
    if (grid->pressure) {
        void *pp = nullptr;
        SHKZ_TRY(AR.get("h_p", M * 4, &pp));
        dp = (float *)pp;
    }
    if (grid->pressure64) {
        void *pp = nullptr;
        SHKZ_TRY(AR.get("h_p64", M * 8, &pp));
        dp64 = (double *)pp;
    }
    dg.pressure = dp;
    dg.pressure64 = dp64;
    shkz_params def;
    shkz_default_params(&def);
    if (export_host) {
        double *ex[5];
        for (int q = 0; q < 5; ++q) {
            void *pp = nullptr;
            SHKZ_TRY(AR.get("h_export" + std::to_string(q), M * 8, &pp));
            ex[q] = (double *)pp;
        }
        ExportOut eo{ex[0], ex[1], ex[2], ex[3], ex[4]};
        int rc = project_impl(C, &dg, params ? params : &def, result, nullptr, &eo);
        if (rc != SHKZ_OK) return rc;
        for (int q = 0; q < 5; ++q) SHKZ_HIP(hipMemcpy(export_host[q], ex[q], M * 8, hipMemcpyDeviceToHost));
        return SHKZ_OK;
    }
    int rc = project_impl(C, &dg, params ? params : &def, result, nullptr);
    result->status = rc;
    if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) return rc;
    for (int dim = 0; dim < 3; ++dim) SHKZ_HIP(hipMemcpy(hv[dim], dv[dim], nf[dim] * 4, hipMemcpyDeviceToHost));
    if (grid->pressure) SHKZ_HIP(hipMemcpy(grid->pressure, dp, M * 4, hipMemcpyDeviceToHost));
    if (grid->pressure64) SHKZ_HIP(hipMemcpy(grid->pressure64, dp64, M * 8, hipMemcpyDeviceToHost));
    return rc;
}

int shkz_release_scene(shkz_ctx *C, uint64_t scene_id) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    auto it = C->scenes.find(scene_id);
    if (it == C->scenes.end()) return fail(SHKZ_ERR_NOT_FOUND, "scene %llu not found", (unsigned long long)scene_id);
    C->scenes.erase(it);
    C->arena.release("scene" + std::to_string(scene_id));
    return SHKZ_OK;
}

uint64_t shkz_device_bytes(const shkz_ctx *C) { return C ? (uint64_t)C->arena.total : 0; }

int shkz_set_kernel_timing(shkz_ctx *C, int enable) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    C->timing = enable != 0;
    return SHKZ_OK;
}

int shkz_get_kernel_stats(shkz_ctx *C, shkz_kernel_stat *out, int max_entries, int *n_out) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    if (!C->pending.empty()) {
        SHKZ_HIP(hipDeviceSynchronize());
        resolve_timing(C);
    }
    int n = 0;
    for (const KernelTotals &K : C->totals) {
        if (K.launches == 0) continue;
        if (out && n < max_entries) {
            std::memset(&out[n], 0, sizeof(shkz_kernel_stat));
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", K.name.c_str());
            out[n].launches = K.launches;
            out[n].total_ms = K.ms;
            out[n].bytes = K.bytes;
        }
        ++n;
    }
    if (n_out) *n_out = n;
    return SHKZ_OK;
}

int shkz_reset_kernel_stats(shkz_ctx *C) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    if (!C->pending.empty()) {
        SHKZ_HIP(hipDeviceSynchronize());
        resolve_timing(C);
    }
    C->totals.clear();
    return SHKZ_OK;
}

}  // extern "C"
