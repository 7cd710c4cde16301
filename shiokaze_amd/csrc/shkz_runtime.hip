// shkz_runtime.hip -- host runtime of the MI355X pressure projection and its C ABI.
//
// One projection of one grid is one macpressuresolver3::project (src/projection/
// macpressuresolver3.cpp:50-272):
//   fractions + labelling          k_flags (+ k_any_negative with surface tension) (:61-156)
//   assembly (matrix-free)         k_operator, k_coarsen, k_idiag        (:159-199)
//   volume correction, warm start  host PI controller, scene state       (:203-230, :239-242)
//   linear solve                   CG / MGPCG / mixed-precision MGPCG    (:233-237 -> pcg.cpp:45-73)
//   scatter + velocity update      k_scatter_pressure, k_velocity_update (:245-268)
// A call may project a batch of independent grids at once.  Their stored bricks are concatenated
// into one tile table per multigrid level, so every solver launch covers every grid of the batch,
// while reductions, scalars, convergence and the multigrid depth stay per grid: each grid's result
// is the one it gets when projected alone.  A single projection is a batch of one.
//
// The context owns every device buffer (grown on demand, never freed per call), the hierarchy of
// the last batch, per-scene resident state, the replayable iteration graphs and the kernel timers.
#include "../../include/shkz_mi355x.h"
#include "shkz_kernels.hpp"
#include "shkz_extrap.hpp"
#include "shkz_advect.hpp"
#include "shkz_serve.hpp"
#include "shkz_error.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
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

}  // namespace shkz

// Error recording for the library's other translation units (shkz_error.hpp).
int shkz_set_error(int code, const char *msg) {
    shkz::g_last_error = msg;
    return code;
}

namespace shkz {

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

// Named device buffers, grown with 1/8 headroom in 2 MiB steps and kept across calls, under a
// byte budget shared with what `extra` reports (resident scene state, other workspaces).
struct Arena {
    std::unordered_map<std::string, DevBuf> bufs;
    std::atomic<uint64_t> total{0};  // read by other threads of a server
    uint64_t budget = 0;
    std::function<uint64_t()> extra;
    // test hook (shkz_inject_alloc_failure): the allocation after `fail_after` more successful ones
    // fails as a failed hipMalloc would; -1 = off
    int64_t fail_after = -1;

    int get(const std::string &name, size_t bytes, void **out) {
        DevBuf &b = bufs[name];
        if (bytes == 0) bytes = 256;
        if (b.cap < bytes) {
            size_t want = bytes + bytes / 8;
            want = (want + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
            const uint64_t other = extra ? extra() : 0;
            if (total - b.cap + want + other > budget)
                return fail(SHKZ_ERR_BUDGET, "device budget exceeded: %s needs %zu bytes (held %llu + %llu of %llu)",
                            name.c_str(), want, (unsigned long long)total.load(), (unsigned long long)other,
                            (unsigned long long)budget);
            if (b.ptr) {
                SHKZ_HIP(hipFree(b.ptr));
                total -= b.cap;
                b.ptr = nullptr;
                b.cap = 0;
            }
            if (fail_after == 0) {
                fail_after = -1;
                return fail(SHKZ_ERR_NOMEM, "hipMalloc(%zu) for %s failed: injected allocation failure", want,
                            name.c_str());
            }
            if (fail_after > 0) --fail_after;
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

// ------------------------------------------------------------------------ hierarchy (host side)
// One grid at one level, before concatenation.  A grid that ended on a coarser level is empty here.
struct GridLevelHost {
    int ni = 0, nj = 0, nk = 0, nbx = 0, nby = 0, nbz = 0;
    std::vector<uint8_t> active;  // per brick position: holds unknowns
    bool single() const { return nbx <= 1 && nby <= 1 && nbz <= 1; }
    bool any_active() const {
        for (uint8_t a : active)
            if (a) return true;
        return false;
    }
    // a grid is coarsened further while it has unknowns and spans more than one brick
    bool coarsens() const { return !single() && any_active(); }
};

static void grid_level_dims(GridLevelHost &L, int ni, int nj, int nk) {
    L.ni = ni;
    L.nj = nj;
    L.nk = nk;
    L.nbx = (ni + BRICK - 1) / BRICK;
    L.nby = (nj + BRICK - 1) / BRICK;
    L.nbz = (nk + BRICK - 1) / BRICK;
}

// One multigrid level of the batch: every grid's stored bricks, concatenated.
struct LevelHost {
    int nbricks = 0;
    bool any_coarsest = false;  // some grid ends on this level
    std::vector<int> bmap, bxyz, bgrid;
    std::vector<GridLevel> gl;
    std::vector<BrickMeta> meta;
    std::vector<int> child64, parent8;  // restriction / prolongation brick tables (multigrid only)
    std::vector<int> nb27;              // 3x3x3 brick neighbourhoods (coarse levels only)
    LevelDesc desc{};
    // operator and V-cycle vectors (element type = level precision)
    void *dir = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr, *idiag = nullptr;
    void *rhs = nullptr, *z = nullptr, *tmp = nullptr, *res = nullptr;
    int16_t *ccode = nullptr;  // coarsest bricks: component codes (k_coarse_components)
    int64_t unknowns = 0;
    int64_t uniform_cells = 0;  // cells of BM_UNIFORM bricks (level 0)
    std::vector<int64_t> unknowns_grid;
};

// n default levels, whose host tables reuse the storage of the levels in `levels` (emptied).
static void recycle_levels(std::vector<LevelHost> &levels, int n) {
    std::vector<LevelHost> out(n);
    for (int l = 0; l < n && l < (int)levels.size(); ++l) {
        LevelHost &F = out[l], &O = levels[l];
        auto take = [](auto &dst, auto &src) {
            dst = std::move(src);
            dst.clear();
        };
        take(F.bmap, O.bmap);
        take(F.bxyz, O.bxyz);
        take(F.bgrid, O.bgrid);
        take(F.gl, O.gl);
        take(F.meta, O.meta);
        take(F.child64, O.child64);
        take(F.parent8, O.parent8);
        take(F.nb27, O.nb27);
        take(F.unknowns_grid, O.unknowns_grid);
    }
    levels.swap(out);
}

// Concatenates the grids' active bricks (Morton order inside each grid) and builds the dense maps
// and the neighbour table of the level.
// Depth-first walk of the octree over one grid's brick positions, children in Morton order (x the
// most significant bit of each triple): assigns consecutive brick ids to the active positions.
// An octant of 2^m x 2^m x 2^m positions is skipped when the grid's mask m levels coarser (when
// the hierarchy has it) shows no active position in it: a coarse position is active exactly when
// one of its children is.
struct MortonWalk {
    const std::vector<std::vector<GridLevelHost>> &levels;
    int level;  // the level of H in levels
    const GridLevelHost &H;
    int *bmap;   // the grid's dense map
    int *xyz;    // the level's brick positions (3 per brick)
    int *bgrid;  // the level's brick -> grid
    int grid, next;
    void leaf(int x, int y, int z) {
        const int b = (x * H.nby + y) * H.nbz + z;
        if (!H.active[b]) return;
        bmap[b] = next;
        xyz[3 * next] = x;
        xyz[3 * next + 1] = y;
        xyz[3 * next + 2] = z;
        bgrid[next] = grid;
        ++next;
    }
    void walk(int x0, int y0, int z0, int m) {
        if (x0 >= H.nbx || y0 >= H.nby || z0 >= H.nbz) return;
        if (m == 0) {
            leaf(x0, y0, z0);
            return;
        }
        if (level + m < (int)levels.size()) {
            const GridLevelHost &C = levels[level + m][grid];
            if (!C.active.empty() && !C.active[((x0 >> m) * C.nby + (y0 >> m)) * C.nbz + (z0 >> m)]) return;
        }
        if (m == 1) {  // the eight positions of the octant, in Morton order
            for (int c = 0; c < 8; ++c) {
                const int x = x0 + ((c >> 2) & 1), y = y0 + ((c >> 1) & 1), z = z0 + (c & 1);
                if (x < H.nbx && y < H.nby && z < H.nbz) leaf(x, y, z);
            }
            return;
        }
        const int h = 1 << (m - 1);
        for (int c = 0; c < 8; ++c) walk(x0 + ((c >> 2) & 1) * h, y0 + ((c >> 1) & 1) * h, z0 + (c & 1) * h, m - 1);
    }
};

static void build_level(LevelHost &L, const std::vector<std::vector<GridLevelHost>> &levels, int level) {
    const std::vector<GridLevelHost> &grids = levels[level];
    const int G = (int)grids.size();
    L.gl.assign(G, GridLevel{});
    L.any_coarsest = false;
    // sizes first: the dense maps of all grids, and the active positions (the level's bricks)
    size_t nmap = 0, nact = 0;
    for (int g = 0; g < G; ++g) {
        const GridLevelHost &H = grids[g];
        nmap += (size_t)H.nbx * H.nby * H.nbz;
        for (uint8_t a : H.active) nact += a != 0;
    }
    L.bmap.assign(nmap, -1);
    L.bxyz.resize(3 * nact);
    L.bgrid.resize(nact);
    int nb = 0;
    size_t off = 0;
    for (int g = 0; g < G; ++g) {
        const GridLevelHost &H = grids[g];
        GridLevel &gl = L.gl[g];
        gl.ni = H.ni;
        gl.nj = H.nj;
        gl.nk = H.nk;
        gl.nbx = H.nbx;
        gl.nby = H.nby;
        gl.nbz = H.nbz;
        gl.bmap_off = (int)off;
        gl.brick_start = nb;
        const int npos = H.nbx * H.nby * H.nbz;
        // active positions in Morton order (x the most significant bit of each triple): a depth-
        // first walk of the octree over the brick grid, skipping octants outside it or with no
        // active position, visits them in that order (no sort)
        int m = 0;
        while ((1 << m) < H.nbx || (1 << m) < H.nby || (1 << m) < H.nbz) ++m;
        MortonWalk W{levels, level, H, L.bmap.data() + off, L.bxyz.data(), L.bgrid.data(), g, nb};
        if (npos > 0) W.walk(0, 0, 0, m);
        nb = W.next;
        off += (size_t)npos;
        gl.brick_end = nb;
        gl.coarsest = (npos > 0 && H.single()) ? 1 : 0;
        if (gl.coarsest && gl.brick_end > gl.brick_start) L.any_coarsest = true;
    }
    L.nbricks = nb;
    // the neighbour table, grid by grid (raw pointers: no aliasing between the map and the records)
    L.meta.resize(nb);
    const int *__restrict__ xyz = L.bxyz.data();
    BrickMeta *__restrict__ meta = L.meta.data();
    for (int g = 0; g < G; ++g) {
        const GridLevel &gl = L.gl[g];
        const int *__restrict__ gm = L.bmap.data() + gl.bmap_off;
        const int sy = gl.nbz, sx = gl.nby * gl.nbz;
        const int flags = gl.coarsest ? BM_COARSEST : 0;
        for (int id = gl.brick_start; id < gl.brick_end; ++id) {
            const int x = xyz[3 * id], y = xyz[3 * id + 1], z = xyz[3 * id + 2];
            const int p = x * sx + y * sy + z;
            BrickMeta m;
            m.nbr[NB_XM] = x > 0 ? gm[p - sx] : -1;
            m.nbr[NB_XP] = x + 1 < gl.nbx ? gm[p + sx] : -1;
            m.nbr[NB_YM] = y > 0 ? gm[p - sy] : -1;
            m.nbr[NB_YP] = y + 1 < gl.nby ? gm[p + sy] : -1;
            m.nbr[NB_ZM] = z > 0 ? gm[p - 1] : -1;
            m.nbr[NB_ZP] = z + 1 < gl.nbz ? gm[p + 1] : -1;
            m.grid = g;
            m.flags = flags;
            meta[id] = m;
        }
    }
    L.child64.clear();
    L.parent8.clear();
    L.nb27.clear();
}

// Brick of grid g at brick coordinates (x, y, z) of level L, -1 outside the grid or unstored.
static int brick_at(const LevelHost &L, int g, int x, int y, int z) {
    const GridLevel &gl = L.gl[g];
    if (x < 0 || y < 0 || z < 0 || x >= gl.nbx || y >= gl.nby || z >= gl.nbz) return -1;
    return L.bmap[gl.bmap_off + (x * gl.nby + y) * gl.nbz + z];
}

// Restriction and prolongation tables between level F and the next coarser level Cl: every
// coarse brick B lists the fine bricks 2B-1 .. 2B+2 per axis, every fine brick B the coarse
// bricks (B+1)/2 - 1 + {0,1} per axis (bricks of the same grid; -1 where none is stored).
static void link_levels(LevelHost &F, LevelHost &Cl) {
    Cl.child64.assign((size_t)64 * Cl.nbricks, -1);
    for (int id = 0; id < Cl.nbricks; ++id) {
        const int g = Cl.bgrid[id];
        const int X = Cl.bxyz[3 * id], Y = Cl.bxyz[3 * id + 1], Z = Cl.bxyz[3 * id + 2];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                for (int c = 0; c < 4; ++c)
                    Cl.child64[64 * (size_t)id + (a * 4 + b) * 4 + c] =
                        brick_at(F, g, 2 * X - 1 + a, 2 * Y - 1 + b, 2 * Z - 1 + c);
    }
    F.parent8.assign((size_t)8 * F.nbricks, -1);
    for (int id = 0; id < F.nbricks; ++id) {
        const int g = F.bgrid[id];
        if (F.gl[g].coarsest) continue;
        const int X = F.bxyz[3 * id], Y = F.bxyz[3 * id + 1], Z = F.bxyz[3 * id + 2];
        const int x0 = (X + 1) / 2 - 1, y0 = (Y + 1) / 2 - 1, z0 = (Z + 1) / 2 - 1;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c)
                    F.parent8[8 * (size_t)id + (a * 2 + b) * 2 + c] = brick_at(Cl, g, x0 + a, y0 + b, z0 + c);
    }
}

// Every brick's 3x3x3 neighbourhood (bricks B-1 .. B+1 per axis of its grid, -1 where none is
// stored), in the order k_smooth_k stages it: ((dx + 1) * 3 + (dy + 1)) * 3 + (dz + 1).
static void neighbourhoods(LevelHost &L) {
    L.nb27.assign((size_t)27 * L.nbricks, -1);
    for (int id = 0; id < L.nbricks; ++id) {
        const int g = L.bgrid[id];
        const int X = L.bxyz[3 * id], Y = L.bxyz[3 * id + 1], Z = L.bxyz[3 * id + 2];
        for (int t = 0; t < 27; ++t)
            L.nb27[27 * (size_t)id + t] = brick_at(L, g, X + t / 9 - 1, Y + (t / 3) % 3 - 1, Z + t % 3 - 1);
    }
}

// ------------------------------------------------------------------------ timing
struct TimedLaunch {
    int kernel;  // index into names
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
// State a scene keeps on the device between projections: the previous pressure (warm start,
// macpressuresolver3.cpp:239-242, kept cell-indexed) and the volume-correction integrator (:208).
struct SceneState {
    int w = 0, h = 0, d = 0;
    bool has_pressure = false;
    double y_prev = 0.0;
    double *p = nullptr;  // device, w*h*d
    size_t cap = 0;       // bytes held at p
};

// Scene states by id.  A context owns one; the contexts of a server share the server's (a scene is
// used by one projection at a time: the server never runs two requests of a scene together).
struct SceneStore {
    std::mutex mu;
    std::unordered_map<uint64_t, SceneState> map;  // node-based: element addresses are stable
    std::atomic<uint64_t> bytes{0};
    std::function<bool(uint64_t)> may_grow;  // budget test for `want` more bytes

    // The scene's state with a pressure buffer for a w*h*d grid (reset when the shape changed).
    int acquire(uint64_t id, int w, int h, int d, SceneState **out) {
        std::lock_guard<std::mutex> g(mu);
        SceneState &S = map[id];
        if (S.w != w || S.h != h || S.d != d) {
            S.w = w;
            S.h = h;
            S.d = d;
            S.has_pressure = false;
        }
        const size_t want = (size_t)w * h * d * sizeof(double);
        if (S.cap < want) {
            if (may_grow && !may_grow(want - S.cap))
                return fail(SHKZ_ERR_BUDGET, "device budget exceeded by the state of scene %llu (%zu bytes)",
                            (unsigned long long)id, want);
            if (S.p) {
                SHKZ_HIP(hipFree(S.p));
                bytes -= S.cap;
                S.p = nullptr;
                S.cap = 0;
            }
            SHKZ_HIP(hipMalloc(&S.p, want));
            S.cap = want;
            bytes += want;
        }
        *out = &S;
        return SHKZ_OK;
    }
    int release(uint64_t id) {
        std::lock_guard<std::mutex> g(mu);
        auto it = map.find(id);
        if (it == map.end()) return fail(SHKZ_ERR_NOT_FOUND, "scene %llu not found", (unsigned long long)id);
        if (it->second.p) {
            (void)hipFree(it->second.p);
            bytes -= it->second.cap;
        }
        map.erase(it);
        return SHKZ_OK;
    }
    void clear() {
        std::lock_guard<std::mutex> g(mu);
        for (auto &kv : map)
            if (kv.second.p) (void)hipFree(kv.second.p);
        map.clear();
        bytes = 0;
    }
};

}  // namespace shkz

using namespace shkz;

struct shkz_ctx {
    int device = 0;
    Arena arena;
    std::mutex mu;  // one projection call at a time per context
    SceneStore own_scenes;
    SceneStore *scenes = &own_scenes;  // a server points its contexts at its shared store
    double y_prev = 0.0;  // stateless calls share the module-level integrator, as m_y_prev does
    // hierarchy of the current batch
    std::vector<LevelHost> levels;
    // pinned read-back / upload of per-grid scalars
    SolveScalars *host_scalars = nullptr;
    int host_scalars_cap = 0;
    BatchCtl *host_ctl = nullptr;
    // pinned staging for the per-call transfers (brick flags down, level tables up, counts down)
    struct Pinned {
        void *p = nullptr;
        size_t cap = 0;
    } pin_flags, pin_meta0, pin_meta, pin_links, pin_counts;
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
    bool host_links = false;  // SHKZ_HOST_LINKS=1: transfer tables built on the host (a cross-check)
    // mixed precision: each defect-correction pass after the first continues the previous pass's
    // search direction (p = z + beta p with the replaced residual) instead of restarting CG
    bool keep_direction = false;
    // FP32 Krylov iterations with the direction update fused into the operator apply and the
    // first finest-level sweep fused into the residual update (SHKZ_NO_FUSE=1: separate kernels;
    // the results are the same)
    bool fuse_krylov = true;
    // coarse levels of at most this many stored bricks smooth with multi-sweep launches unless the
    // autotuner picks another limit (SHKZ_MULTI_MAX_BRICKS overrides the default; the results do
    // not depend on it)
    int multi_max_bricks = 1024;
    // rocprofv3's tool library is loaded (it exports its options to the process environment): the
    // autotuner does not run.  Two kernel-trace runs that tuned under the profiler died with a
    // segmentation fault inside the tool library's copy routine, called from a HIP call of the
    // solve; the same tuning runs cleanly without the profiler.
    bool under_profiler = false;
    std::string device_key;  // device identity of the tuning cache
    // replayable PCG iteration graphs, keyed by everything their kernel arguments depend on; an
    // exec is destroyed only to make room (least recently used, after the stream has drained) or
    // with the context, so trying many controls (the autotuner) re-instantiates nothing it keeps
    struct GraphSlot {
        std::string key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    std::vector<GraphSlot> graphs;
    uint64_t graph_clock = 0;
    // iterations each solve of a shape class took (per defect-correction pass, slowest grid): the
    // next solve of the class replays that many iterations before its first poll
    std::unordered_map<std::string, std::vector<int>> iter_hist;
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
    // Launches kernel<<<grid, block>>> on the work stream.  With kernel timing on (and outside a
    // graph capture) the launch carries a start and a stop event that the dispatch itself stamps
    // (hipExtLaunchKernelGGL), so the elapsed time is the kernel's own execution, as a kernel trace
    // reports it, without the cost of separate event packets.  The byte model (per unknown of
    // `level`, plus fixed) feeds the per-kernel achieved-bandwidth statistics.
    template <class K, class... Args>
    void go(const char *name, int level, double bytes_per_unknown, double bytes_fixed, dim3 grid, dim3 block, K kernel,
            Args... args) {
        if (grid.x == 0) return;
        if (C->timing && !capturing) {
            hipEvent_t e0 = event(), e1 = event();
            if (e0 && e1) {
                hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
                C->pending.push_back({kernel_id(name), level, bytes_per_unknown, bytes_fixed, e0, e1});
                return;
            }
            if (e0) C->event_pool.push_back(e0);
            if (e1) C->event_pool.push_back(e1);
        }
        kernel<<<grid, block, 0, s>>>(args...);
    }
};

static inline unsigned brick_grid(int nbricks) { return (unsigned)(((nbricks + 7) / 8) * 8); }
// Blocks of `per` threads over n items, a multiple of 8 (kernels remap block ids XCD-aware and
// bounds-check the index).
static inline unsigned xcd_blocks(size_t n, unsigned per) { return (unsigned)((((n + per - 1) / per) + 7) / 8 * 8); }

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

// Byte model: the bytes a level-0 kernel does not stream on BM_UNIFORM bricks (`per_cell` per
// cell), subtracted from its per-unknown figure; 0 on other levels.
static double usave(const shkz_ctx *C, int level, double per_cell) {
    return level == 0 ? per_cell * (double)C->levels[0].uniform_cells : 0.0;
}

// ------------------------------------------------------------------------ solve plumbing
struct Problem {
    int ngrids = 0;
    int nlev = 0;
    bool mixed = false;
    bool fp64_levels = true;  // V-cycle element type
    // level-0 FP64 arrays
    double *dir64 = nullptr, *gx64 = nullptr, *gy64 = nullptr, *gz64 = nullptr;
    double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr;
    // Krylov vectors of the level-0 solve (element type of the solve)
    void *kr = nullptr, *kp = nullptr, *kq = nullptr, *kz = nullptr, *ktmp = nullptr;
    void *kp2 = nullptr;  // FP32: the other direction buffer of the fused iteration (ping-pong)
    // per-grid scalars: outer (FP64 defect / FP64 solve), inner (FP32 solve), final check
    SolveScalars *S_out = nullptr, *S_in = nullptr, *S_fin = nullptr;
    BatchCtl *ctl = nullptr;  // [0] outer / FP64 solve, [1] inner solve
    RedBuf red{};
    int pre = 4, post = 4, pre_c = 4, post_c = 4, coarse_sweeps = 48;
    int smoother = SHKZ_SMOOTH_JACOBI;  // finest level (SHKZ_SMOOTH_BRICK only with an FP32 V-cycle)
    int local_k = LOCAL_SGS_SWEEPS;      // brick-local smoother: sweeps per finest-level application (2, 3)
    double omega = 0.9;
    int multi_max_bricks = 1024;  // coarse levels up to this size smooth with multi-sweep launches
    double eta = 0.0;  // mixed precision: relative defect reduction asked of each inner solve (0: to the tolerance)
};

template <class T>
static OpDesc<T> op_of(const LevelHost &L) {
    return OpDesc<T>{(const T *)L.dir, (const T *)L.gx, (const T *)L.gy, (const T *)L.gz};
}

// Which of the level's two vectors (z or tmp) holds the smoothed result at the end of the level's
// post-smoothing, given the ping-pong of the sweeps: 0 = z, 1 = tmp.
// multi_pre / multi_post: the level smooths with one k_smooth_k launch (which writes the other
// vector once) instead of per-sweep launches.
static int final_slot(int pre, int post, bool multi_pre, bool multi_post, bool finest) {
    // pre-smoothing ends after pre - 1 swaps on the finest level (one sweep from zero, then plain
    // sweeps), after pre - 2 elsewhere (the fused two-sweep kernel, then plain sweeps)
    const int after_pre = (multi_pre || pre <= 1) ? 0 : (finest ? ((pre - 1) & 1) : ((pre - 2) & 1));
    const int post_swaps = multi_post ? (post > 0 ? 1 : 0) : post;
    return after_pre ^ (post_swaps & 1);
}

// Sweep counts k_smooth_k is built for (its cube of cells must fit LDS).
template <class T>
static bool smooth_k_supported(int k) {
    return k >= 2 && k <= (sizeof(T) == 4 ? 5 : 3);
}

// Sweep counts k_smooth_k's residual variant is built for (one more ghost layer in LDS).
template <class T>
static bool smooth_k_res_supported(int k) {
    return k >= 2 && k <= (sizeof(T) == 4 ? 4 : 2);
}

template <class T, int KK, int FZ, int RES, int PRO = 0>
static void smooth_k_go(Launcher &K, int l, dim3 g, const LevelDesc &L, OpDesc<T> A, const T *id, const T *rhs,
                        const T *zin, T *zout, T om, Gate Q, T *res, const T *zcoarse = nullptr,
                        double coarse_unknowns = 0) {
    // per unknown, the arrays it reads once and writes once (dir and res with the residual, dir
    // with the prolongation, which also reads the coarse correction once)
    const double bytes = (RES ? 9.0 : PRO ? 8.0 : 7.0) * sizeof(T);
    K.go(RES ? "smooth_k_res_Lc" : PRO ? "smooth_k_pro_Lc" : "smooth_k_Lc", l, bytes, coarse_unknowns * sizeof(T), g,
         dim3(1024), k_smooth_k<T, KK, FZ, RES, PRO>, L, A, id, rhs, zin, zout, om, Q, res, zcoarse);
}

// k damped-Jacobi sweeps of one coarse level in one launch (from zero or from zin); with res
// (pre-smoothing from zero, k within smooth_k_res_supported), also the residual of the result;
// with zcoarse (post-smoothing), zin first receives the prolongation of the coarser level's
// correction zcoarse (coarse_unknowns of them, for the byte model).
template <class T>
static void smooth_k(Launcher &K, int l, int k, bool from_zero, const LevelDesc &L, OpDesc<T> A, const T *id,
                     const T *rhs, const T *zin, T *zout, T om, Gate Q, T *res = nullptr,
                     const T *zcoarse = nullptr, double coarse_unknowns = 0) {
    const dim3 g(brick_grid(L.nbricks));
    if (zcoarse && !from_zero) {
        if (k == 2) smooth_k_go<T, 2, 0, 0, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr, zcoarse, coarse_unknowns);
        if (k == 3) smooth_k_go<T, 3, 0, 0, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr, zcoarse, coarse_unknowns);
        if constexpr (sizeof(T) == 4) {
            if (k == 4) smooth_k_go<T, 4, 0, 0, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr, zcoarse, coarse_unknowns);
            if (k == 5) smooth_k_go<T, 5, 0, 0, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr, zcoarse, coarse_unknowns);
        }
        return;
    }
    if (res && from_zero) {
        if (k == 2) smooth_k_go<T, 2, 1, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, res);
        if constexpr (sizeof(T) == 4) {
            if (k == 3) smooth_k_go<T, 3, 1, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, res);
            if (k == 4) smooth_k_go<T, 4, 1, 1>(K, l, g, L, A, id, rhs, zin, zout, om, Q, res);
        }
        return;
    }
    switch (k) {
    case 2:
        if (from_zero) smooth_k_go<T, 2, 1, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
        else smooth_k_go<T, 2, 0, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
        break;
    case 3:
        if (from_zero) smooth_k_go<T, 3, 1, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
        else smooth_k_go<T, 3, 0, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
        break;
    default:
        if constexpr (sizeof(T) == 4) {
            if (k == 4) {
                if (from_zero) smooth_k_go<T, 4, 1, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
                else smooth_k_go<T, 4, 0, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
            } else if (k == 5) {
                if (from_zero) smooth_k_go<T, 5, 1, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
                else smooth_k_go<T, 5, 0, 0>(K, l, g, L, A, id, rhs, zin, zout, om, Q, nullptr);
            }
        }
        break;
    }
}

// Workgroups of the row kernels that take four bricks each (a multiple of 8 of them).
static inline unsigned row4_grid(int nbricks) { return (unsigned)((((nbricks + 3) / 4) + 7) / 8 * 8); }
// Launch shape of the row kernels on a level: four bricks per 512-thread workgroup on the finest
// level, one brick per 128-thread workgroup on the (smaller) coarse levels, for more workgroups.
struct Row4Launch {
    dim3 grid, block;
};
static inline Row4Launch row4_launch(int nbricks, bool finest) {
    if (finest) return Row4Launch{dim3(row4_grid(nbricks)), dim3(BLOCK)};
    return Row4Launch{dim3(brick_grid(nbricks)), dim3(ROW4_THREADS)};
}

// One damped-Jacobi sweep zout = J(zin) of level l (dot: with the rhs.zout partials of the finest
// level's last sweep).  FP32 levels sweep four cells per thread (k_jacobi4*), FP64 levels one.
template <class T>
static void jacobi_sweep(shkz_ctx *C, Launcher &K, int l, bool dot, const LevelDesc &L, OpDesc<T> A, const T *id,
                         const T *rhs, const T *zin, T *zout, T om, Gate Q, RedBuf red) {
    const bool f = l == 0;
    const double es = (double)sizeof(T);
    const char *name = dot ? "jacobi_dot_L0" : f ? "jacobi_L0" : "jacobi_Lc";
    const unsigned nb = brick_grid(L.nbricks);
    if constexpr (sizeof(T) == 4) {
        if (dot) {
            K.go(name, l, 7 * es, -usave(C, l, 4 * es), dim3(nb), dim3(ROW4_THREADS), k_jacobi4_dot, L, A, id, rhs, zin,
                 zout, om, Q, red);
        } else {
            const Row4Launch R4 = row4_launch(L.nbricks, f);
            K.go(name, l, 7 * es, -usave(C, l, 4 * es), R4.grid, R4.block, f ? k_jacobi4<1> : k_jacobi4<0>, L, A, id,
                 rhs, zin, zout, om, Q, (int)R4.grid.x);
        }
    } else if (dot) {
        // one brick per workgroup: the reduction of the dot variant costs two barriers per brick
        K.go(name, l, 7 * es, -usave(C, l, 4 * es), dim3(nb), dim3(BLOCK), k_jacobi<T, 1, 1>, L, A, id, rhs, zin, zout,
             om, Q, red, (int)nb);
    } else {
        K.go(name, l, 7 * es, -usave(C, l, 4 * es), dim3(std::min<unsigned>(nb, PERSIST_BLOCKS)), dim3(BLOCK),
             f ? k_jacobi<T, 1, 0> : k_jacobi<T, 0, 0>, L, A, id, rhs, zin, zout, om, Q, red, (int)nb);
    }
}

// One V-cycle z = M r on the level-0 vectors of every grid not done; returns the pointer holding
// z.  Each grid descends to its own coarsest level (one brick), solved there by k_coarse_solve;
// the smoothing kernels skip the bricks of grids that end on their level.  With dot_partials, the
// last post-smoothing sweep on the finest level (or the coarse solve of a one-brick grid) also
// stores r.z partials in P.red.
template <class T>
static T *vcycle(shkz_ctx *C, Launcher &K, Problem &P, const T *rhs0, T *z0, T *tmp0, Gate Q, bool dot_partials,
                 bool first_done = false) {
    const int nl = P.nlev;
    const bool brick_smoother = sizeof(T) == 4 && P.smoother == SHKZ_SMOOTH_BRICK;
    std::vector<T *> zcur(nl), zoth(nl);
    const T om = (T)P.omega;
    const double es = (double)sizeof(T);
    // coarse levels small enough smooth with one multi-sweep launch per smoothing step (Jacobi);
    // with the brick-local smoother every level uses it, whatever its size
    std::vector<char> mpre(nl, 0), mpost(nl, 0);
    for (int l = 1; l < nl && !brick_smoother; ++l) {
        const bool small = C->levels[l].nbricks <= P.multi_max_bricks;
        mpre[l] = small && smooth_k_supported<T>(P.pre_c);
        mpost[l] = small && smooth_k_supported<T>(P.post_c);
    }
    for (int l = 0; l < nl; ++l) {
        LevelHost &L = C->levels[l];
        const T *rhs = l == 0 ? rhs0 : (const T *)L.rhs;
        T *z = l == 0 ? z0 : (T *)L.z;
        T *tmp = l == 0 ? tmp0 : (T *)L.tmp;
        const OpDesc<T> A = op_of<T>(L);
        const T *id = (const T *)L.idiag;
        const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
        const dim3 gp(std::min<unsigned>(g.x, PERSIST_BLOCKS));  // persistent kernels
        const bool f = l == 0;
        const int pre = f ? P.pre : P.pre_c, post = f ? P.post : P.post_c;
        const bool last = l == nl - 1;
        if (L.any_coarsest) {
            // grids ending here: their z goes where the other grids' smoothed z ends up
            // (the brick-local smoother swaps as the finest level's Jacobi does on every level)
            T *dst = (last || final_slot(pre, post, mpre[l], mpost[l], f || brick_smoother) == 0) ? z : tmp;
            // one cell per thread: measured faster per sweep than two cells per thread (48 sweeps: 13 us
            // against 19.3 us)
            constexpr int CPT = 1;
            const dim3 cb(BRICK_CELLS / CPT);
            if (f && dot_partials)
                K.go("coarse_solve_dot", l, 6 * es, 0, dim3((unsigned)P.ngrids), cb, k_coarse_solve<T, 1, CPT>, L.desc,
                     A, rhs, dst, P.coarse_sweeps, om, Q, P.red, (const int16_t *)L.ccode);
            else
                K.go("coarse_solve", l, 6 * es, 0, dim3((unsigned)P.ngrids), cb, k_coarse_solve<T, 0, CPT>, L.desc, A,
                     rhs, dst, P.coarse_sweeps, om, Q, P.red,
                     (const int16_t *)L.ccode);
        }
        if (last) {
            zcur[l] = z;
            zoth[l] = tmp;
            break;
        }
        T *cur = z, *oth = tmp;
        bool res_done = false;
        if (mpre[l]) {
            res_done = smooth_k_res_supported<T>(pre);
            smooth_k<T>(K, l, pre, true, L.desc, A, id, rhs, (const T *)nullptr, z, om, Q,
                        res_done ? (T *)L.res : nullptr);
        } else if (brick_smoother) {
            // brick-local smoother: z = M_b rhs from zero (on the finest level done by the residual
            // update when first_done; on coarse levels by the restriction that formed rhs), then
            // further applications z += M_b (rhs - A z)
            if constexpr (sizeof(T) == 4) {
                const bool k3 = P.local_k == 3;
                if (f && !first_done)
                    K.go("pre_local_L0", 0, 6 * es, -usave(C, 0, 4 * es), dim3(brick_grid(L.nbricks)), dim3(ROW4_THREADS),
                         k3 ? k_pre_local4<3> : k_pre_local4<2>, L.desc, A, id, rhs, z, Q);
                for (int s = 1; s < pre; ++s) {
                    if (f)
                        K.go("post_local_L0", 0, 8 * es, -usave(C, 0, 5 * es), dim3(brick_grid(L.nbricks)),
                             dim3(ROW4_THREADS), k3 ? k_post_local4<0, 3> : k_post_local4<0, 2>, L.desc, A, id, rhs,
                             (const T *)cur, oth, Q, P.red);
                    else
                        K.go("post_local_Lc", l, 8 * es, 0, dim3(brick_grid(L.nbricks)), dim3(ROW4_THREADS),
                             k_post_local4<0, LOCAL_SGS_SWEEPS_COARSE, 0>, L.desc, A, id, rhs, (const T *)cur, oth, Q, P.red);
                    std::swap(cur, oth);
                }
            }
        } else if (pre <= 1 || f) {
            // finest level: one sweep from zero, then plain sweeps (cheaper there than the fused
            // two-sweep kernel, whose on-the-fly neighbour values cost 12 extra loads per cell)
            if (f && first_done) {
                // z = omega * idiag * rhs was written by the residual update (k_cg_update4z)
            } else if constexpr (sizeof(T) == 4)
                K.go(f ? "smooth_first_L0" : "smooth_first_Lc", l, 3 * es, -usave(C, l, es),
                     row4_launch(L.nbricks, f).grid, row4_launch(L.nbricks, f).block,
                     f ? k_smooth_first4<1> : k_smooth_first4<0>, L.desc, id, rhs, z, om, Q);
            else
                K.go(f ? "smooth_first_L0" : "smooth_first_Lc", l, 3 * es, -usave(C, l, es), g, b,
                     f ? k_smooth_first<T, 1> : k_smooth_first<T, 0>, L.desc, id, rhs, z, om, Q);
            for (int s = 1; s < pre; ++s) {
                jacobi_sweep<T>(C, K, l, false, L.desc, A, id, rhs, (const T *)cur, oth, om, Q, P.red);
                std::swap(cur, oth);
            }
        } else {
            K.go(f ? "smooth2_L0" : "smooth2_Lc", l, 6 * es, -usave(C, l, 4 * es), g, b, f ? k_smooth2<T, 1> : k_smooth2<T, 0>, L.desc, A,
                 id, rhs, z, om, Q);
            for (int s = 2; s < pre; ++s) {
                jacobi_sweep<T>(C, K, l, false, L.desc, A, id, rhs, (const T *)cur, oth, om, Q, P.red);
                std::swap(cur, oth);
            }
        }
        if (!res_done) {
            if constexpr (sizeof(T) == 4)
                K.go(f ? "residual_L0" : "residual_Lc", l, 7 * es, -usave(C, l, 4 * es), row4_launch(L.nbricks, f).grid,
                     row4_launch(L.nbricks, f).block, f ? k_residual4<1> : k_residual4<0>, L.desc, A, rhs,
                     (const T *)cur, (T *)L.res, Q);
            else
                K.go(f ? "residual_L0" : "residual_Lc", l, 7 * es, -usave(C, l, 4 * es), g, b,
                     f ? k_residual<T, 1> : k_residual<T, 0>, L.desc, A, rhs, (const T *)cur, (T *)L.res, Q);
        }
        LevelHost &Lc = C->levels[l + 1];
        if constexpr (sizeof(T) == 4) {
            if (brick_smoother) {
                // the coarse level's pre-smoothing from zero joins its restriction
                K.go(f ? "restrict_pre_L0" : "restrict_pre_Lc", l, es, 4 * es * (double)Lc.unknowns,
                     dim3(brick_grid(Lc.nbricks)), b, f ? k_restrict_pre<1> : k_restrict_pre<0>, Lc.desc, op_of<T>(Lc),
                     (const T *)Lc.idiag, (const T *)L.res, (T *)Lc.rhs, (T *)Lc.z, Q);
                zcur[l] = cur;
                zoth[l] = oth;
                continue;
            }
        }
        K.go(f ? "restrict_L0" : "restrict_Lc", l, es, 2 * es * (double)Lc.unknowns, dim3(brick_grid(Lc.nbricks)), b,
             f ? k_restrict<T, 1> : k_restrict<T, 0>, Lc.desc, op_of<T>(Lc), (const T *)L.res, (T *)Lc.rhs, Q);
        zcur[l] = cur;
        zoth[l] = oth;
    }
    for (int l = nl - 2; l >= 0; --l) {
        LevelHost &L = C->levels[l];
        LevelHost &Lc = C->levels[l + 1];
        const T *rhs = l == 0 ? rhs0 : (const T *)L.rhs;
        const OpDesc<T> A = op_of<T>(L);
        const T *id = (const T *)L.idiag;
        const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
        const dim3 gp(std::min<unsigned>(g.x, PERSIST_BLOCKS));  // persistent kernels
        const bool f = l == 0;
        const int post = f ? P.post : P.post_c;
        T *cur = zcur[l], *oth = zoth[l];
        if (mpost[l]) {
            // the prolongation joins the multi-sweep launch
            smooth_k<T>(K, l, post, false, L.desc, A, id, rhs, (const T *)cur, oth, om, Q, nullptr,
                        (const T *)zcur[l + 1], (double)Lc.unknowns);
            std::swap(cur, oth);
            zcur[l] = cur;
            zoth[l] = oth;
            continue;
        }
        if constexpr (sizeof(T) == 4)
            K.go(f ? "prolong_L0" : "prolong_Lc", l, 3 * es, es * (double)Lc.unknowns - usave(C, l, es),
                 row4_launch(L.nbricks, f).grid, row4_launch(L.nbricks, f).block, f ? k_prolong4<1> : k_prolong4<0>,
                 L.desc, A, (const T *)zcur[l + 1], zcur[l], Q);
        else
            K.go(f ? "prolong_L0" : "prolong_Lc", l, 3 * es, es * (double)Lc.unknowns - usave(C, l, es), g, b,
                 f ? k_prolong<T, 1> : k_prolong<T, 0>, L.desc, A, (const T *)zcur[l + 1], zcur[l], Q);
        for (int s = 0; s < post; ++s) {
            const bool last_dot = f && dot_partials && s == post - 1;
            if constexpr (sizeof(T) == 4) {
                if (brick_smoother) {
                    const dim3 gb(brick_grid(L.nbricks));
                    const bool k3 = P.local_k == 3;
                    if (last_dot)
                        K.go("post_local_dot_L0", 0, 8 * es, -usave(C, 0, 5 * es), gb, dim3(ROW4_THREADS),
                             k3 ? k_post_local4<1, 3> : k_post_local4<1, 2>, L.desc, A, id, rhs, (const T *)cur, oth, Q,
                             P.red);
                    else if (f)
                        K.go("post_local_L0", 0, 8 * es, -usave(C, 0, 5 * es), gb, dim3(ROW4_THREADS),
                             k3 ? k_post_local4<0, 3> : k_post_local4<0, 2>, L.desc, A, id, rhs, (const T *)cur, oth, Q,
                             P.red);
                    else
                        K.go("post_local_Lc", l, 8 * es, 0, gb, dim3(ROW4_THREADS),
                             k_post_local4<0, LOCAL_SGS_SWEEPS_COARSE, 0>, L.desc, A, id, rhs, (const T *)cur, oth, Q,
                             P.red);
                    std::swap(cur, oth);
                    continue;
                }
            }
            jacobi_sweep<T>(C, K, l, last_dot, L.desc, A, id, rhs, (const T *)cur, oth, om, Q, P.red);
            std::swap(cur, oth);
        }
        zcur[l] = cur;
        zoth[l] = oth;
    }
    return zcur[0];
}

static int ensure_host_scalars(shkz_ctx *C, int n) {
    if (C->host_scalars_cap >= n) return SHKZ_OK;
    if (C->host_scalars) (void)hipHostFree(C->host_scalars);
    C->host_scalars = nullptr;
    C->host_scalars_cap = 0;
    SHKZ_HIP(hipHostMalloc((void **)&C->host_scalars, (size_t)n * sizeof(SolveScalars), hipHostMallocDefault));
    C->host_scalars_cap = n;
    return SHKZ_OK;
}

// Grows a pinned staging buffer (callers only grow it between calls' synchronisations, so no
// copy is in flight from the old one).
static int ensure_pinned(shkz_ctx::Pinned &b, size_t bytes, void **out) {
    if (b.cap < bytes) {
        if (b.p) (void)hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        SHKZ_HIP(hipHostMalloc(&b.p, want, hipHostMallocDefault));
        b.cap = want;
    }
    *out = b.p;
    return SHKZ_OK;
}

// Reads n per-grid scalars (and optionally a control block, and a second block of n scalars
// into out2) back to the host in one synchronisation.
static int read_scalars(shkz_ctx *C, hipStream_t s, const SolveScalars *dS, int n, std::vector<SolveScalars> &out,
                        const BatchCtl *dctl = nullptr, BatchCtl *ctl_out = nullptr, const SolveScalars *dS2 = nullptr,
                        std::vector<SolveScalars> *out2 = nullptr) {
    SHKZ_TRY(ensure_host_scalars(C, 2 * n));
    SHKZ_HIP(hipMemcpyAsync(C->host_scalars, dS, (size_t)n * sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    if (dS2)
        SHKZ_HIP(hipMemcpyAsync(C->host_scalars + n, dS2, (size_t)n * sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    if (dctl) SHKZ_HIP(hipMemcpyAsync(C->host_ctl, dctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    SHKZ_HIP(hipGetLastError());
    out.assign(C->host_scalars, C->host_scalars + n);
    if (dS2 && out2) out2->assign(C->host_scalars + n, C->host_scalars + 2 * n);
    if (ctl_out) *ctl_out = *C->host_ctl;
    if (C->timing) resolve_timing(C);
    return SHKZ_OK;
}

static void finalize(Launcher &K, const Problem &P, const LevelDesc &L0, SolveScalars *S, BatchCtl *ctl, int op,
                     double param = 0.0, int iparam = 0, const SolveScalars *other = nullptr) {
    K.go("finalize", -1, 0, 0, dim3((unsigned)P.ngrids), dim3(1024), k_finalize, (const double *)P.red.partials, L0, S,
         ctl, op, param, iparam, other);
}

template <class T>
static std::string graph_key(shkz_ctx *C, const Problem &P, const void *const *vecs, int nvec, bool precond) {
    std::string k;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "T%zu pc%d g%d nl%d pre%d post%d prc%d poc%d cs%d mm%d sm%d lk%d om%.17g red%p|",
                  sizeof(T), precond ? 1 : 0, P.ngrids, P.nlev, P.pre, P.post, P.pre_c, P.post_c, P.coarse_sweeps,
                  P.multi_max_bricks, P.smoother, P.local_k, P.omega, (void *)P.red.partials);
    k += buf;
    for (int i = 0; i < nvec; ++i) {
        std::snprintf(buf, sizeof(buf), "%p,", vecs[i]);
        k += buf;
    }
    const int nl = precond ? P.nlev : 1;
    for (int l = 0; l < nl; ++l) {
        const LevelHost &L = C->levels[l];
        // the transfer kernels branch on flat (bilinear in 2D) and the reductions on ngrids
        std::snprintf(buf, sizeof(buf), "|%d %d f%d g%d %p %p %p %p %p %p %p %p %p %p %p %p %p %p %p %p %p %p",
                      L.nbricks, L.any_coarsest ? 1 : 0, L.desc.flat, L.desc.ngrids, (const void *)L.desc.bmap, L.dir,
                      L.gx, L.gy, L.gz, L.idiag, L.rhs, L.z,
                      L.tmp, L.res, (const void *)L.desc.bxyz, (const void *)L.desc.meta, (const void *)L.desc.gl,
                      (const void *)L.desc.child64, (const void *)L.desc.parent8, (const void *)L.desc.uni,
                      (const void *)L.desc.nb27, (const void *)L.ccode);
        k += buf;
    }
    return k;
}

// Krylov loop on level 0 for every grid of the batch: T = element type; precond = V-cycle or
// identity (plain CG).  Expects r (residual) and per-grid S with tol/maxit/iter set and the control
// block counting the grids already done.  x is updated in place.  One iteration is captured once
// into a hipGraph and replayed; every kernel returns at once when all grids are done and skips the
// bricks of a grid that is done, so a chunk of replays may overshoot by a few no-op iterations
// without changing any result.  `last` receives the final per-grid scalars.
//
// first_chunk: iterations replayed before the first poll (the caller's prediction; results do not
// depend on it).  S2/last2: another scalar block read back with every poll.
//
// keep_dir: the first direction continues the search direction p already in place (beta from the
// new r.z over the previous one in S) instead of starting with p = z.
template <class T>
static int krylov(shkz_ctx *C, Launcher &K, Problem &P, T *x, T *r, T *p, T *q, T *z, T *tmp, SolveScalars *S,
                  BatchCtl *ctl, bool precond, std::vector<SolveScalars> &last, int first_chunk = 4,
                  const SolveScalars *S2 = nullptr, std::vector<SolveScalars> *last2 = nullptr,
                  bool keep_dir = false) {
    LevelHost &L0 = C->levels[0];
    const unsigned nb = brick_grid(L0.nbricks);
    const dim3 g(nb), b(BLOCK);
    const OpDesc<T> A = op_of<T>(L0);
    const double es = (double)sizeof(T);
    const Gate Q{ctl, S};
    // FP32 (the inner solve of the mixed-precision mode): the iterate is the FP64 pressure and the
    // operator apply uses the FP64 operator in FP64 arithmetic (k_apply_dot4, k_cg_update4)
    double *const x64 = P.x64;
    const OpDesc<double> A64{P.dir64, P.gx64, P.gy64, P.gz64};
    // FP32: the row kernels (four cells per thread)
    auto update_p = [&](double bytes, const T *zv, int first) {
        if constexpr (sizeof(T) == 4)
            K.go("update_p", 0, bytes, 0, dim3(row4_grid(L0.nbricks)), b, k_update_p4, L0.desc, zv, p,
                 (const SolveScalars *)S, (const BatchCtl *)ctl, first);
        else
            K.go("update_p", 0, bytes, 0, g, b, k_update_p<T>, L0.desc, zv, p, (const SolveScalars *)S,
                 (const BatchCtl *)ctl, first);
    };
    // the finest level's last post-smoothing sweep forms r.z when there is one
    const bool fused_dot = precond && P.nlev > 1 && P.post > 0;
    // first direction
    const T *zp = r;
    if (precond) zp = vcycle<T>(C, K, P, r, z, tmp, Q, fused_dot);
    if (!fused_dot)
        K.go("dot_rz", 0, 2 * es, 0, g, b, k_dot_rz<T>, L0.desc, (const T *)r, zp, (const SolveScalars *)S,
             (const BatchCtl *)ctl, P.red);
    finalize(K, P, L0.desc, S, ctl, keep_dir ? FIN_RZ : FIN_RZ_FIRST);
    update_p(keep_dir ? 3 * es : 2 * es, zp, keep_dir ? 0 : 1);

    // FP32 preconditioned form with fused kernels: the loop is rotated so that each iteration
    // ends with the operator apply of the next direction, formed in the same kernel
    // (k_apply_dot4p), and starts with the residual update that also performs the V-cycle's first
    // sweep (k_cg_update4z); the direction alternates between two buffers, so one captured graph
    // holds two iterations.  Same kernels' arithmetic, same results as the separate kernels.
    const bool fused = sizeof(T) == 4 && precond && !keep_dir && C->fuse_krylov && P.kp2;
    T *const p2 = (T *)P.kp2;
    auto half = [&](T *pc, T *pn) {
        if constexpr (sizeof(T) == 4) {
            const bool zfirst = P.nlev > 1;
            const LevelHost &F0 = C->levels[0];
            const OpDesc<float> A32 = op_of<float>(F0);
            if (zfirst && P.smoother == SHKZ_SMOOTH_BRICK)
                K.go("cg_update_z", 0, 52, -usave(C, 0, 16), g, dim3(ROW4_THREADS),
                     P.local_k == 3 ? k_cg_update4<2, 3> : k_cg_update4<2, 2>, L0.desc, x64, r,
                     (const T *)pc, (const T *)q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red,
                     (const T *)F0.idiag, z, (T)P.omega, A32);
            else if (zfirst)
                K.go("cg_update_z", 0, 40, -usave(C, 0, 4), g, dim3(ROW4_THREADS), k_cg_update4<1>, L0.desc, x64, r,
                     (const T *)pc, (const T *)q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red,
                     (const T *)F0.idiag, z, (T)P.omega, A32);
            else
                K.go("cg_update", 0, 32, 0, g, dim3(ROW4_THREADS), k_cg_update4<0>, L0.desc, x64, r, (const T *)pc,
                     (const T *)q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red, (const T *)nullptr,
                     (T *)nullptr, (T)0, A32);
            finalize(K, P, L0.desc, S, ctl, FIN_UPDATE);
            const T *zz = vcycle<T>(C, K, P, r, z, tmp, Q, fused_dot, zfirst);
            if (!fused_dot)
                K.go("dot_rz", 0, 2 * es, 0, g, b, k_dot_rz<T>, L0.desc, (const T *)r, zz, (const SolveScalars *)S,
                     (const BatchCtl *)ctl, P.red);
            finalize(K, P, L0.desc, S, ctl, FIN_RZ);
            K.go("apply_dot_p", 0, 48, -usave(C, 0, 32), g, dim3(ROW4_THREADS), k_apply_dot4p, L0.desc, A64, zz,
                 (const T *)pc, pn, q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red);
            finalize(K, P, L0.desc, S, ctl, FIN_PQ);
        }
    };
    if (fused) {
        // the first direction's operator apply (the loop starts with the residual update)
        if constexpr (sizeof(T) == 4)
            K.go("apply_dot", 0, 40, -usave(C, 0, 32), g, dim3(ROW4_THREADS), k_apply_dot4, L0.desc, A64, (const T *)p,
                 q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red);
        finalize(K, P, L0.desc, S, ctl, FIN_PQ);
    }
    auto iteration = [&]() {
        if (fused) {
            half(p, p2);
            half(p2, p);
            return;
        }
        if constexpr (sizeof(T) == 4)
            K.go("apply_dot", 0, 40, -usave(C, 0, 32), g, dim3(ROW4_THREADS), k_apply_dot4, L0.desc, A64, (const T *)p,
                 q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red);
        else
            K.go("apply_dot", 0, 6 * es, -usave(C, 0, 4 * es), g, b, k_apply_dot<T>, L0.desc, A, (const T *)p, q,
                 (const SolveScalars *)S, (const BatchCtl *)ctl, P.red);
        finalize(K, P, L0.desc, S, ctl, FIN_PQ);
        if constexpr (sizeof(T) == 4)
            K.go("cg_update", 0, 32, 0, g, dim3(ROW4_THREADS), k_cg_update4<0>, L0.desc, x64, r, (const T *)p,
                 (const T *)q, (const SolveScalars *)S, (const BatchCtl *)ctl, P.red, (const T *)nullptr, (T *)nullptr,
                 (T)0, op_of<float>(L0));
        else
            K.go("cg_update", 0, 6 * es, 0, g, b, k_cg_update<T>, L0.desc, x, r, (const T *)p, (const T *)q,
                 (const SolveScalars *)S, (const BatchCtl *)ctl, P.red);
        finalize(K, P, L0.desc, S, ctl, precond ? FIN_UPDATE : FIN_UPDATE_CG);
        if (precond) {
            const T *zz = vcycle<T>(C, K, P, r, z, tmp, Q, fused_dot);
            if (!fused_dot)
                K.go("dot_rz", 0, 2 * es, 0, g, b, k_dot_rz<T>, L0.desc, (const T *)r, zz, (const SolveScalars *)S,
                     (const BatchCtl *)ctl, P.red);
            finalize(K, P, L0.desc, S, ctl, FIN_RZ);
            update_p(3 * es, zz, 0);
        } else {
            update_p(3 * es, (const T *)r, 0);
        }
    };
    // the iteration (fused: each of the two half-iterations, which alternate the direction
    // buffers) captured once into a hipGraph and replayed
    auto graph_for = [&](const std::string &key, auto &&body, hipGraphExec_t *out) -> int {
        for (auto &g : C->graphs)
            if (g.exec && g.key == key) {
                g.used = ++C->graph_clock;
                *out = g.exec;
                return SHKZ_OK;
            }
        constexpr size_t kMaxGraphs = 24;
        if (C->graphs.size() >= kMaxGraphs) {
            // evict the least recently used exec once every launch of it has completed
            size_t lru = 0;
            for (size_t i = 1; i < C->graphs.size(); ++i)
                if (C->graphs[i].used < C->graphs[lru].used) lru = i;
            SHKZ_HIP(hipStreamSynchronize(K.s));
            SHKZ_HIP(hipGraphExecDestroy(C->graphs[lru].exec));
            C->graphs.erase(C->graphs.begin() + (long)lru);
        }
        hipGraph_t graph = nullptr;
        SHKZ_HIP(hipStreamBeginCapture(K.s, hipStreamCaptureModeThreadLocal));
        K.capturing = true;
        body();
        K.capturing = false;
        SHKZ_HIP(hipStreamEndCapture(K.s, &graph));
        hipGraphExec_t made = nullptr;
        hipError_t e = hipGraphInstantiateWithFlags(&made, graph, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(SHKZ_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        C->graphs.push_back(shkz_ctx::GraphSlot{key, made, ++C->graph_clock});
        *out = made;
        return SHKZ_OK;
    };
    hipGraphExec_t exec = nullptr, exec2 = nullptr;
    if (C->use_graphs && !C->timing) {
        const void *vecs[] = {sizeof(T) == 4 ? (const void *)x64 : (const void *)x, r, p, q, z, tmp, S, ctl,
                              fused ? (const void *)p2 : nullptr, P.dir64, P.gx64, P.gy64, P.gz64};
        const std::string key = std::string(sizeof(T) == 8 ? "d" : "f") + (fused ? "F" : "") +
                                graph_key<T>(C, P, vecs, 13, precond);
        if (fused) {
            // one graph per half-iteration (the second one a later find cannot evict: the first
            // was used last but one), so that a solve replays exactly the iterations it asks for
            SHKZ_TRY(graph_for(key + "A", [&]() { half(p, p2); }, &exec));
            SHKZ_TRY(graph_for(key + "B", [&]() { half(p2, p); }, &exec2));
        } else {
            SHKZ_TRY(graph_for(key, iteration, &exec));
        }
    }
    int chunk = std::max(1, std::min(first_chunk, 64));
    int phase = 0;  // fused: the half-iteration that comes next (0: direction in p, 1: in p2)
    BatchCtl hc{};
    for (;;) {
        for (int it = 0; it < chunk; ++it) {
            if (!fused) {
                if (exec) SHKZ_HIP(hipGraphLaunch(exec, K.s));
                else iteration();
                continue;
            }
            if (exec) SHKZ_HIP(hipGraphLaunch(phase ? exec2 : exec, K.s));
            else if (phase) half(p2, p);
            else half(p, p2);
            phase ^= 1;
        }
        SHKZ_TRY(read_scalars(C, K.s, S, P.ngrids, last, ctl, &hc, S2, last2));
        if (hc.all_done) break;
        // size the next chunk from the slowest grid's observed contraction rate
        int next = 1;
        for (const SolveScalars &s : last) {
            if (s.done) continue;
            const double cur = s.norm == 0 ? s.rinf : s.rl2;
            const double r0 = s.norm == 0 ? s.r0inf : s.r0l2;
            int need = chunk;
            if (s.iter > 0 && cur > 0 && r0 > 0 && cur < r0) {
                const double rate = std::pow(cur / r0, 1.0 / s.iter);
                if (rate > 0 && rate < 1) need = (int)std::ceil(std::log(s.tol / cur) / std::log(rate));
            }
            need = std::min(need, s.maxit - s.iter);
            next = std::max(next, need);
        }
        chunk = std::max(1, std::min(next, 64));
    }
    return SHKZ_OK;
}

// ------------------------------------------------------------------------ the projection (batch)
struct ExportOut {
    double *diag, *gx, *gy, *gz, *rhs;  // device, cell-indexed (grid 0)
};

struct GridJob {
    const shkz_grid *in = nullptr;
    GridDesc G{};
    size_t M = 0;
    int *exist = nullptr;
    uint8_t *cflag = nullptr, *bflag = nullptr;
    GridLevelHost l0;
    SceneState *scene = nullptr;
    double rhs_correct = 0.0;
};

// A 2D grid (w, h) travels inside the runtime as a 3D grid with d = 1 and no w velocity; its solid
// level set is the 2D nodal array (w+1, h+1).
static inline bool is_flat(const shkz_grid *grid) { return grid->w_vel == nullptr; }

static int validate_grid(const shkz_grid *grid) {
    const int w = grid->w, h = grid->h, d = grid->d;
    if (w <= 0 || h <= 0 || d <= 0) return fail(SHKZ_ERR_INVALID, "grid dimensions must be positive (%d,%d,%d)", w, h, d);
    if (!grid->fluid || !grid->u || !grid->v || (!grid->w_vel && d != 1))
        return fail(SHKZ_ERR_INVALID, "fluid and velocities are required");
    if (!(grid->dt > 0)) return fail(SHKZ_ERR_INVALID, "dt must be positive");
    if ((int64_t)w * h * d > (int64_t)1 << 31) return fail(SHKZ_ERR_INVALID, "grid too large for 32-bit brick ids");
    if ((int64_t)(w + 1) * (h + 1) * (d + 1) >= (int64_t)1 << 32)
        return fail(SHKZ_ERR_INVALID, "grid too large for 32-bit node indices");
    return SHKZ_OK;
}

static int project_stages(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *prm, shkz_result *res,
                          hipStream_t s, const ExportOut *exp);

// Runs the projection on the context's work stream and orders the caller's stream after it, so
// that work the caller queues next (and synchronous copies on the null stream) sees the outputs.
static int project_impl(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *prm, shkz_result *res,
                        hipStream_t caller, const ExportOut *exp = nullptr) {
    int rc = project_stages(C, grids, ngrids, prm, res, caller, exp);
    if (C->work && caller != C->work) {
        if (hipEventRecord(C->ev_out, C->work) != hipSuccess || hipStreamWaitEvent(caller, C->ev_out, 0) != hipSuccess) {
            (void)hipStreamSynchronize(C->work);
            if (rc == SHKZ_OK) rc = fail(SHKZ_ERR_DEVICE, "ordering the caller's stream after the projection failed");
        } else if (rc != SHKZ_OK) {
            (void)hipStreamSynchronize(C->work);
        }
    }
    if (rc != SHKZ_OK)
        for (int g = 0; g < ngrids; ++g)
            if (res[g].status == SHKZ_OK && rc != SHKZ_ERR_NUMERIC) res[g].status = rc;
    return rc;
}

// Output of one solve phase, per grid.
struct SolveOut {
    std::vector<int> total_iters, outer_passes, status;
};

// The linear solve of every grid of the built batch (macpressuresolver3.cpp:233-237): FP64
// CG / MGPCG, or FP64 defect correction around FP32 MGPCG.  Starts from P.x64 and leaves the
// solution there; fills the per-grid residual and convergence fields of res.
static std::string tune_key(const shkz_ctx *C, const Problem &P, const shkz_params *prm);

// Key of a solve's iteration history: the shape class and the controls in use.
static std::string hist_key(const shkz_ctx *C, const Problem &P, const shkz_params *prm) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "|%d %d %d %d %d %d %d %d %.6g %.3g", P.pre, P.post, P.pre_c, P.post_c,
                  P.coarse_sweeps, P.multi_max_bricks, P.smoother, P.local_k, P.omega, P.eta);
    return tune_key(C, P, prm) + buf;
}

static int solve_phase(shkz_ctx *C, Launcher &K, Problem &P, const shkz_params *prm, bool multigrid, SolveOut &O,
                       shkz_result *res) {
    const int ngrids = P.ngrids;
    // iterations the last solve of this class took per defect-correction pass (slowest grid): the
    // first chunk of each Krylov solve, and whether a pass is expected to run a Krylov solve at all
    // (then the host does not wait for the pass's set-up before queueing it).  Every kernel is
    // gated on the device scalars, so a wrong guess only costs time, never a different result.
    const std::string hkey = hist_key(C, P, prm);
    const std::vector<int> hist = C->iter_hist.count(hkey) ? C->iter_hist[hkey] : std::vector<int>();
    std::vector<int> seen;
    auto seen_pass = [&](const std::vector<SolveScalars> &inner) {
        int mx = 0;
        for (const SolveScalars &v : inner) mx = std::max(mx, (int)v.iter);
        seen.push_back(mx);
    };
    hipStream_t s = K.s;
    LevelHost &F0 = C->levels[0];
    const int norm = prm->norm == SHKZ_NORM_L2 ? 1 : 0;
    const int maxit = prm->max_iterations > 0 ? prm->max_iterations : 30000;
    const double residual = prm->residual < 1e-30 ? 1e-30 : prm->residual;
    OpDesc<double> A64{P.dir64, P.gx64, P.gy64, P.gz64};
    const unsigned g0 = brick_grid(F0.nbricks);
    const dim3 gb(g0);
    std::vector<int> &total_iters = O.total_iters, &outer_passes = O.outer_passes, &status = O.status;
    total_iters.assign(ngrids, 0);
    outer_passes.assign(ngrids, 0);
    status.assign(ngrids, SHKZ_OK);
    std::vector<SolveScalars> hs(ngrids), hi, hf;
    // set when the last outer defect was taken at the returned pressure of every grid
    bool final_measured = false;
    if (F0.nbricks > 0) {
        K.go("init_scalars", -1, 0, 0, dim3((unsigned)((ngrids + 255) / 256)), dim3(256), k_init_scalars, P.S_out,
             ngrids, norm, maxit);
        K.go("init_ctl", -1, 0, 0, dim3(1), dim3(64), k_init_ctl, P.ctl, ngrids);
    }
    if (F0.nbricks == 0) {
        // no unknowns anywhere: nothing to solve
    } else if (!P.mixed) {
        // one FP64 Krylov solve per grid: CG (the reference iteration) or MGPCG
        K.go("defect", 0, 7 * 8, -usave(C, 0, 32), gb, dim3(ROW4_THREADS), k_defect4, F0.desc, A64,
             (const double *)P.b64, (const double *)P.x64, P.r64, P.red);
        finalize(K, P, F0.desc, P.S_out, P.ctl, FIN_DEFECT_FIRST, residual);
        BatchCtl hc{};
        if (hist.empty() || hist[0] == 0) SHKZ_TRY(read_scalars(C, s, P.S_out, ngrids, hs, P.ctl, &hc));
        if (!hc.all_done) {
            SHKZ_TRY(krylov<double>(C, K, P, P.x64, P.r64, (double *)P.kp, (double *)P.kq, (double *)P.kz,
                                    (double *)P.ktmp, P.S_out, P.ctl, multigrid, hs, hist.empty() ? 4 : hist[0]));
            seen_pass(hs);
        }
        for (int g = 0; g < ngrids; ++g) {
            total_iters[g] = hs[g].iter;
            outer_passes[g] = F0.gl[g].brick_end > F0.gl[g].brick_start ? 1 : 0;
            if (hs[g].status) status[g] = SHKZ_ERR_NUMERIC;
        }
    } else {
        // FP64 defect correction around FP32 MGPCG solves, per grid: the outer scalars decide on
        // the device when a grid is done (converged, out of iterations, or no longer reduced)
        const double eta = P.eta;
        K.go("init_scalars", -1, 0, 0, dim3((unsigned)((ngrids + 255) / 256)), dim3(256), k_init_scalars, P.S_in,
             ngrids, norm, maxit);
        for (int outer = 0; outer < 64; ++outer) {
            K.go("defect", 0, 7 * 8, -usave(C, 0, 32), gb, dim3(ROW4_THREADS), k_defect4, F0.desc, A64,
                 (const double *)P.b64, (const double *)P.x64, P.r64, P.red);
            if (outer == 0)
                finalize(K, P, F0.desc, P.S_out, nullptr, FIN_DEFECT_FIRST, residual);
            else
                finalize(K, P, F0.desc, P.S_out, nullptr, FIN_OUTER, 0.0, 0, P.S_in);
            K.go("init_ctl", -1, 0, 0, dim3(1), dim3(64), k_init_ctl, P.ctl + 1, ngrids);
            K.go("inner_init", 0, 12, 0, gb, dim3(ROW4_THREADS), k_inner_init4, F0.desc, (const double *)P.r64,
                 (float *)P.kr, P.red);
            finalize(K, P, F0.desc, P.S_in, P.ctl + 1, FIN_INNER_INIT, eta, 0, P.S_out);
            BatchCtl hc{};
            // the history expects an inner solve in this pass: queue it at once; the outer scalars
            // come back with the Krylov solve's polls (a grid found done runs no-op kernels and
            // adds a zero correction)
            const bool expect_inner = outer < (int)hist.size() && hist[outer] > 0;
            if (!expect_inner) {
                SHKZ_TRY(read_scalars(C, s, P.S_out, ngrids, hs, P.ctl + 1, &hc));
                if (outer > 0)
                    for (int g = 0; g < ngrids; ++g) total_iters[g] = hs[g].iter;
                if (hc.all_done) {
                    final_measured = true;
                    break;
                }
            }
            SHKZ_TRY(krylov<float>(C, K, P, (float *)nullptr, (float *)P.kr, (float *)P.kp, (float *)P.kq, (float *)P.kz,
                                   (float *)P.ktmp, P.S_in, P.ctl + 1, true, hi, expect_inner ? hist[outer] : 4,
                                   expect_inner ? P.S_out : nullptr, expect_inner ? &hs : nullptr,
                                   C->keep_direction && outer > 0));
            if (expect_inner) {
                if (outer > 0)
                    for (int g = 0; g < ngrids; ++g) total_iters[g] = hs[g].iter;
                bool all = true;
                for (int g = 0; g < ngrids; ++g) all = all && hs[g].done;
                if (all) {
                    // every grid was done at this pass's defect: what the poll would have found
                    final_measured = true;
                    break;
                }
            }
            seen_pass(hi);
            for (int g = 0; g < ngrids; ++g)
                if (!hs[g].done) outer_passes[g] = outer + 1;
            if (C->trace)
                for (int g = 0; g < ngrids; ++g)
                    if (!hs[g].done)
                        std::fprintf(stderr, "[shkz] grid %d outer %d: defect %.6e (tol %.3e) inner it %d rinf %.6e\n",
                                     g, outer, norm == 0 ? hs[g].rinf : hs[g].rl2, hs[g].tol, hi[g].iter, hi[g].rinf);
            if (outer == 63)
                for (int g = 0; g < ngrids; ++g)
                    if (!hs[g].done) total_iters[g] += hi[g].iter;
        }
        for (int g = 0; g < ngrids; ++g)
            if (hs[g].status) status[g] = SHKZ_ERR_NUMERIC;
    }
    if (F0.nbricks > 0) C->iter_hist[hkey] = seen;
    // true FP64 residual of the returned pressure, per grid.  When the defect correction ended
    // on a pass whose defect found every grid done, that defect was taken at the returned
    // pressure (a grid done earlier only gained zero corrections since), so its norms are kept.
    if (final_measured) {
        hf = hs;
    } else if (F0.nbricks > 0) {
        K.go("defect", 0, 7 * 8, -usave(C, 0, 32), gb, dim3(ROW4_THREADS), k_defect4, F0.desc, A64,
             (const double *)P.b64, (const double *)P.x64, P.r64, P.red);
        finalize(K, P, F0.desc, P.S_fin, nullptr, FIN_DEFECT);
        SHKZ_TRY(read_scalars(C, s, P.S_fin, ngrids, hf));
    }
    for (int g = 0; g < ngrids; ++g) {
        if (F0.gl[g].brick_end == F0.gl[g].brick_start) {
            res[g].converged = 1;  // no unknowns
            continue;
        }
        const SolveScalars &o = hs[g];
        res[g].reresid_inf = o.r0inf > 0 ? hf[g].rinf / o.r0inf : 0.0;
        res[g].reresid_l2 = o.r0l2 > 0 ? hf[g].rl2 / o.r0l2 : 0.0;
        res[g].reresid = norm == 0 ? res[g].reresid_inf : res[g].reresid_l2;
        const double cur = norm == 0 ? hf[g].rinf : hf[g].rl2;
        const double r0 = norm == 0 ? o.r0inf : o.r0l2;
        res[g].converged = (r0 == 0.0 || cur <= o.tol * (1.0 + 1e-12)) ? 1 : 0;
        if (!P.mixed && !o.converged && r0 != 0.0) res[g].converged = 0;
    }
    return SHKZ_OK;
}

// ------------------------------------------------------------------------ autotuner
// Multigrid controls picked per (device, shape class) by timing repeated solves of the built
// system.  Coordinate search from the current controls over each control the caller left at 0:
// finest-level sweeps (pre = post, keeping the V-cycle symmetric), coarse-level sweeps, the Jacobi
// weight, the coarse-solve sweeps, and (mixed precision) the inner-solve reduction.  A candidate
// counts only if every grid of the batch converges; it replaces the current one when it is at
// least 2% faster (best of two timed solves, so graph capture is not timed).
struct Knobs {
    int pre, post, pre_c, post_c, coarse, multi;
    double omega, eta;
    int smoother;
    int local_k;
};

static Knobs knobs_of(const Problem &P) {
    return Knobs{P.pre,   P.post, P.pre_c,    P.post_c, P.coarse_sweeps, P.multi_max_bricks,
                 P.omega, P.eta,  P.smoother, P.local_k};
}

// Sets P's controls to k except where the caller fixed a control in prm.
static void apply_knobs(Problem &P, const Knobs &k, const shkz_params *prm) {
    if (prm->mg_pre_sweeps <= 0) P.pre = k.pre;
    if (prm->mg_post_sweeps <= 0) P.post = k.post;
    if (prm->mg_coarse_pre_sweeps <= 0) P.pre_c = k.pre_c;
    if (prm->mg_coarse_post_sweeps <= 0) P.post_c = k.post_c;
    if (prm->mg_coarse_sweeps <= 0) P.coarse_sweeps = k.coarse;
    if (!(prm->mg_omega > 0)) P.omega = k.omega;
    if (!(prm->inner_residual > 0)) P.eta = k.eta;
    if (prm->mg_smoother <= 0 && (k.smoother == SHKZ_SMOOTH_JACOBI || P.mixed)) P.smoother = k.smoother;
    if (prm->mg_local_sweeps <= 0) P.local_k = k.local_k;
    P.multi_max_bricks = k.multi;
}

struct TuneCache {
    std::mutex mu;
    std::unordered_map<std::string, Knobs> map;
    bool loaded = false;
    std::string path;
};
static TuneCache g_tune;

static void tune_load_locked() {
    if (g_tune.loaded) return;
    g_tune.loaded = true;
    const char *p = std::getenv("SHKZ_TUNE_CACHE");
    if (!p || !*p) return;
    g_tune.path = p;
    FILE *f = std::fopen(p, "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
        char key[320];
        Knobs k;
        if (std::sscanf(line, "%319s %d %d %d %d %d %d %lf %lf %d %d", key, &k.pre, &k.post, &k.pre_c, &k.post_c,
                        &k.coarse, &k.multi, &k.omega, &k.eta, &k.smoother, &k.local_k) == 11 &&
            k.pre > 0 && k.post > 0 && k.pre_c > 0 && k.post_c > 0 && k.coarse > 0 && k.multi >= 0 && k.omega > 0 &&
            k.omega <= 1.0 && k.eta >= 0 && (k.smoother == SHKZ_SMOOTH_JACOBI || k.smoother == SHKZ_SMOOTH_BRICK) &&
            (k.local_k == 2 || k.local_k == 3))
            g_tune.map[key] = k;
    }
    std::fclose(f);
}

static bool tune_find(const std::string &key, Knobs *k) {
    std::lock_guard<std::mutex> g(g_tune.mu);
    tune_load_locked();
    auto it = g_tune.map.find(key);
    if (it == g_tune.map.end()) return false;
    *k = it->second;
    return true;
}

static void tune_store(const std::string &key, const Knobs &k) {
    std::lock_guard<std::mutex> g(g_tune.mu);
    tune_load_locked();
    g_tune.map[key] = k;
    if (g_tune.path.empty()) return;
    FILE *f = std::fopen(g_tune.path.c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %d %d %d %.17g %.17g %d %d\n", key.c_str(), k.pre, k.post, k.pre_c, k.post_c,
                 k.coarse, k.multi, k.omega, k.eta, k.smoother, k.local_k);
    std::fclose(f);
}

// Shape class: device, solver, precision, norm, residual decade, stored bricks (half-octaves),
// multigrid depth, batch size class, and which controls the caller fixed.
static std::string tune_key(const shkz_ctx *C, const Problem &P, const shkz_params *prm) {
    const int nb = C->levels[0].nbricks;
    const int bclass = (int)std::lround(2.0 * std::log2((double)std::max(nb, 1)));
    const int rclass = (int)std::lround(std::log10(std::max(prm->residual, 1e-30)));
    const int gclass = P.ngrids == 1 ? 0 : (int)std::ceil(std::log2((double)P.ngrids) / 3.0);
    const int fixed = (prm->mg_pre_sweeps > 0) | (prm->mg_post_sweeps > 0) << 1 | (prm->mg_coarse_pre_sweeps > 0) << 2 |
                      (prm->mg_coarse_post_sweeps > 0) << 3 | (prm->mg_coarse_sweeps > 0) << 4 |
                      (prm->mg_omega > 0) << 5 | (prm->inner_residual > 0) << 6 | (prm->mg_smoother > 0) << 7 |
                      (prm->mg_local_sweeps > 0) << 8;
    char buf[320];
    std::snprintf(buf, sizeof(buf), "%s|s%d|p%d|n%d|r%d|b%d|l%d|g%d|f%d", C->device_key.c_str(), prm->solver,
                  prm->precision, prm->norm, rclass, bclass, P.nlev, gclass, fixed);
    return buf;
}

static int autotune(shkz_ctx *C, Launcher &K, Problem &P, const shkz_params *prm, bool multigrid) {
    const size_t n0 = (size_t)std::max(C->levels[0].nbricks, 1) * BRICK_CELLS;
    double *x0 = nullptr;
    SHKZ_TRY(arena_get(C->arena, "tune_x0", n0, &x0));
    SHKZ_HIP(hipMemcpyAsync(x0, P.x64, n0 * sizeof(double), hipMemcpyDeviceToDevice, K.s));
    std::vector<shkz_result> scratch(P.ngrids);
    auto trial = [&](const Knobs &k, double *ms) -> int {
        apply_knobs(P, k, prm);
        double best = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            SHKZ_HIP(hipMemcpyAsync(P.x64, x0, n0 * sizeof(double), hipMemcpyDeviceToDevice, K.s));
            SHKZ_HIP(hipStreamSynchronize(K.s));
            SolveOut o;
            const double t0 = now_ms();
            SHKZ_TRY(solve_phase(C, K, P, prm, multigrid, o, scratch.data()));
            const double t1 = now_ms();
            for (int g = 0; g < P.ngrids; ++g)
                if (!scratch[g].converged || o.status[g] != SHKZ_OK) {
                    *ms = 1e300;
                    return SHKZ_OK;
                }
            best = std::min(best, t1 - t0);
        }
        *ms = best;
        return SHKZ_OK;
    };
    Knobs cur = knobs_of(P);
    double cur_ms = 1e300;
    SHKZ_TRY(trial(cur, &cur_ms));
    auto consider = [&](Knobs k) -> int {
        double ms = 1e300;
        SHKZ_TRY(trial(k, &ms));
        if (C->trace)
            std::fprintf(stderr,
                         "[shkz] tune sm %d pre/post %d/%d (x%d) coarse %d/%d cs %d multi %d omega %.2f eta %.0e: %.3f "
                         "ms\n",
                         k.smoother, k.pre, k.post, k.local_k, k.pre_c, k.post_c, k.coarse, k.multi, k.omega, k.eta, ms);
        if (ms < 0.99 * cur_ms) {
            cur = k;
            cur_ms = ms;
        }
        return SHKZ_OK;
    };
    const bool sweeps_free = prm->mg_pre_sweeps <= 0 && prm->mg_post_sweeps <= 0;
    if (P.mixed && prm->mg_smoother <= 0) {
        // the other smoother family, with its own sweep count
        Knobs k = cur;
        k.smoother = cur.smoother == SHKZ_SMOOTH_BRICK ? SHKZ_SMOOTH_JACOBI : SHKZ_SMOOTH_BRICK;
        if (sweeps_free) k.pre = k.post = k.smoother == SHKZ_SMOOTH_BRICK ? 1 : 3;
        if (prm->mg_coarse_pre_sweeps <= 0 && prm->mg_coarse_post_sweeps <= 0)
            k.pre_c = k.post_c = k.smoother == SHKZ_SMOOTH_BRICK ? 1 : 4;
        SHKZ_TRY(consider(k));
    }
    if (sweeps_free)
        for (int v : {1, 2, 3, 4, 6}) {
            if (v == cur.pre && v == cur.post) continue;
            if (cur.smoother == SHKZ_SMOOTH_BRICK ? v > 2 : v < 2) continue;
            Knobs k = cur;
            k.pre = k.post = v;
            SHKZ_TRY(consider(k));
        }
    if (prm->mg_coarse_pre_sweeps <= 0 && prm->mg_coarse_post_sweeps <= 0 && P.nlev > 2)
        for (int v : {1, 2, 3, 4, 5, 6}) {
            if (v == cur.pre_c && v == cur.post_c) continue;
            if (cur.smoother == SHKZ_SMOOTH_BRICK ? v > 2 : v < 2) continue;
            Knobs k = cur;
            k.pre_c = k.post_c = v;
            SHKZ_TRY(consider(k));
        }
    if (!(prm->mg_omega > 0))
        for (double v : {0.8, 0.9, 1.0}) {
            if (v == cur.omega) continue;
            Knobs k = cur;
            k.omega = v;
            SHKZ_TRY(consider(k));
        }
    if (prm->mg_coarse_sweeps <= 0)
        // the bottom grid is one brick: a few sweeps can be enough there (at 256^3 four sweeps keep
        // the iteration count of sixteen), and the tuner times the whole solve
        for (int v : {4, 8, 16, 32, 48, 96}) {
            if (v == cur.coarse) continue;
            Knobs k = cur;
            k.coarse = v;
            SHKZ_TRY(consider(k));
        }
    if (P.nlev > 2)
        for (int v : {0, 256, 4096}) {
            if (v == cur.multi) continue;
            Knobs k = cur;
            k.multi = v;
            SHKZ_TRY(consider(k));
        }
    if (cur.smoother == SHKZ_SMOOTH_BRICK && prm->mg_local_sweeps <= 0) {
        // sweeps per finest-level application of the brick-local smoother, then the sweep counts
        // once more around the choice
        Knobs k = cur;
        k.local_k = cur.local_k == 2 ? 3 : 2;
        SHKZ_TRY(consider(k));
        if (sweeps_free && cur.local_k == k.local_k)
            for (int v : {1, 2}) {
                if (v == cur.pre && v == cur.post) continue;
                Knobs k2 = cur;
                k2.pre = k2.post = v;
                SHKZ_TRY(consider(k2));
            }
    }
    if (P.mixed && !(prm->inner_residual > 0))
        for (double v : {0.0, 1e-2, 1e-3, 1e-4}) {
            if (v == cur.eta) continue;
            Knobs k = cur;
            k.eta = v;
            SHKZ_TRY(consider(k));
        }
    apply_knobs(P, cur, prm);
    SHKZ_HIP(hipMemcpyAsync(P.x64, x0, n0 * sizeof(double), hipMemcpyDeviceToDevice, K.s));
    return SHKZ_OK;
}

static int project_stages(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *prm, shkz_result *res,
                          hipStream_t s, const ExportOut *exp) {
    const double t_start = now_ms();
    for (int g = 0; g < ngrids; ++g) std::memset(&res[g], 0, sizeof(shkz_result));
    if (ngrids <= 0) return fail(SHKZ_ERR_INVALID, "empty batch");
    for (int g = 0; g < ngrids; ++g) SHKZ_TRY(validate_grid(&grids[g]));
    if (prm->solver != SHKZ_SOLVER_MGPCG && prm->solver != SHKZ_SOLVER_CG) return fail(SHKZ_ERR_INVALID, "unknown solver %d", prm->solver);
    if (prm->precision != SHKZ_PREC_MIXED && prm->precision != SHKZ_PREC_FP64)
        return fail(SHKZ_ERR_INVALID, "unknown precision %d", prm->precision);
    if (prm->mg_local_sweeps != 0 && prm->mg_local_sweeps != 2 && prm->mg_local_sweeps != 3)
        return fail(SHKZ_ERR_INVALID, "mg_local_sweeps must be 0, 2 or 3 (got %d)", prm->mg_local_sweeps);
    const bool flat = is_flat(&grids[0]);
    for (int g = 1; g < ngrids; ++g)
        if (is_flat(&grids[g]) != flat) return fail(SHKZ_ERR_INVALID, "a batch cannot mix 2D and 3D grids");
    for (int a = 0; a < ngrids; ++a)
        for (int b = a + 1; b < ngrids; ++b)
            if (grids[a].scene_id && grids[a].scene_id == grids[b].scene_id)
                return fail(SHKZ_ERR_INVALID, "scene %llu appears twice in one batch", (unsigned long long)grids[a].scene_id);
    SHKZ_HIP(hipSetDevice(C->device));
    // all work runs on the context's stream, ordered after what the caller queued on its stream
    if (s != C->work) {
        SHKZ_HIP(hipEventRecord(C->ev_in, s));
        SHKZ_HIP(hipStreamWaitEvent(C->work, C->ev_in, 0));
        s = C->work;
    }
    Launcher K{C, s};
    Arena &AR = C->arena;

    Problem P;
    P.ngrids = ngrids;
    P.mixed = (prm->solver == SHKZ_SOLVER_MGPCG && prm->precision == SHKZ_PREC_MIXED);
    P.fp64_levels = !P.mixed;
    // finest-level smoother: the brick-local one needs the FP32 V-cycle of the mixed mode
    P.smoother = prm->mg_smoother > 0 ? prm->mg_smoother : (P.mixed ? SHKZ_SMOOTH_BRICK : SHKZ_SMOOTH_JACOBI);
    if (!P.mixed) P.smoother = SHKZ_SMOOTH_JACOBI;
    const bool brick = P.smoother == SHKZ_SMOOTH_BRICK;
    P.pre = prm->mg_pre_sweeps > 0 ? prm->mg_pre_sweeps : (brick ? 1 : 4);
    P.post = prm->mg_post_sweeps > 0 ? prm->mg_post_sweeps : (brick ? 1 : 4);
    P.local_k = prm->mg_local_sweeps > 0 ? prm->mg_local_sweeps : LOCAL_SGS_SWEEPS;
    P.pre_c = prm->mg_coarse_pre_sweeps > 0 ? prm->mg_coarse_pre_sweeps : (brick ? 1 : P.pre);
    P.post_c = prm->mg_coarse_post_sweeps > 0 ? prm->mg_coarse_post_sweeps : (brick ? 1 : P.post);
    P.coarse_sweeps = prm->mg_coarse_sweeps > 0 ? prm->mg_coarse_sweeps : 48;
    P.omega = prm->mg_omega > 0 ? prm->mg_omega : 0.9;
    // 0: each inner solve goes to the final tolerance (one defect-correction pass unless the FP32
    // recursive residual drifted from the FP64 defect)
    P.eta = prm->inner_residual > 0 ? prm->inner_residual : 0.0;
    P.multi_max_bricks = C->multi_max_bricks;
    const bool multigrid = prm->solver == SHKZ_SOLVER_MGPCG;

    // ---------------------------------------------------------------- per grid: fractions + labelling
    std::vector<GridJob> jobs(ngrids);
    // per-grid dense flags live in three combined buffers (256-byte aligned slices)
    size_t cflag_total = 0, bflag_total = 0;
    std::vector<size_t> cflag_off(ngrids), bflag_off(ngrids);
    for (int g = 0; g < ngrids; ++g) {
        const size_t M = (size_t)grids[g].w * grids[g].h * grids[g].d;
        const size_t npos = (size_t)((grids[g].w + BRICK - 1) / BRICK) * ((grids[g].h + BRICK - 1) / BRICK) *
                            ((grids[g].d + BRICK - 1) / BRICK);
        cflag_off[g] = cflag_total;
        bflag_off[g] = bflag_total;
        cflag_total += (M + 255) & ~(size_t)255;
        bflag_total += (npos + 255) & ~(size_t)255;
    }
    uint8_t *cflag_all = nullptr, *bflag_all = nullptr;
    int *exist_all = nullptr;
    SHKZ_TRY(arena_get(AR, "cflag", cflag_total, &cflag_all));
    SHKZ_TRY(arena_get(AR, "bflag", bflag_total, &bflag_all));
    SHKZ_TRY(arena_get(AR, "exist", (size_t)4 * ngrids, &exist_all));
    K.go("reset", -1, 0, 16.0 * ngrids, dim3((unsigned)((4 * ngrids + 255) / 256)), dim3(256), k_reset_ints, exist_all,
         4 * ngrids);
    for (int g = 0; g < ngrids; ++g) {
        GridJob &J = jobs[g];
        const shkz_grid *grid = &grids[g];
        J.in = grid;
        const int w = grid->w, h = grid->h, d = grid->d;
        J.M = (size_t)w * h * d;
        GridDesc &G = J.G;
        G.w = w;
        G.h = h;
        G.d = d;
        G.phi64 = grid->fluid_dtype == SHKZ_F64;
        G.solid64 = grid->solid_dtype == SHKZ_F64;
        G.dx = grid->dx > 0 ? grid->dx : std::min(1.0 / w, std::min(1.0 / h, 1.0 / d));
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
        G.active[2] = flat ? nullptr : grid->active_w;
        G.active_out[0] = flat ? nullptr : grid->active_out_u;
        G.active_out[1] = flat ? nullptr : grid->active_out_v;
        G.active_out[2] = flat ? nullptr : grid->active_out_w;
        G.flat = flat ? 1 : 0;
        J.exist = exist_all + 4 * g;
        G.exist = J.exist;
        // levelset_exist() of the liquid: the labelling finds it on the way (read back with the brick
        // flags), unless surface tension needs it first (the solid needs no scan, as the face
        // fractions show)
        const bool scan_liquid = prm->surface_tension != 0.0;
        if (scan_liquid) {
            const unsigned nbk = (unsigned)std::min<size_t>((J.M + 255) / 256, 4096);
            K.go("any_negative", -1, 0, (double)J.M * (G.phi64 ? 8 : 4), dim3(nbk), dim3(256), k_any_negative,
                 grid->fluid, (int)G.phi64, J.M, J.exist + 0);
        }
        if (prm->surface_tension != 0.0) {
            // sized for the batch's largest grid once, so no buffer is reallocated under queued work
            size_t maxM = 1;
            for (int q = 0; q < ngrids; ++q) maxM = std::max(maxM, (size_t)grids[q].w * grids[q].h * grids[q].d);
            float *curv = nullptr;
            SHKZ_TRY(arena_get(AR, "curvature", maxM, &curv));
            K.go("curvature", -1, 0, (double)J.M * ((G.phi64 ? 8 : 4) + 4), dim3(xcd_blocks(J.M, 256)), dim3(256),
                 k_curvature, G, curv);
            for (int dim = 0; dim < (flat ? 2 : 3); ++dim) {
                const size_t nf = (size_t)(w + (dim == 0)) * (h + (dim == 1)) * (d + (dim == 2));
                K.go("surface_tension", -1, 0, (double)nf * 8, dim3(xcd_blocks(nf, 256)), dim3(256),
                     k_surface_tension, G, dim, (const float *)curv, prm->surface_tension);
            }
        }
        grid_level_dims(J.l0, w, h, d);
        const int npos = J.l0.nbx * J.l0.nby * J.l0.nbz;
        J.cflag = cflag_all + cflag_off[g];
        J.bflag = bflag_all + bflag_off[g];
        K.go("flags", -1, 0, (double)J.M * ((G.phi64 ? 8 : 4) + 1 + (grid->solid ? 4 : 0)), dim3(brick_grid(npos)),
             dim3(BLOCK), k_flags, G, J.l0.nbx, J.l0.nby, J.l0.nbz, J.cflag, J.bflag);
    }
    const double t_queued = now_ms();
    {
        void *hflags = nullptr;
        SHKZ_TRY(ensure_pinned(C->pin_flags, bflag_total, &hflags));
        SHKZ_HIP(hipMemcpyAsync(hflags, bflag_all, bflag_total, hipMemcpyDeviceToHost, s));
        SHKZ_HIP(hipStreamSynchronize(s));
        SHKZ_HIP(hipGetLastError());
        for (int g = 0; g < ngrids; ++g) {
            GridJob &J = jobs[g];
            const uint8_t *src = (const uint8_t *)hflags + bflag_off[g];
            const size_t npos = (size_t)J.l0.nbx * J.l0.nby * J.l0.nbz;
            J.l0.active.resize(npos);
            bool liquid = false;
            for (size_t b = 0; b < npos; ++b) {
                J.l0.active[b] = src[b] & 1;
                liquid |= (src[b] & 2) != 0;
            }
            // the later kernels take it from here (with surface tension, the scan has set exist[0])
            if (prm->surface_tension == 0.0) J.G.liquid = liquid ? 2 : 1;
        }
    }
    const double t_frac = now_ms();

    // ---------------------------------------------------------------- hierarchy (host)
    // each grid is coarsened (cells halved, a coarse brick active when any child brick is) until
    // it fits one brick; the batch has as many levels as its deepest grid
    std::vector<std::vector<GridLevelHost>> glevels(1);
    for (int g = 0; g < ngrids; ++g) glevels[0].push_back(jobs[g].l0);
    if (multigrid) {
        for (;;) {
            const std::vector<GridLevelHost> &F = glevels.back();
            bool any = false;
            for (const GridLevelHost &H : F) any |= H.coarsens();
            if (!any) break;
            if ((int)glevels.size() >= MAX_LEVELS) return fail(SHKZ_ERR_INVALID, "more than %d multigrid levels", MAX_LEVELS);
            std::vector<GridLevelHost> next(ngrids);
            for (int g = 0; g < ngrids; ++g) {
                const GridLevelHost &H = F[g];
                GridLevelHost &N = next[g];
                if (!H.coarsens()) continue;  // ended on this or a finer level
                grid_level_dims(N, (H.ni + 1) / 2, (H.nj + 1) / 2, (H.nk + 1) / 2);
                N.active.assign((size_t)N.nbx * N.nby * N.nbz, 0);
                const uint8_t *a = H.active.data();
                for (int bx = 0; bx < H.nbx; ++bx)
                    for (int by = 0; by < H.nby; ++by) {
                        uint8_t *row = N.active.data() + ((size_t)(bx / 2) * N.nby + by / 2) * N.nbz;
                        for (int bz = 0; bz < H.nbz; ++bz, ++a) row[bz / 2] |= *a;
                    }
            }
            glevels.push_back(std::move(next));
        }
    }
    P.nlev = (int)glevels.size();
    // fresh levels that keep the previous projection's table storage: the tables of a 256^3 grid
    // take about a megabyte, whose pages would otherwise be mapped anew (faulted in) on every call
    recycle_levels(C->levels, P.nlev);
    // Level 0 first: its tables go up and the assembly starts on the device while the host builds
    // the coarser levels and the transfer tables.
    auto finish_level = [&](int l) {
        LevelHost &L = C->levels[l];
        build_level(L, glevels, l);
        if (!multigrid) {
            // plain CG: no coarse solve anywhere
            L.any_coarsest = false;
            for (GridLevel &gl : L.gl) gl.coarsest = 0;
            for (BrickMeta &m : L.meta) m.flags = 0;
        }
    };
    // Uploads the tables of levels [l0, l1) (bmap | bxyz | meta | gl, and with links also child64 |
    // parent8 | nb27, each 32-byte aligned) from the pinned staging pin into the arena buffer name.
    // A staging buffer stays untouched until this call's next synchronisation.
    auto upload_tables = [&](int l0, int l1, bool links, const char *name, shkz_ctx::Pinned &pin,
                             bool base) -> int {
        auto al = [](size_t n) { return (n + 7) & ~(size_t)7; };
        struct Offs {
            size_t base, xyz, meta, gl, c64, p8, n27;
        };
        std::vector<Offs> offs(P.nlev);
        size_t total = 0;
        for (int l = l0; l < l1; ++l) {
            const LevelHost &L = C->levels[l];
            Offs &o = offs[l];
            o.base = total;
            o.xyz = o.base + (base ? al(L.bmap.size()) : 0);
            o.meta = o.xyz + (base ? al(L.bxyz.size()) : 0);
            o.gl = o.meta + (base ? al(L.meta.size() * (sizeof(BrickMeta) / sizeof(int))) : 0);
            o.c64 = o.gl + (base ? al(L.gl.size() * (sizeof(GridLevel) / sizeof(int))) : 0);
            o.p8 = o.c64 + (links ? al(L.child64.size()) : 0);
            o.n27 = o.p8 + (links ? al(L.parent8.size()) : 0);
            total = o.n27 + (links ? al(L.nb27.size()) : 0);
        }
        total += 8;
        int *dmeta = nullptr;
        SHKZ_TRY(arena_get(AR, name, total, &dmeta));
        void *hmeta = nullptr;
        SHKZ_TRY(ensure_pinned(pin, total * sizeof(int), &hmeta));
        int *host = (int *)hmeta;  // every table is written in full; only alignment padding is left unset
        for (int l = l0; l < l1; ++l) {
            LevelHost &L = C->levels[l];
            const Offs &o = offs[l];
            if (base) {
                std::copy(L.bmap.begin(), L.bmap.end(), host + o.base);
                std::copy(L.bxyz.begin(), L.bxyz.end(), host + o.xyz);
                std::memcpy(host + o.meta, L.meta.data(), L.meta.size() * sizeof(BrickMeta));
                std::memcpy(host + o.gl, L.gl.data(), L.gl.size() * sizeof(GridLevel));
                L.desc.nbricks = L.nbricks;
                L.desc.ngrids = ngrids;
                L.desc.flat = flat ? 1 : 0;
                L.desc.bmap = dmeta + o.base;
                L.desc.bxyz = dmeta + o.xyz;
                L.desc.meta = reinterpret_cast<const BrickMeta *>(dmeta + o.meta);
                L.desc.gl = reinterpret_cast<const GridLevel *>(dmeta + o.gl);
                L.desc.uni = nullptr;
            }
            if (links) {
                std::copy(L.child64.begin(), L.child64.end(), host + o.c64);
                std::copy(L.parent8.begin(), L.parent8.end(), host + o.p8);
                std::copy(L.nb27.begin(), L.nb27.end(), host + o.n27);
                L.desc.child64 = L.child64.empty() ? nullptr : dmeta + o.c64;
                L.desc.parent8 = L.parent8.empty() ? nullptr : dmeta + o.p8;
                L.desc.nb27 = L.nb27.empty() ? nullptr : dmeta + o.n27;
            }
        }
        SHKZ_HIP(hipMemcpyAsync(dmeta, host, total * sizeof(int), hipMemcpyHostToDevice, s));
        return SHKZ_OK;
    };
    const double t_hier = now_ms();
    finish_level(0);
    const double t_built0 = now_ms();
    SHKZ_TRY(upload_tables(0, 1, false, "meta0", C->pin_meta0, true));
    const double t_level0 = now_ms();

    // ---------------------------------------------------------------- operator + vectors
    LevelHost &F0 = C->levels[0];
    const size_t n0 = (size_t)std::max(F0.nbricks, 1) * BRICK_CELLS;
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
    // the assembly also writes level 0's inverse diagonal (the V-cycle's), the brick counts and the
    // BM_UNIFORM records
    const size_t es = P.fp64_levels ? 8 : 4;
    // stored bricks of all levels (the coarse levels are built later: their active positions)
    int count_total = 0;
    for (const std::vector<GridLevelHost> &lv : glevels)
        for (const GridLevelHost &H : lv)
            for (uint8_t a : H.active) count_total += a != 0;
    int *counts = nullptr;
    SHKZ_TRY(arena_get(AR, "counts", (size_t)std::max(count_total, 1), &counts));
    {
        void *pp = nullptr;
        SHKZ_TRY(AR.get("idiag_l0", n0 * es, &pp));
        F0.idiag = multigrid ? pp : nullptr;
        UniCoef *uni = nullptr;
        SHKZ_TRY(arena_get(AR, "uni", (size_t)std::max(F0.nbricks, 1), &uni));
        F0.desc.uni = uni;
    }
    const OperatorFinish FIN{multigrid && !P.fp64_levels ? (float *)F0.idiag : nullptr,
                             multigrid && P.fp64_levels ? (double *)F0.idiag : nullptr,
                             const_cast<BrickMeta *>(F0.desc.meta), const_cast<UniCoef *>(F0.desc.uni), counts};
    // per grid: volume correction (macpressuresolver3.cpp:203-217), warm start, assembly
    for (int g = 0; g < ngrids; ++g) {
        GridJob &J = jobs[g];
        const shkz_grid *grid = J.in;
        if (grid->scene_id) SHKZ_TRY(C->scenes->acquire(grid->scene_id, grid->w, grid->h, grid->d, &J.scene));
        if (prm->gain && prm->target_volume) {
            double &yprev = J.scene ? J.scene->y_prev : C->y_prev;
            const double xv = (prm->current_volume - prm->target_volume) / prm->target_volume;
            const double y = yprev + xv * grid->dt;
            yprev = y;
            const double kp = prm->gain * 2.3 / (25.0 * 0.01);
            const double ki = kp * kp / 16.0;
            J.rhs_correct = -(kp * xv + ki * y) / (xv + 1.0);
        }
        res[g].rhs_correct = J.rhs_correct;
        const double *warm = (prm->warm_start && J.scene && J.scene->has_pressure) ? J.scene->p : nullptr;
        const GridLevel &GL = F0.gl[g];
        const int nbg = GL.brick_end - GL.brick_start;
        K.go("operator", -1, 0,
             (double)nbg * BRICK_CELLS *
                 ((J.G.phi64 ? 8 : 4) + (grid->solid ? 4 : 0) + 12 + 1 + 6 * 8 + (P.mixed ? 16 : 0)),
             dim3(brick_grid(nbg)), dim3(BLOCK), k_operator, J.G, F0.desc, GL, (const uint8_t *)J.cflag, P.dir64,
             P.gx64, P.gy64, P.gz64, dir32, gx32, gy32, gz32, P.b64, P.x64, warm, J.rhs_correct, FIN);
    }
    // the coarser levels and the transfer tables, built on the host while the assembly runs
    const double t_assembly = now_ms();
    int max_bricks = 8;
    for (int l = 1; l < P.nlev; ++l) finish_level(l);
    for (int l = 0; l < P.nlev; ++l) max_bricks = std::max(max_bricks, (int)brick_grid(C->levels[l].nbricks));
    // levels >= 1: their base tables
    SHKZ_TRY(upload_tables(1, P.nlev, false, "meta", C->pin_meta, true));
    if (C->host_links) {
        for (int l = 0; l + 1 < P.nlev; ++l) link_levels(C->levels[l], C->levels[l + 1]);
        for (int l = 1; l < P.nlev; ++l) neighbourhoods(C->levels[l]);
        SHKZ_TRY(upload_tables(0, P.nlev, true, "meta_links", C->pin_links, false));
    } else {
        // every level's transfer and neighbourhood tables, built on the device from the base
        // tables: child64 on levels >= 1, parent8 on all but the coarsest, nb27 on levels >= 1
        std::vector<size_t> o64(P.nlev), o8(P.nlev), o27(P.nlev);
        size_t total = 0;
        for (int l = 0; l < P.nlev; ++l) {
            const size_t nb = (size_t)C->levels[l].nbricks;
            o64[l] = total;
            total += l >= 1 ? 64 * nb : 0;
            o8[l] = total;
            total += l + 1 < P.nlev ? 8 * nb : 0;
            o27[l] = total;
            total += l >= 1 ? 27 * nb : 0;
        }
        int *dl = nullptr;
        SHKZ_TRY(arena_get(AR, "links", std::max<size_t>(total, 1), &dl));
        for (int l = 0; l < P.nlev; ++l) {
            LevelDesc &D = C->levels[l].desc;
            const bool any = C->levels[l].nbricks > 0;
            D.child64 = (l >= 1 && any) ? dl + o64[l] : nullptr;
            D.parent8 = (l + 1 < P.nlev && any) ? dl + o8[l] : nullptr;
            D.nb27 = (l >= 1 && any) ? dl + o27[l] : nullptr;
        }
        for (int l = 0; l < P.nlev; ++l) {
            const LevelHost &L = C->levels[l];
            if (L.nbricks == 0 || (!L.desc.child64 && !L.desc.parent8 && !L.desc.nb27)) continue;
            const LevelDesc &Lf = C->levels[l > 0 ? l - 1 : l].desc, &Lc = C->levels[l + 1 < P.nlev ? l + 1 : l].desc;
            K.go("links", -1, 0, 0, dim3((unsigned)(((size_t)L.nbricks * LINK_SLOTS + 255) / 256)), dim3(256), k_links,
                 L.desc, Lf, Lc, const_cast<int *>(L.desc.child64), const_cast<int *>(L.desc.parent8),
                 const_cast<int *>(L.desc.nb27));
        }
    }
    if (C->trace) {
        std::fprintf(stderr,
                     "[shkz] host ms: queue labelling %.3f, wait for brick flags %.3f, level 0 %.3f (masks %.3f, tables %.3f), "
                     "queue assembly %.3f, coarse levels and tables %.3f\n",
                     t_queued - t_start, t_frac - t_queued, t_level0 - t_frac, t_hier - t_frac, t_built0 - t_hier,
                     t_assembly - t_level0,
                     now_ms() - t_assembly);
        std::fprintf(stderr, "[shkz] %d grid(s), levels:", ngrids);
        for (int l = 0; l < P.nlev; ++l) std::fprintf(stderr, " %d%s", C->levels[l].nbricks, C->levels[l].any_coarsest ? "*" : "");
        std::fprintf(stderr, " bricks\n");
    }
    // reduction scratch, scalars, control blocks
    SHKZ_TRY(arena_get(AR, "partials", (size_t)2 * max_bricks, &P.red.partials));
    SolveScalars *Sbuf = nullptr;
    SHKZ_TRY(arena_get(AR, "scalars", (size_t)3 * ngrids, &Sbuf));
    P.S_out = Sbuf;
    P.S_in = Sbuf + ngrids;
    P.S_fin = Sbuf + 2 * ngrids;
    SHKZ_TRY(arena_get(AR, "ctl", 2, &P.ctl));

    for (int l = 1; l < P.nlev; ++l) {
        LevelHost &L = C->levels[l];
        const size_t n = (size_t)std::max(L.nbricks, 1) * BRICK_CELLS;
        const std::string sfx = std::to_string(l);
        void *pp = nullptr;
        SHKZ_TRY(AR.get("dir_l" + sfx, n * es, &pp));
        L.dir = pp;
        SHKZ_TRY(AR.get("gx_l" + sfx, n * es, &pp));
        L.gx = pp;
        SHKZ_TRY(AR.get("gy_l" + sfx, n * es, &pp));
        L.gy = pp;
        SHKZ_TRY(AR.get("gz_l" + sfx, n * es, &pp));
        L.gz = pp;
        SHKZ_TRY(AR.get("idiag_l" + sfx, n * es, &pp));
        L.idiag = pp;
        SHKZ_TRY(AR.get("rhs_l" + sfx, n * es, &pp));
        L.rhs = pp;
        SHKZ_TRY(AR.get("z_l" + sfx, n * es, &pp));
        L.z = pp;
        SHKZ_TRY(AR.get("tmp_l" + sfx, n * es, &pp));
        L.tmp = pp;
        SHKZ_TRY(AR.get("res_l" + sfx, n * es, &pp));
        L.res = pp;
    }
    {
        void *pp = nullptr;
        SHKZ_TRY(AR.get("res_l0", n0 * es, &pp));
        F0.res = pp;
    }
    const size_t ks = P.mixed ? 4 : 8;
    SHKZ_TRY(AR.get("kp", n0 * ks, &P.kp));
    SHKZ_TRY(AR.get("kq", n0 * ks, &P.kq));
    SHKZ_TRY(AR.get("kz", n0 * ks, &P.kz));
    SHKZ_TRY(AR.get("ktmp", n0 * ks, &P.ktmp));
    if (P.mixed) {
        SHKZ_TRY(AR.get("kr32", n0 * 4, &P.kr));
        SHKZ_TRY(AR.get("kp2", n0 * 4, &P.kp2));
    }


    for (int l = 1; l < P.nlev; ++l) {
        LevelHost &Lf = C->levels[l - 1];
        LevelHost &Lc = C->levels[l];
        const dim3 g(brick_grid(Lc.nbricks)), b(BLOCK);
        if (P.fp64_levels) {
            K.go("coarsen", -1, 0, 0, g, b, k_coarsen<double, double>, Lf.desc, op_of<double>(Lf), Lc.desc,
                 (double *)Lc.dir, (double *)Lc.gx, (double *)Lc.gy, (double *)Lc.gz);
        } else {
            // mixed: level 1 too is formed from the FP32 copy of the level-0 operator the assembly
            // writes (the same sums in double of values already rounded to FP32, stored in FP32;
            // half the bytes of the FP64 operator)
            K.go("coarsen", -1, 0, 0, g, b, k_coarsen<float, float>, Lf.desc, op_of<float>(Lf), Lc.desc,
                 (float *)Lc.dir, (float *)Lc.gx, (float *)Lc.gy, (float *)Lc.gz);
        }
    }
    // components of the coarsest bricks (the coarse solve makes a pure-Neumann component's
    // right-hand side consistent)
    if (multigrid) {
        for (int l = 0; l < P.nlev; ++l) {
            LevelHost &L = C->levels[l];
            L.ccode = nullptr;
            if (!L.any_coarsest) continue;
            SHKZ_TRY(arena_get(AR, "ccode_l" + std::to_string(l), (size_t)L.nbricks * BRICK_CELLS, &L.ccode));
            if (P.fp64_levels)
                K.go("coarse_components", -1, 0, 0, dim3((unsigned)P.ngrids), dim3(BRICK_CELLS),
                     k_coarse_components<double>, L.desc, op_of<double>(L), L.ccode);
            else
                K.go("coarse_components", -1, 0, 0, dim3((unsigned)P.ngrids), dim3(BRICK_CELLS),
                     k_coarse_components<float>, L.desc, op_of<float>(L), L.ccode);
        }
    }
    // unknowns per level and grid (results + byte model of the kernel timers): level 0 counted by
    // the assembly, the coarse levels by their inverse-diagonal kernels below
    int total = 0;
    for (int l = 0; l < P.nlev; ++l) total += C->levels[l].nbricks;
    if (total != count_total) return fail(SHKZ_ERR_DEVICE, "internal: %d stored bricks, %d counted", total, count_total);
    // inverse diagonals for the smoothers (level 0: written by the assembly), with the brick counts
    if (multigrid) {
        int off = C->levels[0].nbricks;
        for (int l = 1; l < P.nlev; ++l) {
            LevelHost &L = C->levels[l];
            const dim3 g(brick_grid(L.nbricks)), b(BLOCK);
            if (P.fp64_levels)
                K.go("idiag", -1, 0, 0, g, b, k_idiag<double>, L.desc, op_of<double>(L), (double *)L.idiag, counts + off);
            else
                K.go("idiag", -1, 0, 0, g, b, k_idiag<float>, L.desc, op_of<float>(L), (float *)L.idiag, counts + off);
            off += L.nbricks;
        }
    }
    // level-0 bricks with a constant operator (BM_UNIFORM, marked by the assembly): the smoothers
    // and operator kernels read their constants instead of the operator arrays
    {
        int *cnt = counts;
        void *hcounts = nullptr;
        SHKZ_TRY(ensure_pinned(C->pin_counts, (size_t)std::max(total, 1) * sizeof(int), &hcounts));
        if (total) SHKZ_HIP(hipMemcpyAsync(hcounts, cnt, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, s));
    }
    // the counts arrive with the next synchronisation; the kernel timers' byte model needs them
    // before the solve's first read-back, the results only at the end
    auto take_counts = [&]() -> int {
        SHKZ_HIP(hipStreamSynchronize(s));
        SHKZ_HIP(hipGetLastError());
        const int *hc = (const int *)C->pin_counts.p;
        int off = 0;
        for (int l = 0; l < P.nlev; ++l) {
            LevelHost &L = C->levels[l];
            L.unknowns_grid.assign(ngrids, 0);
            L.unknowns = 0;
            L.uniform_cells = 0;
            for (int bk = 0; bk < L.nbricks; ++bk) {
                const int n = hc[off + bk] & (UNIFORM_COUNT_BIT - 1);
                L.unknowns_grid[L.bgrid[bk]] += n;
                L.unknowns += n;
                if (hc[off + bk] & UNIFORM_COUNT_BIT) L.uniform_cells += BRICK_CELLS;
            }
            off += L.nbricks;
        }
        for (int g = 0; g < ngrids; ++g) {
            res[g].n_unknowns = F0.unknowns_grid[g];
            res[g].n_bricks = F0.gl[g].brick_end - F0.gl[g].brick_start;
            int depth = 0;
            for (int l = 0; l < P.nlev; ++l) depth += C->levels[l].gl[g].brick_end > C->levels[l].gl[g].brick_start;
            res[g].n_levels = depth;
        }
        return SHKZ_OK;
    };
    if (C->timing || exp) SHKZ_TRY(take_counts());
    if (exp) {
        const GridJob &J = jobs[0];
        K.go("export_system", -1, 0, 0, dim3((unsigned)((J.M + 255) / 256)), dim3(256), k_export_system, J.G, F0.desc,
             F0.gl[0], OpDesc<double>{P.dir64, P.gx64, P.gy64, P.gz64}, (const double *)P.b64, exp->diag, exp->gx,
             exp->gy, exp->gz, exp->rhs);
        SHKZ_HIP(hipStreamSynchronize(s));
        SHKZ_HIP(hipGetLastError());
        if (C->timing) resolve_timing(C);
        return SHKZ_OK;
    }
    const double t_build = now_ms();

    // ---------------------------------------------------------------- solve
    int tuned = 0;
    if (prm->autotune && multigrid && F0.nbricks > 0) {
        const std::string key = tune_key(C, P, prm);
        Knobs k;
        if (tune_find(key, &k)) {
            apply_knobs(P, k, prm);
            tuned = 1;
        } else if (C->under_profiler) {
            // no tuning under rocprofv3 (the library defaults are used; controls tuned by an
            // earlier run still apply through SHKZ_TUNE_CACHE)
            if (C->trace) std::fprintf(stderr, "[shkz] autotune skipped: rocprofv3 tool loaded\n");
        } else {
            SHKZ_TRY(autotune(C, K, P, prm, multigrid));
            tune_store(key, knobs_of(P));
            tuned = 2;
        }
    }
    SolveOut SO;
    SHKZ_TRY(solve_phase(C, K, P, prm, multigrid, SO, res));
    if (!C->timing) SHKZ_TRY(take_counts());
    for (int g = 0; g < ngrids; ++g) {
        res[g].mg_pre_sweeps = P.pre;
        res[g].mg_post_sweeps = P.post;
        res[g].mg_coarse_pre_sweeps = P.pre_c;
        res[g].mg_coarse_post_sweeps = P.post_c;
        res[g].mg_coarse_sweeps = P.coarse_sweeps;
        res[g].mg_omega = P.omega;
        res[g].mg_smoother = P.smoother;
        res[g].mg_local_sweeps = P.local_k;
        res[g].inner_residual = P.eta;
        res[g].tuned = tuned;
    }
    const std::vector<int> &total_iters = SO.total_iters, &outer_passes = SO.outer_passes, &status = SO.status;
    const double t_solve = now_ms();

    // ---------------------------------------------------------------- per grid: scatter + update
    float *pdense = nullptr;
    {
        size_t maxM = 1;
        for (const GridJob &J : jobs) maxM = std::max(maxM, J.M);
        SHKZ_TRY(arena_get(AR, "pdense", maxM, &pdense));
    }
    for (int g = 0; g < ngrids; ++g) {
        GridJob &J = jobs[g];
        const shkz_grid *grid = J.in;
        const GridLevel &GL = F0.gl[g];
        // the velocity update reads a dense float pressure: the caller's, or the scratch array
        float *P32 = grid->pressure ? grid->pressure : pdense;
        uint8_t *unk_out = grid->unknown_out;
        K.go("scatter_pressure", -1, 0, (double)J.M * (4 + (grid->pressure64 ? 8 : 0) + (unk_out ? 2 : 0)),
             dim3(xcd_blocks(J.M, 256)), dim3(256), k_scatter_pressure, J.G, F0.desc, GL, (const double *)P.x64, P32,
             grid->pressure64, (const uint8_t *)J.cflag, unk_out);
        if (J.scene)
            K.go("scatter_pressure", -1, 0, (double)J.M * 8, dim3(xcd_blocks(J.M, 256)), dim3(256),
                 k_scatter_pressure, J.G, F0.desc, GL, (const double *)P.x64, (float *)nullptr, J.scene->p,
                 (const uint8_t *)nullptr, (uint8_t *)nullptr);
        const size_t ne = (size_t)(grid->w + 1) * (grid->h + 1) * (grid->d + 1);
        K.go("velocity_update", -1, 0, (double)J.M * (24 + 4 + (J.G.phi64 ? 8 : 4) + (grid->solid ? 4 : 0)),
             dim3(xcd_blocks(ne, 256)), dim3(256), k_velocity_update, J.G, (const float *)P32);
        if (J.scene) J.scene->has_pressure = true;
        res[g].iterations = total_iters[g];
        res[g].outer_iterations = outer_passes[g];
        res[g].status = status[g];
        res[g].batch_size = ngrids;
    }
    if (C->timing) {
        SHKZ_HIP(hipStreamSynchronize(s));
        resolve_timing(C);
    }
    const double t_end = now_ms();
    int rc = SHKZ_OK;
    for (int g = 0; g < ngrids; ++g) {
        shkz_result &r = res[g];
        r.t_solid_fluid_fractions = t_frac - t_start;
        r.t_build_highres_linsystem = t_build - t_frac;
        r.t_linsolve = t_solve - t_build;
        r.t_update_velocity = t_end - t_solve;
        r.t_projection = t_end - t_start;
        if (status[g] != SHKZ_OK) rc = status[g];
    }
    if (rc != SHKZ_OK) return fail(rc, "non-finite value met in the linear solve");
    return SHKZ_OK;
}

// Host-array path: stages every grid's arrays on the device (buffers named per batch slot),
// projects, copies velocities and requested pressures back.
static int project_host(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *params, shkz_result *res,
                        double *const *export_host) {
    for (int g = 0; g < ngrids; ++g) SHKZ_TRY(validate_grid(&grids[g]));
    SHKZ_HIP(hipSetDevice(C->device));
    Arena &AR = C->arena;
    std::vector<shkz_grid> dg(grids, grids + ngrids);
    struct Back {
        float *dv[3];
        uint8_t *dao[3];
        size_t nf[3];
        float *dp;
        double *dp64;
        uint8_t *dunk;
        size_t M;
    };
    std::vector<Back> back(ngrids);
    for (int g = 0; g < ngrids; ++g) {
        const shkz_grid *grid = &grids[g];
        const std::string sfx = std::to_string(g);
        const size_t w = grid->w, h = grid->h, d = grid->d;
        const bool flat = is_flat(grid);
        const size_t M = w * h * d, Nn = (w + 1) * (h + 1) * (flat ? 1 : d + 1);
        const size_t nf[3] = {(w + 1) * h * d, w * (h + 1) * d, flat ? 0 : w * h * (d + 1)};
        void *phi = nullptr, *solid = nullptr;
        const size_t pe = grid->fluid_dtype == SHKZ_F64 ? 8 : 4, se = grid->solid_dtype == SHKZ_F64 ? 8 : 4;
        SHKZ_TRY(AR.get("h_phi" + sfx, M * pe, &phi));
        SHKZ_HIP(hipMemcpy(phi, grid->fluid, M * pe, hipMemcpyHostToDevice));
        dg[g].fluid = phi;
        if (grid->solid) {
            SHKZ_TRY(AR.get("h_solid" + sfx, Nn * se, &solid));
            SHKZ_HIP(hipMemcpy(solid, grid->solid, Nn * se, hipMemcpyHostToDevice));
            dg[g].solid = solid;
        }
        float *hv[3] = {grid->u, grid->v, grid->w_vel};
        const uint8_t *ha[3] = {grid->active_u, grid->active_v, grid->active_w};
        uint8_t *da[3] = {nullptr, nullptr, nullptr};
        const char *vn[3] = {"h_u", "h_v", "h_w"}, *an[3] = {"h_au", "h_av", "h_aw"};
        Back &B = back[g];
        for (int dim = 0; dim < 3; ++dim) {
            B.dv[dim] = nullptr;
            B.nf[dim] = nf[dim];
            if (!hv[dim]) continue;  // no w velocity in 2D
            void *pp = nullptr;
            SHKZ_TRY(AR.get(vn[dim] + sfx, nf[dim] * 4, &pp));
            B.dv[dim] = (float *)pp;
            SHKZ_HIP(hipMemcpy(B.dv[dim], hv[dim], nf[dim] * 4, hipMemcpyHostToDevice));
            if (ha[dim]) {
                SHKZ_TRY(AR.get(an[dim] + sfx, nf[dim], &pp));
                da[dim] = (uint8_t *)pp;
                SHKZ_HIP(hipMemcpy(da[dim], ha[dim], nf[dim], hipMemcpyHostToDevice));
            }
        }
        dg[g].u = B.dv[0];
        dg[g].v = B.dv[1];
        dg[g].w_vel = B.dv[2];
        dg[g].active_u = da[0];
        dg[g].active_v = da[1];
        dg[g].active_w = da[2];
        uint8_t *hao[3] = {grid->active_out_u, grid->active_out_v, grid->active_out_w};
        const char *aon[3] = {"h_aou", "h_aov", "h_aow"};
        for (int dim = 0; dim < 3; ++dim) {
            B.dao[dim] = nullptr;
            if (!hao[dim] || !hv[dim] || flat) continue;
            void *pp = nullptr;
            SHKZ_TRY(AR.get(aon[dim] + sfx, nf[dim], &pp));
            B.dao[dim] = (uint8_t *)pp;
        }
        dg[g].active_out_u = B.dao[0];
        dg[g].active_out_v = B.dao[1];
        dg[g].active_out_w = B.dao[2];
        B.dunk = nullptr;
        if (grid->unknown_out) {
            void *pp = nullptr;
            SHKZ_TRY(AR.get("h_unk" + sfx, M, &pp));
            B.dunk = (uint8_t *)pp;
        }
        dg[g].unknown_out = B.dunk;
        B.dp = nullptr;
        B.dp64 = nullptr;
        B.M = M;
        if (grid->pressure) {
            void *pp = nullptr;
            SHKZ_TRY(AR.get("h_p" + sfx, M * 4, &pp));
            B.dp = (float *)pp;
        }
        if (grid->pressure64) {
            void *pp = nullptr;
            SHKZ_TRY(AR.get("h_p64" + sfx, M * 8, &pp));
            B.dp64 = (double *)pp;
        }
        dg[g].pressure = B.dp;
        dg[g].pressure64 = B.dp64;
    }
    shkz_params def;
    shkz_default_params(&def);
    if (export_host) {
        const size_t M = back[0].M;
        double *ex[5];
        for (int q = 0; q < 5; ++q) {
            void *pp = nullptr;
            SHKZ_TRY(AR.get("h_export" + std::to_string(q), M * 8, &pp));
            ex[q] = (double *)pp;
        }
        ExportOut eo{ex[0], ex[1], ex[2], ex[3], ex[4]};
        const int rc = project_impl(C, dg.data(), 1, params ? params : &def, res, nullptr, &eo);
        if (rc != SHKZ_OK) return rc;
        for (int q = 0; q < 5; ++q) SHKZ_HIP(hipMemcpy(export_host[q], ex[q], M * 8, hipMemcpyDeviceToHost));
        return SHKZ_OK;
    }
    const int rc = project_impl(C, dg.data(), ngrids, params ? params : &def, res, nullptr);
    if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) return rc;
    for (int g = 0; g < ngrids; ++g) {
        const shkz_grid *grid = &grids[g];
        float *hv[3] = {grid->u, grid->v, grid->w_vel};
        uint8_t *hao[3] = {grid->active_out_u, grid->active_out_v, grid->active_out_w};
        for (int dim = 0; dim < 3; ++dim) {
            if (hv[dim]) SHKZ_HIP(hipMemcpy(hv[dim], back[g].dv[dim], back[g].nf[dim] * 4, hipMemcpyDeviceToHost));
            if (back[g].dao[dim]) SHKZ_HIP(hipMemcpy(hao[dim], back[g].dao[dim], back[g].nf[dim], hipMemcpyDeviceToHost));
        }
        if (grid->pressure) SHKZ_HIP(hipMemcpy(grid->pressure, back[g].dp, back[g].M * 4, hipMemcpyDeviceToHost));
        if (back[g].dunk) SHKZ_HIP(hipMemcpy(grid->unknown_out, back[g].dunk, back[g].M, hipMemcpyDeviceToHost));
        if (grid->pressure64)
            SHKZ_HIP(hipMemcpy(grid->pressure64, back[g].dp64, back[g].M * 8, hipMemcpyDeviceToHost));
    }
    return rc;
}

}  // namespace shkz

// ============================================================================ server
namespace shkz {

// One submitted request: its grid (device pointers), parameters and, once done, its result.
struct ServerJob {
    shkz_grid grid;
    shkz_params params;
    shkz_result result;
    std::string error;
};

// Upper bound of the device working set of one grid in a batch: every brick of the box stored,
// level-0 operator and vectors of the solve, coarse levels (at most 1/7 of level 0), dense flags,
// and the scene's resident pressure.
static uint64_t estimate_bytes(const shkz_grid &g, const shkz_params &p) {
    const uint64_t M = (uint64_t)g.w * g.h * g.d;
    const uint64_t nb = (uint64_t)((g.w + BRICK - 1) / BRICK) * ((g.h + BRICK - 1) / BRICK) * ((g.d + BRICK - 1) / BRICK);
    const bool mixed = p.solver == SHKZ_SOLVER_MGPCG && p.precision == SHKZ_PREC_MIXED;
    // FP64: dir,gx,gy,gz,b,x,r (56) + p,q,z,tmp (32) + res,idiag (16) + coarse 9 arrays x 8 / 7
    // mixed: FP64 operator, b, x, r (56) + FP32 operator (16) + r,x,p,q,z,tmp (24) + res,idiag (8)
    //        + coarse 9 arrays x 4 / 7
    const uint64_t per_cell = mixed ? 56 + 16 + 24 + 8 + 6 : 56 + 32 + 16 + 11;
    return nb * BRICK_CELLS * per_cell + nb * 64 + 2 * M + (g.scene_id ? M * 8 : 0) + ((uint64_t)1 << 20);
}

}  // namespace shkz

struct shkz_server {
    int device = 0;
    shkz_server_config cfg{};
    uint64_t budget = 0;
    SceneStore store;
    std::vector<shkz_ctx *> exec;
    std::unique_ptr<ServeCore> core;
    std::mutex mu;  // tickets and parameter groups
    uint64_t next_ticket = 1;
    std::vector<shkz_params> groups;

    uint64_t held_except(int k) const {
        uint64_t b = store.bytes.load();
        for (int j = 0; j < (int)exec.size(); ++j)
            if (j != k) b += exec[j]->arena.total.load();
        return b;
    }
    void run_batch(int k, std::vector<ServeRequest> &batch, std::vector<int> &status) {
        shkz_ctx *C = exec[k];
        const int n = (int)batch.size();
        std::vector<shkz_grid> grids(n);
        std::vector<shkz_result> res(n);
        for (int i = 0; i < n; ++i) grids[i] = ((ServerJob *)batch[i].payload)->grid;
        const shkz_params prm = ((ServerJob *)batch[0].payload)->params;
        int rc;
        {
            std::lock_guard<std::mutex> g(C->mu);
            rc = hipSetDevice(device) == hipSuccess ? SHKZ_OK : fail(SHKZ_ERR_DEVICE, "hipSetDevice failed");
            if (rc == SHKZ_OK) rc = project_impl(C, grids.data(), n, &prm, res.data(), C->work);
            // outputs are complete when the request is reported done
            if (hipStreamSynchronize(C->work) != hipSuccess && rc == SHKZ_OK)
                rc = fail(SHKZ_ERR_DEVICE, "server stream synchronisation failed");
        }
        const std::string err = rc != SHKZ_OK ? shkz_last_error() : "";
        for (int i = 0; i < n; ++i) {
            ServerJob *J = (ServerJob *)batch[i].payload;
            J->result = res[i];
            J->result.batch_size = n;
            if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC) J->result.status = rc;
            status[i] = J->result.status;
            if (status[i] != SHKZ_OK) J->error = err;
        }
    }
};

// ============================================================================ C ABI
// ============================================================================ after the projection
// dx of a face field, and whether it is a power of two far from the exponent range's ends (then
// its inverse is exact and the kernels multiply by it)
static void vf_set_dx(VelField &F, const shkz_velocity_field *f) {
    F.dx = f->dx > 0 ? f->dx : std::min(1.0 / f->w, std::min(1.0 / f->h, 1.0 / f->d));
    int ex = 0;
    const double mant = std::frexp(F.dx, &ex);
    F.dx_pow2 = (mant == 0.5 && ex > -400 && ex < 400) ? 1 : 0;
    F.inv_dx = 1.0 / F.dx;
}

// macutility3::extrapolate_and_constrain_velocity (macutility3.cpp:89-93) on the work stream:
// `width` in-place extrapolation passes over the three face arrays, then the constraint (which
// also normalises the activity codes of the passes).  vel/act/solid are device arrays.
static int extrapolate_stages(shkz_ctx *C, const shkz_velocity_field *f, int width, int constrain, hipStream_t s,
                              float *const vel[3], uint8_t *const act[3], const void *solid) {
    Arena &AR = C->arena;
    Launcher K{C, s};
    VelField F;
    F.w = f->w;
    F.h = f->h;
    F.d = f->d;
    F.solid64 = f->solid_dtype == SHKZ_F64;
    vf_set_dx(F, f);
    F.solid = solid;
    for (int a = 0; a < 3; ++a) {
        F.vel[a] = vel[a];
        F.act[a] = act[a];
        F.n[a] = (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * (f->d + (a == 2));
    }
    const size_t total = F.n[0] + F.n[1] + F.n[2];
    const size_t nodes = (size_t)(f->w + 1) * (f->h + 1) * (f->d + 1);
    const dim3 grid = vf_grid(F.n), block(VF_BLOCK);
    int *exist = nullptr;
    SHKZ_TRY(arena_get(AR, "vx_exist", 1, &exist));
    K.go("reset", -1, 0, 4.0, dim3(1), dim3(256), k_reset_ints, exist, 1);
    const bool with_solid = constrain && solid;
    if (with_solid) {
        const unsigned nbs = (unsigned)std::min<size_t>((nodes + 255) / 256, 4096);
        K.go("any_negative", -1, 0, (double)nodes * (F.solid64 ? 8 : 4), dim3(nbs), dim3(256), k_any_negative, solid,
             (int)F.solid64, nodes, exist);
    }
    // byte models: a pass scans the activity bytes (neighbour reads hit the cache); the
    // constraint scans the activity and reads the nodal solid once
    if (width > 0) {
        // bit-packed activity: the initial bits and two ping-pong copies
        BitField B[3];
        size_t words = 0, wmax = 0;
        int W[3];
        for (int a = 0; a < 3; ++a) {
            W[a] = (f->d + (a == 2) + 31) / 32;
            const size_t nw = (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * W[a];
            words += nw;
            wmax = std::max(wmax, nw);
        }
        uint32_t *bits = nullptr;
        SHKZ_TRY(arena_get(AR, "vx_bits", 3 * words, &bits));
        size_t off = 0;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) {
                B[c].bits[a] = bits + off;
                B[c].W[a] = W[a];
                off += (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * W[a];
            }
        const dim3 gpack((unsigned)((wmax * 32 + 256 * PACK_GROUPS - 1) / (256 * PACK_GROUPS)), 3),
            gword((unsigned)((wmax + 255) / 256), 3);
        // byte models: the pack reads the activity bytes and writes the bits; a pass reads and
        // writes the bits (the values of the faces it activates are few)
        K.go("extrap_pack", -1, 0, (double)total + 4.0 * words, gpack, block, k_pack_activity, F, B[0]);
        for (int p = 1; p <= width; ++p)
            K.go("extrap_pass", -1, 0, 8.0 * words, gword, block, k_extrap_pass, F, B[p == 1 ? 0 : 1 + (p & 1)],
                 B[1 + ((p + 1) & 1)]);
        K.go("extrap_unpack", -1, 0, 8.0 * words, gword, block, k_unpack_activity, F, B[0], B[1 + ((width + 1) & 1)]);
    }
    float *xs[3] = {nullptr, nullptr, nullptr};
    if (with_solid) {
        const char *names[3] = {"vx_side0", "vx_side1", "vx_side2"};
        for (int a = 0; a < 3; ++a) SHKZ_TRY(arena_get(AR, names[a], F.n[a], &xs[a]));
    }
    // the side buffers are written and read only at constrained faces
    K.go("constrain", -1, 0, with_solid ? (double)total + (double)nodes * (F.solid64 ? 8 : 4) : (double)total, grid,
         block, k_constrain, F, xs[0], xs[1], xs[2], (const int *)exist, with_solid ? 1 : 0);
    if (with_solid)
        K.go("constrain_apply", -1, 0, (double)total, grid, block, k_constrain_apply, F, (const float *)xs[0],
             (const float *)xs[1], (const float *)xs[2], (const int *)exist, 1);
    SHKZ_HIP(hipGetLastError());
    return SHKZ_OK;
}

static int validate_vfield(const shkz_velocity_field *f, int width, int constrain, bool device) {
    if (!f) return fail(SHKZ_ERR_INVALID, "NULL field");
    if (f->w <= 0 || f->h <= 0 || f->d <= 0) return fail(SHKZ_ERR_INVALID, "bad grid %dx%dx%d", f->w, f->h, f->d);
    if ((int64_t)(f->w + 1) * (f->h + 1) * (f->d + 1) >= (int64_t)1 << 32)
        return fail(SHKZ_ERR_INVALID, "grid too large for 32-bit indices");
    if (width < 0 || width > 254) return fail(SHKZ_ERR_INVALID, "width %d outside 0..254", width);
    if (!f->u || !f->v || !f->w_vel || !f->active_u || !f->active_v || !f->active_w)
        return fail(SHKZ_ERR_INVALID, "velocity and activity arrays are required");
    if (f->solid_dtype != SHKZ_F32 && f->solid_dtype != SHKZ_F64)
        return fail(SHKZ_ERR_INVALID, "unknown solid dtype %d", f->solid_dtype);
    if (constrain && f->solid && (f->w < 2 || f->h < 2 || f->d < 2))
        return fail(SHKZ_ERR_INVALID, "the constraint needs at least 2 cells per axis (got %dx%dx%d)", f->w, f->h, f->d);
    (void)device;
    return SHKZ_OK;
}

// Runs fn(stream) on the context's work stream ordered after the caller's stream, and orders the
// caller's stream after it (as project_impl does for the projection).
template <class Fn>
static int on_work_stream(shkz_ctx *C, hipStream_t caller, const char *what, Fn fn) {
    hipStream_t s = C->work;
    if (caller != s) {
        SHKZ_HIP(hipEventRecord(C->ev_in, caller));
        SHKZ_HIP(hipStreamWaitEvent(s, C->ev_in, 0));
    }
    int rc = fn(s);
    if (caller != s) {
        if (hipEventRecord(C->ev_out, s) != hipSuccess || hipStreamWaitEvent(caller, C->ev_out, 0) != hipSuccess) {
            (void)hipStreamSynchronize(s);
            if (rc == SHKZ_OK) rc = fail(SHKZ_ERR_DEVICE, "ordering the caller's stream after the %s failed", what);
        }
    }
    if (rc != SHKZ_OK) (void)hipStreamSynchronize(s);
    return rc;
}

static VelField vel_field_of(const shkz_velocity_field *f, float *const vel[3], uint8_t *const act[3]) {
    VelField F;
    F.w = f->w;
    F.h = f->h;
    F.d = f->d;
    F.solid64 = f->solid_dtype == SHKZ_F64;
    vf_set_dx(F, f);
    F.solid = nullptr;
    for (int a = 0; a < 3; ++a) {
        F.vel[a] = vel[a];
        F.act[a] = act[a];
        F.n[a] = (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * (f->d + (a == 2));
    }
    return F;
}

static int validate_advect(const shkz_velocity_field *f, const shkz_advect_params *p, bool scalar) {
    if (!f || !p) return fail(SHKZ_ERR_INVALID, "NULL field or parameters");
    if (f->w < 2 || f->h < 2 || f->d < 2)
        return fail(SHKZ_ERR_INVALID, "advection needs at least 2 cells per axis (got %dx%dx%d)", f->w, f->h, f->d);
    if ((int64_t)(f->w + 1) * (f->h + 1) * (f->d + 1) >= (int64_t)1 << 32)
        return fail(SHKZ_ERR_INVALID, "grid too large for 32-bit indices");
    if (!f->u || !f->v || !f->w_vel || !f->active_u || !f->active_v || !f->active_w)
        return fail(SHKZ_ERR_INVALID, "velocity and activity arrays are required");
    if (!std::isfinite(p->dt)) return fail(SHKZ_ERR_INVALID, "dt is not finite");
    if (p->trim_narrowband < 0) return fail(SHKZ_ERR_INVALID, "negative TrimNarrowBand");
    if (p->weno != 0 && p->weno != 1) return fail(SHKZ_ERR_INVALID, "weno must be 0 or 1 (got %d)", p->weno);
    if (!scalar && !p->fluid) return fail(SHKZ_ERR_INVALID, "velocity advection needs the liquid level set");
    if (p->fluid && p->fluid_dtype != SHKZ_F32 && p->fluid_dtype != SHKZ_F64)
        return fail(SHKZ_ERR_INVALID, "unknown fluid dtype %d", p->fluid_dtype);
    return SHKZ_OK;
}

// macadvection3::advect_vector(u, u, fluid, dt) on device arrays, on stream s.
static int advect_velocity_stages(shkz_ctx *C, const shkz_velocity_field *f, const shkz_advect_params *p,
                                  hipStream_t s, float *const vel[3], uint8_t *const act[3], const void *fluid) {
    Arena &AR = C->arena;
    Launcher K{C, s};
    VelField F = vel_field_of(f, vel, act);
    const size_t total = F.n[0] + F.n[1] + F.n[2];
    const dim3 grid = vf_grid(F.n), block(VF_BLOCK);
    AdvScratch S;
    const char *nv[3] = {"adv_v0", "adv_v1", "adv_v2"}, *nl[3] = {"adv_lo0", "adv_lo1", "adv_lo2"},
               *nh[3] = {"adv_hi0", "adv_hi1", "adv_hi2"}, *nn[3] = {"adv_nb0", "adv_nb1", "adv_nb2"};
    const bool mc = p->maccormack != 0;
    for (int a = 0; a < 3; ++a) {
        SHKZ_TRY(arena_get(AR, nv[a], F.n[a], &S.v0[a]));
        S.lo[a] = S.hi[a] = nullptr;
        S.nb[a] = nullptr;
        if (mc) {
            SHKZ_TRY(arena_get(AR, nl[a], F.n[a], &S.lo[a]));
            SHKZ_TRY(arena_get(AR, nh[a], F.n[a], &S.hi[a]));
            SHKZ_TRY(arena_get(AR, nn[a], F.n[a], &S.nb[a]));
        }
    }
    const int f64 = p->fluid_dtype == SHKZ_F64;
    // byte models per face: forward reads the face and its activity (neighbours hit the cache)
    // and writes v0 (+ bounds and flag); backward reads v0, bounds, flag, the input and writes it
    if (p->weno)
        K.go("adv_face_fwd_weno", -1, 0, (double)total * (mc ? 18 : 9), grid, block, k_adv_face_fwd<1>, F,
             (const float *)vel[0], (const float *)vel[1], (const float *)vel[2], S, fluid, f64, p->dt,
             (int)p->trim_narrowband, mc ? 1 : 0);
    else
        K.go("adv_face_fwd", -1, 0, (double)total * (mc ? 18 : 9), grid, block, k_adv_face_fwd<0>, F,
             (const float *)vel[0], (const float *)vel[1], (const float *)vel[2], S, fluid, f64, p->dt,
             (int)p->trim_narrowband, mc ? 1 : 0);
    if (mc && p->weno)
        K.go("adv_face_bwd_weno", -1, 0, (double)total * 22, grid, block, k_adv_face_bwd<1>, F, S, -p->dt);
    else if (mc)
        K.go("adv_face_bwd", -1, 0, (double)total * 22, grid, block, k_adv_face_bwd<0>, F, S, -p->dt);
    else
        K.go("adv_face_copy", -1, 0, (double)total * 9, grid, block, k_adv_face_copy, F, S);
    SHKZ_HIP(hipGetLastError());
    return SHKZ_OK;
}

// macadvection3::advect_scalar(q, velocity, fluid, dt) on device arrays, on stream s.
static int advect_scalar_stages(shkz_ctx *C, const shkz_velocity_field *f, float *q, const uint8_t *aq,
                                const shkz_advect_params *p, hipStream_t s, float *const vel[3],
                                uint8_t *const act[3], const void *fluid, int fluid64) {
    Arena &AR = C->arena;
    Launcher K{C, s};
    VelField F = vel_field_of(f, vel, act);
    const size_t M = (size_t)f->w * f->h * f->d;
    const dim3 grid(xcd_blocks(M, VF_BLOCK)), block(VF_BLOCK);
    AdvScratch S = {};
    const bool mc = p->maccormack != 0;
    SHKZ_TRY(arena_get(AR, "adv_q0", M, &S.v0[0]));
    if (mc) {
        SHKZ_TRY(arena_get(AR, "adv_qlo", M, &S.lo[0]));
        SHKZ_TRY(arena_get(AR, "adv_qhi", M, &S.hi[0]));
        SHKZ_TRY(arena_get(AR, "adv_qnb", M, &S.nb[0]));
    }
    if (!fluid) {  // the narrow band of the scalar itself (the level-set tracker's call)
        fluid = q;
        fluid64 = 0;
    }
    if (p->weno)
        K.go("adv_cell_fwd_weno", -1, 0, (double)M * (mc ? 18 : 8), grid, block, k_adv_cell_fwd<1>, F,
             (const float *)q, aq, S, fluid, fluid64, p->dt, (int)p->trim_narrowband, p->background, mc ? 1 : 0);
    else
        K.go("adv_cell_fwd", -1, 0, (double)M * (mc ? 18 : 8), grid, block, k_adv_cell_fwd<0>, F, (const float *)q,
             aq, S, fluid, fluid64, p->dt, (int)p->trim_narrowband, p->background, mc ? 1 : 0);
    if (mc && p->weno)
        K.go("adv_cell_bwd_weno", -1, 0, (double)M * 21, grid, block, k_adv_cell_bwd<1>, F, q, aq, S, -p->dt);
    else if (mc)
        K.go("adv_cell_bwd", -1, 0, (double)M * 21, grid, block, k_adv_cell_bwd<0>, F, q, aq, S, -p->dt);
    else
        K.go("adv_cell_copy", -1, 0, (double)M * 8, grid, block, k_adv_cell_copy, M, q, aq, S);
    SHKZ_HIP(hipGetLastError());
    return SHKZ_OK;
}

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
    C->arena.extra = [C] { return C->own_scenes.bytes.load(); };
    C->own_scenes.may_grow = [C](uint64_t want) {
        return C->arena.total.load() + C->own_scenes.bytes.load() + want <= C->arena.budget;
    };
    if (ensure_host_scalars(C, 16) != SHKZ_OK ||
        hipHostMalloc((void **)&C->host_ctl, sizeof(BatchCtl), hipHostMallocDefault) != hipSuccess) {
        if (C->host_scalars) (void)hipHostFree(C->host_scalars);
        delete C;
        return fail(SHKZ_ERR_NOMEM, "pinned scalar buffer allocation failed");
    }
    if (hipStreamCreateWithFlags(&C->work, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_out, hipEventDisableTiming) != hipSuccess) {
        (void)hipHostFree(C->host_scalars);
        (void)hipHostFree(C->host_ctl);
        delete C;
        return fail(SHKZ_ERR_DEVICE, "stream/event creation failed");
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            std::string k = prop.gcnArchName;
            k += "/cu";
            k += std::to_string(prop.multiProcessorCount);
            for (char &c : k)
                if (c == ' ' || c == '|' || c == '\t') c = '_';
            C->device_key = k;
        } else {
            C->device_key = "unknown";
        }
    }
    const char *ng = std::getenv("SHKZ_NO_GRAPHS");
    C->use_graphs = !(ng && ng[0] == '1');
    const char *tr = std::getenv("SHKZ_TRACE");
    C->trace = tr && tr[0] == '1';
    const char *hl = std::getenv("SHKZ_HOST_LINKS");
    C->host_links = hl && hl[0] == '1';
    const char *kd = std::getenv("SHKZ_KEEP_DIRECTION");
    C->keep_direction = kd && kd[0] == '1';
    const char *nf = std::getenv("SHKZ_NO_FUSE");
    C->fuse_krylov = !(nf && nf[0] == '1');
    C->under_profiler = std::getenv("ROCP_TOOL_LIBRARIES") || std::getenv("ROCPROF_OUTPUT_PATH");
    const char *mb = std::getenv("SHKZ_MULTI_MAX_BRICKS");
    if (mb && mb[0]) C->multi_max_bricks = std::atoi(mb);
    *out = C;
    return SHKZ_OK;
}

void shkz_destroy(shkz_ctx *C) {
    if (!C) return;
    (void)hipSetDevice(C->device);
    if (C->work) (void)hipStreamSynchronize(C->work);
    for (auto &t : C->pending) {
        (void)hipEventDestroy(t.ev0);
        (void)hipEventDestroy(t.ev1);
    }
    for (hipEvent_t e : C->event_pool) (void)hipEventDestroy(e);
    for (auto &g : C->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    C->graphs.clear();
    if (C->ev_in) (void)hipEventDestroy(C->ev_in);
    if (C->ev_out) (void)hipEventDestroy(C->ev_out);
    if (C->work) (void)hipStreamDestroy(C->work);
    C->arena.clear();
    C->own_scenes.clear();
    if (C->host_scalars) (void)hipHostFree(C->host_scalars);
    if (C->host_ctl) (void)hipHostFree(C->host_ctl);
    for (shkz_ctx::Pinned *b : {&C->pin_flags, &C->pin_meta0, &C->pin_meta, &C->pin_links, &C->pin_counts})
        if (b->p) (void)hipHostFree(b->p);
    delete C;
}

// The 3D entry points take 3D grids only (a 2D grid goes through the *2d entry points).
static int require_3d(const shkz_grid *grids, int n) {
    for (int g = 0; g < n; ++g)
        if (!grids[g].w_vel) return fail(SHKZ_ERR_INVALID, "the w velocity is required (2D grids: shkz_project2d*)");
    return SHKZ_OK;
}

int shkz_project_batch_device(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *params,
                              shkz_result *results, void *stream) {
    if (!C || !grids || !results || ngrids <= 0) return fail(SHKZ_ERR_INVALID, "NULL argument or empty batch");
    SHKZ_TRY(require_3d(grids, ngrids));
    shkz_params def;
    shkz_default_params(&def);
    const shkz_params *p = params ? params : &def;
    std::lock_guard<std::mutex> lock(C->mu);
    return project_impl(C, grids, ngrids, p, results, (hipStream_t)stream);
}

int shkz_extrapolate_velocity_device(shkz_ctx *C, const shkz_velocity_field *f, int32_t width, int32_t constrain,
                                     void *stream) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_vfield(f, width, constrain, true));
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    float *const vel[3] = {f->u, f->v, f->w_vel};
    uint8_t *const act[3] = {f->active_u, f->active_v, f->active_w};
    return on_work_stream(C, (hipStream_t)stream, "extrapolation",
                          [&](hipStream_t s) { return extrapolate_stages(C, f, width, constrain, s, vel, act, f->solid); });
}

void shkz_advect_default_params(shkz_advect_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->maccormack = 1;       // macadvection3.cpp:291
    p->trim_narrowband = 1;  // :293
    p->fluid_dtype = SHKZ_F32;
}

int shkz_advect_velocity_device(shkz_ctx *C, const shkz_velocity_field *f, const shkz_advect_params *p, void *stream) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_advect(f, p, false));
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    float *const vel[3] = {f->u, f->v, f->w_vel};
    uint8_t *const act[3] = {f->active_u, f->active_v, f->active_w};
    return on_work_stream(C, (hipStream_t)stream, "advection",
                          [&](hipStream_t s) { return advect_velocity_stages(C, f, p, s, vel, act, p->fluid); });
}

int shkz_advect_scalar_device(shkz_ctx *C, const shkz_velocity_field *f, float *q, const uint8_t *active_q,
                              const shkz_advect_params *p, void *stream) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_advect(f, p, true));
    if (!q) return fail(SHKZ_ERR_INVALID, "NULL scalar");
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    float *const vel[3] = {f->u, f->v, f->w_vel};
    uint8_t *const act[3] = {f->active_u, f->active_v, f->active_w};
    return on_work_stream(C, (hipStream_t)stream, "advection", [&](hipStream_t s) {
        return advect_scalar_stages(C, f, q, active_q, p, s, vel, act, p->fluid, p->fluid_dtype == SHKZ_F64);
    });
}

// Host-array forms: stage the arrays in the arena, run, copy the results back.
static int stage_field(shkz_ctx *C, const shkz_velocity_field *f, hipStream_t s, float *vel[3], uint8_t *act[3],
                       size_t n[3]) {
    Arena &AR = C->arena;
    float *const hv[3] = {f->u, f->v, f->w_vel};
    uint8_t *const ha[3] = {f->active_u, f->active_v, f->active_w};
    const char *vn[3] = {"hx_u", "hx_v", "hx_w"}, *an[3] = {"hx_au", "hx_av", "hx_aw"};
    for (int a = 0; a < 3; ++a) {
        n[a] = (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * (f->d + (a == 2));
        SHKZ_TRY(arena_get(AR, vn[a], n[a], &vel[a]));
        SHKZ_TRY(arena_get(AR, an[a], n[a], &act[a]));
        SHKZ_HIP(hipMemcpyAsync(vel[a], hv[a], n[a] * sizeof(float), hipMemcpyHostToDevice, s));
        SHKZ_HIP(hipMemcpyAsync(act[a], ha[a], n[a], hipMemcpyHostToDevice, s));
    }
    return SHKZ_OK;
}

static int stage_cells(shkz_ctx *C, const char *name, const void *host, size_t bytes, hipStream_t s, void **dev) {
    *dev = nullptr;
    if (!host) return SHKZ_OK;
    SHKZ_TRY(C->arena.get(name, bytes, dev));
    SHKZ_HIP(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, s));
    return SHKZ_OK;
}

int shkz_advect_velocity(shkz_ctx *C, const shkz_velocity_field *f, const shkz_advect_params *p) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_advect(f, p, false));
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    hipStream_t s = C->work;
    float *vel[3];
    uint8_t *act[3];
    size_t n[3];
    SHKZ_TRY(stage_field(C, f, s, vel, act, n));
    const size_t M = (size_t)f->w * f->h * f->d;
    void *fluid = nullptr;
    SHKZ_TRY(stage_cells(C, "hx_fluid", p->fluid, M * (p->fluid_dtype == SHKZ_F64 ? 8 : 4), s, &fluid));
    const int rc = advect_velocity_stages(C, f, p, s, vel, act, fluid);
    float *const hv[3] = {f->u, f->v, f->w_vel};
    if (rc == SHKZ_OK)
        for (int a = 0; a < 3; ++a)
            SHKZ_HIP(hipMemcpyAsync(hv[a], vel[a], n[a] * sizeof(float), hipMemcpyDeviceToHost, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    return rc;
}

int shkz_advect_scalar(shkz_ctx *C, const shkz_velocity_field *f, float *q, const uint8_t *active_q,
                       const shkz_advect_params *p) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_advect(f, p, true));
    if (!q) return fail(SHKZ_ERR_INVALID, "NULL scalar");
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    hipStream_t s = C->work;
    float *vel[3];
    uint8_t *act[3];
    size_t n[3];
    SHKZ_TRY(stage_field(C, f, s, vel, act, n));
    const size_t M = (size_t)f->w * f->h * f->d;
    void *dq = nullptr, *daq = nullptr, *fluid = nullptr;
    SHKZ_TRY(stage_cells(C, "hx_q", q, M * 4, s, &dq));
    SHKZ_TRY(stage_cells(C, "hx_aq", active_q, M, s, &daq));
    SHKZ_TRY(stage_cells(C, "hx_fluid", p->fluid, M * (p->fluid_dtype == SHKZ_F64 ? 8 : 4), s, &fluid));
    const int rc = advect_scalar_stages(C, f, (float *)dq, (const uint8_t *)daq, p, s, vel, act, fluid,
                                        p->fluid_dtype == SHKZ_F64);
    if (rc == SHKZ_OK) SHKZ_HIP(hipMemcpyAsync(q, dq, M * 4, hipMemcpyDeviceToHost, s));
    SHKZ_HIP(hipStreamSynchronize(s));
    return rc;
}

int shkz_extrapolate_velocity(shkz_ctx *C, const shkz_velocity_field *f, int32_t width, int32_t constrain) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    SHKZ_TRY(validate_vfield(f, width, constrain, false));
    std::lock_guard<std::mutex> lock(C->mu);
    SHKZ_HIP(hipSetDevice(C->device));
    hipStream_t s = C->work;
    Arena &AR = C->arena;
    float *const hv[3] = {f->u, f->v, f->w_vel};
    uint8_t *const ha[3] = {f->active_u, f->active_v, f->active_w};
    float *vel[3];
    uint8_t *act[3];
    size_t n[3];
    const char *vn[3] = {"hx_u", "hx_v", "hx_w"}, *an[3] = {"hx_au", "hx_av", "hx_aw"};
    for (int a = 0; a < 3; ++a) {
        n[a] = (size_t)(f->w + (a == 0)) * (f->h + (a == 1)) * (f->d + (a == 2));
        SHKZ_TRY(arena_get(AR, vn[a], n[a], &vel[a]));
        SHKZ_TRY(arena_get(AR, an[a], n[a], &act[a]));
        SHKZ_HIP(hipMemcpyAsync(vel[a], hv[a], n[a] * sizeof(float), hipMemcpyHostToDevice, s));
        SHKZ_HIP(hipMemcpyAsync(act[a], ha[a], n[a], hipMemcpyHostToDevice, s));
    }
    void *solid = nullptr;
    if (f->solid && constrain) {
        const size_t nodes = (size_t)(f->w + 1) * (f->h + 1) * (f->d + 1);
        const size_t es = f->solid_dtype == SHKZ_F64 ? 8 : 4;
        SHKZ_TRY(AR.get("hx_solid", nodes * es, &solid));
        SHKZ_HIP(hipMemcpyAsync(solid, f->solid, nodes * es, hipMemcpyHostToDevice, s));
    }
    int rc = extrapolate_stages(C, f, width, constrain, s, vel, act, solid);
    if (rc == SHKZ_OK)
        for (int a = 0; a < 3; ++a) {
            SHKZ_HIP(hipMemcpyAsync(hv[a], vel[a], n[a] * sizeof(float), hipMemcpyDeviceToHost, s));
            SHKZ_HIP(hipMemcpyAsync(ha[a], act[a], n[a], hipMemcpyDeviceToHost, s));
        }
    SHKZ_HIP(hipStreamSynchronize(s));
    return rc;
}

int shkz_project_device(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result,
                        void *stream) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    return shkz_project_batch_device(C, grid, 1, params, result, stream);
}

int shkz_project_batch(shkz_ctx *C, const shkz_grid *grids, int ngrids, const shkz_params *params,
                       shkz_result *results) {
    if (!C || !grids || !results || ngrids <= 0) return fail(SHKZ_ERR_INVALID, "NULL argument or empty batch");
    SHKZ_TRY(require_3d(grids, ngrids));
    for (int g = 0; g < ngrids; ++g) std::memset(&results[g], 0, sizeof(shkz_result));
    std::lock_guard<std::mutex> lock(C->mu);
    const int rc = project_host(C, grids, ngrids, params, results, nullptr);
    for (int g = 0; g < ngrids; ++g)
        if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC && results[g].status == SHKZ_OK) results[g].status = rc;
    return rc;
}

int shkz_project(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, shkz_result *result) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    return shkz_project_batch(C, grid, 1, params, result);
}

int shkz_export_system(shkz_ctx *C, const shkz_grid *grid, const shkz_params *params, double *diag, double *gx,
                       double *gy, double *gz, double *rhs) {
    if (!C || !grid) return fail(SHKZ_ERR_INVALID, "NULL argument");
    if (!diag || !gx || !gy || !gz || !rhs) return fail(SHKZ_ERR_INVALID, "export arrays are required");
    SHKZ_TRY(require_3d(grid, 1));
    double *outs[5] = {diag, gx, gy, gz, rhs};
    shkz_result r;
    std::lock_guard<std::mutex> lock(C->mu);
    return project_host(C, grid, 1, params, &r, outs);
}

int shkz_inject_alloc_failure(shkz_ctx *C, int32_t after) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    C->arena.fail_after = after < 0 ? -1 : after;
    return SHKZ_OK;
}

int shkz_release_scene(shkz_ctx *C, uint64_t scene_id) {
    if (!C) return fail(SHKZ_ERR_INVALID, "NULL context");
    std::lock_guard<std::mutex> lock(C->mu);
    return C->scenes->release(scene_id);
}

uint64_t shkz_device_bytes(const shkz_ctx *C) {
    return C ? (uint64_t)C->arena.total.load() + (C->scenes == &C->own_scenes ? C->own_scenes.bytes.load() : 0) : 0;
}

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
        SHKZ_HIP(hipStreamSynchronize(C->work));
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
        SHKZ_HIP(hipStreamSynchronize(C->work));
        resolve_timing(C);
    }
    C->totals.clear();
    return SHKZ_OK;
}

}  // extern "C"

// ============================================================================ C ABI: 2D
// macpressuresolver2::project (src/projection/macpressuresolver2.cpp:48-232) on the same engine:
// the 2D grid is a one-cell-thick 3D grid (d = 1, no z faces), with macutility2's edge area
// fractions and the 2D module's assembly order.
static shkz_grid grid_of_2d(const shkz_grid2d *g) {
    shkz_grid q;
    std::memset(&q, 0, sizeof(q));
    q.w = g->w;
    q.h = g->h;
    q.d = 1;
    q.dx = g->dx > 0 ? g->dx : std::min(1.0 / g->w, 1.0 / g->h);
    q.dt = g->dt;
    q.fluid = g->fluid;
    q.fluid_dtype = g->fluid_dtype;
    q.solid = g->solid;
    q.solid_dtype = g->solid_dtype;
    q.u = g->u;
    q.v = g->v;
    q.w_vel = nullptr;
    q.active_u = g->active_u;
    q.active_v = g->active_v;
    q.pressure = g->pressure;
    q.pressure64 = g->pressure64;
    q.scene_id = g->scene_id;
    q.unknown_out = g->unknown_out;
    return q;
}

extern "C" {

int shkz_project2d(shkz_ctx *C, const shkz_grid2d *grid, const shkz_params *params, shkz_result *result) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    const shkz_grid q = grid_of_2d(grid);
    std::memset(result, 0, sizeof(*result));
    std::lock_guard<std::mutex> lock(C->mu);
    const int rc = project_host(C, &q, 1, params, result, nullptr);
    if (rc != SHKZ_OK && rc != SHKZ_ERR_NUMERIC && result->status == SHKZ_OK) result->status = rc;
    return rc;
}

int shkz_project2d_batch_device(shkz_ctx *C, const shkz_grid2d *grids, int ngrids, const shkz_params *params,
                                shkz_result *results, void *stream) {
    if (!C || !grids || !results || ngrids <= 0) return fail(SHKZ_ERR_INVALID, "NULL argument or empty batch");
    std::vector<shkz_grid> q(ngrids);
    for (int g = 0; g < ngrids; ++g) q[g] = grid_of_2d(&grids[g]);
    shkz_params def;
    shkz_default_params(&def);
    std::lock_guard<std::mutex> lock(C->mu);
    return project_impl(C, q.data(), ngrids, params ? params : &def, results, (hipStream_t)stream);
}

int shkz_project2d_device(shkz_ctx *C, const shkz_grid2d *grid, const shkz_params *params, shkz_result *result,
                          void *stream) {
    if (!C || !grid || !result) return fail(SHKZ_ERR_INVALID, "NULL argument");
    return shkz_project2d_batch_device(C, grid, 1, params, result, stream);
}

int shkz_export_system2d(shkz_ctx *C, const shkz_grid2d *grid, const shkz_params *params, double *diag, double *gx,
                         double *gy, double *rhs) {
    if (!C || !grid) return fail(SHKZ_ERR_INVALID, "NULL argument");
    if (!diag || !gx || !gy || !rhs) return fail(SHKZ_ERR_INVALID, "export arrays are required");
    const shkz_grid q = grid_of_2d(grid);
    std::vector<double> gz((size_t)grid->w * grid->h);
    double *outs[5] = {diag, gx, gy, gz.data(), rhs};
    shkz_result r;
    std::lock_guard<std::mutex> lock(C->mu);
    return project_host(C, &q, 1, params, &r, outs);
}

}  // extern "C"

// ============================================================================ C ABI: server
extern "C" {

void shkz_server_default_config(shkz_server_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->executors = 2;
    cfg->max_batch_grids = 64;
    cfg->max_batch_cells = (int64_t)1 << 26;
}

int shkz_server_create(const shkz_server_config *cfg_in, shkz_server **out) {
    if (!out) return fail(SHKZ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    shkz_server_config cfg;
    shkz_server_default_config(&cfg);
    if (cfg_in) {
        cfg = *cfg_in;
        if (cfg.executors <= 0) cfg.executors = 2;
        if (cfg.max_batch_grids <= 0) cfg.max_batch_grids = 64;
        if (cfg.max_batch_cells <= 0) cfg.max_batch_cells = (int64_t)1 << 26;
    }
    int ndev = 0;
    SHKZ_HIP(hipGetDeviceCount(&ndev));
    if (cfg.device < 0 || cfg.device >= ndev)
        return fail(SHKZ_ERR_INVALID, "device %d out of range (%d devices)", cfg.device, ndev);
    SHKZ_HIP(hipSetDevice(cfg.device));
    size_t free_b = 0, total_b = 0;
    SHKZ_HIP(hipMemGetInfo(&free_b, &total_b));
    auto *srv = new shkz_server();
    srv->device = cfg.device;
    srv->cfg = cfg;
    srv->budget = cfg.hbm_budget_bytes ? cfg.hbm_budget_bytes : (uint64_t)(0.9 * (double)total_b);
    srv->store.may_grow = [srv](uint64_t want) { return srv->held_except(-1) + want <= srv->budget; };
    for (int k = 0; k < cfg.executors; ++k) {
        shkz_ctx *C = nullptr;
        const int rc = shkz_create(cfg.device, srv->budget, &C);
        if (rc != SHKZ_OK) {
            for (shkz_ctx *E : srv->exec) shkz_destroy(E);
            delete srv;
            return rc;
        }
        C->scenes = &srv->store;
        C->arena.extra = [srv, k] { return srv->held_except(k); };
        srv->exec.push_back(C);
    }
    ServeLimits lim;
    lim.executors = cfg.executors;
    lim.max_batch_grids = cfg.max_batch_grids;
    lim.max_batch_cells = cfg.max_batch_cells;
    lim.budget = srv->budget;
    srv->core.reset(new ServeCore(
        lim, [srv](int k, std::vector<ServeRequest> &b, std::vector<int> &st) { srv->run_batch(k, b, st); },
        [srv] { return srv->store.bytes.load(); }));
    *out = srv;
    return SHKZ_OK;
}

void shkz_server_destroy(shkz_server *srv) {
    if (!srv) return;
    srv->core->shutdown();
    for (const ServeOutcome &o : srv->core->take_all()) delete (ServerJob *)o.payload;  // never collected
    srv->core.reset();
    for (shkz_ctx *C : srv->exec) shkz_destroy(C);
    srv->store.clear();
    delete srv;
}

int shkz_submit(shkz_server *srv, const shkz_grid *grid, const shkz_params *params, uint64_t *ticket) {
    if (!srv || !grid || !ticket) return fail(SHKZ_ERR_INVALID, "NULL argument");
    SHKZ_TRY(require_3d(grid, 1));
    SHKZ_TRY(validate_grid(grid));
    auto *J = new ServerJob();
    J->grid = *grid;
    if (params) J->params = *params;
    else shkz_default_params(&J->params);
    std::memset(&J->result, 0, sizeof(J->result));
    ServeRequest r;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        size_t gi = 0;
        for (; gi < srv->groups.size(); ++gi)
            if (std::memcmp(&srv->groups[gi], &J->params, sizeof(shkz_params)) == 0) break;
        if (gi == srv->groups.size()) srv->groups.push_back(J->params);
        r.ticket = srv->next_ticket++;
        r.group = gi;
    }
    r.scene_id = grid->scene_id;
    r.est_bytes = estimate_bytes(*grid, J->params);
    r.cells = (int64_t)grid->w * grid->h * grid->d;
    r.payload = J;
    if (srv->core->submit(r) != 0) {
        delete J;
        return fail(SHKZ_ERR_INVALID, "server is shutting down");
    }
    *ticket = r.ticket;
    return SHKZ_OK;
}

static int server_take(shkz_server *srv, int got, const ServeOutcome &o, shkz_result *result) {
    if (got == 1) return SHKZ_PENDING;
    if (got != 0) return fail(SHKZ_ERR_NOT_FOUND, "unknown ticket");
    auto *J = (ServerJob *)o.payload;
    shkz_result r = J->result;
    if (o.status == SHKZ_ERR_BUDGET && r.status == SHKZ_OK) {
        r.status = SHKZ_ERR_BUDGET;
        J->error = "the request's estimated working set exceeds the server's device budget";
    }
    if (o.status == SHKZ_ERR_CANCELLED && r.status == SHKZ_OK) {
        r.status = SHKZ_ERR_CANCELLED;
        J->error = "the request was cancelled before it started";
    }
    r.t_queue = o.t_start_ms - o.t_submit_ms;
    r.t_latency = o.t_end_ms - o.t_submit_ms;
    if (result) *result = r;
    const int st = r.status;
    if (st != SHKZ_OK) g_last_error = J->error;
    delete J;
    return st;
}

int shkz_cancel(shkz_server *srv, uint64_t ticket) {
    if (!srv) return fail(SHKZ_ERR_INVALID, "NULL server");
    const int rc = srv->core->cancel(ticket);
    if (rc == -5) return fail(SHKZ_ERR_NOT_FOUND, "unknown ticket");
    return rc == 0 ? SHKZ_OK : SHKZ_PENDING;
}
int shkz_poll(shkz_server *srv, uint64_t ticket, shkz_result *result) {
    if (!srv) return fail(SHKZ_ERR_INVALID, "NULL server");
    ServeOutcome o;
    const int got = srv->core->poll(ticket, &o);
    return server_take(srv, got, o, result);
}

int shkz_wait(shkz_server *srv, uint64_t ticket, double timeout_ms, shkz_result *result) {
    if (!srv) return fail(SHKZ_ERR_INVALID, "NULL server");
    ServeOutcome o;
    const int got = srv->core->wait(ticket, timeout_ms, &o);
    return server_take(srv, got, o, result);
}

int shkz_server_release_scene(shkz_server *srv, uint64_t scene_id) {
    if (!srv) return fail(SHKZ_ERR_INVALID, "NULL server");
    if (srv->core->scene_busy(scene_id))
        return fail(SHKZ_ERR_INVALID, "scene %llu has queued or running requests", (unsigned long long)scene_id);
    return srv->store.release(scene_id);
}

int shkz_server_get_stats(shkz_server *srv, shkz_server_stats *out) {
    if (!srv || !out) return fail(SHKZ_ERR_INVALID, "NULL argument");
    const ServeStats s = srv->core->stats();
    std::memset(out, 0, sizeof(*out));
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->rejected = s.rejected;
    out->batches = s.batches;
    out->batched_requests = s.batched_requests;
    out->max_batch = s.max_batch;
    out->max_in_flight_batches = s.max_in_flight_batches;
    out->max_in_flight_bytes = s.max_in_flight_bytes;
    out->resident_bytes = srv->store.bytes.load();
    out->device_bytes = srv->held_except(-1);
    out->cancelled = s.cancelled;
    return SHKZ_OK;
}

}  // extern "C"
