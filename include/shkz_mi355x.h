/*
 * shkz_mi355x.h -- C ABI of the MI355X pressure projection for Shiokaze.
 *
 * One call of shkz_project() is one call of the reference's
 *   macproject3_interface::project(dt, velocity, solid, fluid, surface_tension, rigidbodies)
 *   (include/shiokaze/projection/macproject3_interface.h:67-72, implemented by
 *    macpressuresolver3::project, src/projection/macpressuresolver3.cpp:50-272)
 * on dense arrays.  Everything crosses the boundary as plain pointers and sizes.
 *
 * Array layouts (all dense, z index fastest):
 *   cell  (w, h, d)            index (i*h + j)*d + k          fluid level set, pressure
 *   u     (w+1, h, d)          index (i*h + j)*d + k
 *   v     (w, h+1, d)          index (i*(h+1) + j)*d + k
 *   w     (w, h, d+1)          index (i*h + j)*(d+1) + k
 *   nodal (w+1, h+1, d+1)      index (i*(h+1) + j)*(d+1) + k  solid level set
 * Face (i,j,k) of dimension dim separates cells (i,j,k) and (i,j,k) - e_dim
 * (include/shiokaze/math/vec.h:353-381).
 *
 * Errors: every entry point returns an int status (SHKZ_OK = 0, negative = error) and
 * records a message readable with shkz_last_error() on the calling thread.  A solve that
 * reaches max_iterations is not an error (as in the reference, pcg_solver.h:292-294): the
 * result carries converged = 0.
 */
#ifndef SHKZ_MI355X_H
#define SHKZ_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHKZ_ABI_VERSION 8

/* ---------------------------------------------------------------- enums */
enum shkz_status {
    SHKZ_OK = 0,
    SHKZ_PENDING = 1,        /* shkz_poll / shkz_wait: the request is queued or running */
    SHKZ_ERR_INVALID = -1,   /* bad argument or shape */
    SHKZ_ERR_DEVICE = -2,    /* HIP runtime error */
    SHKZ_ERR_NOMEM = -3,     /* device allocation failed */
    SHKZ_ERR_BUDGET = -4,    /* request exceeds the context's HBM budget */
    SHKZ_ERR_NOT_FOUND = -5, /* unknown scene or ticket */
    SHKZ_ERR_NUMERIC = -6,   /* NaN/Inf met in the solve (reference: pcg_solver.h:267-271) */
    SHKZ_ERR_CANCELLED = -7  /* server: the request was cancelled before it started */
};

enum shkz_dtype { SHKZ_F32 = 0, SHKZ_F64 = 1 };

/* Linear solver behind the projection. */
enum shkz_solver {
    SHKZ_SOLVER_MGPCG = 0, /* multigrid-preconditioned CG (V-cycle preconditioner) */
    SHKZ_SOLVER_CG = 1     /* unpreconditioned CG: the iteration the reference "pcg" module
                              actually runs (its MIC(0) output is discarded,
                              local/include/pcgsolver/pcg_solver.h:379-384) */
};

enum shkz_precision {
    SHKZ_PREC_MIXED = 0, /* FP32 V-cycle / inner CG inside an FP64 defect-correction loop */
    SHKZ_PREC_FP64 = 1   /* every vector and operator in FP64 */
};

/* Finest-level smoother of the V-cycle.  Coarser levels always use damped Jacobi. */
enum shkz_smoother {
    SHKZ_SMOOTH_DEFAULT = 0, /* brick-local in mixed precision, Jacobi in FP64 (or the tuner's choice) */
    SHKZ_SMOOTH_JACOBI = 1,  /* damped Jacobi, one sweep per mg_pre/post_sweeps */
    SHKZ_SMOOTH_BRICK = 2    /* brick-local symmetric Gauss-Seidel (block Jacobi across 8^3 bricks, two
                                red-black sweeps inside each), one application per mg_pre/post_sweeps;
                                mixed precision only (FP64 solves use Jacobi) */
};

enum shkz_norm {
    SHKZ_NORM_INF = 0, /* reference criterion: ||r||_inf <= residual * ||r0||_inf (pcg_solver.h:252-283) */
    SHKZ_NORM_L2 = 1   /* ||r||_2 <= residual * ||r0||_2 */
};

/* ---------------------------------------------------------------- grid */
typedef struct shkz_grid {
    int32_t w, h, d;          /* cells per axis */
    double dx;                /* 0 selects min(1/w,1/h,1/d) (shape3::dx, shape.h:755-759) */
    double dt;                /* time step */
    const void *fluid;        /* cell level set (w*h*d), negative inside liquid */
    int32_t fluid_dtype;      /* SHKZ_F32 or SHKZ_F64 */
    int32_t solid_dtype;
    const void *solid;        /* nodal solid level set ((w+1)(h+1)(d+1)), negative inside the
                                 solid, or NULL for walls only (no-solid branch,
                                 macutility3.cpp:149-164) */
    float *u, *v, *w_vel;     /* face velocities, updated in place */
    const uint8_t *active_u;  /* optional per-face activity (NULL = every face active); inactive */
    const uint8_t *active_v;  /* faces are neither read as nonzero by the caller's contract nor */
    const uint8_t *active_w;  /* written, matching macarray3::parallel_actives */
    float *pressure;          /* optional out: m_pressure (cells, float; 0 off the unknowns) */
    double *pressure64;       /* optional out: the FP64 solution (cells) */
    uint64_t scene_id;        /* 0 = stateless; otherwise keys warm-start pressure and the
                                 volume-correction integrator kept on the device across calls */
    uint8_t *active_out_u;    /* optional out (3D grids): 1 where the face is active after the */
    uint8_t *active_out_v;    /* projection, 0 where it is switched off or was inactive: open */
    uint8_t *active_out_w;    /* faces and closed boundary faces with liquid inside stay active
                                 (macpressuresolver3.cpp:252-266).  The input of the velocity
                                 extrapolation (shkz_extrapolate_velocity). */
    uint8_t *unknown_out;     /* optional out (cells): 1 on the cells that are unknowns of the
                                 system (the reference's index_map actives, macpressuresolver3.cpp:
                                 121-156), 0 elsewhere: the cells whose m_pressure the reference
                                 sets (:245-248), zeros included */
} shkz_grid;

/* ---------------------------------------------------------------- parameters */
typedef struct shkz_params {
    double residual;              /* relative tolerance; reference key Residual (pcg.cpp:76, 1e-4) */
    int32_t norm;                 /* shkz_norm */
    int32_t max_iterations;       /* MaxIterations (pcg.cpp:77, 30000); counts CG iterations */
    int32_t solver;               /* shkz_solver */
    int32_t precision;            /* shkz_precision */
    double eps_fluid;             /* EpsFluid (macutility3.cpp:409, 1e-2) */
    double eps_solid;             /* EpsSolid (macutility3.cpp:410, 1e-2) */
    int32_t second_order_fluid;   /* SecondOrderAccurateFluid (macpressuresolver3.cpp:275, 1) */
    int32_t second_order_solid;   /* SecondOrderAccurateSolid (:276, 1) */
    double surface_tension;       /* kappa; 0 = off (macpressuresolver3.cpp:85-115) */
    double gain;                  /* Gain (:277, 1.0): volume-correction PI gain */
    double current_volume;        /* set_target_volume() arguments (:45-48); target 0 = off */
    double target_volume;
    int32_t warm_start;           /* WarmStart (:278, 0): start from the scene's previous pressure */
    /* multigrid controls (0 = library default / autotuned) */
    int32_t mg_pre_sweeps;
    int32_t mg_post_sweeps;
    int32_t mg_coarse_sweeps;
    double mg_omega;
    double inner_residual;        /* mixed precision: relative reduction asked of each inner FP32
                                     solve (0: the final tolerance, one pass) */
    int32_t mg_coarse_pre_sweeps; /* sweeps on the coarser levels (mg_pre/post_sweeps: finest) */
    int32_t mg_coarse_post_sweeps;
    int32_t mg_bottom_bricks;     /* unused (kept for layout): every grid's coarsest level is one
                                     brick, solved by one workgroup */
    int32_t autotune;             /* 1: multigrid controls left at 0 come from the autotuner: looked
                                     up per (device, shape class), tuned on the first call of a class
                                     (repeated solves of the built system; the returned projection
                                     uses the best controls).  Cached per process, and in the file
                                     named by the environment variable SHKZ_TUNE_CACHE if set. */
    int32_t mg_smoother;          /* finest-level smoother (shkz_smoother; 0 = library default / tuned) */
    int32_t mg_local_sweeps;      /* brick-local smoother: symmetric Gauss-Seidel sweeps per finest-level
                                     application, 2 or 3 (0 = library default 2 / tuned) */
    int32_t reserved[2];
} shkz_params;

/* ---------------------------------------------------------------- result */
typedef struct shkz_result {
    int32_t status;               /* shkz_status of the call */
    int32_t converged;            /* 1 when the criterion was met */
    int32_t iterations;           /* CG iterations (all inner iterations in mixed mode) */
    int32_t outer_iterations;     /* defect-correction passes (mixed mode), else 1 */
    double reresid;               /* final relative residual in the requested norm, FP64 */
    double reresid_inf;           /* ||b - A x||_inf / ||r0||_inf */
    double reresid_l2;            /* ||b - A x||_2 / ||r0||_2 */
    int64_t n_unknowns;           /* liquid unknowns (the reference's matrix size) */
    int64_t n_bricks;             /* active 8^3 bricks on the finest level */
    int32_t n_levels;             /* multigrid levels */
    int32_t batch_size;           /* grids in the batch this projection ran in */
    double rhs_correct;           /* volume-correction term added to every row */
    /* stage times in milliseconds, named after the reference's scoped_timer records
       (macpressuresolver3.cpp:82,201,237,269,271) */
    double t_solid_fluid_fractions;
    double t_build_highres_linsystem;
    double t_linsolve;
    double t_update_velocity;
    double t_projection;
    /* server requests only (shkz_submit): milliseconds from submission to the start of its batch,
       and to its completion */
    double t_queue;
    double t_latency;
    /* the multigrid controls the solve used (defaults, the caller's, or autotuned) */
    int32_t mg_pre_sweeps, mg_post_sweeps, mg_coarse_pre_sweeps, mg_coarse_post_sweeps, mg_coarse_sweeps;
    int32_t tuned;                /* 0 = defaults / given, 1 = from the tuning cache, 2 = tuned now */
    double mg_omega;
    double inner_residual;
    int32_t mg_smoother;          /* shkz_smoother used on the finest level */
    int32_t mg_local_sweeps;      /* its sweeps per finest-level application (brick-local smoother) */
} shkz_result;

/* ---------------------------------------------------------------- context */
typedef struct shkz_ctx shkz_ctx;

/* Library ABI version (SHKZ_ABI_VERSION). */
int shkz_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *shkz_last_error(void);

/* Defaults equal the reference's configuration defaults (see the field comments). */
void shkz_default_params(shkz_params *p);

/* Creates a context on HIP device `device`.  hbm_budget_bytes bounds the device memory the
 * context may hold (0 = 90% of the device).  Replaces module::alloc_module + the driver's
 * initialize() (src/core/module.cpp:103-175, macproject3_interface.h:90-98). */
int shkz_create(int device, uint64_t hbm_budget_bytes, shkz_ctx **out);
void shkz_destroy(shkz_ctx *ctx);

/* macproject3_interface::project on host arrays (macpressuresolver3.cpp:50-272).  Copies the
 * inputs to the device, projects, copies velocities (and pressure if requested) back. */
int shkz_project(shkz_ctx *ctx, const shkz_grid *grid, const shkz_params *params, shkz_result *result);

/* The same projection on arrays already resident in device memory (every pointer in `grid`
 * is a device pointer).  Work is issued on `stream` (a hipStream_t, NULL = the null stream);
 * the call returns once the result is known. */
int shkz_project_device(shkz_ctx *ctx, const shkz_grid *grid, const shkz_params *params,
                        shkz_result *result, void *stream);

/* A batch of `ngrids` independent projections in one call (no counterpart in the reference,
 * whose driver projects one grid per time step, macliquid3.cpp:321-367): every solver launch
 * covers the stored bricks of all grids, while each grid keeps its own reductions, multigrid
 * depth, defect-correction state and convergence test, so results[g] and grid g's outputs equal
 * those of projecting grid g alone.  Grids may differ in size; scene ids must be distinct (0 =
 * no scene).  Host arrays, like shkz_project. */
int shkz_project_batch(shkz_ctx *ctx, const shkz_grid *grids, int ngrids, const shkz_params *params,
                       shkz_result *results);

/* The batch on device-resident arrays, issued on `stream`, like shkz_project_device. */
int shkz_project_batch_device(shkz_ctx *ctx, const shkz_grid *grids, int ngrids, const shkz_params *params,
                              shkz_result *results, void *stream);

/* Runs fractions, labelling and assembly only and writes the assembled FP64 system on dense
 * cell arrays (host): diag (0 off the unknowns), the couplings of each cell's low x/y/z face
 * (A[c, c-e] = -g) and the right-hand side.  The counterpart of inspecting the RCMatrix the
 * reference assembles (macpressuresolver3.cpp:159-199, RCMatrix_utility::report at :219). */
int shkz_export_system(shkz_ctx *ctx, const shkz_grid *grid, const shkz_params *params, double *diag, double *gx,
                       double *gy, double *gz, double *rhs);

/* Drops a scene's resident state (warm-start pressure, volume integrator). */
int shkz_release_scene(shkz_ctx *ctx, uint64_t scene_id);

/* Device bytes currently held by the context. */
uint64_t shkz_device_bytes(const shkz_ctx *ctx);

/* Test hook: the context's device allocation after `after` more successful ones fails once, as a
 * failed hipMalloc does (SHKZ_ERR_NOMEM from the call that needed it; the device buffers the
 * context already holds, its scene state and the caller's arrays are left as they were, and later
 * calls allocate normally); after < 0 clears it.  The reference has no recovery path here (an
 * allocation failure ends the process, src/core/module.cpp:140-147); this exercises the library's. */
int shkz_inject_alloc_failure(shkz_ctx *ctx, int32_t after);

/* ------------------------------------------------------- after the projection
 * Velocity extrapolation and the solid constraint: macutility3::extrapolate_and_constrain_velocity
 * (src/utility/macutility3.cpp:89-93), as macliquid3::extend_both (src/liquid/macliquid3.cpp:309-319)
 * and macsmoke3 (src/smoke/macsmoke3.cpp:294) call it on the velocity the projection left.
 *  - `width` passes of macarray_extrapolator3::extrapolate (include/shiokaze/array/
 *    array_extrapolator3.h:47-80): each pass gives every inactive face with an active
 *    6-neighbour the mean of its active neighbours as they were before the pass, and activates it;
 *  - with `constrain`, macutility3::constrain_velocity (:61-88): where the solid level set has a
 *    negative value, an active face whose centre samples the solid below 0 loses the component of
 *    the velocity along the solid's normal when the velocity points into the solid, and active
 *    boundary faces pointing out of the domain become 0.  Nothing is constrained when `solid` is
 *    NULL or has no negative value.
 * Inactive faces read as 0 wherever the reference reads an inactive face (its background value).
 * The activity arrays are the projection's active_out_* (or the caller's own), updated in place.
 * Results are bit-identical to the reference's arithmetic (float velocities and level set,
 * double normals).  With `constrain`, every axis needs at least 2 cells. */
typedef struct shkz_velocity_field {
    int32_t w, h, d;              /* cells */
    int32_t solid_dtype;          /* SHKZ_F32 (the reference's Real) or SHKZ_F64 */
    double dx;                    /* 0 = min(1/w, 1/h, 1/d) */
    const void *solid;            /* nodal solid level set ((w+1)(h+1)(d+1)) or NULL */
    float *u, *v, *w_vel;         /* face velocities, updated in place */
    uint8_t *active_u, *active_v, *active_w; /* face activity (0/1), updated in place */
} shkz_velocity_field;

/* Host arrays.  width: 0..254 passes. */
int shkz_extrapolate_velocity(shkz_ctx *ctx, const shkz_velocity_field *field, int32_t width, int32_t constrain);
/* Device arrays, ordered on `stream` like shkz_project_device (so it may directly follow a
 * projection of the same arrays on the same stream). */
int shkz_extrapolate_velocity_device(shkz_ctx *ctx, const shkz_velocity_field *field, int32_t width,
                                     int32_t constrain, void *stream);

/* Advection between projections: macadvection3 (src/advection/macadvection3.cpp) with its
 * defaults (MacCormack, trilinear interpolation, TrimNarrowBand 1) and its options (:56-60):
 * semi-Lagrangian instead of MacCormack, TrimNarrowBand, and WENO sampling (weno = 1: the value
 * passes sample with WENO3::interpolate, order 6, include/shiokaze/math/WENO3.h:105-135, in
 * double; MacCormack's bounds and the narrow-band test stay trilinear, :96-118,207-231).
 *  - shkz_advect_velocity*: advect_vector(u, u, fluid, dt) (:47-53 -> advect_u, :140-177): each
 *    active face moves back along its own full velocity (convert_to_full, macarray3.h:350-372);
 *    MacCormack's correction is clamped to the forward stencil's min/max and replaced by the
 *    semi-Lagrangian value where the liquid level set at the face exceeds
 *    -dx * trim_narrowband.  Inactive faces read 0 and are not written.
 *  - shkz_advect_scalar*: advect_scalar(q, velocity, fluid, dt) (:38-45 -> advect_cell,
 *    :239-288) on a float cell array; the cell's full velocity is the mean of its two faces per
 *    axis when all six are active, else zero (convert_to_full, macarray3.h:305-343).  With
 *    fluid NULL the narrow band is measured on q itself, as the level-set tracker calls it
 *    (maclevelsetsurfacetracker3.cpp:50-51).  active_q (optional) marks the cells to advect;
 *    the others are not written, the forward pass reads their stored value and the backward
 *    pass reads `background` (a fresh array's background value, as in the reference).
 * Results are bit-identical to the reference's arithmetic.  Every axis needs at least 2 cells. */
typedef struct shkz_advect_params {
    double dt;
    int32_t maccormack;           /* 1 (default) MacCormack, 0 semi-Lagrangian */
    int32_t trim_narrowband;      /* TrimNarrowBand (default 1) */
    const void *fluid;            /* liquid level set (cells) for the narrow-band test */
    int32_t fluid_dtype;          /* SHKZ_F32 or SHKZ_F64 */
    float background;             /* scalars: see above */
    int32_t weno;                 /* 0 (default) trilinear, 1 WENO (config key "WENO", :59) */
    int32_t reserved[3];
} shkz_advect_params;

void shkz_advect_default_params(shkz_advect_params *params);
int shkz_advect_velocity(shkz_ctx *ctx, const shkz_velocity_field *field, const shkz_advect_params *params);
int shkz_advect_velocity_device(shkz_ctx *ctx, const shkz_velocity_field *field, const shkz_advect_params *params,
                                void *stream);
int shkz_advect_scalar(shkz_ctx *ctx, const shkz_velocity_field *field, float *q, const uint8_t *active_q,
                       const shkz_advect_params *params);
int shkz_advect_scalar_device(shkz_ctx *ctx, const shkz_velocity_field *field, float *q, const uint8_t *active_q,
                              const shkz_advect_params *params, void *stream);

/* ------------------------------------------------------- 2D (macpressuresolver2)
 * macproject2_interface::project (include/shiokaze/projection/macproject2_interface.h:67-72, implemented
 * by macpressuresolver2::project, src/projection/macpressuresolver2.cpp:48-232) on dense 2D arrays
 * (y fastest): cells (w, h) index i*h + j; u (w+1, h); v (w, h+1); nodal solid (w+1, h+1).  Face
 * areas follow macutility2 (src/utility/macutility2.cpp:89-125: 1 - fraction of the two solid
 * samples at the ends of the face); assembly, solver and update are the 2D module's.  The
 * update writes 0 to every face the reference sets to 0 or switches off; the reference keeps a
 * closed boundary face with liquid inside active only while its velocity points out of the
 * domain (:224-226), which a caller of the dense API applies to its own activity flags. */
typedef struct shkz_grid2d {
    int32_t w, h;
    double dx;                    /* 0 = min(1/w, 1/h) */
    double dt;
    const void *fluid;
    int32_t fluid_dtype, solid_dtype;
    const void *solid;            /* nodal (w+1, h+1) or NULL */
    float *u, *v;                 /* faces, updated in place */
    const uint8_t *active_u, *active_v;
    float *pressure;              /* optional outputs (cells) */
    double *pressure64;
    uint64_t scene_id;
    uint8_t *unknown_out;         /* optional out (cells): 1 on the unknowns (macpressuresolver2.cpp
                                     sets m_pressure on each of them, zeros included) */
} shkz_grid2d;

int shkz_project2d(shkz_ctx *ctx, const shkz_grid2d *grid, const shkz_params *params, shkz_result *result);
int shkz_project2d_device(shkz_ctx *ctx, const shkz_grid2d *grid, const shkz_params *params, shkz_result *result,
                          void *stream);
/* A batch of 2D grids (device arrays), like shkz_project_batch_device. */
int shkz_project2d_batch_device(shkz_ctx *ctx, const shkz_grid2d *grids, int ngrids, const shkz_params *params,
                                shkz_result *results, void *stream);
/* The assembled FP64 2D system on dense cell arrays (host): diag, low-face couplings, rhs. */
int shkz_export_system2d(shkz_ctx *ctx, const shkz_grid2d *grid, const shkz_params *params, double *diag, double *gx,
                         double *gy, double *rhs);

/* ------------------------------------------------------- LinSolver drop-in (CSR systems)
 * Stands in for the pcg LinSolver module, RCMatrix_solver_interface<size_t,double>::solve
 * (include/shiokaze/linsolver/RCMatrix_solver.h:77, implemented by pcg_solver::solve,
 * src/linsolver/pcg.cpp:45-73 -> PCGSolver<double>::solve, pcg_solver.h:246-295), for any
 * symmetric positive (semi)definite system the caller assembled: x0 = 0, stop at
 * ||r||_inf <= residual * ||b||_inf or max_iterations.  SHKZ_LS_NONE is the reference's
 * iteration (its MIC(0) output is discarded, :379-384: plain CG); SHKZ_LS_JACOBI applies the
 * inverse diagonal.  FP64 throughout. */
enum shkz_ls_preconditioner { SHKZ_LS_NONE = 0, SHKZ_LS_JACOBI = 1 };

typedef struct shkz_linsolve_params {
    double residual;          /* Residual (pcg.cpp:76, 1e-4) */
    int32_t max_iterations;   /* MaxIterations (:77, 30000) */
    int32_t preconditioner;   /* shkz_ls_preconditioner */
    int32_t reserved[6];
} shkz_linsolve_params;

typedef struct shkz_linsolve_result {
    int32_t status;
    int32_t converged;
    int32_t iterations;       /* Result::count */
    int32_t pad;
    double reresid;           /* Result::reresid: ||r||_inf / ||b||_inf */
    double t_upload, t_solve; /* milliseconds */
} shkz_linsolve_result;

typedef struct shkz_linsolver shkz_linsolver;

void shkz_linsolve_default_params(shkz_linsolve_params *p);
int shkz_linsolver_create(int device, shkz_linsolver **out);
void shkz_linsolver_destroy(shkz_linsolver *solver);
/* Solves A x = b for an n x n CSR matrix on host arrays (rows may be in any column order; indices
 * are checked).  x receives the solution. */
int shkz_linsolve_csr(shkz_linsolver *solver, int64_t n, const int64_t *rowptr, const int64_t *col,
                      const double *val, const double *b, double *x, const shkz_linsolve_params *params,
                      shkz_linsolve_result *result);

/* Allocation-free variant of shkz_linsolve_csr.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  The system is given in device memory: rowptr (n+1) and col (nnz) as 32-bit
 * indices, val (nnz), b (n); x (n, device) receives the solution.  The indices are trusted (they
 * are not read back to be checked).  Same iteration, kernels and launch shapes as
 * shkz_linsolve_csr, so the same result, launched directly on `stream` (no graph).
 * The working vectors live in one fixed device array inside the library, sized for at most
 * SHKZ_LS_NOALLOC_MAX_ROWS rows (about 10.6 MB; each GPU the library runs on holds its own copy
 * until the process ends); larger systems fail with SHKZ_ERR_INVALID.  No call allocates or frees
 * device memory.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from any solver object or thread) race and must be serialised by the caller.  Each call
 * initialises every part of the array it reads, so nothing carries over between calls.
 * Returns when the solve is complete (the scalars are read back between chunks of iterations). */
#define SHKZ_LS_NOALLOC_MAX_ROWS (1 << 18)
int shkz_linsolve_csr_device_noalloc(shkz_linsolver *solver, int32_t n, const int32_t *rowptr, const int32_t *col,
                                     const double *val, const double *b, double *x,
                                     const shkz_linsolve_params *params, shkz_linsolve_result *result, void *stream);

/* ------------------------------------------------------- server (continuous batching)
 * No counterpart in the reference (its driver projects one grid per time step on one thread,
 * ui.cpp:703-710).  A server owns `executors` workspaces, each with its own HIP stream; every
 * executor that is free takes the next batch of queued requests, so requests that arrive while
 * others run join the next batch and batches overlap on the device.  Admission: the estimated
 * working set of the batches in flight plus the resident scene state stays within
 * hbm_budget_bytes; memory is granted in arrival order; each scene's requests run in order and
 * never two at once; only requests with equal shkz_params share a batch.  Every request's result
 * equals that of projecting it alone with shkz_project_device.
 * Requests carry device pointers (as shkz_project_device): inputs must be complete when submitted
 * and stay valid and unmodified until the request completes; outputs are complete when
 * shkz_poll / shkz_wait report it done.  All server calls are thread-safe. */
typedef struct shkz_server shkz_server;

typedef struct shkz_server_config {
    int32_t device;
    int32_t executors;            /* batches in flight at once (0 = 2) */
    int32_t max_batch_grids;      /* 0 = 64 */
    int32_t reserved0;
    int64_t max_batch_cells;      /* box cells per batch (0 = 2^26) */
    uint64_t hbm_budget_bytes;    /* workspaces + resident scene state (0 = 90% of the device) */
    int32_t reserved[8];
} shkz_server_config;

typedef struct shkz_server_stats {
    uint64_t submitted, completed, rejected;
    uint64_t batches, batched_requests;
    int32_t max_batch, max_in_flight_batches;
    uint64_t max_in_flight_bytes;  /* largest estimated working set in flight */
    uint64_t resident_bytes;       /* scene state held now */
    uint64_t device_bytes;         /* workspaces + scene state held now */
    uint64_t cancelled;            /* requests cancelled before they started */
} shkz_server_stats;

void shkz_server_default_config(shkz_server_config *cfg);
int shkz_server_create(const shkz_server_config *cfg, shkz_server **out);
/* Completes every queued request, then frees everything. */
void shkz_server_destroy(shkz_server *srv);
/* Queues one projection (params NULL = defaults); *ticket identifies it. */
int shkz_submit(shkz_server *srv, const shkz_grid *grid, const shkz_params *params, uint64_t *ticket);
/* SHKZ_PENDING while queued or running; SHKZ_OK once done (result filled, ticket retired); else the
 * request's error (result filled, ticket retired) or SHKZ_ERR_NOT_FOUND for an unknown ticket. */
int shkz_poll(shkz_server *srv, uint64_t ticket, shkz_result *result);
/* shkz_poll after waiting up to timeout_ms (< 0: no limit) for the request. */
int shkz_wait(shkz_server *srv, uint64_t ticket, double timeout_ms, shkz_result *result);
/* Cancels a queued request: SHKZ_OK when it was still queued (it then completes at once with
 * SHKZ_ERR_CANCELLED, collected by shkz_poll / shkz_wait, and none of its outputs is written);
 * SHKZ_PENDING when it has already started or completed (it runs to its end, result unchanged);
 * SHKZ_ERR_NOT_FOUND for an unknown or collected ticket.  Later requests of the same scene no
 * longer wait for it.  No counterpart in the reference (its driver has no request queue). */
int shkz_cancel(shkz_server *srv, uint64_t ticket);
/* Drops a scene's resident state; SHKZ_ERR_INVALID while a request of the scene is queued or running. */
int shkz_server_release_scene(shkz_server *srv, uint64_t scene_id);
int shkz_server_get_stats(shkz_server *srv, shkz_server_stats *stats);

/* ------------------------------------------------------- kernel timing (measurement) */
typedef struct shkz_kernel_stat {
    char name[48];
    int64_t launches;
    double total_ms;          /* summed event-measured durations */
    double bytes;             /* summed algorithmic bytes of those launches */
} shkz_kernel_stat;

/* When enabled, each kernel launch is bracketed by HIP events on its stream and its
 * algorithmic byte count is recorded; shkz_get_kernel_stats() resolves the events. */
int shkz_set_kernel_timing(shkz_ctx *ctx, int enable);
int shkz_get_kernel_stats(shkz_ctx *ctx, shkz_kernel_stat *out, int max_entries, int *n_out);
int shkz_reset_kernel_stats(shkz_ctx *ctx);

/* ------------------------------------------------------- synthetic workloads
 * Scene synthesis for benchmarks and tests (no counterpart in the reference, whose scenes
 * are dylibs sampled by macutility3::assign_initial_variables, macutility3.cpp:308-400; the
 * sampling positions follow it: level sets at cell centres (i+1/2)dx, solids at nodes i*dx,
 * velocities at face centres).  RNG: std::mt19937_64 with a Box-Muller normal. */
enum shkz_scene_kind {
    SHKZ_SCENE_DROP = 0,        /* sphere r=0.15 at (0.5,0.7,0.5) over a pool of depth 0.3, walls */
    SHKZ_SCENE_DROP_SOLIDS = 1, /* the same plus 3 solid spheres r=0.1, centres U[0.2,0.8]^3 */
    SHKZ_SCENE_SPHERES_BOX = 2, /* 8 spheres r~U[0.05,0.15] over a pool of depth 0.25 and one
                                   random solid box of volume 0.1 */
    SHKZ_SCENE_RANDOM_DROP = 3, /* seeded sphere-drop (pool depth and sphere drawn from seed) */
    SHKZ_SCENE_SMOKE = 4        /* all-fluid domain (fluid = -1 everywhere, as macsmoke3 initialises
                                   it, macsmoke3.cpp:106) around 2 solid spheres r~U[0.08,0.15],
                                   centres U[0.25,0.75]^3; the liquid is not pushed out of the solid */
};

typedef struct shkz_scene_desc {
    int32_t kind;           /* shkz_scene_kind */
    int32_t w, h, d;
    uint64_t seed;
    double dt;
    double gravity;         /* added as gravity*dt to v (e.g. -9.8); 0 = none */
    double shift[3];        /* rigid translation of the liquid level set (phi(x - shift)) */
    int32_t fluid_dtype;    /* dtype written to fluid */
    int32_t velocity_mode;  /* 0 = i.i.d. N(0,1), 1 = zero */
    uint64_t velocity_seed; /* 0 = use seed */
} shkz_scene_desc;

/* Fills host arrays for a scene: fluid (cells, fluid_dtype), solid (nodal float, may be NULL;
 * *has_solid tells whether the scene has a solid), u, v, w (faces, may be NULL). */
int shkz_scene_fill(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid,
                    float *u, float *v, float *w);

/* 2D scenes for shkz_grid2d workloads: the same seeded shapes in the plane z = sphere centre
 * (circles, the pool, the box's x-y rectangle); fluid (w, h), solid nodal (w+1, h+1) float (may be
 * NULL), u (w+1, h) and v (w, h+1) (may be NULL).  d and shift[2] are ignored. */
int shkz_scene_fill2d(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid, float *u, float *v);

/* The same scenes generated on the device into device buffers, on `stream` (synthetic inputs
 * for serving workloads, where host generation would dominate).  The level sets equal
 * shkz_scene_fill's; velocities are i.i.d. N(0,1) from a counter-based generator (splitmix64 of
 * seed, array and face index, Box-Muller), so they differ from shkz_scene_fill's mt19937_64
 * stream but are reproducible for a seed. */
int shkz_scene_fill_device(const shkz_scene_desc *desc, void *fluid, float *solid, int32_t *has_solid, float *u,
                           float *v, float *w, void *stream);

#ifdef __cplusplus
}
#endif
#endif
