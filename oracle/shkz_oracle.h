/*
 * shkz_oracle.h -- CPU restatement of Shiokaze's MAC pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * tests, the smoke check and the benchmark's CPU-baseline leg may call it;
 * the product library never links it and never falls back to it.
 *
 * It restates, in plain C, the reference path
 *   macpressuresolver3::project      src/projection/macpressuresolver3.cpp:50-272
 *   macutility3 face fractions       src/utility/macutility3.cpp:94-194
 *   utility::fraction / get_area     include/shiokaze/utility/utility.h:162-214
 *   pcg_solver / PCGSolver<double>   src/linsolver/pcg.cpp:45-80,
 *                                    local/include/pcgsolver/pcg_solver.h:89-295,379-384
 * on dense arrays (z fastest):
 *   cell  (w,h,d)        idx = (i*h + j)*d + k
 *   u     (w+1,h,d)      v (w,h+1,d)      w (w,h,d+1)
 *   nodal (w+1,h+1,d+1)  (solid level set)
 */
#ifndef SHKZ_ORACLE_H
#define SHKZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int w, h, d;            /* cells */
    double dx;              /* min(1/w,1/h,1/d) as shape3::dx (shape.h:755-759) */
    double dt;
    const void *fluid;      /* cell level set, float or double */
    int fluid_f64;
    const void *solid;      /* nodal level set or NULL (walls only) */
    int solid_f64;
    float *vel[3];          /* face velocities (in/out) */
    const uint8_t *active[3]; /* optional face activity (NULL = all active) */
} oracle_grid;

enum { ORACLE_MIC_FAITHFUL = 0, ORACLE_MIC_APPLIED = 1, ORACLE_MIC_NONE = 2 };

typedef struct {
    double residual;            /* pcg.cpp:76 default 1e-4 */
    int max_iterations;         /* pcg.cpp:77 default 30000 */
    int mic_mode;               /* ORACLE_MIC_* */
    double mic_omega;           /* 0.97 (pcg_solver.h:89) */
    double mic_min_diag_ratio;  /* 0.25 */
    double eps_fluid, eps_solid;/* macutility3.cpp:408-422 defaults 1e-2 */
    int second_order_fluid;     /* macpressuresolver3.cpp:302-303 default 1 */
    int second_order_solid;
    double surface_tension;     /* 0 = off */
    double rhs_correct;         /* volume-correction term added to every row (0 = off) */
    const double *warm_start;   /* cell-indexed previous pressure or NULL */
    int threads;                /* threaded stages (fractions/assembly/update); 0 = all */
} oracle_params;

typedef struct {
    int iterations;
    double reresid;             /* ||r||_inf / ||r0||_inf as pcg_solver.h:283,293 */
    int converged;
    int64_t n_unknowns;
    double t_fractions, t_build, t_linsolve, t_update, t_projection; /* seconds */
} oracle_result;

void oracle_default_params(oracle_params *p);

/* utility.h:162-170 */
double oracle_fraction(double phi0, double phi1);
/* utility.h:179-214; q is isosurf[2][2] row-major: q[0]=[0][0] q[1]=[0][1] q[2]=[1][0] q[3]=[1][1] */
double oracle_get_area(const double q[4]);
/* utility.h:113-120 */
double oracle_box_sdf(const double p[3], const double p0[3], const double p1[3]);

/* macutility3.cpp:94-194 plus the first-order overrides of macpressuresolver3.cpp:70-80.
 * area[dim], rho[dim] are dense face arrays (float). */
void oracle_face_fractions(const oracle_grid *g, const oracle_params *p, float *area[3], float *rho[3]);

/* macpressuresolver3.cpp:85-115: adds the surface-tension impulse to g->vel. */
void oracle_surface_tension(const oracle_grid *g, double kappa, float *rho[3]);

/* macpressuresolver3.cpp:121-156: index_map[c] = unknown number or -1. Returns N. */
int64_t oracle_label(const oracle_grid *g, float *area[3], float *rho[3], int64_t *index_map);

/* macpressuresolver3.cpp:159-199: CSR rows sorted by column (as SparseMatrix keeps them,
 * sparse_matrix.h:75-90).  rowptr has N+1 entries; col/val have rowptr[N] entries. */
int64_t oracle_assemble_count(const oracle_grid *g, const int64_t *index_map, float *area[3], float *rho[3]);
void oracle_assemble(const oracle_grid *g, const int64_t *index_map, float *area[3], float *rho[3],
                     int64_t *rowptr, int64_t *col, double *val, double *b);

/* macpressuresolver3.cpp:207-211 PI controller; returns the term added to every rhs row and
 * updates *y_prev. */
double oracle_volume_correction(double current_volume, double target_volume, double dt,
                                double gain, double *y_prev);

/* sparse_matrix.h:265-275: y = A x */
void oracle_spmv(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
                 const double *x, double *y);

/* PCGSolver<double>::solve (pcg_solver.h:246-295) with the MIC(0) factor of :89-189 and the
 * triangular solves of :196-223; mic_mode selects the reference behaviour (factor and solves
 * computed, result discarded, :379-384), the applied preconditioner, or neither. */
int oracle_pcg(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
               const double *b, double *x, double residual, int max_iterations, int mic_mode,
               double omega, double min_diag_ratio, int *iterations, double *reresid);
/* oracle_pcg that also keeps the iterate at the first iteration meeting snap_residual (relative to
 * ||b||_inf) in x_snap, and that iteration count in *snap_iterations (-1 if never met). */
int oracle_pcg_snap(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val, const double *b,
                    double *x, double residual, int max_iterations, int mic_mode, double omega,
                    double min_diag_ratio, int *iterations, double *reresid, double snap_residual, double *x_snap,
                    int *snap_iterations);

/* macpressuresolver3.cpp:252-268 velocity update with pressure stored in float. */
void oracle_velocity_update(const oracle_grid *g, const float *pressure, float *area[3], float *rho[3],
                            uint8_t *active_out[3]);

/* Divergence sum_f sgn_f area_f u_f / dx over the faces of every unknown (negated rhs formula of
 * macpressuresolver3.cpp:193).  div may be NULL; returns max |div|. */
double oracle_divergence(const oracle_grid *g, const int64_t *index_map, float *area[3], double *div);

/* The whole macpressuresolver3::project (macpressuresolver3.cpp:50-272).  pressure (float, cells)
 * receives m_pressure; pressure64 (optional, cells) receives the FP64 solution. */
int oracle_project(const oracle_grid *g, const oracle_params *p, float *pressure,
                   double *pressure64, oracle_result *res);

/* ------------------------------------------------------------------ 2D
 * The 2D module macpressuresolver2 (src/projection/macpressuresolver2.cpp:48-232) with
 * macutility2's fractions (src/utility/macutility2.cpp:89-150), on dense arrays (y fastest):
 *   cell (w,h) idx = i*h + j;  u (w+1,h);  v (w,h+1);  nodal solid (w+1,h+1). */
typedef struct {
    int w, h;
    double dx;              /* min(1/w,1/h) (shape2::dx) */
    double dt;
    const void *fluid;
    int fluid_f64;
    const void *solid;      /* nodal level set or NULL (walls only) */
    int solid_f64;
    float *vel[2];
    const uint8_t *active[2];
} oracle_grid2;

void oracle2_face_fractions(const oracle_grid2 *g, const oracle_params *p, float *area[2], float *rho[2]);
void oracle2_surface_tension(const oracle_grid2 *g, double kappa, float *rho[2]);
int64_t oracle2_label(const oracle_grid2 *g, float *area[2], float *rho[2], int64_t *index_map);
int64_t oracle2_assemble_count(const oracle_grid2 *g, const int64_t *index_map, float *area[2], float *rho[2]);
void oracle2_assemble(const oracle_grid2 *g, const int64_t *index_map, float *area[2], float *rho[2],
                      int64_t *rowptr, int64_t *col, double *val, double *b);
/* active_out (optional, per face): 1 where the face stays active after the update. */
void oracle2_velocity_update(const oracle_grid2 *g, const float *pressure, float *area[2], float *rho[2],
                             uint8_t *active_out[2]);
double oracle2_divergence(const oracle_grid2 *g, const int64_t *index_map, float *area[2], double *div);
int oracle2_project(const oracle_grid2 *g, const oracle_params *p, float *pressure, double *pressure64,
                    oracle_result *res);

/* ---------------------------------------- velocity extrapolation and solid constraint
 * (shkz_oracle_vel.c): the step after a projection, macutility3.cpp:61-93. */
typedef struct {
    int w, h, d;            /* cells */
    double dx;
    const void *solid;      /* nodal level set or NULL */
    int solid_f64;          /* 0: float (Real), 1: double */
    float *vel[3];          /* face velocities (in/out) */
    uint8_t *active[3];     /* face activity (in/out), required */
} oracle_vfield;

/* array_interpolator3.h:40-62: the 8 corner indices (idx[3n+a]) and weights at p, index space of
 * an array of the given shape */
void oracle_interp_coef(const int shape[3], const double p[3], int idx[24], double coef[8]);
/* array_derivative3.h:40-85: corner indices and the three derivative weight sets coef[8*dim+n] */
void oracle_deriv_coef(const int shape[3], const double p[3], int idx[24], double coef[24]);
/* macarray_extrapolator3::extrapolate(velocity, width); returns the faces activated */
int64_t oracle_extrapolate(const oracle_vfield *f, int width);
/* macutility3::constrain_velocity; returns 1 when the solid level set exists */
int oracle_constrain_velocity(const oracle_vfield *f);

/* ---------------------------------------- advection (shkz_oracle_adv.c): macadvection3.cpp
 * advect_vector(u, u, fluid, dt) in place on f->vel (f->active required; solid unused); weno: the
 * WENO option (six-point WENO sampling instead of trilinear) */
void oracle_advect_velocity(const oracle_vfield *f, const void *fluid, int fluid_f64, double dt, int maccormack,
                            int trim, int weno);
/* advect_scalar(q, velocity f, fluid, dt) in place on q (cells); active_q NULL = all active;
 * fluid NULL = a copy of q; background = the value inactive cells of the backward pass read */
void oracle_advect_scalar(const oracle_vfield *f, float *q, const uint8_t *active_q, const void *fluid, int fluid_f64,
                          double dt, int maccormack, int trim, float background, int weno);
/* WENO::interp6 (WENO.h:55-88): samples f at nodes -2..3, x in [0, 1] */
double oracle_weno6(double x, const double f[6]);
/* WENO3::interpolate(array, p) with order 6 (WENO3.h:105-135) on a float array of shape s */
double oracle_weno3(const float *U, const int s[3], const double p[3]);

#ifdef __cplusplus
}
#endif
#endif
