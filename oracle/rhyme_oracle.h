/*
 * rhyme_oracle.h -- CPU restatement of the Rhyme hydro hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path (librhyme_hip.so) is compared against, and the CPU baseline that
 * bench.py times beside it.  Nothing in the product links, loads or calls it.
 *
 * Every function restates one reference routine, in the reference's operation
 * order (no FMA contraction: built with -O2 -ffp-contract=off), and cites the
 * Fortran file:line it follows (paths relative to the reference's src/).
 *
 * Box layout (same bytes as the reference's samr_box_t%cells):
 *   cells(i, j, k, c), Fortran column-major, 1-based interior indices
 *   1..n_d, 2 ghost layers on each side of every active axis
 *   (rhyme_samr/src/rhyme_samr_init_box.f90:29-33), so the storage extent of
 *   axis d is n_d + 4 (1 for inactive axes of a 1D/2D box), i is fastest and
 *   each component c is one contiguous plane (struct of arrays).
 *   Components: rho, rho_u, [rho_v, [rho_w,]] e_tot, then the non-hydro ones
 *   (rhyme_units/src/rhyme_units.f90:32-57); only the first ndim+2 are
 *   touched except by inflow boundaries.
 */
#ifndef RHYME_ORACLE_H
#define RHYME_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_CMP 16

/* Same field layout as rhy_params_t in include/rhyme_hip.h. */
typedef struct orc_params {
    int32_t ndim;          /* 1, 2 or 3 (the reference's compile-time NDIM)     */
    int32_t limiter;       /* slid: 1 van_leer, 2 minmod, 3 van_albada, 4 superbee */
    int32_t solver_type;   /* mhid: 1 cpu_intensive, 2 memory_intensive         */
    int32_t n_iteration;   /* irs%n_iteration                                   */
    double gamma;          /* ig_gamma(state_of_matter)                         */
    double kb_amu;         /* rhyme_thermo_base_kb_amu                          */
    double omega;          /* sl%w                                              */
    double tolerance;      /* irs%tolerance                                     */
    double w_vacuum_rho;   /* rp%w_vacuum(cid%rho)                              */
    double w_vacuum_p;     /* rp%w_vacuum(cid%p)                                */
    double courant;        /* cfl%courant_number                                */
    int32_t bc_types[6];   /* bcid: 1 reflective, 2 outflow, 3 periodic, 4 inflow */
    int32_t ncmp;          /* components stored per cell (NCMP)                 */
    int32_t reserved;
    double cons_inflows[6][ORC_MAX_CMP]; /* bc%cons_inflows(:, side)             */
} orc_params;

/* Riemann star/solution record (riemann_problem_solution_t,
 * rhyme_riemann_problem/src/rhyme_riemann_problem.f90:33-61), flattened. */
typedef struct orc_rp_solution {
    double l_rho, l_v[3], l_p, l_cs;
    double r_rho, r_v[3], r_p, r_cs;
    double star_p, star_u;
    int32_t l_is_shock, r_is_shock;
    double l_shock_rho, l_shock_speed;
    double l_fan_rho, l_fan_cs, l_fan_speedH, l_fan_speedT;
    double r_shock_rho, r_shock_speed;
    double r_fan_rho, r_fan_cs, r_fan_speedH, r_fan_speedT;
    double l_f, r_f;
} orc_rp_solution;

/* ---- thermodynamics / ideal gas (unit level) ---- */
double orc_pressure(const orc_params *p, const double *u);
double orc_speed_of_sound(const orc_params *p, const double *u);
void orc_flux(const orc_params *p, const double *u, int axis, double *f);
void orc_prim_vars_to_cons(const orc_params *p, double rho, const double *v,
                           double pr, double *u);

/* ---- slope limiter and half step ---- */
void orc_slope_limiter(const orc_params *p, const double *ul, const double *u,
                       const double *ur, double *delta);
void orc_half_step(const orc_params *p, const double *u, const double *delta,
                   int axis, double dx, double dt, double *l, double *r);

/* ---- iterative Riemann solver ---- */
void orc_irs_guess(const orc_params *p, double l_rho, const double *l_v,
                   double l_p, double l_cs, double r_rho, const double *r_v,
                   double r_p, double r_cs, int axis, double *guess6);
double orc_irs_exec(const orc_params *p, double l_rho, const double *l_v,
                    double l_p, double l_cs, double *l_f, double r_rho,
                    const double *r_v, double r_p, double r_cs, double *r_f,
                    int axis, int32_t *n_newton);
void orc_nonlinear_wave_function(const orc_params *p, double rho, double pk,
                                 double cs, double pstar, double *f,
                                 double *fprime);
void orc_rp_star(const orc_params *p, orc_rp_solution *s, int axis);
void orc_rp_sampling(const orc_params *p, const orc_rp_solution *s, int axis,
                     double dx, double dt, double *u);
void orc_rp_solve(const orc_params *p, const double *l, const double *r,
                  double dx, double dt, int axis, double *u);

/* Batched unit entry points: arrays of n problems, NVAR = ndim+2 values each,
 * problem-major (record q occupies [q*NVAR, (q+1)*NVAR)). */
void orc_rp_solve_batch(const orc_params *p, int64_t n, const double *l,
                        const double *r, const int32_t *axis, double dt,
                        double *u);
void orc_limiter_batch(const orc_params *p, int64_t n, const double *ul,
                       const double *u, const double *ur, double *delta);
void orc_half_step_batch(const orc_params *p, int64_t n, const double *u,
                         const double *delta, const int32_t *axis, double dx,
                         double dt, double *l, double *r);
void orc_flux_batch(const orc_params *p, int64_t n, const double *u,
                    const int32_t *axis, double *f);

/* ---- box level: dims[3] = interior cells per axis (1 for unused axes) ---- */
double orc_cfl_time_step(const orc_params *p, const int32_t *dims,
                         const double *cells, const double *dx);
void orc_bc_set_boundaries(const orc_params *p, const int32_t *dims,
                           double *cells);
void orc_bc_set_face(const orc_params *p, const int32_t *dims, double *cells,
                     int face);
void orc_bc_init_ghosts(const orc_params *p, const int32_t *dims,
                        double *cells);
/* returns 0, or -1 if scratch allocation failed */
int orc_mh_solve(const orc_params *p, const int32_t *dims, double *cells,
                 const double *dx, double dt);
int orc_mh_solve_fluxes(const orc_params *p, const int32_t *dims, double *cells,
                        const double *dx, double dt, double *fr_out);
/* The main-loop step (src/main.f90:109-136): dt = CFL; BC; MUSCL-Hancock.
 * dts[s] receives the dt of step s. Returns 0 on success. */
int orc_run(const orc_params *p, const int32_t *dims, double *cells,
            const double *dx, int32_t nsteps, double *dts);

/* Library introspection */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
