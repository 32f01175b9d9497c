/*
 * rhyme_hip.h -- C ABI of the MI355X (gfx950) hydro hot path of Rhyme.
 *
 * The reference's hot path is the body of its main loop (src/main.f90:106-177):
 *   dt = rhyme_cfl_time_step(...)            (src/main.f90:109)
 *   call rhyme_samr_bc_set_boundaries(...)   (src/main.f90:115)
 *   call rhyme_muscl_hancock_solve(...)      (src/main.f90:132-134)
 * Each entry point below replaces one reference module procedure (cited
 * file:line, relative to the reference's src/).  The reference keeps every
 * array in Fortran allocatables and never returns status; here device memory
 * is owned by the library behind an opaque box handle, host arrays are
 * borrowed mirrors, and every call returns an int status (RHY_OK = 0).
 *
 * Host array layout (identical to samr_box_t%cells, rhyme_samr/src/
 * rhyme_samr.f90:34-42 and rhyme_samr_init_box.f90:29-33): cells(i,j,k,c)
 * column-major with 2 ghost layers on every active axis, i.e. extents
 * (n1+4)[*(n2+4)[*(n3+4)]] * ncmp doubles; components rho, rho_u, [rho_v,
 * [rho_w,]] e_tot, then ncmp-(ndim+2) non-hydro components
 * (rhyme_units/src/rhyme_units.f90:32-57).
 *
 * Threading: calls on one box are serialised on that box's HIP stream; calls
 * on different boxes may come from different host threads.
 */
#ifndef RHYME_HIP_H
#define RHYME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define RHY_OK 0
#define RHY_ERR_ARG (-1)      /* bad argument (NULL, unknown enum value)  */
#define RHY_ERR_HIP (-2)      /* a HIP runtime call failed                 */
#define RHY_ERR_OOM (-3)      /* device allocation failed                  */
#define RHY_ERR_SHAPE (-4)    /* unsupported box shape                     */
#define RHY_ERR_NODEV (-5)    /* no usable gfx950 device                   */
#define RHY_ERR_SOLVER (-6)   /* unknown solver_type (logger%err in ref)   */

/* ---- enumerations (same integer values as the reference) ---- */
/* slope limiter ids: slid, rhyme_slope_limiter/src/rhyme_slope_limiter.f90:6-8 */
#define RHY_VAN_LEER 1
#define RHY_MINMOD 2
#define RHY_VAN_ALBADA 3
#define RHY_SUPERBEE 4
/* solver types: mhid, rhyme_muscl_hancock/src/rhyme_muscl_hancock.f90:10-15 */
#define RHY_CPU_INTENSIVE 1
#define RHY_MEMORY_INTENSIVE 2
/* boundary types: bcid, rhyme_samr_bc/src/rhyme_samr_bc.f90:9-22 */
#define RHY_BC_REFLECTIVE 1
#define RHY_BC_OUTFLOW 2
#define RHY_BC_PERIODIC 3
#define RHY_BC_INFLOW 4
/* box storage precision */
#define RHY_FP64 0   /* state, arithmetic and dt in FP64 (the drop-in)       */
#define RHY_MIXED 1  /* FP32 state storage and bulk arithmetic; FP64 p* Newton
                        solve, wave-speed sampling and dt                     */

#define RHY_MAX_CMP 16

/*
 * Every parameter the reference keeps in process globals or derived-type
 * components, passed explicitly:
 *   gamma, kb_amu           rhyme_thermo_base.f90:21-26,130
 *   n_iteration, tolerance  irs_t (rhyme_irs/src/rhyme_irs.f90:16-19)
 *   w_vacuum_rho/_p         riemann_problem_t%w_vacuum (…riemann_problem.f90:24-31)
 *   limiter, omega          slope_limiter_t (rhyme_slope_limiter.f90:12-15)
 *   solver_type             muscl_hancock_t (rhyme_muscl_hancock.f90:17-19)
 *   courant                 cfl_t (rhyme_cfl/src/rhyme_cfl.f90:7-9)
 *   bc_types, cons_inflows  samr_bc_t (rhyme_samr_bc.f90:24-28); faces in the
 *                           order left, right, bottom, top, back, front.
 */
typedef struct rhy_params {
    int32_t ndim;
    int32_t limiter;
    int32_t solver_type;
    int32_t n_iteration;
    double gamma;
    double kb_amu;
    double omega;
    double tolerance;
    double w_vacuum_rho;
    double w_vacuum_p;
    double courant;
    int32_t bc_types[6];
    int32_t ncmp;
    int32_t reserved;
    double cons_inflows[6][RHY_MAX_CMP];
} rhy_params_t;

typedef struct rhy_box rhy_box_t; /* opaque device-resident box */

/* Select the device and check it is usable; idempotent. */
int rhy_init(int32_t device);
/* Text of the last error on the calling thread ("" if none). */
const char *rhy_last_error(void);
/* Number of kernels / variants compiled in (for "library loads" checks). */
int32_t rhy_version(void);
/* The compile-time options this library was built with, as bits:
 * RHY_OPT_POW_FAST (own FP64 exponentials; clear: device libm pow),
 * RHY_OPT_CONTACT_EXACT (exact contact star states), RHY_OPT_LDS_WALK (LDS
 * ring in the y/z flux walks), RHY_OPT_FAST_RIEMANN (the mixed form's FP64
 * Riemann arithmetic in the FP64 flux kernels). */
#define RHY_OPT_POW_FAST 1
#define RHY_OPT_CONTACT_EXACT 2
#define RHY_OPT_LDS_WALK 4
#define RHY_OPT_FAST_RIEMANN 8
int32_t rhy_build_options(void);

/* ---- box lifetime and host mirror transfers ----
 * Replaces the allocation side of rhyme_samr_init_box (rhyme_samr/src/
 * rhyme_samr_init_box.f90:29-33) and rhyme_mh_workspace_check
 * (rhyme_muscl_hancock/src/rhyme_mh_workspace/src/rhyme_mh_workspace_check.f90:3-104):
 * the device copy and all scratch are created here and resized by
 * rhy_box_resize. dims[d] for d >= ndim must be 1; active dims must be >= 2. */
int rhy_box_create(int32_t ndim, const int32_t dims[3], int32_t ncmp,
                   int32_t precision, rhy_box_t **out);
int rhy_box_resize(rhy_box_t *box, const int32_t dims[3]);
int rhy_box_destroy(rhy_box_t *box);
/* H2D / D2H of the whole ghosted host array (all ncmp components). */
int rhy_box_upload(rhy_box_t *box, const double *cells);
int rhy_box_download(rhy_box_t *box, double *cells);
/* Wait for all work queued on the box. */
int rhy_box_sync(rhy_box_t *box);
/* ---- snapshot payloads (the Chombo writer's data and the restart reader) ----
 * rhyme_chombo_write_level_data (rhyme_chombo/src/rhyme_chombo_write_level_data.f90:41-71)
 * writes per level one FP32 dataset "data:datatype=0": box after box, each
 * box's components one after another, each component's interior cells with
 * i fastest, converted with real() (round to nearest even).  The restart
 * reader (rhyme_initial_condition/src/rhyme_initial_condition_load_rhyme.f90:57-76)
 * widens them back into the boxes.  These calls produce and consume that
 * payload on the device; the HDF5 container (groups, attributes, the boxes
 * table) stays with the host's HDF5 library.
 * rhy_box_snapshot_pack: data[ncmp * nx * ny * nz] from the box's interior.
 * rhy_box_snapshot_unpack: the box's interior from such data (ghosts
 *   untouched; fill them before stepping). */
int rhy_box_snapshot_pack(rhy_box_t *box, float *data);
int rhy_box_snapshot_unpack(rhy_box_t *box, const float *data);

/* ---- the module procedures ---- */

/* rhyme_samr_bc_set_boundaries (rhyme_samr_bc/src/rhyme_samr_bc_set_boundaries.f90:3-22):
 * ghost fill of the 2*ndim faces (faces only, 2 layers). */
int rhy_bc_set_boundaries(rhy_box_t *box, const rhy_params_t *p);
/* rhyme_samr_bc_set_{left,right,bottom,top,back,front}_boundary
 * (…_set_left_boundary.f90:3-92 etc.); face = 0..5 in that order. */
int rhy_bc_set_face(rhy_box_t *box, const rhy_params_t *p, int32_t face);

/* rhyme_cfl_time_step (rhyme_cfl/src/rhyme_cfl_time_step.f90:3-60):
 * dt = courant * minval(dx) / max(|v| + c_s) over the interior. */
int rhy_cfl_time_step(rhy_box_t *box, const rhy_params_t *p, const double dx[3],
                      double *dt);

/* rhyme_muscl_hancock_solve (rhyme_muscl_hancock/src/rhyme_muscl_hancock_solve.f90:3-30),
 * dispatching on p->solver_type to the memory_intensive
 * (…_solve_memory_intensive.f90:3-181) or cpu_intensive
 * (…_solve_cpu_intensive.f90:3-159) update order.  Reads the box's current
 * ghost cells; advances the interior by dt in place. */
int rhy_mh_solve(rhy_box_t *box, const rhy_params_t *p, const double dx[3],
                 double dt);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rhy_mh_solve on a state the caller owns, with every
 * scratch buffer (flux planes, deferred-face list, dt) carved from one static
 * device array of RHY_NOALLOC_SCRATCH_BYTES that the library holds on each
 * GPU for the life of the process.  cells is a device pointer to the box's
 * state laid out as the library's device copy: ncmp planes of the ghosted
 * box, doubles (RHY_FP64) or floats (RHY_MIXED); its ghost cells are read
 * as they are and its interior is advanced by dt in place.  stream is a
 * hipStream_t (NULL: the null stream); the call only enqueues work on it and
 * returns, so the result is there when the stream gets past it.  A box whose
 * scratch would exceed the array is refused with RHY_ERR_SHAPE (the array
 * holds an FP64 3D box of up to 77^3 cells, a mixed one of up to 98^3, an
 * FP64 2D box of up to 1299^2).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
#define RHY_NOALLOC_SCRATCH_BYTES (64 << 20)
int rhy_mh_solve_noalloc(const rhy_params_t *p, int32_t ndim, const int32_t dims[3],
                         int32_t ncmp, int32_t precision, const double dx[3], double dt,
                         void *cells, void *stream);

/* The main-loop step of src/main.f90:109-136 repeated nsteps times, all on
 * the device: dt from the CFL reduction, ghost fill, MUSCL-Hancock.  dts
 * (host, may be NULL) receives the dt of every step.  Blocks until done. */
int rhy_run(rhy_box_t *box, const rhy_params_t *p, const double dx[3],
            int32_t nsteps, double *dts);

/* Same, enqueue only (no host sync); dts are kept on the device and can be
 * fetched after rhy_box_sync with rhy_box_last_dts. */
int rhy_run_async(rhy_box_t *box, const rhy_params_t *p, const double dx[3],
                  int32_t nsteps);
int rhy_box_last_dts(rhy_box_t *box, int32_t n, double *dts);

/* Per-kernel timing with HIP events recorded on the box's stream.  on = 1:
 * around every launch; on = N > 1: around every launch of the steps s of a
 * rhy_run / rhy_run_async call with s % N == 0 (step 0, N, 2N, ...) and of
 * nothing else, so that the events (a few microseconds each) perturb only
 * the sampled steps; on = 0: off.  rhy_box_kernel_stats returns, for kernel
 * index kidx (0: the flux kernels, 1: update, 2: bc, 3: cfl, 4: dt, 5: the
 * deferred-face list reset and fixup), the number of launches timed and their
 * summed milliseconds.  RHY_ERR_ARG for on < 0. */
int rhy_box_set_profiling(rhy_box_t *box, int32_t on);
int rhy_box_kernel_stats(rhy_box_t *box, int32_t kidx, int64_t *launches,
                         double *total_ms);

/* ---- deferred faces ----
 * No counterpart in the reference (its Riemann solve runs the slow paths
 * inline: libm pow in rhyme_irs_nonlinear_wave_function.f90:20-22 and the
 * restarts from guesses 2-6 in rhyme_irs_exec.f90:24-42).  The flux kernels
 * leave faces that need those paths to a fixup kernel, which recomputes them
 * with the full solver in the same step; results are those of the full
 * solver bit for bit.
 * rhy_box_last_deferred (rhy_batch_last_deferred for batches and pools):
 * faces listed by the last flux phase (waits for the stream).  rhy_debug_defer is a process-wide test hook
 * for callers that check the fixup path: every > 0 also lists every face
 * whose low-cell offset is a multiple of `every`, cap >= 0 caps the list (the
 * fixup then recomputes every face of the step); rhy_debug_defer(0, -1)
 * restores normal operation.  Not thread-safe with concurrent runs. */
int rhy_box_last_deferred(rhy_box_t *box, int64_t *n);
int rhy_debug_defer(int32_t every, int64_t cap);

/* ---- diagnostics (rhyme_sanity_check_perform, rhyme_sanity_check/src/
 * rhyme_sanity_check_perform.f90:34-94, and the totals of rhyme_samr_calc_
 * total_mass/energy, rhyme_samr/src/rhyme_samr_calc_total_mass.f90:20-25) ----
 * Over the interior: NaN cells per component, cells with rho < 0 and with
 * e_tot < 0, the minima of rho and e_tot (NaNs skipped) with the 1-based
 * (i,j,k) of their first occurrence in Fortran element order, and the sums of
 * rho and e_tot; plus the reference's totals before its volume factor: the
 * sums of the positive rho and e_tot values over the whole ghosted array.
 * Sums are accumulated in double-double (compensated) and rounded once; the
 * reduction order is fixed, so results repeat bit for bit. */
typedef struct rhy_diag {
    int64_t nan_count[RHY_MAX_CMP];
    int64_t neg_rho, neg_e_tot;
    double min_rho, min_e_tot;
    int32_t min_rho_at[3], min_e_tot_at[3];
    double mass, energy;
    double mass_masked, energy_masked;
} rhy_diag_t;
int rhy_box_diagnostics(rhy_box_t *box, rhy_diag_t *out);

/* ---- initial conditions painted on the device ----
 * No counterpart in the reference, which paints on the host
 * (rhyme_drawing_init, src/rhyme_drawing/src/rhyme_drawing_init.f90:19-78)
 * and uploads nothing.  Two seeded states at rest, written into every
 * component plane (ghosts and the non-hydro components zero, as after
 * rhyme_samr_bc_init), conserved variables formed as conv_prim_vars_to_cons
 * does (E = .5*rho*0 + p/(gamma-1)):
 *   RHY_IC_TWO_STATE  rho_l, p_l for cells whose 0-based index along `axis`
 *                     is below dims[axis]/2, rho_r, p_r above;
 *   RHY_IC_BLAST      rho = 1 + jitter*(-1 + 2u), u the SplitMix64 draw
 *                     (Steele, Lea, Flood 2014) number q of stream `seed`,
 *                     q the 0-based C-order interior index ((k*ny + j)*nx + i),
 *                     u = (draw >> 11) * 2^-53; p = p0, except in the cells
 *                     whose centre offset from the box centre (q_d + 0.5 -
 *                     n_d/2, summed squares in axis order) is <= radius^2,
 *                     where p = (gamma-1)*energy / (N_in * dx^ndim),
 *                     dx = 1/max(dims), N_in the number of such cells. */
#define RHY_IC_TWO_STATE 1
#define RHY_IC_BLAST 2
typedef struct rhy_ic {
    int32_t kind;
    int32_t axis;
    double gamma;
    double rho_l, p_l, rho_r, p_r;
    double radius, jitter, p0, energy;
    uint64_t seed;
} rhy_ic_t;
/* paint the box's device state (the host mirror is not touched) */
int rhy_box_paint(rhy_box_t *box, const rhy_ic_t *ic);

/* ---- device drawing: rhyme_drawing_init (src/rhyme_drawing/src/
 * rhyme_drawing_init.f90:19-78) for a uniform canvas, sharp shapes and
 * harmonic perturbations, in the reference's order ---- */
#define RHY_DRAW_SPHERE 1
#define RHY_DRAW_CUBOID 2
#define RHY_DRAW_MAX_SHAPES 256
#define RHY_DRAW_MAX_HARMONICS 8
/* A shape painted over the canvas (later shapes over earlier ones) with
 * the conserved form of its colour.  Sphere: the cells whose centre
 * (q_d - 1/2 for 1-based index q_d, in cell units) lies within `radius`
 * of `centre`, sum over axes of (x_d - centre_d)^2 <= radius^2 in axis
 * order.  Cuboid: the cells with lo[d] <= q_d <= hi[d] on every axis
 * (1-based, inclusive: the lb..ub range of rhyme_drawing_sharp_cuboid.f90
 * on a base-level box). */
typedef struct rhy_shape {
    int32_t kind;
    int32_t lo[3], hi[3];
    int32_t pad;
    double centre[3];
    double radius;
    double rho, v[3], p;
} rhy_shape_t;
/* A harmonic perturbation (rhyme_drawing_apply_perturbations.f90:51-185):
 * profile[q-1] = A cos(2 pi / lambda * (q - 1/2)), evaluated on the host
 * for the dims[axis] cells along the 0-based axis; every cell then gets
 * conv_prim_to_cons(h) added to its hydro components, with h the sum over
 * the drawing's perturbations (from 0) of profile * (rho, v, p). */
typedef struct rhy_harmonic {
    int32_t axis;
    int32_t pad;
    const double *profile;
    double rho, v[3], p;
} rhy_harmonic_t;
/* The drawing: canvas colour (rhyme_drawing_uniform_canvas.f90), shapes,
 * perturbations; host arrays.  Ghost cells and non-hydro components are
 * written 0. */
typedef struct rhy_drawing {
    double rho, v[3], p;
    int32_t nshapes, nharmonics;
    const rhy_shape_t *shapes;
    const rhy_harmonic_t *harmonics;
} rhy_drawing_t;
/* paint the box's device state from the drawing (gamma from p; the host
 * mirror is not touched).  Blocks until done. */
int rhy_box_draw(rhy_box_t *box, const rhy_params_t *p, const rhy_drawing_t *d);

/* ---- unit-level batches (host arrays, NVAR = ndim+2 values per record,
 * record-major) so the reference's known-answer tests can run on the GPU ---- */

/* rhyme_riemann_problem_solve at dx = 0 (…_solve.f90:3-104): sampled
 * conserved state on each interface; axis[q] in 1..ndim. */
int rhy_riemann_batch(const rhy_params_t *p, int64_t n, const double *l,
                      const double *r, const int32_t *axis, double *u);
/* rhyme_riemann_problem_solve(rp, l, r, dx, dt, axis, u) (…_solve.f90:3-104)
 * at any sample point: record q is sampled at x/t = dx[q] / dt[q] (the
 * reference's dxdt = dx/dt, :17), through the same device routine the flux
 * kernels use at dx = 0.  u receives the conserved state. */
int rhy_riemann_solve_batch(const rhy_params_t *p, int64_t n, const double *l,
                            const double *r, const int32_t *axis, const double *dx,
                            const double *dt, double *u);

/* riemann_problem_solution_t (rhyme_riemann_problem/src/rhyme_riemann_problem.f90:33-61)
 * flattened: the two sides' primitive states (rp_side_t: rho, v(NDIM), p,
 * cs; v holds ndim components) and the star region (rp_star_t: p, u, and per
 * side is_shock, shock%rho/speed, fan%rho/cs/speedH/speedT, and the wave
 * function value f at p* that rhyme_irs_exec returns through l_f / r_f). */
typedef struct rhy_rp_solution {
    double l_rho, l_v[3], l_p, l_cs;
    double r_rho, r_v[3], r_p, r_cs;
    double star_p, star_u;
    int32_t l_is_shock, r_is_shock;
    double l_shock_rho, l_shock_speed;
    double l_fan_rho, l_fan_cs, l_fan_speedH, l_fan_speedT;
    double r_shock_rho, r_shock_speed;
    double r_fan_rho, r_fan_cs, r_fan_speedH, r_fan_speedT;
    double l_f, r_f;
} rhy_rp_solution_t;

/* rhyme_riemann_problem_star(rp, sol, axis) (…_star.f90:3-66): sol[q]'s sides
 * are read, its star members written (p* floored at w_vacuum(p)). */
int rhy_riemann_star_batch(const rhy_params_t *p, int64_t n, rhy_rp_solution_t *sol,
                           const int32_t *axis);
/* rhyme_riemann_problem_sampling(solution, axis, dx, dt, u) (…_sampling.f90:3-48):
 * the conserved state of sol[q] at x/t = dx[q] / dt[q]. */
int rhy_riemann_sampling_batch(const rhy_params_t *p, int64_t n, const rhy_rp_solution_t *sol,
                               const int32_t *axis, const double *dx, const double *dt,
                               double *u);
/* The region states the sampling chooses between, one reference function
 * each: riemann_problem_w_k of the left / right side (…_w_k.f90:3-11),
 * w_starL_sho / w_starR_sho (…_w_starL_sho.f90:3-15, …_w_starR_sho.f90),
 * w_starL_fan / w_starR_fan (…_w_starL_fan.f90, …_w_starR_fan.f90) and
 * w_kfan(side, dxdt, axis, is_right) (…_w_kfan.f90:3-32). */
#define RHY_W_L 1
#define RHY_W_R 2
#define RHY_W_STARL_SHO 3
#define RHY_W_STARR_SHO 4
#define RHY_W_STARL_FAN 5
#define RHY_W_STARR_FAN 6
#define RHY_W_LFAN 7
#define RHY_W_RFAN 8
int rhy_riemann_region_batch(const rhy_params_t *p, int64_t n, const rhy_rp_solution_t *sol,
                             const int32_t *region, const int32_t *axis, const double *dxdt,
                             double *u);

/* rhyme_irs_exec(irs, l_rho, l_v, l_p, l_cs, l_f, r_rho, r_v, r_p, r_cs, r_f, axis)
 * (rhyme_irs/src/rhyme_irs_exec.f90:3-45): l and r are [n][4] records
 * (rho, v(axis), p, cs); p_star, l_f, r_f receive p* and the wave functions'
 * values at the last iterate (the reference's intent(inout) l_f, r_f). */
int rhy_irs_exec_batch(const rhy_params_t *p, int64_t n, const double *l, const double *r,
                       double *p_star, double *l_f, double *r_f);
/* rhyme_irs_guess_p_star (rhyme_irs/src/rhyme_irs_guess_p_star.f90:3-72):
 * the six starting guesses, [n][6], from the same records. */
int rhy_irs_guess_batch(const rhy_params_t *p, int64_t n, const double *l, const double *r,
                        double *guesses);

/* x**y on the device, for pinning the exponentiation of the Riemann solve
 * (Fortran ** on real(8), i.e. libm pow in the reference): policy
 * RHY_POW_PATH is the FP64 path's own (this build's flux/fixup policy),
 * RHY_POW_LIBM the device math library's pow, RHY_POW_MIXED the mixed form's. */
#define RHY_POW_PATH 0
#define RHY_POW_LIBM 1
#define RHY_POW_MIXED 2
int rhy_pow_batch(int32_t policy, int64_t n, const double *x, const double *y, double *out);

/* rhyme_slope_limiter_run (rhyme_slope_limiter/src/rhyme_slope_limiter_run.f90:3-21) */
int rhy_limiter_batch(const rhy_params_t *p, int64_t n, const double *ul,
                      const double *u, const double *ur, double *delta);
/* rhyme_muscl_hancock_half_step_extrapolation (…_half_step_extrapolation.f90:3-24) */
int rhy_half_step_batch(const rhy_params_t *p, int64_t n, const double *u,
                        const double *delta, const int32_t *axis, double dx,
                        double dt, double *l, double *r);
/* calc_flux -> rhyme_ideal_gas_flux (rhyme_ideal_gas/src/rhyme_ideal_gas_flux.f90:3-24) */
int rhy_flux_batch(const rhy_params_t *p, int64_t n, const double *u,
                   const int32_t *axis, double *f);

/* ---- batched form: many independent boxes of uneven shapes ----
 * No counterpart in the reference, which advances its SAMR boxes one call
 * at a time (src/main.f90:130-136) and has no per-box ghost exchange or dt.
 * Here every box keeps its own ghost fill (p->bc_types on its faces) and its
 * own CFL dt, so each box evolves exactly as rhy_run would advance it alone.
 * One launch per phase (dt, all fluxes, all updates) covers every box through
 * a tile table walked by persistent workgroups; the step is captured once in
 * a hipGraph and replayed when use_graph is set.  dims and dx are
 * [nboxes][3] (row-major); host arrays per box as for rhy_box_upload. */
typedef struct rhy_batch rhy_batch_t;
int rhy_batch_create(int32_t ndim, int32_t nboxes, const int32_t *dims, const double *dx,
                     int32_t ncmp, int32_t precision, rhy_batch_t **out);
int rhy_batch_destroy(rhy_batch_t *batch);
/* device bytes held (state + flux scratch) */
int rhy_batch_bytes(rhy_batch_t *batch, int64_t *bytes);
int rhy_batch_upload(rhy_batch_t *batch, int32_t box, const double *cells);
int rhy_batch_download(rhy_batch_t *batch, int32_t box, double *cells);
/* tile shape: threads per workgroup (64..256, multiple of 64), steps of 64
 * row cells per wave along x, faces walked per thread along y/z, cap on
 * persistent workgroups (0: auto) */
int rhy_batch_set_tiles(rhy_batch_t *batch, int32_t block, int32_t nf_x, int32_t nf_yz,
                        int32_t grid_cap);
int rhy_batch_run_async(rhy_batch_t *batch, const rhy_params_t *p, int32_t nsteps,
                        int32_t use_graph);
int rhy_batch_sync(rhy_batch_t *batch);
/* dts of the last run, [n][nboxes] (for a pool: [n][max_boxes], indexed by
 * slot; entries of slots that were not active are left as they were) */
int rhy_batch_last_dts(rhy_batch_t *batch, int32_t n, double *dts);
const char *rhy_batch_last_error(void);
/* rhy_box_diagnostics of one box (or pool slot) of a batch */
int rhy_batch_diagnostics(rhy_batch_t *batch, int32_t box, rhy_diag_t *out);
/* faces the last flux phase left to the fixup kernel (see rhy_box_last_deferred) */
int rhy_batch_last_deferred(rhy_batch_t *batch, int64_t *n);

/* ---- two-level hierarchy (refinement ratio 2) ----
 * No counterpart in the reference, whose main loop advances every box of
 * every level alone with the level's dt (src/main.f90:118-136) and has no
 * coarse-fine coupling.  The coarse level is a box, the fine level a fixed
 * batch whose boxes use face type 0 ("set by the hierarchy": no ghost
 * fill).  lo[3*i+d] is the coarse cell (1-based) that holds fine cells 1
 * and 2 of box i along d; fine extents are even, and every fine box with
 * its 2 ghost layers lies in the coarse interior.
 * rhy_batch_cfl_dts: the CFL dt of every box for its current state (into
 *   dts[nboxes], no step).
 * rhy_batch_run_dt_async: nsteps steps with the given dt for every box.
 * rhy_amr_prolong: the fine boxes' face ghosts from their coarse parents,
 *   piecewise constant.  half = 0 sets each ghost to its parent: call it
 *   before the coarse step.  half = 1 sets each ghost to the mean of its
 *   value and its parent's: call it after the coarse step and the first fine
 *   substep (fine steps with face type 0 leave face ghosts alone), so the
 *   ghost becomes 0.5 * U_old + 0.5 * U_new.  Both then copy the face ghosts
 *   that lie in another fine box's interior from that box.
 * rhy_amr_restrict: every coarse cell under a fine box takes the mean of
 *   its 2^ndim children (summed with i fastest, times 2^-ndim).
 * rhy_amr_flux_register: after a fine substep, for every coarse face on the
 *   boundary of every fine box, the sum S of the fluxes of its 2^(ndim-1)
 *   fine faces (the last step's flux planes), pairwise: (F_00 + F_10) +
 *   (F_01 + F_11), first index on the lower transverse axis, is stored in
 *   the batch's flux register (reset = 1: after the first substep of a
 *   coarse step) or added to it (reset = 0).
 * rhy_amr_reflux: after the coarse step (rhy_mh_solve, whose flux planes
 *   hold the coarse fluxes, or rhy_amr_coarse_step) and both substeps,
 *   every coarse cell next to a fine box takes the correction of its face
 *   flux, F_c replaced by the fine fluxes' mean over the two substeps and
 *   the 2^(ndim-1) fine faces: U += dt/dx_a (F_c - <F_f>) on the low side
 *   of a box, dt/dx_a (<F_f> - F_c) on the high side, applied axis by axis,
 *   low side then high side.  Cells under another fine box get it too and
 *   are then restricted over.
 * rhy_amr_coarse_step: no counterpart in the reference (it has no coupled
 *   levels); for callers of the hierarchy.  The coarse step as the ghost
 *   fill and one step with the given dt (rhy_bc_set_boundaries +
 *   rhy_mh_solve, bit for bit) through the accumulating walks of rhy_run
 *   instead of the flux planes: only the x fluxes stay in the flux planes,
 *   and the box keeps the state it stepped from in its second state buffer,
 *   from which rhy_amr_reflux forms the y/z coarse fluxes it needs (as the
 *   walks formed them, bit for bit).  Any later call that changes the box
 *   (upload, step, paint, draw, resize, unpack) ends that; the box's ghost
 *   faces are not filled after the step.  1D boxes take rhy_mh_solve.
 * The prolong, restrict, register and reflux calls wait for both sides'
 * streams; rhy_amr_coarse_step returns when its step is done. */
int rhy_batch_cfl_dts(rhy_batch_t *batch, const rhy_params_t *p, double *dts);
int rhy_batch_run_dt_async(rhy_batch_t *batch, const rhy_params_t *p, double dt,
                           int32_t nsteps);
int rhy_amr_prolong(rhy_box_t *coarse, rhy_batch_t *fine, const int32_t *lo, int32_t half);
int rhy_amr_restrict(rhy_box_t *coarse, rhy_batch_t *fine, const int32_t *lo);
int rhy_amr_flux_register(rhy_batch_t *fine, const int32_t *lo, int32_t reset);
int rhy_amr_reflux(rhy_box_t *coarse, rhy_batch_t *fine, const int32_t *lo, double dt,
                   const double dx[3]);
int rhy_amr_coarse_step(rhy_box_t *coarse, const rhy_params_t *p, const double dx[3], double dt);
/* rhy_amr_regrid: no counterpart in the reference (it has no coupled
 * levels); for callers of the hierarchy that move the fine level.  Fills
 * the fine batch `to` (its boxes placed at to_lo, same rules as lo above)
 * from the current coarse box and the old fine batch `from` (at from_lo):
 * every cell of every new box, both ghost layers included, takes the old
 * fine cell at the same fine-level position when an old box's interior
 * holds it, else its coarse parent's value (the hydro components).  Call
 * it after a restriction, so that coarse cells under old boxes hold their
 * children's mean and a newly refined cell's children restrict back to it.
 * Waits for both batches' and the coarse box's streams. */
int rhy_amr_regrid(rhy_box_t *coarse, rhy_batch_t *from, const int32_t *from_lo,
                   rhy_batch_t *to, const int32_t *to_lo);
/* The same calls with a box of a fixed batch as the coarse side (`slot` of
 * `coarse`): a level between the coarsest box and the finest batch, whose
 * fine boxes all lie in that one box (lo relative to it).  No counterpart in
 * the reference; for hierarchies of more than two levels.  The coarse
 * fluxes of rhy_amr_reflux_slot are that box's flux planes from its batch's
 * last step. */
int rhy_amr_prolong_slot(rhy_batch_t *coarse, int32_t slot, rhy_batch_t *fine,
                         const int32_t *lo, int32_t half);
int rhy_amr_restrict_slot(rhy_batch_t *coarse, int32_t slot, rhy_batch_t *fine,
                          const int32_t *lo);
int rhy_amr_reflux_slot(rhy_batch_t *coarse, int32_t slot, rhy_batch_t *fine, const int32_t *lo,
                        double dt, const double dx[3]);
int rhy_amr_regrid_slot(rhy_batch_t *coarse, int32_t slot, rhy_batch_t *from,
                        const int32_t *from_lo, rhy_batch_t *to, const int32_t *to_lo);

/* Snapshot payload of a whole level (a batch or pool): the active boxes in
 * slot order, each laid out as rhy_box_snapshot_pack lays out one box; the
 * length is in floats.  Unpacking marks every box for a ghost fill before
 * its next step. */
int rhy_batch_snapshot_length(rhy_batch_t *batch, int64_t *n);
int rhy_batch_snapshot_pack(rhy_batch_t *batch, float *data);
int rhy_batch_snapshot_unpack(rhy_batch_t *batch, const float *data);

/* ---- pool: a batch whose boxes come and go (continuous batching) ----
 * No counterpart in the reference.  A pool holds a device arena of
 * budget_bytes; rhy_pool_add places a box (state + flux scratch,
 * rhy_pool_box_bytes of it) first-fit in the arena and returns its slot, or
 * RHY_ERR_OOM when no free block is large enough or all max_boxes slots are
 * taken.  Each box carries its own slope limiter and face types (bc_types
 * NULL: reflective); p->limiter and p->bc_types are ignored by
 * rhy_batch_run_async on a pool, every other parameter applies to all its
 * boxes.  Add and remove between runs (both wait for the pool's stream);
 * rhy_batch_upload/download/last_dts take the slot as the box index. */
int rhy_pool_create(int32_t ndim, int32_t ncmp, int32_t precision, int64_t budget_bytes,
                    int32_t max_boxes, rhy_batch_t **out);
int rhy_pool_box_bytes(int32_t ndim, const int32_t dims[3], int32_t ncmp, int32_t precision,
                       int64_t *bytes);
int rhy_pool_add(rhy_batch_t *pool, const int32_t dims[3], const double dx[3],
                 const double *cells, int32_t limiter, const int32_t bc_types[6],
                 int32_t *slot);
/* as rhy_pool_add, with the state painted on the device from `ic` */
int rhy_pool_add_ic(rhy_batch_t *pool, const int32_t dims[3], const double dx[3],
                    const rhy_ic_t *ic, int32_t limiter, const int32_t bc_types[6],
                    int32_t *slot);
int rhy_pool_remove(rhy_batch_t *pool, int32_t slot);
/* total free bytes of the arena and the largest free block */
int rhy_pool_free_bytes(rhy_batch_t *pool, int64_t *free_bytes, int64_t *largest_bytes);

#ifdef __cplusplus
}
#endif
#endif
