/*
 * rhyme_oracle.c -- CPU restatement of the Rhyme hydro hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see
 * rhyme_oracle.h.  Built with gcc -O2 -ffp-contract=off -fopenmp, i.e. no FMA
 * contraction and IEEE division/sqrt, matching gfortran -O2 on x86-64.
 * Fortran `x**y` with a real exponent is libm pow(); `x**2` and `x**2.d0`
 * are x*x.  Each expression keeps the Fortran parse (left to right for
 * * and /, parentheses as written).  File:line citations are relative to the
 * reference's src/ directory.
 *
 * Pinning (what checks each function): see tests/test_oracle_golden.py and
 * tests/test_oracle_invariants.py.
 */
#include "rhyme_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TINY DBL_MIN /* tiny(0d0) = 2.2250738585072014e-308 */
#define NV_MAX 5

/* gamma-derived constants: rhyme_riemann_problem_init.f90:20-27 and
 * rhyme_irs/src/rhyme_irs_init.f90:21-27 (identical in both modules). */
typedef struct gconst {
    double g, gm1, gp1, gm1_gp1, gm1_2g, gp1_2g, g_inv;
} gconst;

static gconst make_gconst(double g)
{
    gconst c;
    c.g = g;
    c.gm1 = g - 1.0;
    c.gp1 = g + 1.0;
    c.gm1_gp1 = c.gm1 / c.gp1;
    c.gm1_2g = c.gm1 / (2.0 * g);
    c.gp1_2g = c.gp1 / (2.0 * g);
    c.g_inv = 1.0 / g;
    return c;
}

/* ------------------------------------------------------------------------ */
/* hydro_base / ideal_gas                                                   */
/* ------------------------------------------------------------------------ */

/* rhyme_hydro_base_specific_kinetic_energy.f90:9
 *   .5d0*sum(u(rho_u:rho_u+NDIM-1)**2)/u(rho)**2 */
static double sp_kin_e(int nd, const double *u)
{
    double s = 0.0;
    for (int d = 0; d < nd; ++d)
        s += u[1 + d] * u[1 + d];
    return .5 * s / (u[0] * u[0]);
}

/* rhyme_hydro_base_specific_internal_energy.f90:9 */
static double sp_int_e(int nd, const double *u)
{
    return u[nd + 1] / u[0] - sp_kin_e(nd, u);
}

/* rhyme_ideal_gas_pressure.f90:9: u(rho)*e_int*(gamma - 1) */
static double pressure(int nd, double g, const double *u)
{
    return u[0] * sp_int_e(nd, u) * (g - 1.0);
}

/* rhyme_ideal_gas_temperature_per_mu.f90:9-10 and
 * rhyme_ideal_gas_speed_of_sound.f90:9:
 *   sqrt(gamma*kb_amu*(e_int*(gamma-1)/kb_amu)) */
static double speed_of_sound(int nd, double g, double kb_amu, const double *u)
{
    double t_mu = sp_int_e(nd, u) * (g - 1.0) / kb_amu;
    return sqrt(g * kb_amu * t_mu);
}

/* rhyme_ideal_gas_flux.f90:12-23 (axis is 1-based) */
static void flux(int nd, double g, const double *u, int axis, double *f)
{
    const int ne = nd + 1;
    if (fabs(u[0]) < TINY) {
        for (int c = 0; c <= ne; ++c)
            f[c] = 0.0;
        return;
    }
    double p = pressure(nd, g, u);
    f[0] = u[axis];
    for (int d = 0; d < nd; ++d)
        f[1 + d] = u[1 + d] * f[0] / u[0];
    f[axis] = f[axis] + p;
    f[ne] = f[0] / u[0] * (u[ne] + p);
}

/* rhyme_ideal_gas_primitive_vars_to_conserved.f90:9-11 */
static void prim_vars_to_cons(int nd, double g, double rho, const double *v,
                              double p, double *u)
{
    double s = 0.0;
    u[0] = rho;
    for (int d = 0; d < nd; ++d)
        u[1 + d] = rho * v[d];
    for (int d = 0; d < nd; ++d)
        s += v[d] * v[d];
    u[nd + 1] = .5 * rho * s + p / (g - 1.0);
}

/* rhyme_ideal_gas_primitive_to_conserved.f90 applied to rp%w_vacuum
 * (whose velocity components are never set, so they are 0). */
static void vacuum_to_cons(int nd, double g, double w_rho, double w_p,
                           double *u)
{
    double v[3] = {0.0, 0.0, 0.0};
    prim_vars_to_cons(nd, g, w_rho, v, w_p, u);
}

/* ------------------------------------------------------------------------ */
/* slope limiter: rhyme_slope_limiter/src (all limiters)                  */
/* ------------------------------------------------------------------------ */

/* rhyme_slope_limiter.f90:62-82 */
static double limiter_r(double ul, double u, double ur)
{
    double denom = ur - u;
    if (fabs(denom) < TINY)
        return 0.0;
    return (u - ul) / denom;
}

/* rhyme_slope_limiter_xi_r.f90:11 */
static double xi_r(double w, double r)
{
    return 2.0 / (1.0 - w + (1.0 + w) * r);
}

static double fmin_f(double a, double b) { return a < b ? a : b; }

static void slope_limiter(int nd, int type, double w, const double *ul,
                          const double *u, const double *ur, double *delta)
{
    for (int c = 0; c < nd + 2; ++c) {
        double dL = u[c] - ul[c];
        double dR = ur[c] - u[c];
        /* d = .5d0*(1.d0 + sl%w)*d_L + .5d0*(1.d0 + sl%w)*d_R */
        double d = .5 * (1.0 + w) * dL + .5 * (1.0 + w) * dR;
        double r = limiter_r(ul[c], u[c], ur[c]);
        double out;
        switch (type) {
        case 1: /* rhyme_slope_limiter_van_leer.f90:13-27 */
            out = r > 0.0 ? fmin_f(2.0 * r / (1.0 + r), xi_r(w, r)) * d : 0.0;
            break;
        case 2: /* rhyme_slope_limiter_minmod.f90:13-29 */
            if (r > 1.0)
                out = fmin_f(1.0, xi_r(w, r)) * d;
            else if (r > 0.0)
                out = r * d;
            else
                out = 0.0;
            break;
        case 3: /* rhyme_slope_limiter_van_albada.f90:13-27 */
            out = r > 0.0
                      ? fmin_f(r * (1.0 + r) / (1.0 + r * r), xi_r(w, r)) * d
                      : 0.0;
            break;
        case 4: /* rhyme_slope_limiter_superbee.f90:13-31 */
            if (r > 1.0)
                out = fmin_f(fmin_f(r, xi_r(w, r)), 2.0) * d;
            else if (r > .5)
                out = d;
            else if (r > 0.0)
                out = 2.0 * r * d;
            else
                out = 0.0;
            break;
        default: /* rhyme_slope_limiter_run.f90:10-19: no case, delta unset */
            out = 0.0;
        }
        delta[c] = out;
    }
}

/* rhyme_muscl_hancock_half_step_extrapolation.f90:15-21 */
static void half_step(int nd, double g, const double *u, const double *delta,
                      int axis, double dx, double dt, double *l, double *r)
{
    double fl[NV_MAX], fr[NV_MAX];
    const int nv = nd + 2;
    for (int c = 0; c < nv; ++c) {
        l[c] = u[c] - .5 * delta[c];
        r[c] = u[c] + .5 * delta[c];
    }
    flux(nd, g, l, axis, fl);
    flux(nd, g, r, axis, fr);
    for (int c = 0; c < nv; ++c) {
        double df = fl[c] - fr[c];
        l[c] = l[c] + .5 * dt / dx * df;
        r[c] = r[c] + .5 * dt / dx * df;
    }
}

/* ------------------------------------------------------------------------ */
/* iterative Riemann solver: rhyme_riemann_problem/src/rhyme_irs/src         */
/* ------------------------------------------------------------------------ */

/* rhyme_irs_nonlinear_wave_function.f90:13-23 */
static void wave_function(const gconst *k, double rho, double pk, double cs,
                          double p, double *f, double *fprime)
{
    if (p > pk) {
        double Ak = 2.0 / (k->gp1 * rho);
        double Bk = k->gm1 * pk / k->gp1;
        double factor = sqrt(Ak / (Bk + p));
        *f = (p - pk) * factor;
        *fprime = factor * (1.0 - (p - pk) / (2.0 * (Bk + p)));
    } else {
        *f = 2.0 * cs / k->gm1 * (pow(p / pk, k->gm1_2g) - 1.0);
        *fprime = 1.0 / (rho * cs) * pow(p / pk, -k->gp1_2g);
    }
}

/* rhyme_irs_guess_p_star.f90:69-76 */
static double guess_g_K(const gconst *k, double rho, double p)
{
    return sqrt((2.0 / (k->gp1 * rho)) / (p + k->gm1_gp1 * p));
}

/* Fortran MAX/MIN on valid data (NaN never occurs on the path) */
static double fmax_f(double a, double b) { return a > b ? a : b; }

/* rhyme_irs_guess_p_star.f90:3-72 (axis 1-based; v arrays hold NDIM values) */
static void irs_guess(const gconst *k, double l_rho, const double *l_v,
                      double l_p, double l_cs, double r_rho, const double *r_v,
                      double r_p, double r_cs, int axis, double *ps)
{
    const double lv = l_v[axis - 1], rv = r_v[axis - 1];
    const double q_user = 2.0;
    /* :16-20; the 0.25/0.5/1.0/2.0 default-real literals are exact */
    double cup = 0.25 * (l_rho + r_rho) * (l_cs + r_cs);
    double ppv = fmax_f(0.5 * (l_p + r_p) + 0.5 * (lv - rv) * cup, 0.0);
    double p_min = fmin_f(l_p, r_p);
    double p_max = fmax_f(l_p, r_p);
    double q_max = p_max / p_min;

    if (q_max <= q_user && (p_min < ppv && ppv < p_max)) {
        ps[0] = ppv; /* PVRS */
    } else if (ppv < p_min) {
        /* two-rarefaction: :28-33 */
        double pq = pow(l_p / r_p, k->gm1_2g);
        double um = (pq * lv / l_cs + rv / r_cs + 2.0 / k->gm1 * (pq - 1.0)) /
                    (pq / l_cs + 1.0 / r_cs);
        double ptl = 1.0 + k->gm1 / 2.0 * (lv - um) / l_cs;
        double ptr = 1.0 + k->gm1 / 2.0 * (um - rv) / r_cs;
        ps[0] = 0.5 * (l_p * pow(ptl, 1.0 / k->gm1_2g) +
                       r_p * pow(ptr, 1.0 / k->gm1_2g));
    } else {
        /* two-shock: :35-38 */
        double gel = sqrt(((2.0 / (k->gp1)) / l_rho) / (k->gm1_gp1 * l_p + ppv));
        double ger = sqrt(((2.0 / (k->gp1)) / r_rho) / (k->gm1_gp1 * r_p + ppv));
        ps[0] = (gel * l_p + ger * r_p - (rv - lv)) / (gel + ger);
    }

    /* Ramses guess :42-46 */
    ps[1] = (r_rho * r_cs * l_p + l_rho * l_cs * r_p +
             r_cs * l_cs * (l_rho * lv - r_rho * rv)) /
            (r_rho * r_cs + l_rho * l_cs);

    /* two-rarefaction approximation :49-55 */
    ps[2] = pow((l_cs + r_cs - .5 * k->gm1 * (rv - lv)) /
                    (pow(l_cs / l_p, k->gm1_2g) + pow(r_cs / r_p, k->gm1_2g)),
                1.0 / k->gm1_2g);

    /* two-shock approximation :58-60 */
    double l_g = guess_g_K(k, l_rho, l_p);
    double r_g = guess_g_K(k, r_rho, r_p);
    ps[3] = (l_g * l_p + r_g * r_p - (rv - lv)) / (l_g + r_g);

    /* p_PV :63-64 */
    ps[4] = .5 * (l_p + r_p) - .125 * (rv - lv) * (l_rho + r_rho) * (l_cs + r_cs);

    /* arithmetic mean :67 */
    ps[5] = .5 * (l_p + r_p);

    /* floor :71 */
    for (int i = 0; i < 6; ++i)
        ps[i] = fmax_f(ps[i], TINY);
}

/* rhyme_irs_exec.f90:3-45 */
static double irs_exec(const gconst *k, int n_iteration, double tolerance,
                       double l_rho, const double *l_v, double l_p,
                       double l_cs, double *l_f, double r_rho,
                       const double *r_v, double r_p, double r_cs, double *r_f,
                       int axis, int32_t *n_newton)
{
    double guessed[6];
    double p_star = 0.0, p_prev, l_fp = 0.0, r_fp = 0.0;
    int32_t its = 0;
    const double du = r_v[axis - 1] - l_v[axis - 1];

    irs_guess(k, l_rho, l_v, l_p, l_cs, r_rho, r_v, r_p, r_cs, axis, guessed);

    for (int gid = 0; gid < 6; ++gid) {
        p_star = guessed[gid];
        p_prev = p_star;
        for (int i = 1; i <= n_iteration; ++i) {
            ++its;
            wave_function(k, l_rho, l_p, l_cs, p_star, l_f, &l_fp);
            wave_function(k, r_rho, r_p, r_cs, p_star, r_f, &r_fp);
            p_star = p_prev - (*l_f + *r_f + du) / (l_fp + r_fp);
            if (p_star < 0.0)
                break;
            if (2.0 * fabs(p_star - p_prev) / (p_star + p_prev) < tolerance)
                break;
            p_prev = p_star;
        }
        if (p_star > 0.0)
            break;
    }
    if (p_star < 0.0)
        p_star = 0.0;
    if (n_newton)
        *n_newton = its;
    return p_star;
}

/* ------------------------------------------------------------------------ */
/* Riemann problem: rhyme_riemann_problem/src                                */
/* ------------------------------------------------------------------------ */

typedef struct side {
    double rho, v[3], p, cs;
} side;

/* rhyme_riemann_problem_star.f90:3-66 (the ExactRS branch) */
static void rp_star(const gconst *k, const orc_params *pp, orc_rp_solution *s,
                    int axis, int32_t *n_newton)
{
    const int a = axis - 1;
    s->star_p = irs_exec(k, pp->n_iteration, pp->tolerance, s->l_rho, s->l_v,
                         s->l_p, s->l_cs, &s->l_f, s->r_rho, s->r_v, s->r_p,
                         s->r_cs, &s->r_f, axis, n_newton);
    if (s->star_p < pp->w_vacuum_p)
        s->star_p = pp->w_vacuum_p;

    s->star_u = 0.5 * ((s->r_v[a] + s->l_v[a]) + (s->r_f - s->l_f));

    double ps_pl = s->star_p / s->l_p;
    double ps_pr = s->star_p / s->r_p;

    if (s->star_p > s->l_p) {
        s->l_is_shock = 1;
        s->l_shock_rho = s->l_rho * (k->gm1_gp1 + ps_pl) / (k->gm1_gp1 * ps_pl + 1.0);
        s->l_shock_speed = s->l_v[a] - s->l_cs * sqrt(k->gp1_2g * ps_pl + k->gm1_2g);
    } else {
        s->l_is_shock = 0;
        s->l_fan_rho = s->l_rho * pow(ps_pl, k->g_inv);
        s->l_fan_cs = s->l_cs * pow(ps_pl, k->gm1_2g);
        s->l_fan_speedH = s->l_v[a] - s->l_cs;
        s->l_fan_speedT = s->star_u - s->l_fan_cs;
    }

    if (s->star_p > s->r_p) {
        s->r_is_shock = 1;
        s->r_shock_rho = s->r_rho * (k->gm1_gp1 + ps_pr) / (k->gm1_gp1 * ps_pr + 1.0);
        s->r_shock_speed = s->r_v[a] + s->r_cs * sqrt(k->gp1_2g * ps_pr + k->gm1_2g);
    } else {
        s->r_is_shock = 0;
        s->r_fan_rho = s->r_rho * pow(ps_pr, k->g_inv);
        s->r_fan_cs = s->r_cs * pow(ps_pr, k->gm1_2g);
        s->r_fan_speedH = s->r_v[a] + s->r_cs;
        s->r_fan_speedT = s->star_u + s->r_fan_cs;
    }
}

/* rhyme_riemann_problem_w_kfan.f90:14-31 */
static void w_kfan(int nd, const gconst *k, const side *s, double dxdt,
                   int axis, int is_right, double *u)
{
    const int a = axis - 1;
    double cs = is_right ? -s->cs : s->cs;
    double v[3] = {s->v[0], s->v[1], s->v[2]};
    double rho = s->rho * pow(2.0 / k->gp1 + k->gm1_gp1 / cs * (s->v[a] - dxdt),
                              2.0 / k->gm1);
    v[a] = 2.0 / k->gp1 * (cs + k->gm1 / 2.0 * s->v[a] + dxdt);
    double p = s->p * pow(2.0 / k->gp1 + k->gm1_gp1 / cs * (s->v[a] - dxdt),
                          1.0 / k->gm1_2g);
    prim_vars_to_cons(nd, k->g, rho, v, p, u);
}

static side sol_left(const orc_rp_solution *s)
{
    side x = {s->l_rho, {s->l_v[0], s->l_v[1], s->l_v[2]}, s->l_p, s->l_cs};
    return x;
}

static side sol_right(const orc_rp_solution *s)
{
    side x = {s->r_rho, {s->r_v[0], s->r_v[1], s->r_v[2]}, s->r_p, s->r_cs};
    return x;
}

/* w_star{L,R}_{sho,fan}: rhyme_riemann_problem_w_starL_sho.f90:12-15 etc. */
static void w_star(int nd, const gconst *k, const double *side_v, double rho,
                   double star_u, double star_p, int axis, double *u)
{
    double v[3] = {side_v[0], side_v[1], side_v[2]};
    v[axis - 1] = star_u;
    prim_vars_to_cons(nd, k->g, rho, v, star_p, u);
}

/* rhyme_riemann_problem_sampling.f90:13-47 */
static void rp_sampling(int nd, const gconst *k, const orc_rp_solution *s,
                        int axis, double dx, double dt, double *u)
{
    double dxdt = dx / dt;
    if (dxdt > s->star_u) {
        side R = sol_right(s);
        if (s->r_is_shock) {
            if (dxdt < s->r_shock_speed)
                w_star(nd, k, s->r_v, s->r_shock_rho, s->star_u, s->star_p, axis, u);
            else
                prim_vars_to_cons(nd, k->g, R.rho, R.v, R.p, u);
        } else {
            if (dxdt < s->r_fan_speedT)
                w_star(nd, k, s->r_v, s->r_fan_rho, s->star_u, s->star_p, axis, u);
            else if (dxdt < s->r_fan_speedH)
                w_kfan(nd, k, &R, dxdt, axis, 1, u);
            else
                prim_vars_to_cons(nd, k->g, R.rho, R.v, R.p, u);
        }
    } else {
        side L = sol_left(s);
        if (s->l_is_shock) {
            if (dxdt > s->l_shock_speed)
                w_star(nd, k, s->l_v, s->l_shock_rho, s->star_u, s->star_p, axis, u);
            else
                prim_vars_to_cons(nd, k->g, L.rho, L.v, L.p, u);
        } else {
            if (dxdt > s->l_fan_speedT)
                w_star(nd, k, s->l_v, s->l_fan_rho, s->star_u, s->star_p, axis, u);
            else if (dxdt > s->l_fan_speedH)
                w_kfan(nd, k, &L, dxdt, axis, 0, u);
            else
                prim_vars_to_cons(nd, k->g, L.rho, L.v, L.p, u);
        }
    }
}

/* rhyme_riemann_problem_solve.f90:3-104 */
static void rp_solve(const orc_params *pp, const gconst *k, const double *l,
                     const double *r, double dx, double dt, int axis,
                     double *u, int32_t *n_newton)
{
    const int nd = pp->ndim;
    const int a = axis - 1;
    const double w_rho = pp->w_vacuum_rho, w_p = pp->w_vacuum_p;
    orc_rp_solution s;
    memset(&s, 0, sizeof(s));
    double dxdt = dx / dt;
    if (n_newton)
        *n_newton = 0;

    s.l_rho = fmax_f(l[0], w_rho);
    s.r_rho = fmax_f(r[0], w_rho);

    double e_rho = nextafter(w_rho, INFINITY);
    int vacuum_right = s.r_rho < e_rho && s.l_rho > e_rho;
    int vacuum_left = s.l_rho < e_rho && s.r_rho > e_rho;
    int vacuum_both = s.l_rho < e_rho && s.r_rho < e_rho;

    if (vacuum_right) {
        for (int d = 0; d < nd; ++d)
            s.l_v[d] = l[1 + d] / l[0];
        s.l_p = fmax_f(pressure(nd, k->g, l), w_p);
        s.l_cs = fmax_f(speed_of_sound(nd, k->g, pp->kb_amu, l), TINY);
        s.r_p = w_p;
        s.r_cs = 0.0;
        double s_starl = s.l_v[a] - 2 * s.l_cs / k->gm1;
        side L = sol_left(&s);
        if (dxdt > s_starl)
            vacuum_to_cons(nd, k->g, w_rho, w_p, u);
        else if (dxdt > s.l_v[a] - s.l_cs)
            w_kfan(nd, k, &L, dxdt, axis, 0, u);
        else
            prim_vars_to_cons(nd, k->g, L.rho, L.v, L.p, u);
    } else if (vacuum_left) {
        s.l_p = w_p;
        s.l_cs = 0.0;
        for (int d = 0; d < nd; ++d)
            s.r_v[d] = r[1 + d] / r[0];
        s.r_p = fmax_f(pressure(nd, k->g, r), w_p);
        s.r_cs = fmax_f(speed_of_sound(nd, k->g, pp->kb_amu, r), TINY);
        double s_starr = s.r_v[a] + 2 * s.r_cs / k->gm1;
        side R = sol_right(&s);
        if (dxdt > s.r_v[a] + s.r_cs)
            prim_vars_to_cons(nd, k->g, R.rho, R.v, R.p, u);
        else if (dxdt > s_starr)
            w_kfan(nd, k, &R, dxdt, axis, 1, u);
        else
            vacuum_to_cons(nd, k->g, w_rho, w_p, u);
    } else if (vacuum_both) {
        vacuum_to_cons(nd, k->g, w_rho, w_p, u);
    } else {
        for (int d = 0; d < nd; ++d)
            s.l_v[d] = l[1 + d] / l[0];
        s.l_p = fmax_f(pressure(nd, k->g, l), w_p);
        s.l_cs = fmax_f(speed_of_sound(nd, k->g, pp->kb_amu, l), TINY);
        for (int d = 0; d < nd; ++d)
            s.r_v[d] = r[1 + d] / r[0];
        s.r_p = fmax_f(pressure(nd, k->g, r), w_p);
        s.r_cs = fmax_f(speed_of_sound(nd, k->g, pp->kb_amu, r), TINY);

        double s_starl = s.l_v[a] - 2 * s.l_cs / k->gm1;
        double s_starr = s.r_v[a] + 2 * s.r_cs / k->gm1;
        double du = s.r_v[a] - s.l_v[a];
        double du_crit = 2 * (s.l_cs + s.r_cs) / k->gm1;

        if (du_crit < du) { /* generation of vacuum :83-98 */
            side L = sol_left(&s), R = sol_right(&s);
            if (dxdt > s_starr) {
                if (dxdt > s.r_v[a] + s.r_cs)
                    prim_vars_to_cons(nd, k->g, R.rho, R.v, R.p, u);
                else
                    w_kfan(nd, k, &R, dxdt, axis, 1, u);
            } else if (dxdt > s_starl) {
                vacuum_to_cons(nd, k->g, w_rho, w_p, u);
            } else {
                if (dxdt > s.l_v[a] - s.l_cs)
                    w_kfan(nd, k, &L, dxdt, axis, 0, u);
                else
                    prim_vars_to_cons(nd, k->g, L.rho, L.v, L.p, u);
            }
        } else {
            rp_star(k, pp, &s, axis, n_newton);
            rp_sampling(nd, k, &s, axis, dx, dt, u);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* public unit-level wrappers                                                */
/* ------------------------------------------------------------------------ */

double orc_pressure(const orc_params *p, const double *u)
{
    return pressure(p->ndim, p->gamma, u);
}

double orc_speed_of_sound(const orc_params *p, const double *u)
{
    return speed_of_sound(p->ndim, p->gamma, p->kb_amu, u);
}

void orc_flux(const orc_params *p, const double *u, int axis, double *f)
{
    flux(p->ndim, p->gamma, u, axis, f);
}

void orc_prim_vars_to_cons(const orc_params *p, double rho, const double *v,
                           double pr, double *u)
{
    prim_vars_to_cons(p->ndim, p->gamma, rho, v, pr, u);
}

void orc_slope_limiter(const orc_params *p, const double *ul, const double *u,
                       const double *ur, double *delta)
{
    slope_limiter(p->ndim, p->limiter, p->omega, ul, u, ur, delta);
}

void orc_half_step(const orc_params *p, const double *u, const double *delta,
                   int axis, double dx, double dt, double *l, double *r)
{
    half_step(p->ndim, p->gamma, u, delta, axis, dx, dt, l, r);
}

void orc_irs_guess(const orc_params *p, double l_rho, const double *l_v,
                   double l_p, double l_cs, double r_rho, const double *r_v,
                   double r_p, double r_cs, int axis, double *guess6)
{
    gconst k = make_gconst(p->gamma);
    irs_guess(&k, l_rho, l_v, l_p, l_cs, r_rho, r_v, r_p, r_cs, axis, guess6);
}

double orc_irs_exec(const orc_params *p, double l_rho, const double *l_v,
                    double l_p, double l_cs, double *l_f, double r_rho,
                    const double *r_v, double r_p, double r_cs, double *r_f,
                    int axis, int32_t *n_newton)
{
    gconst k = make_gconst(p->gamma);
    return irs_exec(&k, p->n_iteration, p->tolerance, l_rho, l_v, l_p, l_cs,
                    l_f, r_rho, r_v, r_p, r_cs, r_f, axis, n_newton);
}

void orc_nonlinear_wave_function(const orc_params *p, double rho, double pk,
                                 double cs, double pstar, double *f,
                                 double *fprime)
{
    gconst k = make_gconst(p->gamma);
    wave_function(&k, rho, pk, cs, pstar, f, fprime);
}

void orc_rp_star(const orc_params *p, orc_rp_solution *s, int axis)
{
    gconst k = make_gconst(p->gamma);
    rp_star(&k, p, s, axis, NULL);
}

void orc_rp_sampling(const orc_params *p, const orc_rp_solution *s, int axis,
                     double dx, double dt, double *u)
{
    gconst k = make_gconst(p->gamma);
    rp_sampling(p->ndim, &k, s, axis, dx, dt, u);
}

void orc_rp_solve(const orc_params *p, const double *l, const double *r,
                  double dx, double dt, int axis, double *u)
{
    gconst k = make_gconst(p->gamma);
    rp_solve(p, &k, l, r, dx, dt, axis, u, NULL);
}

void orc_rp_solve_batch(const orc_params *p, int64_t n, const double *l,
                        const double *r, const int32_t *axis, double dt,
                        double *u)
{
    gconst k = make_gconst(p->gamma);
    const int nv = p->ndim + 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q)
        rp_solve(p, &k, l + q * nv, r + q * nv, 0.0, dt, axis[q], u + q * nv,
                 NULL);
}

void orc_limiter_batch(const orc_params *p, int64_t n, const double *ul,
                       const double *u, const double *ur, double *delta)
{
    const int nv = p->ndim + 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q)
        slope_limiter(p->ndim, p->limiter, p->omega, ul + q * nv, u + q * nv,
                      ur + q * nv, delta + q * nv);
}

void orc_half_step_batch(const orc_params *p, int64_t n, const double *u,
                         const double *delta, const int32_t *axis, double dx,
                         double dt, double *l, double *r)
{
    const int nv = p->ndim + 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q)
        half_step(p->ndim, p->gamma, u + q * nv, delta + q * nv, axis[q], dx,
                  dt, l + q * nv, r + q * nv);
}

void orc_flux_batch(const orc_params *p, int64_t n, const double *u,
                    const int32_t *axis, double *f)
{
    const int nv = p->ndim + 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q)
        flux(p->ndim, p->gamma, u + q * nv, axis[q], f + q * nv);
}

/* ------------------------------------------------------------------------ */
/* box geometry                                                              */
/* ------------------------------------------------------------------------ */

typedef struct geom {
    int nd;
    int n[3];       /* interior cells per axis                       */
    int g[3];       /* ghost layers per axis (2 active, 0 inactive)  */
    int64_t s[3];   /* strides                                       */
    int64_t plane;  /* elements per component                        */
} geom;

static geom make_geom(int nd, const int32_t *dims)
{
    geom G;
    G.nd = nd;
    int64_t st = 1;
    for (int d = 0; d < 3; ++d) {
        G.n[d] = d < nd ? dims[d] : 1;
        G.g[d] = d < nd ? 2 : 0;
        G.s[d] = st;
        st *= (int64_t)(G.n[d] + 2 * G.g[d]);
    }
    G.plane = st;
    return G;
}

/* storage offset of Fortran cell (i, j, k), 1-based interior indices */
static inline int64_t off(const geom *G, int i, int j, int k)
{
    return (int64_t)(i - 1 + G->g[0]) * G->s[0] +
           (int64_t)(j - 1 + G->g[1]) * G->s[1] +
           (int64_t)(k - 1 + G->g[2]) * G->s[2];
}

static inline void load(const geom *G, const double *cells, int64_t o,
                        double *u)
{
    for (int c = 0; c < G->nd + 2; ++c)
        u[c] = cells[o + c * G->plane];
}

/* ------------------------------------------------------------------------ */
/* CFL: rhyme_cfl/src/rhyme_cfl_time_step.f90:3-60                           */
/* ------------------------------------------------------------------------ */

double orc_cfl_time_step(const orc_params *p, const int32_t *dims,
                         const double *cells, const double *dx)
{
    geom G = make_geom(p->ndim, dims);
    const int nd = p->ndim;
    double max_u = 0.0;
    for (int k = 1; k <= G.n[2]; ++k)
        for (int j = 1; j <= G.n[1]; ++j)
            for (int i = 1; i <= G.n[0]; ++i) {
                double u[NV_MAX];
                load(&G, cells, off(&G, i, j, k), u);
                double s = 0.0;
                for (int d = 0; d < nd; ++d)
                    s += u[1 + d] * u[1 + d];
                /* calc_u :52-59 */
                double v = sqrt(s) / u[0] +
                           speed_of_sound(nd, p->gamma, p->kb_amu, u);
                if (v > max_u)
                    max_u = v;
            }
    double mdx = dx[0];
    for (int d = 1; d < nd; ++d)
        mdx = dx[d] < mdx ? dx[d] : mdx;
    return p->courant * mdx / max_u;
}

/* ------------------------------------------------------------------------ */
/* boundaries: rhyme_samr_bc/src                                             */
/* ------------------------------------------------------------------------ */

/* One face. face: 0 left, 1 right, 2 bottom, 3 top, 4 back, 5 front
 * (bcid / samrid, rhyme_samr/src/rhyme_samr.f90:23-31).  The face loops
 * span interior indices 1..n of the other axes only
 * (rhyme_samr_bc_set_left_boundary.f90:47-90 and the five siblings). */
void orc_bc_set_face(const orc_params *p, const int32_t *dims, double *cells,
                     int face)
{
    geom G = make_geom(p->ndim, dims);
    const int ax = face / 2;       /* 0 x, 1 y, 2 z */
    const int hi = face % 2;       /* 0 low side, 1 high side */
    const int n = G.n[ax];
    const int nv = p->ndim + 2;
    const int type = p->bc_types[face];
    if (ax >= p->ndim)
        return;
    int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
    if (o1 > o2) {
        int t = o1;
        o1 = o2;
        o2 = t;
    }
    for (int b = 1; b <= G.n[o2]; ++b)
        for (int a = 1; a <= G.n[o1]; ++a) {
            int idx[3];
            idx[o1] = a;
            idx[o2] = b;
            /* ghost g0 (0 or n+1) and g1 (-1 or n+2), and their sources */
            int g0 = hi ? n + 1 : 0, g1 = hi ? n + 2 : -1;
            int s0, s1;
            if (type == 3) { /* periodic */
                s0 = hi ? 1 : n;
                s1 = hi ? 2 : n - 1;
            } else { /* reflective / outflow: mirror */
                s0 = hi ? n : 1;
                s1 = hi ? n - 1 : 2;
            }
            idx[ax] = g0;
            int64_t og0 = off(&G, idx[0], idx[1], idx[2]);
            idx[ax] = g1;
            int64_t og1 = off(&G, idx[0], idx[1], idx[2]);
            if (type == 4) { /* inflow: all NCMP components */
                for (int c = 0; c < p->ncmp; ++c) {
                    cells[og0 + c * G.plane] = p->cons_inflows[face][c];
                    cells[og1 + c * G.plane] = p->cons_inflows[face][c];
                }
                continue;
            }
            if (type < 1 || type > 3)
                continue; /* unset: select case has no branch */
            idx[ax] = s0;
            int64_t os0 = off(&G, idx[0], idx[1], idx[2]);
            idx[ax] = s1;
            int64_t os1 = off(&G, idx[0], idx[1], idx[2]);
            for (int c = 0; c < nv; ++c) {
                cells[og0 + c * G.plane] = cells[os0 + c * G.plane];
                cells[og1 + c * G.plane] = cells[os1 + c * G.plane];
            }
            if (type == 1) { /* normal momentum negated */
                int c = 1 + ax;
                cells[og0 + c * G.plane] = -cells[os0 + c * G.plane];
                cells[og1 + c * G.plane] = -cells[os1 + c * G.plane];
            }
        }
}

/* rhyme_samr_bc_set_boundaries.f90:9-20: L, R, B, T, back, front */
void orc_bc_set_boundaries(const orc_params *p, const int32_t *dims,
                           double *cells)
{
    for (int f = 0; f < 2 * p->ndim; ++f)
        orc_bc_set_face(p, dims, cells, f);
}

/* rhyme_samr_bc_init.f90:61-87: every ghost cell (vars 1..NDIM+2) := 0 */
void orc_bc_init_ghosts(const orc_params *p, const int32_t *dims,
                        double *cells)
{
    geom G = make_geom(p->ndim, dims);
    for (int k = 1 - G.g[2]; k <= G.n[2] + G.g[2]; ++k)
        for (int j = 1 - G.g[1]; j <= G.n[1] + G.g[1]; ++j)
            for (int i = 1 - G.g[0]; i <= G.n[0] + G.g[0]; ++i) {
                int ghost = i < 1 || i > G.n[0] || j < 1 || j > G.n[1] ||
                            k < 1 || k > G.n[2];
                if (!ghost)
                    continue;
                int64_t o = off(&G, i, j, k);
                for (int c = 0; c < p->ndim + 2; ++c)
                    cells[o + c * G.plane] = 0.0;
            }
}

/* ------------------------------------------------------------------------ */
/* MUSCL-Hancock                                                             */
/* ------------------------------------------------------------------------ */

/* rhyme_muscl_hancock_solve_memory_intensive.f90:3-181.
 * ul/ur/fr are the workspace arrays ws%ul/ur/fr(i, j, k, var, dir) over the
 * full ghosted box (rhyme_mh_workspace_check.f90:36-75), stored here as
 * [dir][var][cell]. */
static int mh_memory_intensive(const orc_params *p, const geom *G,
                               double *cells, const double *dx, double dt,
                               double *fr_out)
{
    const int nd = p->ndim, nv = nd + 2;
    const gconst k = make_gconst(p->gamma);
    const int64_t P = G->plane;
    size_t sz = (size_t)P * nv * nd;
    double *ul = (double *)malloc(sz * sizeof(double));
    double *ur = (double *)malloc(sz * sizeof(double));
    double *fr = (double *)malloc(sz * sizeof(double));
    if (!ul || !ur || !fr) {
        free(ul);
        free(ur);
        free(fr);
        return -1;
    }
#define WS(arr, dir, c, o) arr[((int64_t)(dir) * nv + (c)) * P + (o)]
    int lb[3], ub[3];
    for (int d = 0; d < 3; ++d) {
        lb[d] = d < nd ? 0 : 1;
        ub[d] = d < nd ? G->n[d] + 1 : 1;
    }

    /* loop A :49-98 */
#pragma omp parallel for schedule(static) collapse(2)
    for (int kk = lb[2]; kk <= ub[2]; ++kk)
        for (int j = lb[1]; j <= ub[1]; ++j)
            for (int i = lb[0]; i <= ub[0]; ++i) {
                int64_t o = off(G, i, j, kk);
                double uc[NV_MAX], um[NV_MAX], up[NV_MAX], delta[NV_MAX];
                double l[NV_MAX], r[NV_MAX];
                load(G, cells, o, uc);
                for (int a = 0; a < nd; ++a) {
                    load(G, cells, o - G->s[a], um);
                    load(G, cells, o + G->s[a], up);
                    slope_limiter(nd, p->limiter, p->omega, um, uc, up, delta);
                    half_step(nd, p->gamma, uc, delta, a + 1, dx[a], dt, l, r);
                    for (int c = 0; c < nv; ++c) {
                        WS(ul, a, c, o) = l[c];
                        WS(ur, a, c, o) = r[c];
                    }
                }
            }

    /* loop B :100-140 */
    for (int d = 0; d < 3; ++d)
        ub[d] = d < nd ? G->n[d] : 1;
#pragma omp parallel for schedule(static) collapse(2)
    for (int kk = lb[2]; kk <= ub[2]; ++kk)
        for (int j = lb[1]; j <= ub[1]; ++j)
            for (int i = lb[0]; i <= ub[0]; ++i) {
                int64_t o = off(G, i, j, kk);
                for (int a = 0; a < nd; ++a) {
                    double L[NV_MAX], R[NV_MAX], ev[NV_MAX], f[NV_MAX];
                    for (int c = 0; c < nv; ++c) {
                        L[c] = WS(ur, a, c, o);
                        R[c] = WS(ul, a, c, o + G->s[a]);
                    }
                    rp_solve(p, &k, L, R, 0.0, dt, a + 1, ev, NULL);
                    flux(nd, p->gamma, ev, a + 1, f);
                    for (int c = 0; c < nv; ++c)
                        WS(fr, a, c, o) = f[c];
                }
            }

    /* loop C :142-176: u + ((A + B) + C) */
#pragma omp parallel for schedule(static) collapse(2)
    for (int kk = 1; kk <= G->n[2]; ++kk)
        for (int j = 1; j <= G->n[1]; ++j)
            for (int i = 1; i <= G->n[0]; ++i) {
                int64_t o = off(G, i, j, kk);
                for (int c = 0; c < nv; ++c) {
                    double acc = dt / dx[0] *
                                 (WS(fr, 0, c, o - G->s[0]) - WS(fr, 0, c, o));
                    if (nd > 1)
                        acc = acc + dt / dx[1] *
                                        (WS(fr, 1, c, o - G->s[1]) - WS(fr, 1, c, o));
                    if (nd > 2)
                        acc = acc + dt / dx[2] *
                                        (WS(fr, 2, c, o - G->s[2]) - WS(fr, 2, c, o));
                    cells[o + c * P] = cells[o + c * P] + acc;
                }
            }
#undef WS
    if (fr_out)
        memcpy(fr_out, fr, sz * sizeof(double));
    free(ul);
    free(ur);
    free(fr);
    return 0;
}

/* rhyme_muscl_hancock_solve_cpu_intensive.f90:3-159 */
static int mh_cpu_intensive(const orc_params *p, const geom *G, double *cells,
                            const double *dx, double dt)
{
    const int nd = p->ndim, nv = nd + 2;
    const gconst k = make_gconst(p->gamma);
    const int64_t P = G->plane;
    const int64_t nint = (int64_t)G->n[0] * G->n[1] * G->n[2];
    double *ws = (double *)malloc((size_t)nint * nv * sizeof(double));
    if (!ws)
        return -1;

#pragma omp parallel for schedule(static) collapse(2)
    for (int kk = 1; kk <= G->n[2]; ++kk)
        for (int j = 1; j <= G->n[1]; ++j)
            for (int i = 1; i <= G->n[0]; ++i) {
                int64_t o = off(G, i, j, kk);
                int64_t w = (int64_t)(i - 1) +
                            (int64_t)G->n[0] * ((j - 1) + (int64_t)G->n[1] * (kk - 1));
                double u[NV_MAX];
                double hl[3][3][NV_MAX], hr[3][3][NV_MAX]; /* [axis][idx+1][var] */
                load(G, cells, o, u);
                for (int a = 0; a < nd; ++a)
                    for (int idx = -1; idx <= 1; ++idx) {
                        double um[NV_MAX], uc[NV_MAX], up[NV_MAX], delta[NV_MAX];
                        int64_t oc = o + idx * G->s[a];
                        load(G, cells, oc - G->s[a], um);
                        load(G, cells, oc, uc);
                        load(G, cells, oc + G->s[a], up);
                        slope_limiter(nd, p->limiter, p->omega, um, uc, up, delta);
                        half_step(nd, p->gamma, uc, delta, a + 1, dx[a], dt,
                                  hl[a][idx + 1], hr[a][idx + 1]);
                    }
                for (int a = 0; a < nd; ++a) {
                    double fl[2][NV_MAX];
                    for (int idx = -1; idx <= 0; ++idx) {
                        double ev[NV_MAX];
                        rp_solve(p, &k, hr[a][idx + 1], hl[a][idx + 2], 0.0, dt,
                                 a + 1, ev, NULL);
                        flux(nd, p->gamma, ev, a + 1, fl[idx + 1]);
                    }
                    for (int c = 0; c < nv; ++c) {
                        double df = fl[0][c] - fl[1][c];
                        u[c] = u[c] + dt / dx[a] * df;
                    }
                }
                for (int c = 0; c < nv; ++c)
                    ws[w + c * nint] = u[c];
            }

    /* copy back :154-157 */
#pragma omp parallel for schedule(static) collapse(2)
    for (int kk = 1; kk <= G->n[2]; ++kk)
        for (int j = 1; j <= G->n[1]; ++j)
            for (int i = 1; i <= G->n[0]; ++i) {
                int64_t o = off(G, i, j, kk);
                int64_t w = (int64_t)(i - 1) +
                            (int64_t)G->n[0] * ((j - 1) + (int64_t)G->n[1] * (kk - 1));
                for (int c = 0; c < nv; ++c)
                    cells[o + c * P] = ws[w + c * nint];
            }
    free(ws);
    return 0;
}

/* rhyme_muscl_hancock_solve.f90:17-26 */
int orc_mh_solve(const orc_params *p, const int32_t *dims, double *cells,
                 const double *dx, double dt)
{
    geom G = make_geom(p->ndim, dims);
    if (p->solver_type == 1)
        return mh_cpu_intensive(p, &G, cells, dx, dt);
    if (p->solver_type == 2)
        return mh_memory_intensive(p, &G, cells, dx, dt, NULL);
    return -2; /* logger%err('Unknown solver type') */
}

/* orc_mh_solve that also returns the intercell fluxes of the step in
 * fr_out[dir][var][cell] (ndim * nvar planes of the ghosted box, the face
 * between cells o and o + s_dir at the low cell o): the workspace array
 * ws%fr of the memory_intensive solve (…_memory_intensive.f90:110-116).
 * Both solver types solve the same interfaces, so for cpu_intensive the
 * fluxes come from a memory_intensive solve of a copy of the state. */
int orc_mh_solve_fluxes(const orc_params *p, const int32_t *dims, double *cells,
                        const double *dx, double dt, double *fr_out)
{
    geom G = make_geom(p->ndim, dims);
    if (p->solver_type == 2)
        return mh_memory_intensive(p, &G, cells, dx, dt, fr_out);
    if (p->solver_type != 1)
        return -2;
    const size_t n = (size_t)G.plane * (size_t)(p->ndim + 2);
    double *copy = (double *)malloc(n * sizeof(double));
    if (!copy)
        return -1;
    memcpy(copy, cells, n * sizeof(double));
    int rc = mh_memory_intensive(p, &G, copy, dx, dt, fr_out);
    free(copy);
    return rc ? rc : mh_cpu_intensive(p, &G, cells, dx, dt);
}

int orc_run(const orc_params *p, const int32_t *dims, double *cells,
            const double *dx, int32_t nsteps, double *dts)
{
    for (int32_t s = 0; s < nsteps; ++s) {
        double dt = orc_cfl_time_step(p, dims, cells, dx);
        if (dts)
            dts[s] = dt;
        orc_bc_set_boundaries(p, dims, cells);
        int rc = orc_mh_solve(p, dims, cells, dx, dt);
        if (rc)
            return rc;
    }
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
