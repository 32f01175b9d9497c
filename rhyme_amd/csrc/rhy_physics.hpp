// rhy_physics.hpp -- device-side ideal-gas hydro physics of the hot path:
// EOS and fluxes, slope limiters, half-step extrapolation and the exact
// iterative Riemann solver, all evaluated per lane in registers.
//
// Arithmetic order is part of the contract: the file must be compiled with
// -ffp-contract=off, and each expression keeps the parse of the reference
// Fortran routine it restates (cited file:line, relative to the reference's
// src/), so that FP64 results agree with the reference CPU path to rounding
// of pow() only.  Division and sqrt of doubles are IEEE correctly rounded on
// gfx950 under the default (non fast-math) lowering.
//
// Template parameter ND is the box dimensionality (1, 2, 3); a state vector
// holds NV = ND + 2 conserved values: rho, rho*v[0..ND), e_tot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhy {

constexpr double kTiny = 2.2250738585072014e-308;  // tiny(0d0)

// Problem constants that the reference keeps in module-private globals
// (rhyme_riemann_problem.f90:11-16, rhyme_irs.f90:9-14,
// rhyme_thermo_base.f90:25-26) and derived-type defaults.
struct Consts {
  double g, gm1, gp1, gm1_gp1, gm1_2g, gp1_2g, g_inv;  // gamma and derived
  double kb_amu;
  double omega;          // sl%w
  double tol;            // irs%tolerance
  double w_rho, w_p;     // rp%w_vacuum(rho), rp%w_vacuum(p)
  double e_rho;          // nearest(w_rho, +1), computed on the host
  int32_t limiter;       // slid
  int32_t n_iter;        // irs%n_iteration
  int32_t defer_every;   // > 0: test hook, defer every such face (by offset)
};

// tiny() of the arithmetic type (tiny(0d0) / tiny(0e0))
template <typename R>
__device__ __forceinline__ R tiny_of();
template <>
__device__ __forceinline__ double tiny_of<double>() { return kTiny; }
template <>
__device__ __forceinline__ float tiny_of<float>() { return 1.17549435e-38f; }

// a / b and sqrt in the bulk type: IEEE for double (the FP64 path, whose
// results match the oracle bit for bit), hardware reciprocal times a and
// hardware square root for float (the mixed form; <= 1.5 ulp).
template <typename R>
__device__ __forceinline__ R rdiv(R a, R b);
template <>
__device__ __forceinline__ double rdiv<double>(double a, double b) { return a / b; }
template <>
__device__ __forceinline__ float rdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

// FP64 quotients that share one divisor b, each equal to the IEEE a / b bit
// for bit.  gfx950 lowers a / b to: div_scale of b and of a, rcp of the
// scaled b, two Newton steps on that reciprocal (fma(-b, y, 1), fma(y, e,
// y)), q = a * y, r = fma(-b, q, a), div_fmas(r, y, q), div_fixup(q, b, a).
// div_scale returns its operand unchanged, and div_fmas is then a plain fma,
// unless |b| is denormal or near overflow, a / b is denormal, a is below
// 2^-969 or the exponents of a and b differ by 768 or more.  With
// 2^-64 <= |b| <= 2^64 and a zero or 2^-601 <= |a| < 2^399 none of that
// applies, so the reciprocal is formed once here and each quotient is
// finished with the same mul, fma, fma and div_fixup (a zero a: div_fixup
// returns the signed zero).  `ok` collects the range checks; a caller redoes
// its quotients with `/` on the lanes where one failed.
struct SharedDiv {
  double b, y;
  __device__ __forceinline__ explicit SharedDiv(double d) : b(d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    y = __builtin_fma(r, e, r);
  }
  __device__ __forceinline__ bool divisor_ok() const {
    return fabs(b) >= 0x1p-64 && fabs(b) <= 0x1p64;
  }
  // a / b; clears ok unless a is zero or frexp(a)'s exponent is in [-600, 400)
  __device__ __forceinline__ double operator()(double a, bool& ok) const {
    ok = ok && static_cast<unsigned>(__builtin_amdgcn_frexp_exp(a) + 600) < 1000u;
    double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    q = __builtin_fma(r, y, q);
    return __builtin_amdgcn_div_fixup(q, b, a);
  }
};

template <typename R>
__device__ __forceinline__ R root(R x);
template <>
__device__ __forceinline__ double root<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float root<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// True on every active lane of the wave when p holds on all of them (a
// wave-uniform value: a branch on it skips its body for the whole wave).
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

// ---------------------------------------------------------------- EOS ----

// rhyme_hydro_base_specific_kinetic_energy.f90:9.  For a state at rest the
// numerator is +0 and, for rho*rho >= the smallest normal number of R (the
// FP32 quotient multiplies by a reciprocal that must stay finite), the
// quotient is +0: a wave whose lanes all hold such states skips the
// division, and every other wave divides on every lane, as before.
template <int ND, typename R = double>
__device__ __forceinline__ R sp_kin_e(const R* u) {
  R s = R(0);
#pragma unroll
  for (int d = 0; d < ND; ++d) s += u[1 + d] * u[1 + d];
  const R num = R(.5) * s, den = u[0] * u[0];
  if (wave_all(num == R(0) && den >= tiny_of<R>())) return R(0);
  return rdiv<R>(num, den);
}

// rhyme_hydro_base_specific_internal_energy.f90:9
template <int ND, typename R = double>
__device__ __forceinline__ R sp_int_e(const R* u) {
  return rdiv<R>(u[ND + 1], u[0]) - sp_kin_e<ND, R>(u);
}

// rhyme_ideal_gas_pressure.f90:9.  FP64 keeps the reference's
// rho * (e / rho - kin) order.  The mixed form (FP32) computes the same
// quantity as (E - |m|^2 / (2 rho)) * (gamma - 1): a gas at rest with
// uniform E then has bitwise uniform p whatever its density, so such
// interfaces take the exact contact solution in the Riemann solve.
template <int ND, typename R = double>
__device__ __forceinline__ R pressure(const Consts& k, const R* u) {
  if constexpr (sizeof(R) == sizeof(float)) {
    R s = R(0);
#pragma unroll
    for (int d = 0; d < ND; ++d) s += u[1 + d] * u[1 + d];
    return (u[ND + 1] - rdiv<R>(R(.5) * s, u[0])) * R(k.g - 1.0);
  } else {
    return u[0] * sp_int_e<ND, R>(u) * R(k.g - 1.0);
  }
}

// rhyme_ideal_gas_speed_of_sound.f90:9 with temperature_per_mu.f90:9-10
template <int ND, typename R = double>
__device__ __forceinline__ R speed_of_sound(const Consts& k, const R* u) {
  R t_mu = rdiv<R>(sp_int_e<ND, R>(u) * R(k.g - 1.0), R(k.kb_amu));
  return root<R>(R(k.g) * R(k.kb_amu) * t_mu);
}

// rhyme_ideal_gas_flux.f90:12-23 for rho >= tiny; A is the 0-based axis.
template <int ND, int A, typename R = double>
__device__ __forceinline__ void flux_of(const Consts& k, const R* u, R* f) {
  R p = pressure<ND, R>(k, u);
  f[0] = u[1 + A];
#pragma unroll
  for (int d = 0; d < ND; ++d) f[1 + d] = rdiv<R>(u[1 + d] * f[0], u[0]);
  f[1 + A] = f[1 + A] + p;
  f[ND + 1] = rdiv<R>(f[0], u[0]) * (u[ND + 1] + p);
}

// rhyme_ideal_gas_flux.f90:12-23; A is the 0-based axis.  In FP64 the five
// divisions by rho (e / rho in the pressure, m * F_rho / rho, F_rho / rho)
// share one reciprocal (SharedDiv), with the same results as flux_of.
template <int ND, int A, typename R = double>
__device__ __forceinline__ void flux(const Consts& k, const R* u, R* f) {
  if (fabs(u[0]) < tiny_of<R>()) {
#pragma unroll
    for (int c = 0; c < ND + 2; ++c) f[c] = R(0);
    return;
  }
  if constexpr (sizeof(R) == sizeof(double)) {
    const SharedDiv by_rho(u[0]);
    bool ok = by_rho.divisor_ok();
    const double p = u[0] * (by_rho(u[ND + 1], ok) - sp_kin_e<ND, R>(u)) * (k.g - 1.0);
    f[0] = u[1 + A];
#pragma unroll
    for (int d = 0; d < ND; ++d) f[1 + d] = by_rho(u[1 + d] * f[0], ok);
    f[1 + A] = f[1 + A] + p;
    f[ND + 1] = by_rho(f[0], ok) * (u[ND + 1] + p);
    if (!ok) flux_of<ND, A, R>(k, u, f);
  } else {
    flux_of<ND, A, R>(k, u, f);
  }
}

// rhyme_ideal_gas_primitive_vars_to_conserved.f90:9-11
template <int ND, typename R = double>
__device__ __forceinline__ void prim_to_cons(const Consts& k, R rho, const R* v, R p, R* u) {
  R s = R(0);
  u[0] = rho;
#pragma unroll
  for (int d = 0; d < ND; ++d) u[1 + d] = rho * v[d];
#pragma unroll
  for (int d = 0; d < ND; ++d) s += v[d] * v[d];
  u[ND + 1] = R(.5) * rho * s + rdiv<R>(p, R(k.g - 1.0));
}

// ------------------------------------------------------- slope limiter ----

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

// rhyme_slope_limiter_run.f90:10-19 and the four limiters
// (…_van_leer.f90:13-27, …_minmod.f90:13-29, …_van_albada.f90:13-27,
// …_superbee.f90:13-31), r from rhyme_slope_limiter.f90:62-82 and
// xi_R from …_xi_r.f90:11.  LIM selects the limiter at compile time
// (1..4, slid values); LIM = 0 dispatches on `type` at run time.
template <int LIM, typename R = double>
__device__ __forceinline__ R limit_one(int type, R w, R ul, R u, R ur) {
  const int t = LIM ? LIM : type;
  R dL = u - ul;
  R dR = ur - u;
  R d = R(.5) * (R(1) + w) * dL + R(.5) * (R(1) + w) * dR;
  R denom = ur - u;
  // r = 0 for |denom| < tiny, and every limiter gives 0 for r <= 0: return
  // before the division (a wave whose lanes all see a flat variable, such
  // as the momenta of a medium at rest, skips it)
  if (fabs(denom) < tiny_of<R>()) return R(0);
  R r = rdiv<R>(u - ul, denom);
  if (!(r > R(0))) return R(0);
  R xi = rdiv<R>(R(2), R(1) - w + (R(1) + w) * r);
  switch (t) {
    case 1:  // van Leer
      return rmin(rdiv<R>(R(2) * r, R(1) + r), xi) * d;
    case 2:  // minmod
      return r > R(1) ? rmin(R(1), xi) * d : r * d;
    case 3:  // van Albada
      return rmin(rdiv<R>(r * (R(1) + r), R(1) + r * r), xi) * d;
    case 4:  // superbee
      if (r > R(1)) return rmin(rmin(r, xi), R(2)) * d;
      if (r > R(.5)) return d;
      return R(2) * r * d;
    default:
      return R(0);
  }
}

// rhyme_muscl_hancock_half_step_extrapolation.f90:15-21: slope from the
// three cells um, u, up along axis A, then the half-step states l and r.
template <int ND, int A, int LIM = 0, typename R = double>
__device__ __forceinline__ void half_step(const Consts& k, const R* um, const R* u, const R* up,
                                          R hdtdx, R* l, R* r) {
  constexpr int NV = ND + 2;
  R fl[NV], fr[NV];
  const R w = R(k.omega);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    R delta = limit_one<LIM, R>(k.limiter, w, um[c], u[c], up[c]);
    l[c] = u[c] - R(.5) * delta;
    r[c] = u[c] + R(.5) * delta;
  }
  flux<ND, A, R>(k, l, fl);
  flux<ND, A, R>(k, r, fr);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    R df = fl[c] - fr[c];
    l[c] = l[c] + hdtdx * df;  // hdtdx = (.5d0*dt)/dx, evaluated once
    r[c] = r[c] + hdtdx * df;
  }
}

// -------------------------------------------------- Riemann (IRS) ----

// Arithmetic policies of the FP64 Riemann solve: pw(x, y) = x**y, pw2 gives
// x**a and x**b of one base, dv(a, b) = a / b, sq(x) = sqrt(x); kFast lets
// the wave function use algebraically equal forms with fewer divisions and
// powers (the reference's forms otherwise).
//
// PowLib: x**y as the reference evaluates it (libm pow; ocml on the device),
// IEEE division and sqrt.
struct PowLib {
  static constexpr bool kFast = false;
  static constexpr bool kDefer = false;
  __device__ __forceinline__ static double dv(double a, double b) { return a / b; }
  __device__ __forceinline__ static double sq(double x) { return sqrt(x); }
  __device__ __forceinline__ static double pw(double x, double y) { return pow(x, y); }
  __device__ __forceinline__ static void pw2(double x, double a, double b, double& xa,
                                             double& xb) {
    xa = pow(x, a);
    xb = pow(x, b);
  }
};

// PowFast: IEEE division and sqrt; x**y = exp(y log x) in plain FP64 for
// x > 0 (relative error a few 1e-16, not correctly rounded): log via
// 2 atanh((m-1)/(m+1)) on m in [1/sqrt2, sqrt2), exp via a Cody-Waite
// reduction and a degree-13 Taylor polynomial; other x, and |y log x| >= 700,
// go to pow().  pw2 evaluates the log once for both exponents, with the same
// results as two pw calls.
struct PowFast {
  static constexpr bool kFast = false;
  static constexpr bool kDefer = false;
  __device__ __forceinline__ static double dv(double a, double b) { return a / b; }
  __device__ __forceinline__ static double sq(double x) { return sqrt(x); }
  static constexpr double LN2_HI = 6.93147180369123816490e-01;
  static constexpr double LN2_LO = 1.90821492927058770002e-10;

  // ln x, as e*LN2_HI + (atanh series + e*LN2_LO), for 0 < x < inf
  __device__ __forceinline__ static double ln(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
      m = m * 2.0;
      e -= 1;
    }
    // s = (m - 1) / (m + 1) with a refined hardware reciprocal (m + 1 lies
    // in [1.7, 2.5]; two Newton steps reach full precision)
    const double d = m + 1.0;
    double rd = __builtin_amdgcn_rcp(d);
    rd = fma(fma(-d, rd, 1.0), rd, rd);
    rd = fma(fma(-d, rd, 1.0), rd, rd);
    const double s = (m - 1.0) * rd;
    const double z = s * s;
    double q = 2.0 / 23.0;
    q = fma(q, z, 2.0 / 21.0);
    q = fma(q, z, 2.0 / 19.0);
    q = fma(q, z, 2.0 / 17.0);
    q = fma(q, z, 2.0 / 15.0);
    q = fma(q, z, 2.0 / 13.0);
    q = fma(q, z, 2.0 / 11.0);
    q = fma(q, z, 2.0 / 9.0);
    q = fma(q, z, 2.0 / 7.0);
    q = fma(q, z, 2.0 / 5.0);
    q = fma(q, z, 2.0 / 3.0);
    const double lm = fma(s * z, q, 2.0 * s);
    const double de = (double)e;
    return de * LN2_HI + (lm + de * LN2_LO);
  }

  // e**t for |t| < 700
  __device__ __forceinline__ static double ex(double t) {
    const double kd = rint(t * 1.44269504088896338700e+00);
    const double r = (t - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r * r, r);
    return ldexp(1.0 + p, (int)kd);
  }

  __device__ __forceinline__ static double pw(double x, double y) {
    if (!(x > 0.0) || !(x < INFINITY)) return pow(x, y);
    const double t = y * ln(x);
    if (!(fabs(t) < 700.0)) return pow(x, y);
    return ex(t);
  }

  __device__ __forceinline__ static void pw2(double x, double a, double b, double& xa,
                                             double& xb) {
    if (wave_all(x == 1.0)) {  // ln 1 = 0 and e**(+-0) = 1 exactly below
      xa = xb = 1.0;
      return;
    }
    if (!(x > 0.0) || !(x < INFINITY)) {
      xa = pow(x, a);
      xb = pow(x, b);
      return;
    }
    const double l = ln(x);
    const double ta = a * l, tb = b * l;
    xa = fabs(ta) < 700.0 ? ex(ta) : pow(x, a);
    xb = fabs(tb) < 700.0 ? ex(tb) : pow(x, b);
  }
};

// MixedMath (the mixed form): FP64 throughout, with PowFast exponentials;
// a / b as a times the hardware reciprocal of b refined by one Newton step
// (within ~1e-15 of the IEEE quotient), sqrt from the hardware reciprocal
// square root refined by a Newton and a Goldschmidt step (within a few ulp);
// arguments outside [2^-1000, 2^1000], zero and non-finite ones included,
// take the IEEE operation.
struct MixedMath : PowFast {
  static constexpr bool kFast = true;
  __device__ __forceinline__ static double dv(double a, double b) {
    const double ab = fabs(b);
    if (!(ab > 0x1p-1000 && ab < 0x1p+1000)) return a / b;
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    return a * r;
  }
  __device__ __forceinline__ static double sq(double x) {
    if (!(x > 0x1p-1000 && x < 0x1p+1000)) return sqrt(x);
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    return fma(fma(-g, g, x), h, g);
  }
};

// Exponentials of the FP64 path: PowFast unless RHY_FP64_POW_FAST=0 selects
// libm pow (last-bit agreement with a libm host, ~20% slower flux kernels).
#ifndef RHY_FP64_POW_FAST
#define RHY_FP64_POW_FAST 1
#endif
#if RHY_FP64_POW_FAST
using PowF64 = PowFast;
#else
using PowF64 = PowLib;
#endif

// Deferred fallbacks.  The slow paths of the Riemann solve -- libm pow for
// x <= 0, non-finite x or |y log x| >= 700, and the Newton restarts from
// guesses 2-6 -- are rare, but compiled into a kernel they hold ~90 VGPRs
// and cut it from 3 to 2 waves per SIMD.  The flux kernels therefore use
// Deferring<P>: where P would take a slow path it raises the lane's defer
// flag and returns a placeholder, and the kernel lists the face; a fixup
// kernel then recomputes every listed face with P itself.  A face's flux
// depends only on the cells, so the fixed-up flux is the one P gives, bit
// for bit, and on every other face Deferring<P> computes what P computes.
//
// The flag is one LDS word per lane (a lane only touches its own word), so
// policies need no state; kernels using it have at most 256 threads per
// workgroup and clear the flag when they start.
__device__ __forceinline__ uint32_t& defer_flag() {
  __shared__ uint32_t flags[256];
  return flags[threadIdx.x];
}
__device__ __forceinline__ void defer_mark() { defer_flag() = 1u; }

template <class P>
struct Deferring : P {
  static constexpr bool kDefer = true;
  __device__ __forceinline__ static double pw(double x, double y) {
    if (!(x > 0.0) || !(x < INFINITY)) {
      defer_mark();
      return 1.0;
    }
    const double t = y * P::ln(x);
    if (!(fabs(t) < 700.0)) {
      defer_mark();
      return 1.0;
    }
    return P::ex(t);
  }
  __device__ __forceinline__ static void pw2(double x, double a, double b, double& xa,
                                             double& xb) {
    xa = xb = 1.0;
    if (wave_all(x == 1.0)) return;  // P's result at x = 1, as in P::pw2
    if (!(x > 0.0) || !(x < INFINITY)) {
      defer_mark();
      return;
    }
    const double l = P::ln(x);
    const double ta = a * l, tb = b * l;
    if (!(fabs(ta) < 700.0) || !(fabs(tb) < 700.0)) {
      defer_mark();
      return;
    }
    xa = P::ex(ta);
    xb = P::ex(tb);
  }
};

// Policies of the flux kernels and of their fixup kernels.
// Build option RHY_FP64_FAST_RIEMANN=1: the FP64 flux kernels solve their
// Riemann problems with the mixed form's FP64 policy (refined reciprocal
// division and square root, the algebraic wave-function forms): about 4%
// faster at 256^3, trajectories ~1e-14 from the oracle instead of ~1e-15.
#ifndef RHY_FP64_FAST_RIEMANN
#define RHY_FP64_FAST_RIEMANN 0
#endif
// Build option RHY_CONTACT_EXACT (default 1): interfaces with equal
// pressures and normal velocities take the exact contact star state
// (irs_exec) instead of the Newton iteration.
#ifndef RHY_CONTACT_EXACT
#define RHY_CONTACT_EXACT 1
#endif
#if RHY_FP64_FAST_RIEMANN
using PowF64Flux = Deferring<MixedMath>;
using PowF64Fix = MixedMath;
#elif RHY_FP64_POW_FAST
using PowF64Flux = Deferring<PowFast>;
using PowF64Fix = PowFast;
#else
using PowF64Flux = PowLib;
using PowF64Fix = PowLib;
#endif
using MixedFlux = Deferring<MixedMath>;

// rhyme_irs_nonlinear_wave_function.f90:13-23.  The terms that do not
// depend on p (A_K, B_K of the shock branch, 2 c_K / (gamma-1) and
// 1 / (rho_K c_K) of the rarefaction branch) are evaluated once per
// interface side, with the reference's expressions, instead of in every
// Newton iteration.
struct WaveSide {
  double pk, Ak, Bk, c2, irc;
};

template <class PW = PowLib>
__device__ __forceinline__ WaveSide wave_side(const Consts& k, double rho, double pk,
                                              double cs) {
  WaveSide w;
  w.pk = pk;
  w.Ak = PW::dv(2.0, k.gp1 * rho);
  w.Bk = PW::dv(k.gm1 * pk, k.gp1);
  w.c2 = PW::dv(2.0 * cs, k.gm1);
  w.irc = PW::dv(1.0, rho * cs);
  return w;
}

template <class PW = PowLib>
__device__ __forceinline__ void wave_fn(const Consts& k, const WaveSide& w, double p, double& f,
                                        double& fp) {
  if constexpr (PW::kFast) {
    // one reciprocal of B_K + p for both shock terms, and
    // (p/p_K)**(-(gamma+1)/2gamma) = (p/p_K)**((gamma-1)/2gamma) / (p/p_K)
    if (p > w.pk) {
      const double ib = PW::dv(1.0, w.Bk + p);
      const double factor = PW::sq(w.Ak * ib);
      f = (p - w.pk) * factor;
      fp = factor * (1.0 - 0.5 * (p - w.pk) * ib);
    } else {
      const double x = PW::dv(p, w.pk);
      const double xa = PW::pw(x, k.gm1_2g);
      f = w.c2 * (xa - 1.0);
      fp = w.irc * PW::dv(xa, x);
    }
  } else {
    if (p > w.pk) {
      double factor = sqrt(PW::dv(w.Ak, w.Bk + p));
      f = (p - w.pk) * factor;
      fp = factor * (1.0 - PW::dv(p - w.pk, 2.0 * (w.Bk + p)));
    } else {
      double xa, xb;
      PW::pw2(PW::dv(p, w.pk), k.gm1_2g, -k.gp1_2g, xa, xb);
      f = w.c2 * (xa - 1.0);
      fp = w.irc * xb;
    }
  }
}

// rhyme_irs_guess_p_star.f90 guess 1 (:14-39): PVRS / two-rarefaction /
// two-shock selection.
template <class PW = PowLib>
__device__ __forceinline__ double guess1(const Consts& k, double lr, double lv, double lp,
                                         double lc, double rr, double rv, double rp, double rc) {
  double cup = 0.25 * (lr + rr) * (lc + rc);
  double ppv = dmax(0.5 * (lp + rp) + 0.5 * (lv - rv) * cup, 0.0);
  double p_min = dmin(lp, rp);
  double p_max = dmax(lp, rp);
  double q_max = PW::dv(p_max, p_min);
  double g;
  if (q_max <= 2.0 && (p_min < ppv && ppv < p_max)) {
    g = ppv;
  } else if (ppv < p_min) {
    double pq = PW::pw(PW::dv(lp, rp), k.gm1_2g);
    double um = PW::dv(PW::dv(pq * lv, lc) + PW::dv(rv, rc) + PW::dv(2.0, k.gm1) * (pq - 1.0),
                       PW::dv(pq, lc) + PW::dv(1.0, rc));
    double ptl = 1.0 + PW::dv(PW::dv(k.gm1, 2.0) * (lv - um), lc);
    double ptr = 1.0 + PW::dv(PW::dv(k.gm1, 2.0) * (um - rv), rc);
    g = 0.5 * (lp * PW::pw(ptl, PW::dv(1.0, k.gm1_2g)) + rp * PW::pw(ptr, PW::dv(1.0, k.gm1_2g)));
  } else {
    double gel = PW::sq(PW::dv(PW::dv(PW::dv(2.0, k.gp1), lr), k.gm1_gp1 * lp + ppv));
    double ger = PW::sq(PW::dv(PW::dv(PW::dv(2.0, k.gp1), rr), k.gm1_gp1 * rp + ppv));
    g = PW::dv(gel * lp + ger * rp - (rv - lv), gel + ger);
  }
  return dmax(g, kTiny);
}

// Guesses 2..6 (rhyme_irs_guess_p_star.f90:42-71).  The reference always
// computes all six; they are only read when the Newton iteration from the
// previous guess ends at p* <= 0 (rhyme_irs_exec.f90:24-42), so computing
// guess i on demand yields the same values.
template <class PW = PowLib>
__device__ __noinline__ double guess_n(const Consts& k, int i, double lr, double lv, double lp,
                                       double lc, double rr, double rv, double rp, double rc) {
  double g;
  switch (i) {
    case 1:
      g = (rr * rc * lp + lr * lc * rp + rc * lc * (lr * lv - rr * rv)) / (rr * rc + lr * lc);
      break;
    case 2:
      g = PW::pw((lc + rc - .5 * k.gm1 * (rv - lv)) /
                     (PW::pw(lc / lp, k.gm1_2g) + PW::pw(rc / rp, k.gm1_2g)),
                 1.0 / k.gm1_2g);
      break;
    case 3: {
      double lg = sqrt((2.0 / (k.gp1 * lr)) / (lp + k.gm1_gp1 * lp));
      double rg = sqrt((2.0 / (k.gp1 * rr)) / (rp + k.gm1_gp1 * rp));
      g = (lg * lp + rg * rp - (rv - lv)) / (lg + rg);
      break;
    }
    case 4:
      g = .5 * (lp + rp) - .125 * (rv - lv) * (lr + rr) * (lc + rc);
      break;
    default:
      g = .5 * (lp + rp);
      break;
  }
  return dmax(g, kTiny);
}

// Newton iteration from one starting guess (rhyme_irs_exec.f90:25-39).
// Leaves lf/rf at the wave-function values of the last evaluated iterate.
template <class PW = PowLib>
__device__ __forceinline__ double newton(const Consts& k, double p_star, const WaveSide& L,
                                         const WaveSide& R, double du, double& lf, double& rf) {
  double p_prev = p_star;
  for (int i = 1; i <= k.n_iter; ++i) {
    double lfp, rfp;
    wave_fn<PW>(k, L, p_star, lf, lfp);
    wave_fn<PW>(k, R, p_star, rf, rfp);
    p_star = p_prev - PW::dv(lf + rf + du, lfp + rfp);
    if (p_star < 0.0) break;
    if (PW::dv(2.0 * fabs(p_star - p_prev), p_star + p_prev) < k.tol) break;
    p_prev = p_star;
  }
  return p_star;
}

// rhyme_irs_exec.f90:3-45
template <class PW = PowLib>
__device__ __forceinline__ double irs_exec(const Consts& k, double lr, double lv, double lp,
                                           double lc, double rr, double rv, double rp, double rc,
                                           double& lf, double& rf) {
  // Equal pressures and normal velocities: the interface carries at most a
  // contact, whose exact star state is p* = p and u* = u (f_L(p) = f_R(p) = 0).
  // The reference's Newton iteration reaches the same state to within a few
  // ulps; taking it exactly skips the guesses and the iteration on every
  // face of a region at uniform pressure and velocity.  A build with
  // -DRHY_CONTACT_EXACT=0 iterates on these faces as well.
#if RHY_CONTACT_EXACT
  if (lp == rp && lv == rv) {
    lf = 0.0;
    rf = 0.0;
    return lp;
  }
#endif
  const double du = rv - lv;
  const WaveSide L = wave_side<PW>(k, lr, lp, lc);
  const WaveSide R = wave_side<PW>(k, rr, rp, rc);
  double p = newton<PW>(k, guess1<PW>(k, lr, lv, lp, lc, rr, rv, rp, rc), L, R, du, lf, rf);
  if (!(p > 0.0)) {
    if constexpr (PW::kDefer) {
      defer_mark();  // guesses 2-6: the fixup kernel's work
    } else {
      for (int gi = 1; gi < 6; ++gi) {
        p = newton<PW>(k, guess_n<PW>(k, gi, lr, lv, lp, lc, rr, rv, rp, rc), L, R, du, lf, rf);
        if (p > 0.0) break;
      }
    }
  }
  if (p < 0.0) p = 0.0;
  return p;
}

// rhyme_riemann_problem_w_kfan.f90:14-31 (is_right selects the sign of
// cs): the primitive state inside a rarefaction fan, velocity in v.
template <int ND, int A, class PW = PowLib>
__device__ __forceinline__ void w_kfan_prim(const Consts& k, double srho, const double* sv,
                                            double sp, double scs, double dxdt, bool is_right,
                                            double& rho, double* v, double& p) {
  double cs = is_right ? -scs : scs;
#pragma unroll
  for (int d = 0; d < ND; ++d) v[d] = sv[d];
  double xa, xb;
  PW::pw2(PW::dv(2.0, k.gp1) + PW::dv(k.gm1_gp1, cs) * (sv[A] - dxdt), PW::dv(2.0, k.gm1),
          PW::dv(1.0, k.gm1_2g), xa, xb);
  rho = srho * xa;
  v[A] = PW::dv(2.0, k.gp1) * (cs + PW::dv(k.gm1, 2.0) * sv[A] + dxdt);
  p = sp * xb;
}

// Primitive variables of one side of an interface (…_solve.f90:70-75):
// v = m / rho, p = max(calc_p, w_vacuum(p)), c = max(calc_cs, tiny), formed
// in the type of the half states (FP32 in the mixed form) and widened.
template <int ND, typename RI>
__device__ __forceinline__ void side_prims(const Consts& k, const RI* s, double* v, double& p,
                                           double& c) {
  if constexpr (sizeof(RI) == sizeof(double)) {
    // the divisions by rho (v and e / rho) share one reciprocal (SharedDiv)
    const SharedDiv by_rho(s[0]);
    bool ok = by_rho.divisor_ok();
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = by_rho(s[1 + d], ok);
    const double ie = by_rho(s[ND + 1], ok) - sp_kin_e<ND, double>(s);
    p = dmax(s[0] * ie * (k.g - 1.0), k.w_p);
    c = dmax(sqrt(k.g * k.kb_amu * ((ie * (k.g - 1.0)) / k.kb_amu)), kTiny);
    if (ok) return;
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) v[d] = (double)rdiv<RI>(s[1 + d], s[0]);
  p = dmax((double)pressure<ND, RI>(k, s), k.w_p);
  c = dmax((double)speed_of_sound<ND, RI>(k, s), kTiny);
}

// rhyme_riemann_problem_solve.f90:3-104: the primitive state (rho, v, p)
// sampled at x/t = dxdt (dx / dt of the reference's arguments, :17).  The
// MUSCL-Hancock loops call it with dx = 0, i.e. dxdt = 0, the Godunov state
// on the interface (…_solve_memory_intensive.f90:110-114); the flux kernels
// pass the constant 0, which folds away every comparison's other operand.
// l is the right half-step state of the cell on the low side of the
// interface, r the left half-step state of the cell on the high side.  Every
// branch of the reference ends in conv_prim_vars_to_cons of the state chosen
// here.
template <int ND, int A, class PW, typename RI>
__device__ __forceinline__ void riemann_sample(const Consts& k, const RI* l, const RI* r,
                                               double dxdt, double& srho, double* sv,
                                               double& sp) {
  const double l_rho = dmax((double)l[0], k.w_rho);
  const double r_rho = dmax((double)r[0], k.w_rho);
  const double e_rho = k.e_rho;
  const bool vac_r = r_rho < e_rho && l_rho > e_rho;
  const bool vac_l = l_rho < e_rho && r_rho > e_rho;
  const bool vac_b = l_rho < e_rho && r_rho < e_rho;
  double lv[ND], rv[ND], lp, lc, rp, rc;
  // (the reference converts only the non-vacuum side in the one-sided
  // vacuum branches; the other side's values are unused there)
  side_prims<ND, RI>(k, l, lv, lp, lc);
  side_prims<ND, RI>(k, r, rv, rp, rc);
  // default: the vacuum state w_vacuum, at rest
  srho = k.w_rho;
  sp = k.w_p;
#pragma unroll
  for (int d = 0; d < ND; ++d) sv[d] = 0.0;
  auto take = [&](double rho, const double* v, double p) {
    srho = rho;
    sp = p;
#pragma unroll
    for (int d = 0; d < ND; ++d) sv[d] = v[d];
  };

  if (vac_r) {
    double s_starl = lv[A] - PW::dv(2 * lc, k.gm1);
    if (dxdt > s_starl) {
      // vacuum
    } else if (dxdt > lv[A] - lc) {
      w_kfan_prim<ND, A, PW>(k, l_rho, lv, lp, lc, dxdt, false, srho, sv, sp);
    } else {
      take(l_rho, lv, lp);
    }
  } else if (vac_l) {
    double s_starr = rv[A] + PW::dv(2 * rc, k.gm1);
    if (dxdt > rv[A] + rc)
      take(r_rho, rv, rp);
    else if (dxdt > s_starr)
      w_kfan_prim<ND, A, PW>(k, r_rho, rv, rp, rc, dxdt, true, srho, sv, sp);
  } else if (vac_b) {
    // vacuum
  } else {
    const double du = rv[A] - lv[A];
    // du_crit = 2 (c_L + c_R) / (gamma - 1) > 0 (c_K >= tiny), so
    // du_crit < du needs du > 0: a wave with no such lane skips the division
    bool vacuum_gen = false;
    if (!wave_all(!(du > 0.0))) vacuum_gen = PW::dv(2 * (lc + rc), k.gm1) < du;
    if (vacuum_gen) {  // generation of vacuum (:83-98)
      const double s_starl = lv[A] - PW::dv(2 * lc, k.gm1);
      const double s_starr = rv[A] + PW::dv(2 * rc, k.gm1);
      if (dxdt > s_starr) {
        if (dxdt > rv[A] + rc)
          take(r_rho, rv, rp);
        else
          w_kfan_prim<ND, A, PW>(k, r_rho, rv, rp, rc, dxdt, true, srho, sv, sp);
      } else if (dxdt > s_starl) {
        // vacuum
      } else {
        if (dxdt > lv[A] - lc)
          w_kfan_prim<ND, A, PW>(k, l_rho, lv, lp, lc, dxdt, false, srho, sv, sp);
        else
          take(l_rho, lv, lp);
      }
    } else {
      // star state: rhyme_riemann_problem_star.f90:3-66
      double lf = 0.0, rf = 0.0;
      double ps = irs_exec<PW>(k, l_rho, lv[A], lp, lc, r_rho, rv[A], rp, rc, lf, rf);
      if (ps < k.w_p) ps = k.w_p;
      double us = 0.5 * ((rv[A] + lv[A]) + (rf - lf));
      // sampling at x/t = 0 (rhyme_riemann_problem_sampling.f90:13-47).  The
      // reference's right-of-contact and left-of-contact branches mirror each
      // other; here one path serves both, on the side K of the sampled point
      // with the signed sound speed sc = +c_R or -c_L (a - c == a + (-c) and
      // (-c)*x == -(c*x) exactly, so every value matches its branch).
      const bool right = dxdt > us;
      const double kr = right ? r_rho : l_rho;
      const double kp = right ? rp : lp;
      const double kc = right ? rc : lc;
      const double sc = right ? rc : -lc;
      double kv[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) kv[d] = right ? rv[d] : lv[d];
      // p* / p_K is exactly 1 at a contact (p* = p_K, finite and > 0) under
      // IEEE division: a wave of contacts skips the division
      const double ps_pk = !PW::kFast && wave_all(ps == kp && kp > 0.0 && kp < INFINITY)
                               ? 1.0
                               : PW::dv(ps, kp);
      take(kr, kv, kp);
      if (ps > kp) {  // shock on side K
        double speed = kv[A] + sc * PW::sq(k.gp1_2g * ps_pk + k.gm1_2g);
        if (right ? dxdt < speed : dxdt > speed) {
          srho = PW::dv(kr * (k.gm1_gp1 + ps_pk), k.gm1_gp1 * ps_pk + 1.0);
          sv[A] = us;
          sp = ps;
        }
      } else {  // rarefaction on side K
        double xa, xb;
        PW::pw2(ps_pk, k.gm1_2g, k.g_inv, xa, xb);
        double fan_cs = kc * xa;
        double speedH = kv[A] + sc;
        double speedT = us + (right ? fan_cs : -fan_cs);
        if (right ? dxdt < speedT : dxdt > speedT) {
          srho = kr * xb;
          sv[A] = us;
          sp = ps;
        } else if (right ? dxdt < speedH : dxdt > speedH) {
          w_kfan_prim<ND, A, PW>(k, kr, kv, kp, kc, dxdt, right, srho, sv, sp);
        }
      }
    }
  }
}

// The conserved state sampled at x/t = dxdt (FP64 unit-level solve).
template <int ND, int A, class PW = PowLib>
__device__ __forceinline__ void riemann_state(const Consts& k, const double* l, const double* r,
                                              double dxdt, double* u) {
  double srho, sv[ND], sp;
  riemann_sample<ND, A, PW, double>(k, l, r, dxdt, srho, sv, sp);
  prim_to_cons<ND, double>(k, srho, sv, sp, u);
}

// Riemann solve + calc_flux of the sampled state
// (…_solve_memory_intensive.f90:110-116): the intercell flux.  R is the
// type of the half states and of the flux (double: FP64 throughout; float:
// the mixed form, whose conversions to and from primitive variables and
// flux run in FP32 around the FP64 star-state solve and sampling).
template <int ND, int A, class PW, typename R>
__device__ __forceinline__ void riemann_flux(const Consts& k, const R* l, const R* r, R* F) {
  double srho, sv[ND], sp;
  riemann_sample<ND, A, PW, R>(k, l, r, 0.0, srho, sv, sp);
  R v[ND], u[ND + 2];
#pragma unroll
  for (int d = 0; d < ND; ++d) v[d] = (R)sv[d];
  prim_to_cons<ND, R>(k, (R)srho, v, (R)sp, u);
  flux<ND, A, R>(k, u, F);
}

}  // namespace rhy
