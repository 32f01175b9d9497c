// rhy_kernels.hpp -- gfx950 kernels of the hydro step on a device-resident
// box: per-axis intercell flux, conservative update (with the ghost fill and
// the CFL reduction of the next step fused into its epilogue), stand-alone
// ghost fill, CFL reduction, and the unit-level batch kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>
#include <vector>

#include "../../include/rhyme_hip.h"
#include "rhy_physics.hpp"

// Minimum waves per SIMD requested for the flux kernels (caps their VGPRs:
// 4 -> 128, 3 -> 168, 2 -> 256); 0 leaves the choice to the compiler, which
// can exceed 256 registers with AGPRs and drop to one wave per SIMD.  With
// the slow paths of the Riemann solve deferred to the fixup kernel the flux
// kernels need 145-167 VGPRs, so 3 waves fit without spills.
#ifndef RHY_FLUX_WPE
#define RHY_FLUX_WPE 3
#endif
#if RHY_FLUX_WPE > 0
#define RHY_FLUX_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RHY_FLUX_WPE)))
#else
#define RHY_FLUX_OCCUPANCY
#endif

// 1: the y/z walks stage their cells in a lane-private LDS ring (each cell
// loaded once) -- FP64 walks of single boxes and both walks of the batched
// kernel (its FP64 walks without the early load, which would spill there);
// 0: they re-read the three cells of every half step.  Measured: 7% faster
// single-box FP64 walks, 4% for the batched mixed form, 1.4% for the
// batched FP64 form; the single-box FP32 walks keep their register window,
// which is as fast.
#ifndef RHY_LDS_WALK
#define RHY_LDS_WALK 1
#endif

namespace rhy {

// hipMalloc that leaves no error behind when it fails: a refused allocation
// is reported through the return value (and *p = nullptr) and is not also
// picked up by the next hipGetLastError of an unrelated launch.
template <typename T>
inline hipError_t dev_alloc(T** p, size_t bytes) {
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
  }
  return e;
}

// Geometry of a ghosted box: Fortran cell (i, j, k), 1-based interior
// indices, lives at (i-1+g0)*s0 + (j-1+g1)*s1 + (k-1+g2)*s2 inside each
// component plane; inactive axes have n = 1 and g = 0.
struct Geom {
  int32_t n[3];
  int32_t g[3];
  int64_t s[3];
  int64_t plane;
};

__device__ __forceinline__ int64_t goff(const Geom& G, int i, int j, int k) {
  return (int64_t)(i - 1 + G.g[0]) * G.s[0] + (int64_t)(j - 1 + G.g[1]) * G.s[1] +
         (int64_t)(k - 1 + G.g[2]) * G.s[2];
}

template <int ND, typename R = double>
__device__ __forceinline__ void load_state(const R* __restrict__ cells, int64_t plane, int64_t o,
                                           R* u) {
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) u[c] = cells[o + c * plane];
}

// Component c of a state whose hydro planes (c < nv) start at `hydro` and
// whose other components (temperature, tracers) start at `extra`: the
// single-box path keeps two hydro buffers and one set of the others, a
// batch keeps them contiguous (extra = hydro + nv planes).
template <typename R>
__device__ __forceinline__ R* comp_plane(R* hydro, R* extra, int nv, int c, int64_t plane) {
  return c < nv ? hydro + (int64_t)c * plane : extra + (int64_t)(c - nv) * plane;
}

// The LDS ring of the y/z walks: 4 cells of up to 5 doubles (or floats) for
// each lane of a 256-thread workgroup (40 KB), one allocation per kernel
// whatever walks the kernel instantiates.
constexpr int kWalkLanes = 256;
__device__ __forceinline__ double* walk_ring() {
  __shared__ double ring[4 * 5 * kWalkLanes];
  return ring;
}

// The flux of the last face of each lane of a 256-thread workgroup (5
// values), for the accumulating walks: the cell between two faces of a
// walk takes both, and the lower one waits here across the Riemann solve of
// the upper one instead of in registers.
__device__ __forceinline__ double* walk_fprev() {
  __shared__ double fp[5 * kWalkLanes];
  return fp;
}

// Boundary types per face (left, right, bottom, top, back, front).
struct BcTypes {
  int32_t t[6];
};

// Test hook settings (rhy_debug_defer): every > 0 defers every such face,
// cap >= 0 overrides the list capacity.
int32_t defer_every_setting();
int64_t defer_cap_setting();

// A box's device state as the hierarchy kernels see it (box_view waits for
// the box's stream).
struct BoxView {
  void* cells;
  void* flux;  // flux planes of the last rhy_mh_solve (the x planes after rhy_amr_coarse_step)
  void* prev;  // after rhy_amr_coarse_step: the state it stepped from, else nullptr
  Consts k;    // that step's constants, dt and dx (with prev)
  double step_dt, step_dx[3];
  Geom geom;
  int32_t ndim, ncmp, precision;
};
int box_view(rhy_box_t* b, BoxView* v);

// Faces whose Riemann solve needs a slow path (Deferring<P>), listed by the
// flux kernels for the fixup kernel: item = box << 42 | axis << 40 | the
// offset of the face's low cell in its box's planes.  count may pass cap;
// the fixup then recomputes every face.
struct DeferList {
  uint32_t* count;
  uint64_t* items;
  uint32_t cap;
};

__device__ __forceinline__ uint64_t defer_item(uint32_t box, int axis, int64_t of) {
  return ((uint64_t)box << 42) | ((uint64_t)axis << 40) | (uint64_t)of;
}

// After a face's Riemann solve: list the face if the lane's flag is up
// (or, under the test hook k.defer_every, if its offset is a multiple);
// true if it was listed.
template <class PW>
__device__ __forceinline__ bool defer_take(const Consts& k, const DeferList& dl, uint32_t box,
                                           int axis, int64_t of) {
  if constexpr (PW::kDefer) {
    uint32_t& f = defer_flag();
    if (k.defer_every > 0 && of % k.defer_every == 0) f = 1u;
    if (f) {
      f = 0u;
      const uint32_t i = atomicAdd(dl.count, 1u);
      if (i < dl.cap) dl.items[i] = defer_item(box, axis, of);
      return true;
    }
  }
  return false;
}

template <class PW>
__device__ __forceinline__ void defer_clear() {
  if constexpr (PW::kDefer) defer_flag() = 0u;
}

// Writes the ghost images of an updated interior cell, i.e. what
// rhyme_samr_bc_set_<face>_boundary (rhyme_samr_bc/src/…_left_boundary.f90:47-90
// and siblings) would write from this cell on the next step's fill.
// Reflective/outflow mirror cells 1,2 -> ghosts 0,-1 (n, n-1 -> n+1, n+2),
// periodic wraps n, n-1 -> 0, -1 and 1, 2 -> n+1, n+2; reflective negates the
// normal momentum; inflow ghosts are constant and left alone.
template <int ND, typename R = double>
__device__ __forceinline__ void write_ghost_images(const Geom& G, const BcTypes& bc,
                                                   R* __restrict__ cells, int i, int j, int k,
                                                   const R* u) {
  int idx[3] = {i, j, k};
#pragma unroll
  for (int a = 0; a < ND; ++a) {
    const int n = G.n[a];
    const int ia = idx[a];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int type = bc.t[2 * a + side];
      int ghost = 0;
      bool hit = false;
      if (type == 1 || type == 2) {
        if (side == 0) {
          if (ia == 1) { ghost = 0; hit = true; }
          else if (ia == 2) { ghost = -1; hit = true; }
        } else {
          if (ia == n) { ghost = n + 1; hit = true; }
          else if (ia == n - 1) { ghost = n + 2; hit = true; }
        }
      } else if (type == 3) {
        if (side == 0) {
          if (ia == n) { ghost = 0; hit = true; }
          else if (ia == n - 1) { ghost = -1; hit = true; }
        } else {
          if (ia == 1) { ghost = n + 1; hit = true; }
          else if (ia == 2) { ghost = n + 2; hit = true; }
        }
      }
      if (hit) {
        int gi[3] = {i, j, k};
        gi[a] = ghost;
        int64_t o = goff(G, gi[0], gi[1], gi[2]);
#pragma unroll
        for (int c = 0; c < ND + 2; ++c) {
          R v = u[c];
          if (type == 1 && c == 1 + a) v = -v;
          cells[o + c * G.plane] = v;
        }
      }
    }
  }
}

// |v| + c_s of one cell: calc_u of rhyme_cfl_time_step.f90:52-59.  (A
// wave of cells at rest with rho > 0 skips the square root and division:
// sqrt(+0) / rho is +0.)
template <int ND>
__device__ __forceinline__ double cfl_speed(const Consts& k, const double* u) {
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) s += u[1 + d] * u[1 + d];
  const double v = wave_all(s == 0.0 && u[0] > 0.0) ? 0.0 : sqrt(s) / u[0];
  return v + speed_of_sound<ND>(k, u);
}

// Block-wide max of non-negative values; thread 0 of the block writes the
// block's maximum to partial[blockIdx.x].  max is exact and order
// independent, so the reduction is deterministic; NaN speeds are skipped like
// the reference's strict `u > max_u` test (rhyme_cfl_time_step.f90:44).
// Every thread of the block must call it (blockDim.x a multiple of 64, <= 1024).
__device__ __forceinline__ void block_max_to_index(double v, double* partial, int64_t index) {
  __shared__ double wave_max[16];
  v = v > 0.0 ? v : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wave_max[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) m = wave_max[i] > m ? wave_max[i] : m;
    partial[index] = m;
  }
}

__device__ __forceinline__ void block_max_to_partial(double v, double* partial) {
  block_max_to_index(v, partial, blockIdx.x);
}

// Lane l of the wave gets lane l-1's value (lane 0 keeps its own): one DPP
// wave_shr:1 move per 32 bits instead of an LDS permute.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_below(float x) {
  return __builtin_bit_cast(float, wave_shr1(__builtin_bit_cast(uint32_t, x)));
}
__device__ __forceinline__ double lane_below(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint64_t lo = wave_shr1((uint32_t)b), hi = wave_shr1((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, lo | (hi << 32));
}

// Intercell flux on every face normal to axis A that the interior update
// reads: face f in [0, n_A] sits between cells f and f+1, the other indices
// run over the interior (the faces-only range the update consumes,
// rhyme_muscl_hancock_solve_memory_intensive.f90:142-176).  Each thread
// walks NF consecutive faces along A with a sliding window of cells: the
// half-step states of every cell on the walk (loop A, :49-98) are formed
// once and shared by its two faces, then each face gets its Riemann flux
// (loop B, :100-140), stored at the low cell's offset in F (NV planes).
//
// R is the storage/bulk type (double: the FP64 drop-in; float: the mixed
// form, whose half step and the conversions around the Riemann solve run in
// FP32 while the star-state solve and the sampling run in FP64 with the
// MixedMath policy).  LIM is the slope limiter (slid).  PW is the policy of
// the Riemann solve; with Deferring<P> the faces that need P's slow paths
// are listed in dl, tagged with `box`, for the fixup kernel.
//
// flux_xrows: faces normal to x, walked by whole waves.  The cells of all x rows,
// including the ghost cell on each side whose half steps the end faces use,
// are numbered row after row, g = row * (n0 + 2) + c with c = 0..n0+1 the
// cell's x index; at each of `steps` steps the 64 lanes of a wave take 64
// consecutive cells, form their half steps, take the right state of the cell
// below from lane l-1 (lane 0: lane 63's of the previous step, from a
// warm-up step at the start) and solve the face between the two.  A wave's
// loads and flux stores are contiguous along x, and each half step is formed
// once (plus one warm-up step per wave).
template <int ND, int LIM, typename R, class PW>
__device__ __forceinline__ void flux_xrows(const Consts& k, const Geom& G,
                                           const R* __restrict__ cells, R* __restrict__ F,
                                           double dt, double dx, int steps, int64_t t,
                                           const DeferList& dl, uint32_t box) {
  constexpr int NV = ND + 2;
  defer_clear<PW>();
  const int64_t rowlen = G.n[0] + 2;
  const int64_t ncells = rowlen * G.n[1] * G.n[2];
  const int lane = (int)(t & 63);
  const int64_t base = (t >> 6) * 64 * (int64_t)steps;  // first cell of the wave
  if (base >= ncells) return;                            // the whole wave
  const R hdtdx = (R)(.5 * dt / dx);
  R carry[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) carry[c] = R(0);
  // the lane's cell (x index ci, row j, plane kk): located by division at
  // the warm-up step and the first step, then advanced 64 cells per step
  int ci = 0, jj = 1, kk = 1;
#pragma unroll 1
  for (int m = -1; m < steps; ++m) {
    const int64_t g = base + (int64_t)m * 64 + lane;
    const bool valid = g >= 0 && g < ncells;
    if (m <= 0) {
      if (valid) {
        const int64_t row = g / rowlen;
        ci = (int)(g - row * rowlen);
        jj = (int)(row % G.n[1]) + 1;
        kk = (int)(row / G.n[1]) + 1;
      }
    } else {
      ci += 64;
      while (ci >= rowlen) {
        ci -= (int)rowlen;
        if (++jj > G.n[1]) {
          jj = 1;
          ++kk;
        }
      }
    }
    R lw[NV], rn[NV];
    int64_t o = 0;
    if (valid) {
      o = goff(G, ci, jj, kk);
      R um[NV], u0[NV], up[NV];
      load_state<ND, R>(cells, G.plane, o - 1, um);
      load_state<ND, R>(cells, G.plane, o, u0);
      load_state<ND, R>(cells, G.plane, o + 1, up);
      half_step<ND, 0, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
    } else {
#pragma unroll
      for (int c = 0; c < NV; ++c) lw[c] = rn[c] = R(0);
    }
    R rp[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const R below = lane_below(rn[c]);
      rp[c] = lane == 0 ? carry[c] : below;
      carry[c] = __shfl(rn[c], 63);
    }
    if (m >= 0 && valid && ci >= 1) {  // face ci - 1, between cells ci - 1 and ci
      R Fr[NV];
      riemann_flux<ND, 0, PW, R>(k, rp, lw, Fr);
#pragma unroll
      for (int c = 0; c < NV; ++c) F[o - 1 + c * G.plane] = Fr[c];
      defer_take<PW>(k, dl, box, 0, o - 1);
    }
  }
}

// flux_walk: one thread's walk of nf_max consecutive faces normal to A
// (y, z), starting at face chunk `t` of the box (thread t of the axis' index
// space); nothing if t is past the end.  A == 0 goes to flux_xrows.
// WINDOW (single boxes): FP32 cells through a register sliding window, FP64
// cells through the LDS ring (RHY_LDS_WALK).  Otherwise (the batched kernel,
// whose three walks share one register budget): FP32 cells through the LDS
// ring, FP64 cells through the ring without the early load, which keeps the
// kernel within 3 waves per SIMD unspilled.
template <int ND, int A, int LIM, typename R, class PW, bool WINDOW = true>
__device__ __forceinline__ void flux_walk(const Consts& k, const Geom& G,
                                          const R* __restrict__ cells, R* __restrict__ F,
                                          double dt, double dx, int nf_max, int64_t t,
                                          const DeferList& dl, uint32_t box) {
  if constexpr (A == 0) {
    flux_xrows<ND, LIM, R, PW>(k, G, cells, F, dt, dx, nf_max, t, dl, box);
    return;
  }
  constexpr int NV = ND + 2;
  defer_clear<PW>();
  const int nA = G.n[A];
  const int nchunk = (nA + 1 + nf_max - 1) / nf_max;
  const int e0 = A == 0 ? nchunk : G.n[0];
  const int e1 = A == 1 ? nchunk : G.n[1];
  const int e2 = A == 2 ? nchunk : G.n[2];
  if (t >= (int64_t)e0 * e1 * e2) return;
  int idx[3];
  idx[0] = (int)(t % e0);
  const int64_t rr = t / e0;
  idx[1] = (int)(rr % e1);
  idx[2] = (int)(rr / e1);
  const int f0 = idx[A] * nf_max;
  const int nf = nA + 1 - f0 < nf_max ? nA + 1 - f0 : nf_max;
#pragma unroll
  for (int d = 0; d < 3; ++d) idx[d] = d == A ? f0 : idx[d] + 1;
  const int64_t o = goff(G, idx[0], idx[1], idx[2]);
  const int64_t s = G.s[A];
  const R hdtdx = (R)(.5 * dt / dx);
  R um[NV], u0[NV], up[NV], lw[NV], rp[NV], rn[NV];
  load_state<ND, R>(cells, G.plane, o - s, um);
  load_state<ND, R>(cells, G.plane, o, u0);
  load_state<ND, R>(cells, G.plane, o + s, up);
  half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rp);
  if constexpr (sizeof(R) == sizeof(float) && WINDOW) {
    // FP32 cells: a sliding window carries two cells to the next half step
    // and loads one new cell, issued before the Riemann solve so that its
    // latency hides behind it (cell f0 + m + 3 at face m, i.e. at most cell
    // f0 + nf + 1, the last one the half steps read)
    R nx[NV];
    load_state<ND, R>(cells, G.plane, o + 2 * s, nx);
#pragma unroll 1
    for (int m = 0; m < nf; ++m) {
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        um[c] = u0[c];
        u0[c] = up[c];
        up[c] = nx[c];
      }
      half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
      if (m + 1 < nf) load_state<ND, R>(cells, G.plane, o + (m + 3) * s, nx);
      R Fr[NV];
      riemann_flux<ND, A, PW, R>(k, rp, lw, Fr);
      const int64_t of = o + m * s;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        F[of + c * G.plane] = Fr[c];
        rp[c] = rn[c];
      }
      defer_take<PW>(k, dl, box, A, of);
    }
  } else if constexpr (RHY_LDS_WALK && (sizeof(R) == sizeof(double) ? WINDOW : !WINDOW)) {
    // Cells through a lane-private ring of 4 cells in LDS (cell j of the
    // walk, counted from f0, in slot j & 3): each cell is loaded from HBM
    // once.  At face m the load of cell m + 3 is issued before the half step
    // of cell m + 1 (cells m .. m + 2, read from LDS) and written to the ring
    // after it, so it is in flight during the half step and needs no
    // registers across the Riemann solve.  A lane touches only its own slots
    // (no barrier).
    R* ring = reinterpret_cast<R*>(walk_ring());
    const int lane = threadIdx.x;
    auto put = [&](int j, const R* v) {
#pragma unroll
      for (int c = 0; c < NV; ++c) ring[((j & 3) * NV + c) * kWalkLanes + lane] = v[c];
    };
    auto get = [&](int j, R* v) {
#pragma unroll
      for (int c = 0; c < NV; ++c) v[c] = ring[((j & 3) * NV + c) * kWalkLanes + lane];
    };
    put(0, u0);
    put(1, up);
    R nx[NV];
    load_state<ND, R>(cells, G.plane, o + 2 * s, nx);
    put(2, nx);
#pragma unroll 1
    for (int m = 0; m < nf; ++m) {
      if (m + 1 < nf) load_state<ND, R>(cells, G.plane, o + (m + 3) * s, nx);
      get(m, um);
      get(m + 1, u0);
      get(m + 2, up);
      half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
      if (m + 1 < nf) put(m + 3, nx);
      R Fr[NV];
      riemann_flux<ND, A, PW, R>(k, rp, lw, Fr);
      const int64_t of = o + m * s;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        F[of + c * G.plane] = Fr[c];
        rp[c] = rn[c];
      }
      defer_take<PW>(k, dl, box, A, of);
    }
  } else if constexpr (RHY_LDS_WALK && sizeof(R) == sizeof(double)) {
    // Batched FP64 walks: the ring without the early load, so no registers
    // beyond the re-reading walk's (the early load makes this kernel spill):
    // each face loads its newest cell and takes the two below from the ring
    R* ring = reinterpret_cast<R*>(walk_ring());
    const int lane = threadIdx.x;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      ring[(0 * NV + c) * kWalkLanes + lane] = u0[c];
      ring[(1 * NV + c) * kWalkLanes + lane] = up[c];
    }
#pragma unroll 1
    for (int m = 0; m < nf; ++m) {
      load_state<ND, R>(cells, G.plane, o + (m + 2) * s, up);
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        um[c] = ring[((m & 3) * NV + c) * kWalkLanes + lane];
        u0[c] = ring[(((m + 1) & 3) * NV + c) * kWalkLanes + lane];
        ring[(((m + 2) & 3) * NV + c) * kWalkLanes + lane] = up[c];
      }
      half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
      R Fr[NV];
      riemann_flux<ND, A, PW, R>(k, rp, lw, Fr);
      const int64_t of = o + m * s;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        F[of + c * G.plane] = Fr[c];
        rp[c] = rn[c];
      }
      defer_take<PW>(k, dl, box, A, of);
    }
  } else {
    // Without the LDS ring (RHY_LDS_WALK=0): only rp (the right state of the
    // low cell) is carried across the Riemann solve; the three cells of the
    // next half step are re-read (cache hits), which keeps the register
    // count near the single-face one.
#pragma unroll 1
    for (int m = 0; m < nf; ++m) {
      const int64_t oc = o + (m + 1) * s;
      load_state<ND, R>(cells, G.plane, oc - s, um);
      load_state<ND, R>(cells, G.plane, oc, u0);
      load_state<ND, R>(cells, G.plane, oc + s, up);
      half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
      R Fr[NV];
      riemann_flux<ND, A, PW, R>(k, rp, lw, Fr);
      const int64_t of = o + m * s;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        F[of + c * G.plane] = Fr[c];
        rp[c] = rn[c];
      }
      defer_take<PW>(k, dl, box, A, of);
    }
  }
}

// Threads of the flux index space of axis A for faces walked nf at a time.
// (x: whole waves of 64 lanes, each over 64 * nf consecutive row cells)
__host__ __device__ inline int64_t flux_threads(const Geom& G, int A, int nf) {
  if (A == 0) {
    const int64_t cells = (int64_t)(G.n[0] + 2) * G.n[1] * G.n[2];
    return (cells + 64 * (int64_t)nf - 1) / (64 * (int64_t)nf) * 64;
  }
  int64_t n = (G.n[A] + 1 + nf - 1) / nf;
  for (int d = 0; d < 3; ++d)
    if (d != A) n *= G.n[d];
  return n;
}

template <int ND, int A, int LIM, typename R, class PW>
__global__ __launch_bounds__(256) RHY_FLUX_OCCUPANCY void flux_kernel(
    Consts k, Geom G, const R* __restrict__ cells, R* __restrict__ F,
    const double* __restrict__ dtp, double dx, int nf, DeferList dl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  flux_walk<ND, A, LIM, R, PW>(k, G, cells, F, *dtp, dx, nf, t, dl, 0u);
}

// ------------------------------------------------ accumulating walks ----
//
// The single-box step without a separate update pass.  The x walk stores
// its fluxes F_x (flux_kernel).  Walking y and z, a lane takes a chunk of L
// consecutive cells c0 .. c0+L-1 and solves the L+1 faces c0-1 .. c0+L-1
// around them (one face more than the cells, shared with the neighbouring
// chunk), so that every cell of the chunk has both its faces in the walk:
//   kAccMid  (y of a 3D box): Q = A + B (memory_intensive) or (u + A) + B
//            (cpu_intensive), A = dt/dx_x (F_x[c-1] - F_x[c]) from the F_x
//            planes, B = dt/dx_y (F_y[c-1] - F_y[c]) from the walk;
//   kAccLast (z of a 3D box, y of a 2D box): the new state u + (Q + C) or
//            Q + C (3D), u + (A + B) or (u + A) + B (2D), written to the
//            other state buffer (interior only: a ghost fill of the new
//            state follows the step), and its CFL speed folded into the
//            lane's maximum.
// These are the update's expressions operation for operation
// (rhyme_muscl_hancock_solve_memory_intensive.f90:142-176,
// …_solve_cpu_intensive.f90:117-136), so the state equals the flux-plane
// path's bit for bit.  A cell beside a face the walk deferred gets a
// placeholder here and no CFL speed; acc_fixup_kernel recomputes it.
constexpr int kAccMid = 1, kAccLast = 2;

// The accumulation of component value uc of one cell (q: its offset in the
// component's plane) from its lower and upper face fluxes fl, fu along A;
// stored to out[q] and returned.
template <int ND, int A, typename R, int MODE>
__device__ __forceinline__ R acc_one(const Geom& G, const R* __restrict__ fx,
                                     const R* __restrict__ acc, R* __restrict__ out, int64_t q,
                                     R uc, R fl, R fu, R dtdx0, R dtdxA, int order) {
  const R b = dtdxA * (fl - fu);
  R v;
  if constexpr (MODE == kAccMid || ND == 2) {
    const R a0 = dtdx0 * (fx[q - G.s[0]] - fx[q]);
    if constexpr (MODE == kAccMid)
      v = order == RHY_MEMORY_INTENSIVE ? a0 + b : (uc + a0) + b;
    else
      v = order == RHY_MEMORY_INTENSIVE ? uc + (a0 + b) : (uc + a0) + b;
  } else {
    v = order == RHY_MEMORY_INTENSIVE ? uc + (acc[q] + b) : acc[q] + b;
  }
  out[q] = v;
  return v;
}

// All components of one cell (offset oc, state uc, face fluxes fl, fu);
// the new values also into un.
template <int ND, int A, typename R, int MODE>
__device__ __forceinline__ void acc_cell(const Geom& G, const R* __restrict__ fx,
                                         const R* __restrict__ acc, R* __restrict__ out,
                                         int64_t oc, const R* uc, const R* fl, const R* fu,
                                         R dtdx0, R dtdxA, int order, R* un) {
#pragma unroll
  for (int c = 0; c < ND + 2; ++c)
    un[c] = acc_one<ND, A, R, MODE>(G, fx, acc, out, oc + c * G.plane, uc[c], fl[c], fu[c],
                                    dtdx0, dtdxA, order);
}

// What the epilogue of one cell reads from HBM: F_x at the cell and below
// it (kAccMid, and kAccLast of a 2D box), or Q at the cell (kAccLast, 3D).
template <int ND, typename R, int MODE>
struct AccLoads {
  static constexpr bool kFx = MODE == kAccMid || ND == 2;
  R a[ND + 2];  // F_x[c - 1] or Q[c]
  R b[ND + 2];  // F_x[c]
  __device__ __forceinline__ void load(const Geom& G, const R* __restrict__ fx,
                                       const R* __restrict__ acc, int64_t oc) {
#pragma unroll
    for (int c = 0; c < ND + 2; ++c) {
      const int64_t q = oc + c * G.plane;
      if constexpr (kFx) {
        a[c] = fx[q - G.s[0]];
        b[c] = fx[q];
      } else {
        a[c] = acc[q];
      }
    }
  }
};

// acc_one with the cell's loads already in registers (same arithmetic).
template <int ND, typename R, int MODE>
__device__ __forceinline__ R acc_one_loaded(R* __restrict__ out, int64_t q, R uc, R fl, R fu,
                                            R xa, R xb, R dtdx0, R dtdxA, int order) {
  const R b = dtdxA * (fl - fu);
  R v;
  if constexpr (MODE == kAccMid || ND == 2) {
    const R a0 = dtdx0 * (xa - xb);
    if constexpr (MODE == kAccMid)
      v = order == RHY_MEMORY_INTENSIVE ? a0 + b : (uc + a0) + b;
    else
      v = order == RHY_MEMORY_INTENSIVE ? uc + (a0 + b) : (uc + a0) + b;
  } else {
    v = order == RHY_MEMORY_INTENSIVE ? uc + (xa + b) : xa + b;
  }
  out[q] = v;
  return v;
}

template <int ND, int A, int LIM, typename R, class PW, int MODE>
__device__ __forceinline__ double flux_walk_acc(const Consts& k, const Geom& G,
                                                const R* __restrict__ cells,
                                                const R* __restrict__ fx,
                                                const R* __restrict__ acc, R* __restrict__ out,
                                                const double* __restrict__ sc, int L,
                                                int order, int64_t t, const DeferList& dl) {
  static_assert(A > 0 && A < ND, "the accumulating walks run along y or z");
  constexpr int NV = ND + 2;
  defer_clear<PW>();
  double smax = 0.0;
  const int nA = G.n[A];
  const int nchunk = (nA + L - 1) / L;
  const int e0 = G.n[0];
  const int e1 = A == 1 ? nchunk : G.n[1];
  const int e2 = A == 2 ? nchunk : G.n[2];
  if (t >= (int64_t)e0 * e1 * e2) return smax;
  int idx[3];
  idx[0] = (int)(t % e0);
  const int64_t rr = t / e0;
  idx[1] = (int)(rr % e1);
  idx[2] = (int)(rr / e1);
  const int f0 = idx[A] * L;  // faces f0 .. f0+nf-1; cells f0+1 .. f0+nf-1
  const int nf = nA + 1 - f0 < L + 1 ? nA + 1 - f0 : L + 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) idx[d] = d == A ? f0 : idx[d] + 1;
  const int64_t o = goff(G, idx[0], idx[1], idx[2]);
  const int64_t s = G.s[A];
  // dt/dx and (.5 dt)/dx from the step's constants (uniform loads, held in
  // scalar registers)
  const R hdtdx = (R)sc[4 + A];
  const R dtdxA = (R)sc[1 + A];
  const R dtdx0 = (R)sc[1];
  // cells of the walk through the lane's LDS ring (cell j from f0 in slot
  // j & 3, the next one loaded during the half step, as in flux_walk); the
  // lower face's flux of the next cell in the lane's walk_fprev slot
  R* ring = reinterpret_cast<R*>(walk_ring());
  R* fprev = reinterpret_cast<R*>(walk_fprev());
  const int lane = threadIdx.x;
  auto put = [&](int j, const R* v) {
#pragma unroll
    for (int c = 0; c < NV; ++c) ring[((j & 3) * NV + c) * kWalkLanes + lane] = v[c];
  };
  auto get = [&](int j, R* v) {
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = ring[((j & 3) * NV + c) * kWalkLanes + lane];
  };
  R um[NV], u0[NV], up[NV], lw[NV], rp[NV], rn[NV];
  load_state<ND, R>(cells, G.plane, o - s, um);
  load_state<ND, R>(cells, G.plane, o, u0);
  load_state<ND, R>(cells, G.plane, o + s, up);
  half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rp);
  put(0, u0);
  put(1, up);
  R nx[NV];
  load_state<ND, R>(cells, G.plane, o + 2 * s, nx);
  put(2, nx);
  // The walk is software-pipelined: the cell between faces m-2 and m-1 is
  // finished at face m, between that face's half step and its Riemann
  // solve.  Its loads (AccLoads) are issued at the top of the iteration and
  // arrive during the half step; its upper face's flux waits in registers
  // across the half step, its lower one in the lane's walk_fprev slot;
  // nothing more than a plain walk holds is live across a Riemann solve.
  auto finish = [&](int j, const R* fu, const AccLoads<ND, R, MODE>& ld, bool deferred) {
    const int64_t oc = o + (int64_t)j * s;
    R un[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const R fl = fprev[c * kWalkLanes + lane];
      fprev[c * kWalkLanes + lane] = fu[c];
      const R uc = ring[((j & 3) * NV + c) * kWalkLanes + lane];
      un[c] = acc_one_loaded<ND, R, MODE>(out, oc + c * G.plane, uc, fl, fu[c], ld.a[c],
                                          AccLoads<ND, R, MODE>::kFx ? ld.b[c] : R(0), dtdx0,
                                          dtdxA, order);
    }
    if constexpr (MODE == kAccLast) {
      if (!deferred) {  // a deferred face: the fixup recomputes the speed
        double ud[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) ud[c] = (double)un[c];
        const double sp = cfl_speed<ND>(k, ud);
        smax = sp > smax ? sp : smax;
      }
    }
  };
  R fh[NV];         // flux of face m - 1, held for the finish at face m
  bool dh = false;  // face m - 1 deferred
  bool dp = false;  // face m - 2 deferred (its flux in fprev)
#pragma unroll 1
  for (int m = 0; m < nf; ++m) {
    AccLoads<ND, R, MODE> ld;
    if (m >= 2) ld.load(G, fx, acc, o + (int64_t)(m - 1) * s);
    if (m + 1 < nf) load_state<ND, R>(cells, G.plane, o + (m + 3) * s, nx);
    get(m, um);
    get(m + 1, u0);
    get(m + 2, up);
    half_step<ND, A, LIM, R>(k, um, u0, up, hdtdx, lw, rn);
    if (m >= 2) {  // cell f0 + m - 1 (its ring slot is overwritten by the put)
      finish(m - 1, fh, ld, dp || dh);
      dp = dh;
    }
    if (m + 1 < nf) put(m + 3, nx);
    R Fr[NV];
    riemann_flux<ND, A, PW, R>(k, rp, lw, Fr);
    const bool dm = defer_take<PW>(k, dl, 0u, A, o + (int64_t)m * s);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      rp[c] = rn[c];
      if (m == 0)
        fprev[c * kWalkLanes + lane] = Fr[c];
      else
        fh[c] = Fr[c];
    }
    if (m == 0)
      dp = dm;
    else
      dh = dm;
  }
  {  // the last cell, f0 + nf - 1 (faces nf - 2 and nf - 1; nf >= 2)
    AccLoads<ND, R, MODE> ld;
    ld.load(G, fx, acc, o + (int64_t)(nf - 1) * s);
    finish(nf - 1, fh, ld, dp || dh);
  }
  return smax;
}

// Threads of an accumulating walk along A with chunks of L cells.
__host__ __device__ inline int64_t acc_threads(const Geom& G, int A, int L) {
  int64_t n = (G.n[A] + L - 1) / L;
  for (int d = 0; d < 3; ++d)
    if (d != A) n *= G.n[d];
  return n;
}

template <int ND, int A, int LIM, typename R, class PW, int MODE>
__global__ __launch_bounds__(256) RHY_FLUX_OCCUPANCY void acc_walk_kernel(
    Consts k, Geom G, const R* __restrict__ cells, const R* __restrict__ fx,
    const R* __restrict__ acc, R* __restrict__ out, const double* __restrict__ sc,
    int L, int order, DeferList dl, double* __restrict__ partial) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double sp = flux_walk_acc<ND, A, LIM, R, PW, MODE>(k, G, cells, fx, acc, out, sc, L,
                                                          order, t, dl);
  if constexpr (MODE == kAccLast) block_max_to_partial(sp, partial);
}

// One face's flux from scratch: the half steps of its two cells (the
// stencil of four cells along A) and the Riemann solve with policy PW, the
// computation the flux walks make for it.  of: offset of the low cell.
template <int ND, int A, typename R, class PW>
__device__ __forceinline__ void face_flux(const Consts& k, const Geom& G,
                                          const R* __restrict__ cells, double dt, double dx,
                                          int64_t of, R* Fr) {
  constexpr int NV = ND + 2;
  const int64_t s = G.s[A];
  const R hdtdx = (R)(.5 * dt / dx);
  R um[NV], u0[NV], up[NV], l0[NV], r0[NV], l1[NV], r1[NV];
  load_state<ND, R>(cells, G.plane, of - s, um);
  load_state<ND, R>(cells, G.plane, of, u0);
  load_state<ND, R>(cells, G.plane, of + s, up);
  half_step<ND, A, 0, R>(k, um, u0, up, hdtdx, l0, r0);
  load_state<ND, R>(cells, G.plane, of + 2 * s, um);
  half_step<ND, A, 0, R>(k, u0, up, um, hdtdx, l1, r1);
  riemann_flux<ND, A, PW, R>(k, r0, l1, Fr);
}

template <int ND, int A, typename R, class PW>
__device__ __forceinline__ void flux_one_face(const Consts& k, const Geom& G,
                                              const R* __restrict__ cells, R* __restrict__ FA,
                                              double dt, double dx, int64_t of) {
  constexpr int NV = ND + 2;
  R Fr[NV];
  face_flux<ND, A, R, PW>(k, G, cells, dt, dx, of, Fr);
#pragma unroll
  for (int c = 0; c < NV; ++c) FA[of + c * G.plane] = Fr[c];
}

// The same for a face given by axis at run time; F holds the NDIM flux
// plane sets of the box one after the other.
template <int ND, typename R, class PW>
__device__ __noinline__ void flux_face_fix(const Consts& k, const Geom& G,
                                           const R* __restrict__ cells, R* __restrict__ F,
                                           double dt, const double* dx, int a, int64_t of) {
  R* FA = F + (int64_t)a * (ND + 2) * G.plane;
  if (a == 0) flux_one_face<ND, 0, R, PW>(k, G, cells, FA, dt, dx[0], of);
  if constexpr (ND > 1)
    if (a == 1) flux_one_face<ND, 1, R, PW>(k, G, cells, FA, dt, dx[1], of);
  if constexpr (ND > 2)
    if (a == 2) flux_one_face<ND, 2, R, PW>(k, G, cells, FA, dt, dx[2], of);
}

// Number of faces of axis a that a step computes, and the low cell of the
// q-th of them (faces 0..n_a along a, interior along the other axes).
__device__ __forceinline__ int64_t faces_of_axis(const Geom& G, int a) {
  int64_t n = (int64_t)G.n[a] + 1;
  for (int d = 0; d < 3; ++d)
    if (d != a) n *= G.n[d];
  return n;
}
__device__ __forceinline__ int64_t face_low_cell(const Geom& G, int a, int64_t q) {
  int idx[3];
  for (int d = 0; d < 3; ++d) {
    const int64_t e = d == a ? (int64_t)G.n[d] + 1 : G.n[d];
    idx[d] = (int)(q % e) + (d == a ? 0 : 1);
    q /= e;
  }
  return goff(G, idx[0], idx[1], idx[2]);
}

// Recomputes the faces the flux kernels of one box listed, with the full
// policy PW (grid-stride over the list); if the list overflowed, every face
// of the box.  Runs after the last flux kernel of the step.
template <int ND, typename R, class PW>
__global__ __launch_bounds__(256) void flux_fixup_kernel(Consts k, Geom G,
                                                         const R* __restrict__ cells,
                                                         R* __restrict__ F,
                                                         const double* __restrict__ dtp,
                                                         double dx0, double dx1, double dx2,
                                                         DeferList dl, int axis_only) {
  const uint32_t n = *dl.count;
  if (n == 0) return;
  const double dt = *dtp;
  const double dx[3] = {dx0, dx1, dx2};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n <= dl.cap) {
    for (int64_t i = t0; i < (int64_t)n; i += stride) {
      const uint64_t it = dl.items[i];
      flux_face_fix<ND, R, PW>(k, G, cells, F, dt, dx, (int)((it >> 40) & 3),
                               (int64_t)(it & ((1ull << 40) - 1)));
    }
    return;
  }
  for (int a = 0; a < ND; ++a) {
    if (axis_only >= 0 && a != axis_only) continue;
    const int64_t na = faces_of_axis(G, a);
    for (int64_t q = t0; q < na; q += stride)
      flux_face_fix<ND, R, PW>(k, G, cells, F, dt, dx, a, face_low_cell(G, a, q));
  }
}

// The cells an accumulating walk along A left beside deferred faces,
// recomputed with the full policy PW: both faces of each such cell from
// scratch, then the walk's own accumulation (acc_cell); kAccLast also
// folds the CFL speed into *slot (an extra
// partial slot the dt reduction reads; speeds are >= 0, so their bit
// patterns order as integers).  If the list overflowed, every cell of the
// box.  Cells listed twice (two deferred faces) are written twice with the
// same values.
template <int ND, int A, typename R, class PW, int MODE>
__device__ __forceinline__ void acc_fix_cell(const Consts& k, const Geom& G,
                                             const R* __restrict__ cells,
                                             const R* __restrict__ fx, const R* __restrict__ acc,
                                             R* __restrict__ out, double dt,
                                             double dx0, double dxA, int order, int64_t oc,
                                             double* slot) {
  constexpr int NV = ND + 2;
  int ci[3];
  int64_t rem = oc;
#pragma unroll
  for (int d = 2; d >= 0; --d) {
    ci[d] = (int)(rem / G.s[d]) - G.g[d] + 1;
    rem -= (int64_t)(ci[d] - 1 + G.g[d]) * G.s[d];
  }
  if (ci[A] < 1 || ci[A] > G.n[A]) return;
  R fl[NV], fu[NV], uc[NV], un[NV];
  face_flux<ND, A, R, PW>(k, G, cells, dt, dxA, oc - G.s[A], fl);
  face_flux<ND, A, R, PW>(k, G, cells, dt, dxA, oc, fu);
  load_state<ND, R>(cells, G.plane, oc, uc);
  acc_cell<ND, A, R, MODE>(G, fx, acc, out, oc, uc, fl, fu, (R)(dt / dx0), (R)(dt / dxA),
                           order, un);
  if constexpr (MODE == kAccLast) {
    double ud[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) ud[c] = (double)un[c];
    const double sp = cfl_speed<ND>(k, ud);
    if (sp > 0.0)
      atomicMax(reinterpret_cast<unsigned long long*>(slot),
                (unsigned long long)__double_as_longlong(sp));
  }
}

template <int ND, int A, typename R, class PW, int MODE>
__global__ __launch_bounds__(256) void acc_fixup_kernel(
    Consts k, Geom G, const R* __restrict__ cells, const R* __restrict__ fx,
    const R* __restrict__ acc, R* __restrict__ out, const double* __restrict__ dtp,
    double dx0, double dxA, int order, DeferList dl, double* slot) {
  const uint32_t n = *dl.count;
  if (n == 0) return;
  const double dt = *dtp;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n <= dl.cap) {
    for (int64_t i = t0; i < (int64_t)n; i += stride) {
      const int64_t of = (int64_t)(dl.items[i] & ((1ull << 40) - 1));
      acc_fix_cell<ND, A, R, PW, MODE>(k, G, cells, fx, acc, out, dt, dx0, dxA, order, of,
                                       slot);
      acc_fix_cell<ND, A, R, PW, MODE>(k, G, cells, fx, acc, out, dt, dx0, dxA, order,
                                       of + G.s[A], slot);
    }
    return;
  }
  const int64_t total = (int64_t)G.n[0] * G.n[1] * G.n[2];
  for (int64_t q = t0; q < total; q += stride) {
    const int i = (int)(q % G.n[0]) + 1;
    const int64_t r = q / G.n[0];
    acc_fix_cell<ND, A, R, PW, MODE>(k, G, cells, fx, acc, out, dt, dx0, dxA, order,
                                     goff(G, i, (int)(r % G.n[1]) + 1, (int)(r / G.n[1]) + 1),
                                     slot);
  }
}

// Empties the deferred-face counters of a step (and the extra CFL slot).
[[maybe_unused]] static __global__ void step_reset_kernel(uint32_t* counts, int n, double* slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int i = 0; i < n; ++i) counts[i] = 0u;
    if (slot) *slot = 0.0;
  }
}

// Conservative update of every interior cell from the stored fluxes.
// ORDER 2 (memory_intensive, …_solve_memory_intensive.f90:151-176):
//   u + ((dt/dx_x*(F_x[i-1]-F_x[i]) + dt/dx_y*(...)) + dt/dx_z*(...))
// ORDER 1 (cpu_intensive, …_solve_cpu_intensive.f90:117-136):
//   ((u + dt/dx_x*(...)) + dt/dx_y*(...)) + dt/dx_z*(...)
// FUSED: also writes the new state's ghost images for the next step and
// folds its CFL speeds into maxu_next.
template <int ND, int ORDER, bool FUSED, typename R = double>
__global__ __launch_bounds__(256) void update_kernel(Consts k, Geom G, R* __restrict__ cells,
                                                     const R* __restrict__ F0,
                                                     const R* __restrict__ F1,
                                                     const R* __restrict__ F2,
                                                     const double* __restrict__ dtp, double dx0,
                                                     double dx1, double dx2, BcTypes bc,
                                                     double* __restrict__ partial) {
  constexpr int NV = ND + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)G.n[0] * G.n[1] * G.n[2];
  double speed = 0.0;
  if (t < total) {
    const int i = (int)(t % G.n[0]) + 1;
    const int64_t r = t / G.n[0];
    const int j = (int)(r % G.n[1]) + 1;
    const int kk = (int)(r / G.n[1]) + 1;
    const int64_t o = goff(G, i, j, kk);
    const double dt = *dtp;
    const R dtdx0 = (R)(dt / dx0);
    const R dtdx1 = ND > 1 ? (R)(dt / dx1) : R(0);
    const R dtdx2 = ND > 2 ? (R)(dt / dx2) : R(0);
    R u[NV];
    load_state<ND, R>(cells, G.plane, o, u);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const int64_t oc = o + c * G.plane;
      R a0 = dtdx0 * (F0[oc - G.s[0]] - F0[oc]);
      if (ORDER == 2) {
        R acc = a0;
        if (ND > 1) acc = acc + dtdx1 * (F1[oc - G.s[1]] - F1[oc]);
        if (ND > 2) acc = acc + dtdx2 * (F2[oc - G.s[2]] - F2[oc]);
        u[c] = u[c] + acc;
      } else {
        u[c] = u[c] + a0;
        if (ND > 1) u[c] = u[c] + dtdx1 * (F1[oc - G.s[1]] - F1[oc]);
        if (ND > 2) u[c] = u[c] + dtdx2 * (F2[oc - G.s[2]] - F2[oc]);
      }
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) cells[o + c * G.plane] = u[c];
    if (FUSED) {
      write_ghost_images<ND, R>(G, bc, cells, i, j, kk, u);
      double ud[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) ud[c] = (double)u[c];
      speed = cfl_speed<ND>(k, ud);
    }
  }
  if (FUSED) block_max_to_partial(speed, partial);
}

// Stand-alone ghost fill, one thread per face cell; blockIdx.y = face
// (rhyme_samr_bc_set_boundaries.f90:9-20 calls the faces in this order; the
// faces are disjoint and read interior cells only, so they run together).
template <int ND, typename R = double>
__global__ __launch_bounds__(256) void bc_kernel(Geom G, BcTypes bc, const double* __restrict__ inflow,
                                                 int32_t ncmp, R* __restrict__ cells,
                                                 R* __restrict__ extra, int32_t face_only) {
  const int face = face_only >= 0 ? face_only : (int)blockIdx.y;
  const int a = face / 2, side = face % 2;
  if (a >= ND) return;
  const int o1 = a == 0 ? 1 : 0;
  const int o2 = a == 2 ? 1 : 2;
  const int n1 = G.n[o1], n2 = G.n[o2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n1 * n2) return;
  int idx[3];
  idx[o1] = (int)(t % n1) + 1;
  idx[o2] = (int)(t / n1) + 1;
  const int n = G.n[a];
  const int type = bc.t[face];
  const int g0 = side ? n + 1 : 0, g1 = side ? n + 2 : -1;
  int gi[3] = {idx[0], idx[1], idx[2]};
  gi[a] = g0;
  const int64_t og0 = goff(G, gi[0], gi[1], gi[2]);
  gi[a] = g1;
  const int64_t og1 = goff(G, gi[0], gi[1], gi[2]);
  if (type == 4) {
    for (int c = 0; c < ncmp; ++c) {
      R* pc = comp_plane<R>(cells, extra, ND + 2, c, G.plane);
      pc[og0] = (R)inflow[face * 16 + c];
      pc[og1] = (R)inflow[face * 16 + c];
    }
    return;
  }
  if (type < 1 || type > 3) return;
  int s0, s1;
  if (type == 3) {
    s0 = side ? 1 : n;
    s1 = side ? 2 : n - 1;
  } else {
    s0 = side ? n : 1;
    s1 = side ? n - 1 : 2;
  }
  gi[a] = s0;
  const int64_t os0 = goff(G, gi[0], gi[1], gi[2]);
  gi[a] = s1;
  const int64_t os1 = goff(G, gi[0], gi[1], gi[2]);
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) {
    R v0 = cells[os0 + c * G.plane];
    R v1 = cells[os1 + c * G.plane];
    if (type == 1 && c == 1 + a) {
      v0 = -v0;
      v1 = -v1;
    }
    cells[og0 + c * G.plane] = v0;
    cells[og1 + c * G.plane] = v1;
  }
}

// max over the interior of |v| + c_s (rhyme_cfl_time_step.f90:40-47).
template <int ND, typename R = double>
__global__ __launch_bounds__(256) void cfl_kernel(Consts k, Geom G, const R* __restrict__ cells,
                                                  double* __restrict__ partial) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)G.n[0] * G.n[1] * G.n[2];
  double v = 0.0;
  if (t < total) {
    const int i = (int)(t % G.n[0]) + 1;
    const int64_t r = t / G.n[0];
    const int j = (int)(r % G.n[1]) + 1;
    const int kk = (int)(r / G.n[1]) + 1;
    R ur[ND + 2];
    load_state<ND, R>(cells, G.plane, goff(G, i, j, kk), ur);
    double u[ND + 2];
#pragma unroll
    for (int c = 0; c < ND + 2; ++c) u[c] = (double)ur[c];
    v = cfl_speed<ND>(k, u);
  }
  block_max_to_partial(v, partial);
}

// strided_max: the max of p[begin, end) over the threads of a block, thread
// tid taking every stride-th value from begin + tid, from 0 with the strict
// `>` update (NaN never replaces the running value).  Eight loads are in
// flight before they are combined, so a thread does not wait one memory
// latency per value.
__device__ __forceinline__ double strided_max(const double* __restrict__ p, int64_t begin,
                                              int64_t end, int tid, int stride) {
  double v = 0.0;
  int64_t i = begin + tid;
  for (; i + 7 * (int64_t)stride < end; i += 8 * (int64_t)stride) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = p[i + u * (int64_t)stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) v = x[u] > v ? x[u] : v;
  }
  for (; i < end; i += stride) v = p[i] > v ? p[i] : v;
  return v;
}

// Final max over the per-block partials (one block of 1024 threads), then
// dt = courant*minval(dx)/max_u (rhyme_cfl_time_step.f90:49); the product
// courant*minval(dx) is formed on the host.  Writes dt (and logs it) and the
// raw max_u bits to maxu_out when given.
// reset (with stepc): after the reduction has read every partial, also the
// step's deferred-face counters (4) and the fixup's partial slot, as
// step_reset_kernel would before the step's walks.
[[maybe_unused]] static __global__ __launch_bounds__(1024) void dt_kernel(
    const double* __restrict__ partial, int32_t n, double cmdx, double* dt_out, double* dt_log,
    double* maxu_out, double dx0 = 1.0, double dx1 = 1.0, double dx2 = 1.0,
    double* stepc = nullptr, uint32_t* reset_counts = nullptr, double* reset_slot = nullptr) {
  __shared__ double wave_max[16];
  double v = strided_max(partial, 0, n, (int)threadIdx.x, (int)blockDim.x);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = wave_max[i] > m ? wave_max[i] : m;
    if (maxu_out) *maxu_out = m;
    if (dt_out) {
      double dt = cmdx / m;
      *dt_out = dt;
      if (dt_log) *dt_log = dt;
      if (stepc) {
        // the walks' dt/dx and (.5 dt)/dx, as they would form them
        const double dx[3] = {dx0, dx1, dx2};
        stepc[0] = dt;
        for (int a = 0; a < 3; ++a) {
          stepc[1 + a] = dt / dx[a];
          stepc[4 + a] = .5 * dt / dx[a];
        }
      }
    }
    if (reset_counts)
      for (int i = 0; i < 4; ++i) reset_counts[i] = 0u;
    if (reset_slot) *reset_slot = 0.0;
  }
}

[[maybe_unused]] static __global__ void set_scalar_kernel(double* dst, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// A given dt and the step constants the walks read, as dt_kernel forms them.
[[maybe_unused]] static __global__ void set_step_kernel(double* dt_out, double* stepc, double dt,
                                                        double dx0, double dx1, double dx2,
                                                        uint32_t* reset_counts,
                                                        double* reset_slot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < 4; ++i) reset_counts[i] = 0u;  // as step_reset_kernel
  *reset_slot = 0.0;
  const double dx[3] = {dx0, dx1, dx2};
  *dt_out = dt;
  stepc[0] = dt;
  for (int a = 0; a < 3; ++a) {
    stepc[1 + a] = dt / dx[a];
    stepc[4 + a] = .5 * dt / dx[a];
  }
}

[[maybe_unused]] static __global__ void set_u32_kernel(uint32_t* dst, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// -------------------------------------------------------- diagnostics ----

// (hi, lo) += (bh, bl) in double-double (Knuth's two-sum on the high parts;
// needs -ffp-contract=off, which the whole library is built with).
__host__ __device__ inline void dd_add(double& hi, double& lo, double bh, double bl) {
  const double s = hi + bh;
  const double bb = s - hi;
  double e = (hi - (s - bb)) + (bh - bb);
  e = e + (lo + bl);
  hi = s + e;
  lo = e - (hi - s);
}

// A workgroup's share of the diagnostics (see rhy_diag_t).
struct DiagPartial {
  double sum_hi[4], sum_lo[4];  // rho, e_tot (interior); rho > 0, e_tot > 0 (all)
  double min_v[2];              // rho, e_tot (interior, NaN skipped)
  int64_t min_at[2];            // linear interior index of the minimum, -1: none
  int64_t neg[2];               // rho < 0, e_tot < 0 (interior)
  int64_t nan[RHY_MAX_CMP];     // per component (interior)
};

__device__ __forceinline__ void min_take(double& v, int64_t& at, double ov, int64_t oat) {
  // smaller value wins; on a tie the earlier cell (Fortran element order)
  if (oat >= 0 && (at < 0 || ov < v || (ov == v && oat < at))) {
    v = ov;
    at = oat;
  }
}

template <int ND, typename R>
__global__ __launch_bounds__(256) void diag_kernel(Geom G, int32_t ncmp,
                                                   const R* __restrict__ cells,
                                                   const R* __restrict__ extra,
                                                   DiagPartial* __restrict__ out) {
  constexpr int NV = ND + 2;
  DiagPartial p;
#pragma unroll
  for (int s = 0; s < 4; ++s) p.sum_hi[s] = p.sum_lo[s] = 0.0;
  p.min_v[0] = p.min_v[1] = 0.0;
  p.min_at[0] = p.min_at[1] = -1;
  p.neg[0] = p.neg[1] = 0;
#pragma unroll
  for (int c = 0; c < RHY_MAX_CMP; ++c) p.nan[c] = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < G.plane; t += stride) {
    int q[3];
    bool inside = true;
    int64_t rem = t;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int ext = G.n[d] + 2 * G.g[d];
      q[d] = (int)(rem % ext) - G.g[d];
      rem /= ext;
      inside = inside && q[d] >= 0 && q[d] < G.n[d];
    }
    const double rho = (double)cells[t];
    const double e = (double)cells[t + (int64_t)(NV - 1) * G.plane];
    if (rho > 0.0) dd_add(p.sum_hi[2], p.sum_lo[2], rho, 0.0);
    if (e > 0.0) dd_add(p.sum_hi[3], p.sum_lo[3], e, 0.0);
    if (!inside) continue;
    const int64_t lin = ((int64_t)q[2] * G.n[1] + q[1]) * G.n[0] + q[0];
    dd_add(p.sum_hi[0], p.sum_lo[0], rho, 0.0);
    dd_add(p.sum_hi[1], p.sum_lo[1], e, 0.0);
    if (rho < 0.0) p.neg[0] += 1;
    if (e < 0.0) p.neg[1] += 1;
    if (rho == rho) min_take(p.min_v[0], p.min_at[0], rho, lin);
    if (e == e) min_take(p.min_v[1], p.min_at[1], e, lin);
#pragma unroll
    for (int c = 0; c < RHY_MAX_CMP; ++c) {
      if (c < ncmp) {
        const double v = (double)comp_plane<const R>(cells, extra, NV, c, G.plane)[t];
        if (v != v) p.nan[c] += 1;
      }
    }
  }
  // wave, then workgroup reduction in a fixed order
  __shared__ DiagPartial wave_part[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double oh = __shfl_xor(p.sum_hi[s], off, 64);
      const double ol = __shfl_xor(p.sum_lo[s], off, 64);
      dd_add(p.sum_hi[s], p.sum_lo[s], oh, ol);
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const double ov = __shfl_xor(p.min_v[m], off, 64);
      const long long oat = __shfl_xor((long long)p.min_at[m], off, 64);
      min_take(p.min_v[m], p.min_at[m], ov, oat);
      p.neg[m] += __shfl_xor((long long)p.neg[m], off, 64);
    }
#pragma unroll
    for (int c = 0; c < RHY_MAX_CMP; ++c) p.nan[c] += __shfl_xor((long long)p.nan[c], off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wave_part[w] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    DiagPartial r = wave_part[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      const DiagPartial& o = wave_part[i];
      for (int s = 0; s < 4; ++s) dd_add(r.sum_hi[s], r.sum_lo[s], o.sum_hi[s], o.sum_lo[s]);
      for (int m = 0; m < 2; ++m) {
        min_take(r.min_v[m], r.min_at[m], o.min_v[m], o.min_at[m]);
        r.neg[m] += o.neg[m];
      }
      for (int c = 0; c < RHY_MAX_CMP; ++c) r.nan[c] += o.nan[c];
    }
    out[blockIdx.x] = r;
  }
}

// Host: the workgroup partials combined in order, into the C struct.
inline void diag_finish(const std::vector<DiagPartial>& parts, const Geom& G,
                        rhy_diag_t* out) {
  double hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0};
  double mv[2] = {0, 0};
  int64_t mat[2] = {-1, -1}, neg[2] = {0, 0}, nan[RHY_MAX_CMP] = {0};
  for (const DiagPartial& p : parts) {
    for (int s = 0; s < 4; ++s) dd_add(hi[s], lo[s], p.sum_hi[s], p.sum_lo[s]);
    for (int m = 0; m < 2; ++m) {
      if (p.min_at[m] >= 0 &&
          (mat[m] < 0 || p.min_v[m] < mv[m] || (p.min_v[m] == mv[m] && p.min_at[m] < mat[m]))) {
        mv[m] = p.min_v[m];
        mat[m] = p.min_at[m];
      }
      neg[m] += p.neg[m];
    }
    for (int c = 0; c < RHY_MAX_CMP; ++c) nan[c] += p.nan[c];
  }
  *out = rhy_diag_t{};
  for (int c = 0; c < RHY_MAX_CMP; ++c) out->nan_count[c] = nan[c];
  out->neg_rho = neg[0];
  out->neg_e_tot = neg[1];
  out->mass = hi[0] + lo[0];
  out->energy = hi[1] + lo[1];
  out->mass_masked = hi[2] + lo[2];
  out->energy_masked = hi[3] + lo[3];
  double* mins[2] = {&out->min_rho, &out->min_e_tot};
  int32_t* ats[2] = {out->min_rho_at, out->min_e_tot_at};
  for (int m = 0; m < 2; ++m) {
    *mins[m] = mat[m] >= 0 ? mv[m] : NAN;
    int64_t r = mat[m] < 0 ? 0 : mat[m];
    for (int d = 0; d < 3; ++d) {
      ats[m][d] = mat[m] < 0 ? 0 : (int32_t)(r % G.n[d]) + 1;
      r /= G.n[d];
    }
  }
}

// ------------------------------------------------- initial conditions ----

// SplitMix64 (G. Steele, D. Lea, C. Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014, with the output constants of S. Vigna's
// splitmix64.c): draw number `counter` (0-based) of the stream `seed`.
__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A painted initial state (kind 1: two states split along `axis`; kind 2:
// blast over a jittered medium), with the host-side derived values.
struct PaintIC {
  int32_t kind, axis;
  double gm1;  // gamma - 1
  double rho_l, p_l, rho_r, p_r;
  double r2, jitter, p0, p_in;
  uint64_t seed;
};

// Offset of the centre of 0-based cell q from the centre of an n-cell axis.
__host__ __device__ inline double centre_offset(int q, int n) {
  return ((double)q + 0.5) - (double)n / 2.0;
}

// Host: the derived blast pressure, (gamma-1) * energy / (N_in * dx^ndim)
// with dx = 1/max(dims) and N_in the cells within `radius` of the centre
// (counted over the window that can hold them).
inline PaintIC make_paint(int32_t kind, int32_t axis, double gamma, double rho_l, double p_l,
                          double rho_r, double p_r, double radius, double jitter, double p0,
                          double energy, uint64_t seed, int ndim, const int32_t* dims) {
  PaintIC ic{};
  ic.kind = kind;
  ic.axis = axis;
  ic.gm1 = gamma - 1.0;
  ic.rho_l = rho_l;
  ic.p_l = p_l;
  ic.rho_r = rho_r;
  ic.p_r = p_r;
  ic.r2 = radius * radius;
  ic.jitter = jitter;
  ic.p0 = p0;
  ic.seed = seed;
  if (kind == 2) {
    int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
      const double c = dims[d] / 2.0 - 0.5;
      lo[d] = (int)std::floor(c - radius) - 1;
      hi[d] = (int)std::ceil(c + radius) + 2;
      lo[d] = lo[d] < 0 ? 0 : lo[d];
      hi[d] = hi[d] > dims[d] ? dims[d] : hi[d];
    }
    int64_t n_in = 0;
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) {
          const double x = centre_offset(i, dims[0]);
          double rr = x * x;
          if (ndim > 1) {
            const double y = centre_offset(j, dims[1]);
            rr = rr + y * y;
          }
          if (ndim > 2) {
            const double z = centre_offset(k, dims[2]);
            rr = rr + z * z;
          }
          n_in += rr <= ic.r2;
        }
    int m = dims[0];
    for (int d = 1; d < ndim; ++d) m = dims[d] > m ? dims[d] : m;
    const double dx = 1.0 / m;
    ic.p_in = (gamma - 1.0) * energy / ((double)n_in * std::pow(dx, ndim));
  }
  return ic;
}

// Every cell of every component plane: interior cells get the painted
// conserved state (zero velocity; non-hydro components 0), ghosts 0.
template <int ND, typename R>
__global__ __launch_bounds__(256) void paint_kernel(Geom G, PaintIC ic, int32_t ncmp,
                                                    R* __restrict__ cells, R* __restrict__ extra) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  int q[3];
  bool inside = true;
  int64_t rem = t;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int ext = G.n[d] + 2 * G.g[d];
    q[d] = (int)(rem % ext) - G.g[d];
    rem /= ext;
    inside = inside && q[d] >= 0 && q[d] < G.n[d];
  }
  double rho = 0.0, p = 0.0;
  if (inside) {
    if (ic.kind == 1) {
      const bool left = q[ic.axis] < G.n[ic.axis] / 2;
      rho = left ? ic.rho_l : ic.rho_r;
      p = left ? ic.p_l : ic.p_r;
    } else {
      const uint64_t lin = ((uint64_t)q[2] * G.n[1] + q[1]) * G.n[0] + q[0];
      const double u = (double)(splitmix64(ic.seed, lin) >> 11) * 0x1p-53;
      rho = 1.0 + ic.jitter * (-1.0 + 2.0 * u);
      const double x = centre_offset(q[0], G.n[0]);
      double rr = x * x;
      if (ND > 1) {
        const double y = centre_offset(q[1], G.n[1]);
        rr = rr + y * y;
      }
      if (ND > 2) {
        const double z = centre_offset(q[2], G.n[2]);
        rr = rr + z * z;
      }
      p = rr <= ic.r2 ? ic.p_in : ic.p0;
    }
  }
  for (int c = 0; c < ncmp; ++c) {
    double v = 0.0;
    if (inside) {
      if (c == 0)
        v = rho;
      else if (c <= ND)
        v = rho * 0.0;
      else if (c == ND + 1)
        v = .5 * rho * 0.0 + p / ic.gm1;
    }
    comp_plane<R>(cells, extra, ND + 2, c, G.plane)[t] = (R)v;
  }
}

// Drawing (rhyme_drawing_init.f90:19-78): canvas, sharp shapes in order,
// then the harmonic perturbations added in conserved form.
struct DrawShape {
  int32_t kind;
  int32_t lo[3], hi[3];
  double c[3], r2;
  double w[5];  // rho, v[3], p
};
struct DrawHarmonic {
  int32_t axis;
  int64_t off;  // first value of its profile in the profile array
  double w[5];
};
struct DrawCanvas {
  double w[5];
  double gm1;
  int32_t nshapes, nharmonics;
};

// conv_prim_vars_to_cons (rhyme_ideal_gas_primitive_vars_to_conserved.f90:9-11)
// of w = (rho, v[ND], p) into u
template <int ND>
__device__ __forceinline__ void draw_cons(const double* w, double gm1, double* u) {
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) s = s + w[1 + d] * w[1 + d];
  u[0] = w[0];
#pragma unroll
  for (int d = 0; d < ND; ++d) u[1 + d] = w[0] * w[1 + d];
  u[ND + 1] = .5 * w[0] * s + w[4] / gm1;
}

template <int ND, typename R>
__global__ __launch_bounds__(256) void draw_kernel(Geom G, DrawCanvas cv,
                                                   const DrawShape* __restrict__ shapes,
                                                   const DrawHarmonic* __restrict__ harms,
                                                   const double* __restrict__ profiles,
                                                   int32_t ncmp, R* __restrict__ cells,
                                                   R* __restrict__ extra) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  int q[3];
  bool inside = true;
  int64_t rem = t;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int ext = G.n[d] + 2 * G.g[d];
    q[d] = (int)(rem % ext) - G.g[d];
    rem /= ext;
    inside = inside && q[d] >= 0 && q[d] < G.n[d];
  }
  double u[ND + 2];
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) u[c] = 0.0;
  if (inside) {
    double w[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) w[c] = cv.w[c];
    for (int s = 0; s < cv.nshapes; ++s) {
      const DrawShape& S = shapes[s];
      bool in = true;
      if (S.kind == 1) {
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const double x = ((double)(q[d] + 1) - 0.5) - S.c[d];
          r2 = d == 0 ? x * x : r2 + x * x;
        }
        in = r2 <= S.r2;
      } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) in = in && q[d] + 1 >= S.lo[d] && q[d] + 1 <= S.hi[d];
      }
      if (in) {
#pragma unroll
        for (int c = 0; c < 5; ++c) w[c] = S.w[c];
      }
    }
    draw_cons<ND>(w, cv.gm1, u);
    if (cv.nharmonics > 0) {
      double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int m = 0; m < cv.nharmonics; ++m) {
        const DrawHarmonic& H = harms[m];
        const double a = profiles[H.off + q[H.axis]];
#pragma unroll
        for (int c = 0; c < 5; ++c) h[c] = h[c] + a * H.w[c];
      }
      double up[ND + 2];
      draw_cons<ND>(h, cv.gm1, up);
#pragma unroll
      for (int c = 0; c < ND + 2; ++c) u[c] = u[c] + up[c];
    }
  }
  for (int c = 0; c < ncmp; ++c)
    comp_plane<R>(cells, extra, ND + 2, c, G.plane)[t] = (R)(c < ND + 2 ? u[c] : 0.0);
}

// ------------------------------------------------------ unit batches ----

// Run body(integral_constant A) for the 0-based axis of a 1-based axis id;
// false if the id is outside 1..ND.
template <int ND, class F>
__device__ __forceinline__ bool with_axis(int a, F&& body) {
  if (a == 1) {
    body(std::integral_constant<int, 0>{});
    return true;
  }
  if constexpr (ND > 1)
    if (a == 2) {
      body(std::integral_constant<int, 1>{});
      return true;
    }
  if constexpr (ND > 2)
    if (a == 3) {
      body(std::integral_constant<int, 2>{});
      return true;
    }
  return false;
}

// rhyme_riemann_problem_star.f90:3-66 on a flattened solution record: p*
// from rhyme_irs_exec (floored at w_vacuum(p), :137), u* (:139-143), and for
// each side the shock density and speed, or the fan density, sound speed,
// head and tail speeds (:145-170).  The flux kernels form the same values,
// for the sampled side only, inside riemann_sample.
template <int ND, int A, class PW>
__device__ __forceinline__ void rp_star(const Consts& k, rhy_rp_solution_t& s) {
  double lf = 0.0, rf = 0.0;
  double ps =
      irs_exec<PW>(k, s.l_rho, s.l_v[A], s.l_p, s.l_cs, s.r_rho, s.r_v[A], s.r_p, s.r_cs, lf, rf);
  s.l_f = lf;
  s.r_f = rf;
  if (ps < k.w_p) ps = k.w_p;
  s.star_p = ps;
  s.star_u = 0.5 * ((s.r_v[A] + s.l_v[A]) + (rf - lf));
  const double ps_pl = PW::dv(ps, s.l_p);
  const double ps_pr = PW::dv(ps, s.r_p);
  if (ps > s.l_p) {
    s.l_is_shock = 1;
    s.l_shock_rho = PW::dv(s.l_rho * (k.gm1_gp1 + ps_pl), k.gm1_gp1 * ps_pl + 1.0);
    s.l_shock_speed = s.l_v[A] - s.l_cs * PW::sq(k.gp1_2g * ps_pl + k.gm1_2g);
  } else {
    s.l_is_shock = 0;
    double xa, xb;
    PW::pw2(ps_pl, k.gm1_2g, k.g_inv, xa, xb);
    s.l_fan_rho = s.l_rho * xb;
    s.l_fan_cs = s.l_cs * xa;
    s.l_fan_speedH = s.l_v[A] - s.l_cs;
    s.l_fan_speedT = s.star_u - s.l_fan_cs;
  }
  if (ps > s.r_p) {
    s.r_is_shock = 1;
    s.r_shock_rho = PW::dv(s.r_rho * (k.gm1_gp1 + ps_pr), k.gm1_gp1 * ps_pr + 1.0);
    s.r_shock_speed = s.r_v[A] + s.r_cs * PW::sq(k.gp1_2g * ps_pr + k.gm1_2g);
  } else {
    s.r_is_shock = 0;
    double xa, xb;
    PW::pw2(ps_pr, k.gm1_2g, k.g_inv, xa, xb);
    s.r_fan_rho = s.r_rho * xb;
    s.r_fan_cs = s.r_cs * xa;
    s.r_fan_speedH = s.r_v[A] + s.r_cs;
    s.r_fan_speedT = s.star_u + s.r_fan_cs;
  }
}

// The conserved state of one region of a solved Riemann problem (region
// ids RHY_W_*): w_k of a side, the star states behind a shock or a fan
// (tangential velocity of that side, normal velocity u*), or the state
// inside a fan at x/t = dxdt (w_kfan_prim, the flux kernels' routine).
template <int ND, int A, class PW>
__device__ __forceinline__ void rp_region(const Consts& k, const rhy_rp_solution_t& s,
                                          int region, double dxdt, double* u) {
  const bool right = region == RHY_W_R || region == RHY_W_STARR_SHO ||
                     region == RHY_W_STARR_FAN || region == RHY_W_RFAN;
  double v[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) v[d] = right ? s.r_v[d] : s.l_v[d];
  double rho = right ? s.r_rho : s.l_rho;
  double p = right ? s.r_p : s.l_p;
  switch (region) {
    case RHY_W_STARL_SHO:
    case RHY_W_STARR_SHO:
      rho = right ? s.r_shock_rho : s.l_shock_rho;
      v[A] = s.star_u;
      p = s.star_p;
      break;
    case RHY_W_STARL_FAN:
    case RHY_W_STARR_FAN:
      rho = right ? s.r_fan_rho : s.l_fan_rho;
      v[A] = s.star_u;
      p = s.star_p;
      break;
    case RHY_W_LFAN:
    case RHY_W_RFAN: {
      double sv[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) sv[d] = v[d];
      w_kfan_prim<ND, A, PW>(k, rho, sv, p, right ? s.r_cs : s.l_cs, dxdt, right, rho, v, p);
      break;
    }
    default:  // RHY_W_L, RHY_W_R
      break;
  }
  prim_to_cons<ND, double>(k, rho, v, p, u);
}

// rhyme_riemann_problem_sampling.f90:13-47: the region of x/t = dxdt.
template <int ND, int A, class PW>
__device__ __forceinline__ void rp_sampling(const Consts& k, const rhy_rp_solution_t& s,
                                            double dxdt, double* u) {
  int region;
  if (dxdt > s.star_u) {
    if (s.r_is_shock)
      region = dxdt < s.r_shock_speed ? RHY_W_STARR_SHO : RHY_W_R;
    else
      region = dxdt < s.r_fan_speedT ? RHY_W_STARR_FAN
                                     : (dxdt < s.r_fan_speedH ? RHY_W_RFAN : RHY_W_R);
  } else {
    if (s.l_is_shock)
      region = dxdt > s.l_shock_speed ? RHY_W_STARL_SHO : RHY_W_L;
    else
      region = dxdt > s.l_fan_speedT ? RHY_W_STARL_FAN
                                     : (dxdt > s.l_fan_speedH ? RHY_W_LFAN : RHY_W_L);
  }
  rp_region<ND, A, PW>(k, s, region, dxdt, u);
}

// rhyme_riemann_problem_solve at x/t = dx[q] / dt[q] (0: the kernels' case),
// with the flux kernels' full solver policy.
template <int ND>
__global__ __launch_bounds__(256) void riemann_batch_kernel(Consts k, int64_t n,
                                                            const double* __restrict__ l,
                                                            const double* __restrict__ r,
                                                            const int32_t* __restrict__ axis,
                                                            const double* __restrict__ dx,
                                                            const double* __restrict__ dt,
                                                            double* __restrict__ u) {
  constexpr int NV = ND + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double L[NV], R[NV], U[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    L[c] = l[q * NV + c];
    R[c] = r[q * NV + c];
  }
  const double dxdt = dx ? dx[q] / dt[q] : 0.0;
  if (!with_axis<ND>(axis[q], [&](auto a) {
        riemann_state<ND, decltype(a)::value, PowF64Fix>(k, L, R, dxdt, U);
      })) {
#pragma unroll
    for (int c = 0; c < NV; ++c) U[c] = __builtin_nan("");
  }
#pragma unroll
  for (int c = 0; c < NV; ++c) u[q * NV + c] = U[c];
}

template <int ND>
__global__ __launch_bounds__(256) void rp_star_batch_kernel(Consts k, int64_t n,
                                                            rhy_rp_solution_t* __restrict__ sol,
                                                            const int32_t* __restrict__ axis) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  rhy_rp_solution_t s = sol[q];
  if (!with_axis<ND>(axis[q], [&](auto a) { rp_star<ND, decltype(a)::value, PowF64Fix>(k, s); }))
    s.star_p = s.star_u = __builtin_nan("");
  sol[q] = s;
}

// Sampling (dx != nullptr: x/t = dx[q] / dt[q]) or one given region
// (region != nullptr, x/t = dxdt[q] for the fans).
template <int ND>
__global__ __launch_bounds__(256) void rp_sample_batch_kernel(
    Consts k, int64_t n, const rhy_rp_solution_t* __restrict__ sol,
    const int32_t* __restrict__ axis, const double* __restrict__ dx,
    const double* __restrict__ dt, const int32_t* __restrict__ region,
    const double* __restrict__ dxdt, double* __restrict__ u) {
  constexpr int NV = ND + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const rhy_rp_solution_t s = sol[q];
  double U[NV];
  const bool ok = with_axis<ND>(axis[q], [&](auto a) {
    constexpr int A = decltype(a)::value;
    if (region)
      rp_region<ND, A, PowF64Fix>(k, s, region[q], dxdt[q], U);
    else
      rp_sampling<ND, A, PowF64Fix>(k, s, dx[q] / dt[q], U);
  });
  if (!ok) {
#pragma unroll
    for (int c = 0; c < NV; ++c) U[c] = __builtin_nan("");
  }
#pragma unroll
  for (int c = 0; c < NV; ++c) u[q * NV + c] = U[c];
}

// rhyme_irs_exec on records (rho, v(axis), p, cs); guesses != nullptr: the
// six starting guesses instead (rhyme_irs_guess_p_star: guess 1 as the
// flux kernels form it, guesses 2-6 as their fallback forms them).
__global__ __launch_bounds__(256) static void irs_batch_kernel(
    Consts k, int64_t n, const double* __restrict__ l, const double* __restrict__ r,
    double* __restrict__ p_star, double* __restrict__ l_f, double* __restrict__ r_f,
    double* __restrict__ guesses) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double lr = l[4 * q], lv = l[4 * q + 1], lp = l[4 * q + 2], lc = l[4 * q + 3];
  const double rr = r[4 * q], rv = r[4 * q + 1], rp = r[4 * q + 2], rc = r[4 * q + 3];
  if (guesses) {
    guesses[6 * q] = guess1<PowF64Fix>(k, lr, lv, lp, lc, rr, rv, rp, rc);
    for (int i = 1; i < 6; ++i)
      guesses[6 * q + i] = guess_n<PowF64Fix>(k, i, lr, lv, lp, lc, rr, rv, rp, rc);
    return;
  }
  double lf = 0.0, rf = 0.0;
  p_star[q] = irs_exec<PowF64Fix>(k, lr, lv, lp, lc, rr, rv, rp, rc, lf, rf);
  l_f[q] = lf;
  r_f[q] = rf;
}

// x**y under one of the library's policies (RHY_POW_*)
__global__ __launch_bounds__(256) static void pow_batch_kernel(int32_t policy, int64_t n,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v;
  if (policy == RHY_POW_LIBM)
    v = pow(x[q], y[q]);
  else if (policy == RHY_POW_MIXED)
    v = MixedMath::pw(x[q], y[q]);
  else
    v = PowF64Fix::pw(x[q], y[q]);
  out[q] = v;
}

template <int ND>
__global__ __launch_bounds__(256) void limiter_batch_kernel(Consts k, int64_t n,
                                                            const double* __restrict__ ul,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ ur,
                                                            double* __restrict__ d) {
  constexpr int NV = ND + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
#pragma unroll
  for (int c = 0; c < NV; ++c)
    d[q * NV + c] = limit_one<0>(k.limiter, k.omega, ul[q * NV + c], u[q * NV + c], ur[q * NV + c]);
}

// half step from a given slope (the reference routine takes delta as input)
template <int ND, int A>
__device__ __forceinline__ void half_step_given(const Consts& k, const double* u,
                                                const double* delta, double hdtdx, double* l,
                                                double* r) {
  constexpr int NV = ND + 2;
  double fl[NV], fr[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    l[c] = u[c] - .5 * delta[c];
    r[c] = u[c] + .5 * delta[c];
  }
  flux<ND, A>(k, l, fl);
  flux<ND, A>(k, r, fr);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    double df = fl[c] - fr[c];
    l[c] = l[c] + hdtdx * df;
    r[c] = r[c] + hdtdx * df;
  }
}

template <int ND>
__global__ __launch_bounds__(256) void half_step_batch_kernel(Consts k, int64_t n,
                                                              const double* __restrict__ u,
                                                              const double* __restrict__ delta,
                                                              const int32_t* __restrict__ axis,
                                                              double dx, double dt,
                                                              double* __restrict__ l,
                                                              double* __restrict__ r) {
  constexpr int NV = ND + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double U[NV], D[NV], Lo[NV], Ro[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    U[c] = u[q * NV + c];
    D[c] = delta[q * NV + c];
  }
  const double hdtdx = .5 * dt / dx;
  const int a = axis[q];
  if (a == 1) half_step_given<ND, 0>(k, U, D, hdtdx, Lo, Ro);
  else if (ND > 1 && a == 2) half_step_given<ND, (ND > 1 ? 1 : 0)>(k, U, D, hdtdx, Lo, Ro);
  else half_step_given<ND, (ND > 2 ? 2 : 0)>(k, U, D, hdtdx, Lo, Ro);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    l[q * NV + c] = Lo[c];
    r[q * NV + c] = Ro[c];
  }
}

template <int ND>
__global__ __launch_bounds__(256) void flux_batch_kernel(Consts k, int64_t n,
                                                         const double* __restrict__ u,
                                                         const int32_t* __restrict__ axis,
                                                         double* __restrict__ f) {
  constexpr int NV = ND + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double U[NV], Fo[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) U[c] = u[q * NV + c];
  const int a = axis[q];
  if (a == 1) flux<ND, 0>(k, U, Fo);
  else if (ND > 1 && a == 2) flux<ND, (ND > 1 ? 1 : 0)>(k, U, Fo);
  else flux<ND, (ND > 2 ? 2 : 0)>(k, U, Fo);
#pragma unroll
  for (int c = 0; c < NV; ++c) f[q * NV + c] = Fo[c];
}

// Snapshot payload of one box: every component's interior cells, i fastest,
// component after component, as FP32 (round to nearest even) -- the layout
// of a level's data:datatype=0 dataset in the reference's Chombo writer
// (rhyme_chombo_write_level_data.f90:41-71).  A workgroup covers 256 cells
// of one interior row (snapshot_grid); each thread moves all components of
// its cell, so every thread has ncmp loads in flight.
inline dim3 snapshot_grid(const Geom& G) {
  const unsigned chunks = (unsigned)((G.n[0] + 255) / 256);
  return dim3(chunks * (unsigned)G.n[1] * (unsigned)G.n[2]);
}

template <typename R>
__global__ __launch_bounds__(256) void snapshot_pack_kernel(Geom G, int ncmp, int nv,
                                                            const R* __restrict__ cells,
                                                            const R* __restrict__ extra,
                                                            float* __restrict__ out) {
  const int chunks = (G.n[0] + 255) >> 8;
  const int row = (int)blockIdx.x / chunks;
  const int i = ((int)blockIdx.x - row * chunks) * 256 + (int)threadIdx.x;
  if (i >= G.n[0]) return;
  const int j = row % G.n[1], k = row / G.n[1];
  const int64_t blen = (int64_t)G.n[0] * G.n[1] * G.n[2];
  const int64_t src = goff(G, i + 1, j + 1, k + 1), dst = (int64_t)row * G.n[0] + i;
  R v[RHY_MAX_CMP];
#pragma unroll
  for (int c = 0; c < RHY_MAX_CMP; ++c)
    if (c < ncmp) v[c] = comp_plane<const R>(cells, extra, nv, c, G.plane)[src];
#pragma unroll
  for (int c = 0; c < RHY_MAX_CMP; ++c)
    if (c < ncmp) out[c * blen + dst] = (float)v[c];
}

// The reverse (the restart reader, rhyme_initial_condition_load_rhyme.f90:57-76):
// interior cells from an FP32 payload, widened exactly; ghosts untouched.
template <typename R>
__global__ __launch_bounds__(256) void snapshot_unpack_kernel(Geom G, int ncmp, int nv,
                                                              R* __restrict__ cells,
                                                              R* __restrict__ extra,
                                                              const float* __restrict__ in) {
  const int chunks = (G.n[0] + 255) >> 8;
  const int row = (int)blockIdx.x / chunks;
  const int i = ((int)blockIdx.x - row * chunks) * 256 + (int)threadIdx.x;
  if (i >= G.n[0]) return;
  const int j = row % G.n[1], k = row / G.n[1];
  const int64_t blen = (int64_t)G.n[0] * G.n[1] * G.n[2];
  const int64_t dst = goff(G, i + 1, j + 1, k + 1), src = (int64_t)row * G.n[0] + i;
  float v[RHY_MAX_CMP];
#pragma unroll
  for (int c = 0; c < RHY_MAX_CMP; ++c)
    if (c < ncmp) v[c] = in[c * blen + src];
#pragma unroll
  for (int c = 0; c < RHY_MAX_CMP; ++c)
    if (c < ncmp) comp_plane<R>(cells, extra, nv, c, G.plane)[dst] = (R)v[c];
}

}  // namespace rhy
