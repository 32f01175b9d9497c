// rhy_batch.hip -- many independent boxes of uneven shapes on one device.
//
// A batch owns one HBM arena; every box (a "slot") gets a block of it for
// its state planes and its flux scratch.  A step is one launch per phase
// over all active boxes -- dt of every box, the fluxes of every axis of
// every box, the update of every box -- driven by a tile table that
// persistent workgroups stride over; the step can be captured once in a
// hipGraph and replayed.  Each box keeps its own ghost fill and its own CFL
// dt, i.e. it evolves exactly as it would alone (rhy_run on that box).
//
// A batch made by rhy_batch_create has a fixed set of boxes and takes the
// limiter and face types from the run's parameters.  A pool (rhy_pool_create)
// starts empty with a byte budget: boxes are added (first fit in the arena,
// refused when it does not fit) and removed between runs, each with its own
// limiter and face types -- the continuous-batching engine of the serving
// loop admits jobs through it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>
#include <map>
#include <string>
#include <vector>

#include "../../include/rhyme_hip.h"
#include "rhy_kernels.hpp"

using namespace rhy;

namespace {

thread_local std::string b_err;

int bfail(int code, const std::string& msg) {
  b_err = msg;
  return code;
}

#define BHIP_TRY(expr)                                                                \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail(RHY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

// per-box descriptor, device resident (indexed by slot)
struct BBox {
  Geom G;
  int64_t cell_off;  // element offset of the box's state planes in the arena
  int64_t flux_off;  // element offset of its flux planes
  double dx[3];
  double cmdx;       // courant * minval(dx)
  BcTypes bc;        // face types of this box
  int32_t limiter;   // slope limiter of this box
  int32_t part_begin, part_end;  // its update tiles (CFL partials)
};

// A tile is one workgroup's share: `count` consecutive threads, starting
// at `base`, of box `box`'s index space for phase `kind` (0..2: flux along
// that axis, 3: update).
struct BTile {
  int32_t box;
  int32_t kind;
  int64_t base;
  int32_t count;
};

// A segment is the run of tiles that covers one phase of one box: tiles
// [start, next segment's start) take blockDim.x consecutive threads each of
// the `threads` threads of box `box`'s phase `kind`.  Segments are stored in
// tile order (all flux phases, then the updates), a few per box; a per-tile
// table of segment indices (4 bytes per tile) finds a tile's segment.
struct BSeg {
  int32_t start;
  int32_t box;
  int32_t kind;
  int32_t pad;
  int64_t threads;
};

// Tile ti: its segment from the per-tile index (one load), then its share.
__device__ __forceinline__ BTile tile_of(const BSeg* __restrict__ segs,
                                         const int32_t* __restrict__ tile_seg, int32_t ti) {
  const BSeg S = segs[tile_seg[ti]];
  BTile T;
  T.box = S.box;
  T.kind = S.kind;
  T.base = (int64_t)(ti - S.start) * blockDim.x;
  const int64_t left = S.threads - T.base;
  T.count = (int32_t)(left < (int64_t)blockDim.x ? left : (int64_t)blockDim.x);
  return T;
}

// LIM > 0: every box uses that limiter; LIM == 0: each box's own.
template <int ND, int LIM, typename R, class PW>
__global__ __launch_bounds__(256) RHY_FLUX_OCCUPANCY void batch_flux_kernel(Consts k, const BBox* __restrict__ boxes,
                                                         const BSeg* __restrict__ segs,
                                                         const int32_t* __restrict__ tile_seg,
                                                         int32_t ntiles,
                                                         R* __restrict__ arena,
                                                         const double* __restrict__ dt,
                                                         int32_t nf_x, int32_t nf_yz,
                                                         DeferList dl) {
  for (int32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const BTile T = tile_of(segs, tile_seg, ti);
    if ((int32_t)threadIdx.x >= T.count) continue;
    const BBox& B = boxes[T.box];
    Consts kb = k;
    if (LIM == 0) kb.limiter = B.limiter;
    const int64_t t = T.base + threadIdx.x;
    const R* c = arena + B.cell_off;
    R* F = arena + B.flux_off + (int64_t)T.kind * (ND + 2) * B.G.plane;
    const double d = dt[T.box];
    const uint32_t bx = (uint32_t)T.box;
    if (T.kind == 0)
      flux_walk<ND, 0, LIM, R, PW, false>(kb, B.G, c, F, d, B.dx[0], nf_x, t, dl, bx);
    if constexpr (ND > 1)
      if (T.kind == 1)
        flux_walk<ND, 1, LIM, R, PW, false>(kb, B.G, c, F, d, B.dx[1], nf_yz, t, dl, bx);
    if constexpr (ND > 2)
      if (T.kind == 2)
        flux_walk<ND, 2, LIM, R, PW, false>(kb, B.G, c, F, d, B.dx[2], nf_yz, t, dl, bx);
  }
}

// The faces the flux phase listed, recomputed with the full policy PW
// (each with its box's limiter, dt and dx); if the list overflowed, every
// face of every active box.
template <int ND, typename R, class PW>
__global__ __launch_bounds__(256) void batch_fixup_kernel(Consts k, const BBox* __restrict__ boxes,
                                                          const int32_t* __restrict__ active,
                                                          int32_t nactive, R* __restrict__ arena,
                                                          const double* __restrict__ dt,
                                                          DeferList dl) {
  const uint32_t n = *dl.count;
  if (n == 0) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n <= dl.cap) {
    for (int64_t i = t0; i < (int64_t)n; i += stride) {
      const uint64_t it = dl.items[i];
      const int32_t bx = (int32_t)(it >> 42);
      const BBox& B = boxes[bx];
      Consts kb = k;
      kb.limiter = B.limiter;
      flux_face_fix<ND, R, PW>(kb, B.G, arena + B.cell_off, arena + B.flux_off, dt[bx], B.dx,
                               (int)((it >> 40) & 3), (int64_t)(it & ((1ull << 40) - 1)));
    }
    return;
  }
  for (int32_t ai = 0; ai < nactive; ++ai) {
    const int32_t bx = active[ai];
    const BBox& B = boxes[bx];
    Consts kb = k;
    kb.limiter = B.limiter;
    for (int a = 0; a < ND; ++a) {
      const int64_t na = faces_of_axis(B.G, a);
      for (int64_t q = t0; q < na; q += stride)
        flux_face_fix<ND, R, PW>(kb, B.G, arena + B.cell_off, arena + B.flux_off, dt[bx], B.dx,
                                 a, face_low_cell(B.G, a, q));
    }
  }
}

// Update (memory_intensive or cpu_intensive order) with the fused ghost
// images and CFL partial of every box; tile ti writes partial[ti - first].
template <int ND, int ORDER, typename R>
__global__ __launch_bounds__(256) void batch_update_kernel(Consts k, const BBox* __restrict__ boxes,
                                                           const BSeg* __restrict__ segs,
                                                           const int32_t* __restrict__ tile_seg,
                                                           int32_t first,
                                                           int32_t ntiles,
                                                           R* __restrict__ arena,
                                                           const double* __restrict__ dt,
                                                           double* __restrict__ partial) {
  constexpr int NV = ND + 2;
  for (int32_t ti = first + blockIdx.x; ti < first + ntiles; ti += gridDim.x) {
    const BTile T = tile_of(segs, tile_seg, ti);
    const BBox& B = boxes[T.box];
    const Geom& G = B.G;
    double speed = 0.0;
    if ((int32_t)threadIdx.x < T.count) {
      const int64_t t = T.base + threadIdx.x;
      const int i = (int)(t % G.n[0]) + 1;
      const int64_t r = t / G.n[0];
      const int j = (int)(r % G.n[1]) + 1;
      const int kk = (int)(r / G.n[1]) + 1;
      const int64_t o = goff(G, i, j, kk);
      R* c = arena + B.cell_off;
      const R* F0 = arena + B.flux_off;
      const R* F1 = F0 + (int64_t)NV * G.plane;
      const R* F2 = F0 + (int64_t)2 * NV * G.plane;
      const double d = dt[T.box];
      const R dtdx0 = (R)(d / B.dx[0]);
      const R dtdx1 = ND > 1 ? (R)(d / B.dx[1]) : R(0);
      const R dtdx2 = ND > 2 ? (R)(d / B.dx[2]) : R(0);
      R u[NV];
      load_state<ND, R>(c, G.plane, o, u);
#pragma unroll
      for (int cc = 0; cc < NV; ++cc) {
        const int64_t oc = o + cc * G.plane;
        R a0 = dtdx0 * (F0[oc - G.s[0]] - F0[oc]);
        if (ORDER == 2) {
          R acc = a0;
          if (ND > 1) acc = acc + dtdx1 * (F1[oc - G.s[1]] - F1[oc]);
          if (ND > 2) acc = acc + dtdx2 * (F2[oc - G.s[2]] - F2[oc]);
          u[cc] = u[cc] + acc;
        } else {
          u[cc] = u[cc] + a0;
          if (ND > 1) u[cc] = u[cc] + dtdx1 * (F1[oc - G.s[1]] - F1[oc]);
          if (ND > 2) u[cc] = u[cc] + dtdx2 * (F2[oc - G.s[2]] - F2[oc]);
        }
      }
#pragma unroll
      for (int cc = 0; cc < NV; ++cc) c[o + cc * G.plane] = u[cc];
      write_ghost_images<ND, R>(G, B.bc, c, i, j, kk, u);
      double ud[NV];
#pragma unroll
      for (int cc = 0; cc < NV; ++cc) ud[cc] = (double)u[cc];
      speed = cfl_speed<ND>(k, ud);
    }
    block_max_to_index(speed, partial, ti - first);
    __syncthreads();
  }
}

// CFL partials of the current state (prologue), same tiling as the update.
template <int ND, typename R>
__global__ __launch_bounds__(256) void batch_cfl_kernel(Consts k, const BBox* __restrict__ boxes,
                                                        const BSeg* __restrict__ segs,
                                                        const int32_t* __restrict__ tile_seg,
                                                        int32_t first,
                                                        int32_t ntiles,
                                                        const R* __restrict__ arena,
                                                        double* __restrict__ partial) {
  for (int32_t ti = first + blockIdx.x; ti < first + ntiles; ti += gridDim.x) {
    const BTile T = tile_of(segs, tile_seg, ti);
    const BBox& B = boxes[T.box];
    const Geom& G = B.G;
    double v = 0.0;
    if ((int32_t)threadIdx.x < T.count) {
      const int64_t t = T.base + threadIdx.x;
      const int i = (int)(t % G.n[0]) + 1;
      const int64_t r = t / G.n[0];
      const int j = (int)(r % G.n[1]) + 1;
      const int kk = (int)(r / G.n[1]) + 1;
      R ur[ND + 2];
      load_state<ND, R>(arena + B.cell_off, G.plane, goff(G, i, j, kk), ur);
      double u[ND + 2];
#pragma unroll
      for (int c = 0; c < ND + 2; ++c) u[c] = (double)ur[c];
      v = cfl_speed<ND>(k, u);
    }
    block_max_to_index(v, partial, ti - first);
    __syncthreads();
  }
}

// dt of every active box (block b reduces the partials of box active[b]);
// logs it at dtlog[*step * log_stride + box] while the step counter is
// below the log size.
__global__ __launch_bounds__(256) void batch_dt_kernel(const BBox* __restrict__ boxes,
                                                       const int32_t* __restrict__ active,
                                                       const double* __restrict__ partial,
                                                       double* __restrict__ dt,
                                                       double* __restrict__ dtlog,
                                                       const int32_t* __restrict__ step,
                                                       int32_t log_steps, int32_t log_stride) {
  __shared__ double wave_max[4];
  const int b = active[blockIdx.x];
  const BBox& B = boxes[b];
  double v = strided_max(partial, B.part_begin, B.part_end, (int)threadIdx.x, (int)blockDim.x);
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
    const double d = B.cmdx / m;
    dt[b] = d;
    const int32_t s = *step;
    if (dtlog && s < log_steps) dtlog[(int64_t)s * log_stride + b] = d;
  }
}

__global__ void batch_step_counter_kernel(int32_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = *step + 1;
}

__global__ void batch_set_counter_kernel(int32_t* step, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}

// The given dt for every active box (the fine level of a hierarchy steps
// with the dt its coarse level sets), logged like batch_dt_kernel's.
__global__ __launch_bounds__(256) void batch_set_dt_kernel(const int32_t* __restrict__ active,
                                                           int32_t nactive, double value,
                                                           double* __restrict__ dt,
                                                           double* __restrict__ dtlog,
                                                           const int32_t* __restrict__ step,
                                                           int32_t log_steps, int32_t log_stride) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= nactive) return;
  const int b = active[i];
  dt[b] = value;
  const int32_t s = *step;
  if (dtlog && s < log_steps) dtlog[(int64_t)s * log_stride + b] = value;
}

// Two-level hierarchy, refinement ratio 2.  lo[d] is the coarse cell (1-based)
// that holds fine cells 1 and 2 along d, so fine cell f lies in coarse cell
// lo[d] + (f + 1) / 2 - 1 (f >= -1).
__device__ __forceinline__ int coarse_parent(int lo, int f) { return lo + (f + 1) / 2 - 1; }

// One work item of a hierarchy launch, which covers all fine boxes at once:
// a fine box (prolongation, restriction) or a region of fine ghosts inside
// a sibling (copy).  Item i takes threads [first_i, first_{i+1}).
struct AmrItem {
  int64_t first;
  int64_t dst_off;  // arena offset of the fine box written (prolong, copy) or read (restrict)
  int64_t src_off;  // arena offset of the sibling read (copy)
  Geom GD, GS;      // their geometries
  int32_t lo[3];    // placement (prolong, restrict); first dst cell (copy)
  int32_t slo[3];   // first src cell (copy)
  int32_t ext[3];   // region extent (copy)
};

__device__ __forceinline__ int amr_item_of(const AmrItem* __restrict__ items, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Face ghosts (2 layers, faces only: what the flux walks read) of every fine
// box from their coarse parents, piecewise constant in space.  half == 0:
// the parent's value (called before the coarse step, so the start of the
// coarse step).  half == 1: the mean of the ghost's value and the parent's
// (called after the coarse step and the first fine substep, which leaves
// the fine boxes' face ghosts as they were), i.e. 0.5 * U_old + 0.5 * U_new.
template <int ND, typename R>
__global__ __launch_bounds__(256) void amr_prolong_kernel(const AmrItem* __restrict__ items,
                                                          int nitems, int64_t total,
                                                          R* __restrict__ arena, Geom GC,
                                                          const R* __restrict__ coarse,
                                                          int half) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const AmrItem& it = items[amr_item_of(items, nitems, t0)];
  const Geom& GF = it.GD;
  R* fine = arena + it.dst_off;
  int64_t t = t0 - it.first;
  for (int a = 0; a < ND; ++a) {
    int64_t m = 4;  // 2 sides x 2 layers
    for (int d = 0; d < 3; ++d)
      if (d != a) m *= GF.n[d];
    if (t >= m) {
      t -= m;
      continue;
    }
    int idx[3];
    int64_t q = t;
    const int sl = (int)(q % 4);
    q /= 4;
    for (int d = 0; d < 3; ++d) {
      if (d == a) continue;
      idx[d] = (int)(q % GF.n[d]) + 1;
      q /= GF.n[d];
    }
    const int side = sl >> 1, layer = sl & 1;
    idx[a] = side ? GF.n[a] + 1 + layer : -layer;
    int pc[3] = {1, 1, 1};
    for (int d = 0; d < ND; ++d) pc[d] = coarse_parent(it.lo[d], idx[d]);
    const int64_t of = goff(GF, idx[0], idx[1], idx[2]);
    const int64_t oc = goff(GC, pc[0], pc[1], pc[2]);
#pragma unroll
    for (int c = 0; c < ND + 2; ++c) {
      const R v = coarse[oc + c * GC.plane];
      R* g = fine + of + c * GF.plane;
      *g = half ? R(0.5) * *g + R(0.5) * v : v;
    }
    return;
  }
}

// Regridding: every cell (interior and both ghost layers) of every new
// fine box (items [0, nto), threads over their ghosted cells) takes the old
// fine cell at the same fine-level position when an old box's interior
// holds it (items [nto, nto + nfrom): the old boxes, in the old arena),
// else its coarse parent's value; the ND + 2 hydro components.  In
// fine-level indices a box's cell f lies at 2 (lo - 1) + f.
template <int ND, typename R>
__global__ __launch_bounds__(256) void amr_regrid_kernel(const AmrItem* __restrict__ items,
                                                         int nto, int nfrom, int64_t total,
                                                         R* __restrict__ to_arena,
                                                         const R* __restrict__ from_arena,
                                                         Geom GC, const R* __restrict__ coarse) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const AmrItem& it = items[amr_item_of(items, nto, t0)];
  const Geom& GT = it.GD;
  int64_t q = t0 - it.first;
  int idx[3] = {1, 1, 1};
  for (int d = 0; d < ND; ++d) {
    const int e = GT.n[d] + 4;
    idx[d] = (int)(q % e) - 1;
    q /= e;
  }
  int gf[3] = {0, 0, 0};
  for (int d = 0; d < ND; ++d) gf[d] = 2 * (it.lo[d] - 1) + idx[d];
  const int64_t ot = goff(GT, idx[0], idx[1], idx[2]);
  const R* src = nullptr;
  int64_t os = 0, splane = 0;
  for (int j = 0; j < nfrom && !src; ++j) {
    const AmrItem& fj = items[nto + j];
    int sc[3] = {1, 1, 1};
    bool in = true;
    for (int d = 0; d < ND; ++d) {
      sc[d] = gf[d] - 2 * (fj.lo[d] - 1);
      in = in && sc[d] >= 1 && sc[d] <= fj.GD.n[d];
    }
    if (in) {
      src = from_arena + fj.dst_off;
      os = goff(fj.GD, sc[0], sc[1], sc[2]);
      splane = fj.GD.plane;
    }
  }
  if (!src) {
    int pc[3] = {1, 1, 1};
    for (int d = 0; d < ND; ++d) pc[d] = coarse_parent(it.lo[d], idx[d]);
    src = coarse;
    os = goff(GC, pc[0], pc[1], pc[2]);
    splane = GC.plane;
  }
  R* dst = to_arena + it.dst_off;
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) dst[ot + c * GT.plane] = src[os + c * splane];
}

// Regions of fine ghosts that lie in a sibling's interior, copied from it:
// ext[d] cells from src cell slo[d] to dst cell lo[d], the ND + 2 hydro
// components.
template <int ND, typename R>
__global__ __launch_bounds__(256) void amr_copy_kernel(const AmrItem* __restrict__ items,
                                                       int nitems, int64_t total,
                                                       R* __restrict__ arena) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const AmrItem& it = items[amr_item_of(items, nitems, t0)];
  const int64_t t = t0 - it.first;
  const int x = (int)(t % it.ext[0]), y = (int)(t / it.ext[0] % it.ext[1]),
            z = (int)(t / ((int64_t)it.ext[0] * it.ext[1]));
  const int64_t od = it.dst_off + goff(it.GD, it.lo[0] + x, it.lo[1] + y, it.lo[2] + z);
  const int64_t os = it.src_off + goff(it.GS, it.slo[0] + x, it.slo[1] + y, it.slo[2] + z);
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) arena[od + c * it.GD.plane] = arena[os + c * it.GS.plane];
}

// The coarse cells every fine box covers take the mean of their 2^ND
// children: the children summed i fastest, then times 2^-ND.
template <int ND, typename R>
__global__ __launch_bounds__(256) void amr_restrict_kernel(const AmrItem* __restrict__ items,
                                                           int nitems, int64_t total,
                                                           const R* __restrict__ arena, Geom GC,
                                                           R* __restrict__ coarse) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const AmrItem& it = items[amr_item_of(items, nitems, t0)];
  const Geom& GF = it.GD;
  const R* fine = arena + it.dst_off;
  const int64_t t = t0 - it.first;
  const int m0 = GF.n[0] / 2, m1 = ND > 1 ? GF.n[1] / 2 : 1;
  const int c0 = (int)(t % m0), c1 = (int)(t / m0 % m1), c2 = (int)(t / ((int64_t)m0 * m1));
  const int64_t oc = goff(GC, it.lo[0] + c0, ND > 1 ? it.lo[1] + c1 : 1,
                          ND > 2 ? it.lo[2] + c2 : 1);
  const R w = ND == 3 ? R(0.125) : ND == 2 ? R(0.25) : R(0.5);
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) {
    R sum = R(0);
    for (int z = 0; z < (ND > 2 ? 2 : 1); ++z)
      for (int y = 0; y < (ND > 1 ? 2 : 1); ++y)
        for (int x = 0; x < 2; ++x) {
          const int64_t of = goff(GF, 2 * c0 + 1 + x, ND > 1 ? 2 * c1 + 1 + y : 1,
                                  ND > 2 ? 2 * c2 + 1 + z : 1);
          sum = sum + fine[of + c * GF.plane];
        }
    coarse[oc + c * GC.plane] = sum * w;
  }
}

// Coarse faces on the boundary of the fine boxes: an item is one (box,
// axis a = slo[0], side s = slo[1]) with its Pi_{d != a} n_d / 2 coarse faces;
// dst_off is its first register entry (ND + 2 doubles per face), src_off the
// box's flux planes, lo its placement.  Thread t of the item: coarse face
// (t_u, t_v) on the transverse axes u < v, u fastest.
template <int ND>
__device__ __forceinline__ void reflux_face(const AmrItem& it, int64_t t, int& a, int& s,
                                            int tr[3]) {
  a = it.slo[0];
  s = it.slo[1];
  int64_t q = t;
  for (int d = 0; d < 3; ++d) {
    tr[d] = 0;
    if (d == a || d >= ND) continue;
    const int m = it.GD.n[d] / 2;
    tr[d] = (int)(q % m);
    q /= m;
  }
}

// Flux register of one substep: S = the sum of the coarse face's fine-face
// fluxes, pairwise ((F_00 + F_10) + (F_01 + F_11), first index on the lower
// transverse axis), so that equal fluxes sum exactly; reg = S (reset) or
// reg + S.
template <int ND, typename R>
__global__ __launch_bounds__(256) void amr_register_kernel(const AmrItem* __restrict__ items,
                                                           int nitems, int64_t total,
                                                           const R* __restrict__ arena,
                                                           double* __restrict__ reg, int reset) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const AmrItem& it = items[amr_item_of(items, nitems, t0)];
  const int64_t t = t0 - it.first;
  int a, s, tr[3];
  reflux_face<ND>(it, t, a, s, tr);
  const Geom& GF = it.GD;
  const R* F = arena + it.src_off + (int64_t)a * (ND + 2) * GF.plane;
  double* r = reg + it.dst_off + t * (ND + 2);
  const int fa = s ? GF.n[a] : 0;  // the fine face's low cell along a
  // offsets of the fine faces (x, y): x on the lower transverse axis
  int64_t of[2][2];
  for (int y = 0; y < 2; ++y)
    for (int x = 0; x < 2; ++x) {
      int idx[3] = {1, 1, 1};
      int k = 0;
      for (int d = 0; d < ND; ++d) {
        if (d == a) {
          idx[d] = fa;
        } else {
          idx[d] = 2 * tr[d] + 1 + (k == 0 ? x : y);
          ++k;
        }
      }
      of[y][x] = goff(GF, idx[0], idx[1], idx[2]);
    }
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) {
    const R* Fc = F + (int64_t)c * GF.plane;
    double sum;
    if (ND == 1)
      sum = (double)Fc[of[0][0]];
    else if (ND == 2)
      sum = (double)Fc[of[0][0]] + (double)Fc[of[0][1]];
    else
      sum = ((double)Fc[of[0][0]] + (double)Fc[of[0][1]]) +
            ((double)Fc[of[1][0]] + (double)Fc[of[1][1]]);
    r[c] = reset ? sum : r[c] + sum;
  }
}

// Reflux of one (axis, side) pass: the coarse cell outside each listed face
// takes dtdx (F_c - <F_f>) (low side) or dtdx (<F_f> - F_c) (high side),
// <F_f> = register * 2^-ndim (the mean over 2 substeps and 2^(ndim-1) faces).
// F_c from the coarse flux planes (FROM_STATE = 0), or, for a pass along
// axis FROM_STATE > 0 after rhy_amr_coarse_step, formed from the pre-step
// state `prev` with policy P, the step's constants k, dt and dx.
template <int ND, typename R, int FROM_STATE = 0, class P = PowF64Fix>
__global__ __launch_bounds__(256) void amr_reflux_kernel(const AmrItem* __restrict__ items,
                                                         int nitems, int64_t total, int64_t base,
                                                         const double* __restrict__ reg,
                                                         Geom GC, R* __restrict__ coarse,
                                                         const R* __restrict__ cflux,
                                                         double dtdx, Consts k = Consts{},
                                                         const R* __restrict__ prev = nullptr,
                                                         double dt = 0.0, double dx = 1.0) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= total) return;
  const int64_t tg = base + t0;  // the pass' items number threads from base
  const AmrItem& it = items[amr_item_of(items, nitems, tg)];
  const int64_t t = tg - it.first;
  int a, s, tr[3];
  reflux_face<ND>(it, t, a, s, tr);
  int out[3] = {1, 1, 1}, low[3] = {1, 1, 1};
  for (int d = 0; d < ND; ++d) {
    if (d == a) {
      out[d] = s ? it.lo[d] + it.GD.n[d] / 2 : it.lo[d] - 1;
      low[d] = s ? out[d] - 1 : out[d];
    } else {
      out[d] = low[d] = it.lo[d] + tr[d];
    }
  }
  const double w = ND == 3 ? 0.125 : ND == 2 ? 0.25 : 0.5;
  const double* r = reg + it.dst_off + t * (ND + 2);
  const int64_t oo = goff(GC, out[0], out[1], out[2]);
  const int64_t ol = goff(GC, low[0], low[1], low[2]);
  R Fs[ND + 2];
  if constexpr (FROM_STATE > 0) {
    // the coarse face's flux formed again from the state the coarse step
    // started from, as the step's walk formed it
    face_flux<ND, FROM_STATE, R, P>(k, GC, prev, dt, dx, ol, Fs);
  }
#pragma unroll
  for (int c = 0; c < ND + 2; ++c) {
    const double favg = r[c] * w;
    double fc;
    if constexpr (FROM_STATE > 0)
      fc = (double)Fs[c];
    else
      fc = (double)cflux[ol + (int64_t)a * (ND + 2) * GC.plane + c * GC.plane];
    const double d = s ? dtdx * (favg - fc) : dtdx * (fc - favg);
    R* u = coarse + oo + c * GC.plane;
    *u = (R)((double)*u + d);
  }
}

unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

constexpr int64_t kAlign = 64;  // arena blocks start on 64-element (>= 256 B) boundaries
constexpr int kFixupBlocks = 256;  // workgroups of the fixup kernel (one per CU)

struct Slot {
  bool active = false;
  bool needs_bc = false;  // ghosts not yet filled since the state was set
  BBox B{};
  int64_t off = 0, len = 0;  // arena block (elements)
};

}  // namespace

struct rhy_batch {
  int32_t ndim = 0, ncmp = 0, precision = RHY_FP64;
  bool pool = false;
  size_t esz = sizeof(double);
  std::vector<Slot> slots;
  std::vector<int32_t> active;
  std::map<int64_t, int64_t> free_blocks;  // offset -> length, elements
  int64_t arena_elems = 0;
  std::vector<BSeg> hsegs;
  int32_t nflux_tiles = 0, nupd_tiles = 0;
  int32_t block = 256, nf_x = 8, nf_yz = 8, grid_cap = 0;
  bool dirty = true;
  bool partials_valid = false;  // CFL partials match the state and tile table
  rhy_params_t last_params{};
  hipStream_t stream = nullptr;
  void* arena = nullptr;
  BBox* dboxes = nullptr;
  int32_t* dactive = nullptr;
  BSeg* dsegs = nullptr;
  size_t dsegs_cap = 0;
  int32_t* dtile_seg = nullptr;  // segment index of every tile
  size_t dtile_seg_cap = 0;
  double* partial = nullptr;
  size_t partial_cap = 0;
  double* dt = nullptr;
  double* dtlog = nullptr;
  int32_t dtlog_steps = 0;
  int32_t* step = nullptr;
  int32_t last_nsteps = 0;
  double* inflow = nullptr;
  hipGraphExec_t graph = nullptr;
  uint64_t graph_key = 0;
  DiagPartial* diag = nullptr;  // workgroup partials of rhy_batch_diagnostics
  uint32_t* defer_count = nullptr;  // faces the flux phase left to the fixup
  uint64_t* defer_items = nullptr;
  uint32_t defer_cap = 0;
  float* snap = nullptr;  // staging of the FP32 snapshot payload (lazy, grows)
  int64_t snap_elems = 0;
  AmrItem* amr_items = nullptr;      // work items of the hierarchy launches (lazy, grows)
  int64_t amr_cap = 0;
  double* amr_reg = nullptr;         // flux register of the coarse-fine faces (lazy)
  int64_t amr_reg_len = 0;           // doubles in use
  int64_t amr_reg_cap = 0;
  // the register's work items on the device, built once per placement
  AmrItem* reg_items = nullptr;
  int64_t reg_items_cap = 0;
  std::vector<int64_t> reg_key;      // placement and box extents they were built for
  int reg_pb[7] = {0, 0, 0, 0, 0, 0, 0};
  int64_t reg_pt[6] = {0, 0, 0, 0, 0, 0};
  std::vector<AmrItem> amr_host;     // their host copy, kept until the stream is idle
};

namespace {

Consts consts_of(const rhy_params_t* p) {
  Consts k{};
  const double g = p->gamma;
  k.g = g;
  k.gm1 = g - 1.0;
  k.gp1 = g + 1.0;
  k.gm1_gp1 = k.gm1 / k.gp1;
  k.gm1_2g = k.gm1 / (2.0 * g);
  k.gp1_2g = k.gp1 / (2.0 * g);
  k.g_inv = 1.0 / g;
  k.kb_amu = p->kb_amu;
  k.omega = p->omega;
  k.tol = p->tolerance;
  k.w_rho = p->w_vacuum_rho;
  k.w_p = p->w_vacuum_p;
  k.e_rho = std::nextafter(p->w_vacuum_rho, INFINITY);
  k.limiter = p->limiter;
  k.n_iter = p->n_iteration;
  k.defer_every = defer_every_setting();
  return k;
}

// Geometry and arena footprint (elements) of a box.
int box_geom(int ndim, const int32_t* dims, Geom& G) {
  int64_t st = 1;
  for (int d = 0; d < 3; ++d) {
    const int n = dims[d];
    if (d < ndim && n < 2) return bfail(RHY_ERR_SHAPE, "active extents must be >= 2");
    if (d >= ndim && n != 1) return bfail(RHY_ERR_SHAPE, "inactive extents must be 1");
    G.n[d] = n;
    G.g[d] = d < ndim ? 2 : 0;
    G.s[d] = st;
    st *= (int64_t)(n + 2 * G.g[d]);
  }
  G.plane = st;
  return RHY_OK;
}

int64_t round_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int64_t box_elems(const Geom& G, int ndim, int ncmp) {
  return round_up(G.plane * ncmp) + round_up(G.plane * (ndim + 2) * ndim);
}

// first-fit in the free list; returns -1 when nothing fits
int64_t arena_alloc(rhy_batch* b, int64_t len) {
  for (auto it = b->free_blocks.begin(); it != b->free_blocks.end(); ++it) {
    if (it->second < len) continue;
    const int64_t off = it->first;
    const int64_t rest = it->second - len;
    b->free_blocks.erase(it);
    if (rest > 0) b->free_blocks[off + len] = rest;
    return off;
  }
  return -1;
}

void arena_free(rhy_batch* b, int64_t off, int64_t len) {
  auto it = b->free_blocks.emplace(off, len).first;
  // coalesce with the next and the previous block
  auto nx = std::next(it);
  if (nx != b->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_blocks.erase(nx);
  }
  if (it != b->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_blocks.erase(it);
    }
  }
}

int grow(void** ptr, size_t* cap, size_t need, size_t elem) {
  if (need <= *cap) return RHY_OK;
  if (*ptr) (void)hipFree(*ptr);
  *ptr = nullptr;
  *cap = 0;
  size_t n = std::max(need, (size_t)64);
  if (dev_alloc(ptr, n * elem) != hipSuccess) return bfail(RHY_ERR_OOM, "batch table allocation");
  *cap = n;
  return RHY_OK;
}

// Rebuild the tile table over the active boxes for the current tile shape;
// box b's update tiles are contiguous so its CFL partials are
// [part_begin, part_end).
int build_tiles(rhy_batch* b) {
  // the tables below are read by work that may still be in flight
  BHIP_TRY(hipStreamSynchronize(b->stream));
  b->partials_valid = false;
  b->hsegs.clear();
  b->active.clear();
  for (int i = 0; i < (int)b->slots.size(); ++i)
    if (b->slots[i].active) b->active.push_back(i);
  const int nd = b->ndim;
  int32_t start = 0;
  for (int a = 0; a < nd; ++a)
    for (int i : b->active) {
      const Geom& G = b->slots[i].B.G;
      const int64_t n = flux_threads(G, a, a == 0 ? b->nf_x : b->nf_yz);
      b->hsegs.push_back({start, i, a, 0, n});
      start += (int32_t)cdiv(n, b->block);
    }
  b->nflux_tiles = start;
  for (int i : b->active) {
    BBox& B = b->slots[i].B;
    const int64_t n = (int64_t)B.G.n[0] * B.G.n[1] * B.G.n[2];
    B.part_begin = start - b->nflux_tiles;
    b->hsegs.push_back({start, i, 3, 0, n});
    start += (int32_t)cdiv(n, b->block);
    B.part_end = start - b->nflux_tiles;
  }
  b->nupd_tiles = start - b->nflux_tiles;
  int rc;
  std::vector<int32_t> tseg((size_t)std::max(1, start));
  for (size_t i = 0; i < b->hsegs.size(); ++i) {
    const int32_t end = i + 1 < b->hsegs.size() ? b->hsegs[i + 1].start : start;
    for (int32_t t = b->hsegs[i].start; t < end; ++t) tseg[t] = (int32_t)i;
  }
  if ((rc = grow((void**)&b->dtile_seg, &b->dtile_seg_cap, tseg.size(), sizeof(int32_t))))
    return rc;
  BHIP_TRY(hipMemcpy(b->dtile_seg, tseg.data(), sizeof(int32_t) * tseg.size(),
                     hipMemcpyHostToDevice));
  if ((rc = grow((void**)&b->dsegs, &b->dsegs_cap, b->hsegs.size(), sizeof(BSeg))) ||
      (rc = grow((void**)&b->partial, &b->partial_cap, (size_t)std::max(1, b->nupd_tiles),
                 sizeof(double))))
    return rc;
  if (!b->hsegs.empty())
    BHIP_TRY(hipMemcpy(b->dsegs, b->hsegs.data(), sizeof(BSeg) * b->hsegs.size(),
                       hipMemcpyHostToDevice));
  // descriptors up to the highest active slot
  const size_t nb = b->active.empty() ? 0 : (size_t)b->active.back() + 1;
  std::vector<BBox> hb(nb);
  for (size_t i = 0; i < nb; ++i) hb[i] = b->slots[i].B;
  if (nb) BHIP_TRY(hipMemcpy(b->dboxes, hb.data(), sizeof(BBox) * nb, hipMemcpyHostToDevice));
  if (!b->active.empty())
    BHIP_TRY(hipMemcpy(b->dactive, b->active.data(), sizeof(int32_t) * b->active.size(),
                       hipMemcpyHostToDevice));
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->dirty = false;
  return RHY_OK;
}

unsigned persistent_grid(rhy_batch* b, int32_t ntiles) {
  unsigned cap = b->grid_cap > 0 ? (unsigned)b->grid_cap : 256u * 16u;
  return std::max(1u, std::min(cap, (unsigned)ntiles));
}

DeferList defer_list(rhy_batch* b) {
  DeferList dl;
  dl.count = b->defer_count;
  dl.items = b->defer_items;
  const int64_t cap = defer_cap_setting();
  dl.cap = cap >= 0 ? (uint32_t)std::min<int64_t>(cap, b->defer_cap) : b->defer_cap;
  return dl;
}

// Flux kernels with policy PW (Deferring<P>), then the fixup of the faces
// they listed with P; the list is emptied first.
template <int ND, typename R, class PW, class P>
void launch_flux_phase(rhy_batch* b, const Consts& k, int lim) {
  const unsigned g = persistent_grid(b, b->nflux_tiles);
  R* a = static_cast<R*>(b->arena);
  const DeferList dl = defer_list(b);
  hipLaunchKernelGGL(set_u32_kernel, dim3(1), dim3(64), 0, b->stream, b->defer_count, 0u);
#define RHY_BATCH_FLUX(L)                                                                   \
  hipLaunchKernelGGL((batch_flux_kernel<ND, L, R, PW>), dim3(g), dim3(b->block), 0, b->stream, \
                     k, b->dboxes, b->dsegs, b->dtile_seg, b->nflux_tiles, a, b->dt, \
                     b->nf_x, b->nf_yz, dl)
  switch (lim) {
    case 0: RHY_BATCH_FLUX(0); break;
    case 1: RHY_BATCH_FLUX(1); break;
    case 2: RHY_BATCH_FLUX(2); break;
    case 3: RHY_BATCH_FLUX(3); break;
    default: RHY_BATCH_FLUX(4); break;
  }
#undef RHY_BATCH_FLUX
  if constexpr (PW::kDefer)
    hipLaunchKernelGGL((batch_fixup_kernel<ND, R, P>), dim3(kFixupBlocks), dim3(256), 0,
                       b->stream, k, b->dboxes, b->dactive, (int32_t)b->active.size(), a, b->dt,
                       dl);
}

template <int ND, typename R>
void launch_update_phase(rhy_batch* b, const Consts& k, int order) {
  const unsigned g = persistent_grid(b, b->nupd_tiles);
  R* a = static_cast<R*>(b->arena);
  if (order == RHY_MEMORY_INTENSIVE)
    hipLaunchKernelGGL((batch_update_kernel<ND, 2, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                       b->dboxes, b->dsegs, b->dtile_seg, b->nflux_tiles,
                       b->nupd_tiles, a, b->dt, b->partial);
  else
    hipLaunchKernelGGL((batch_update_kernel<ND, 1, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                       b->dboxes, b->dsegs, b->dtile_seg, b->nflux_tiles,
                       b->nupd_tiles, a, b->dt, b->partial);
}

template <int ND, typename R, class PW, class P>
void enqueue_step(rhy_batch* b, const Consts& k, int order, int lim, double fixed_dt = 0.0) {
  if (fixed_dt > 0.0)
    hipLaunchKernelGGL(batch_set_dt_kernel, dim3(cdiv((int64_t)b->active.size(), 256)), dim3(256),
                       0, b->stream, b->dactive, (int32_t)b->active.size(), fixed_dt, b->dt,
                       b->dtlog, b->step, b->dtlog_steps, (int32_t)b->slots.size());
  else
    hipLaunchKernelGGL(batch_dt_kernel, dim3((unsigned)b->active.size()), dim3(256), 0,
                       b->stream, b->dboxes, b->dactive, b->partial, b->dt, b->dtlog, b->step,
                       b->dtlog_steps, (int32_t)b->slots.size());
  launch_flux_phase<ND, R, PW, P>(b, k, lim);
  launch_update_phase<ND, R>(b, k, order);
  hipLaunchKernelGGL(batch_step_counter_kernel, dim3(1), dim3(64), 0, b->stream, b->step);
}

// Ghost fill, then CFL partials of the current state.  A fixed batch fills
// every box at every run, as rhy_run does.  In a pool the fused update keeps
// the ghosts of a box current from step to step, so only boxes whose state
// was just set get a fill, and the partials are recomputed only when the
// tile table, the parameters or a state changed since the last step.
template <int ND, typename R>
int prologue(rhy_batch* b, const Consts& k, bool need_partials) {
  R* a = static_cast<R*>(b->arena);
  for (int i : b->active) {
    Slot& S = b->slots[i];
    if (b->pool && !S.needs_bc) continue;
    S.needs_bc = false;
    const BBox& B = S.B;
    int64_t maxface = 1;
    for (int ax = 0; ax < ND; ++ax) {
      int64_t m = 1;
      for (int d = 0; d < 3; ++d)
        if (d != ax) m *= B.G.n[d];
      maxface = std::max(maxface, m);
    }
    hipLaunchKernelGGL((bc_kernel<ND, R>), dim3(cdiv(maxface, 256), 2 * ND), dim3(256), 0,
                       b->stream, B.G, B.bc, b->inflow, b->ncmp, a + B.cell_off,
                       a + B.cell_off + (int64_t)(ND + 2) * B.G.plane, -1);
  }
  if (need_partials && (!b->pool || !b->partials_valid)) {
    const unsigned g = persistent_grid(b, b->nupd_tiles);
    hipLaunchKernelGGL((batch_cfl_kernel<ND, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                       b->dboxes, b->dsegs, b->dtile_seg, b->nflux_tiles,
                     b->nupd_tiles,
                       static_cast<const R*>(b->arena), b->partial);
  }
  b->partials_valid = true;  // from here on every update rewrites them
  hipLaunchKernelGGL(batch_set_counter_kernel, dim3(1), dim3(64), 0, b->stream, b->step, 0);
  BHIP_TRY(hipGetLastError());
  return RHY_OK;
}

// mode 0: nsteps steps, each box with its own CFL dt; 1: nsteps steps with
// dt = fixed_dt for every box (no graph); 2: only the CFL dt of every box
// for the current state, into the batch's dt array (no step).
template <int ND, typename R, class PW, class P>
int run_t(rhy_batch* b, const rhy_params_t* p, int32_t nsteps, int32_t use_graph, int mode = 0,
          double fixed_dt = 0.0) {
  const Consts k = consts_of(p);
  if (std::memcmp(p, &b->last_params, sizeof(*p)) != 0) {
    // new parameters: inflow states, ghosts and wave speeds all follow them
    BHIP_TRY(hipStreamSynchronize(b->stream));
    BHIP_TRY(hipMemcpy(b->inflow, p->cons_inflows, sizeof(double) * 6 * RHY_MAX_CMP,
                       hipMemcpyHostToDevice));
    b->last_params = *p;
    b->partials_valid = false;
    for (Slot& S : b->slots) S.needs_bc = S.active;
  }
  // a fixed batch takes limiter and faces from the parameters, a pool keeps
  // each box's own
  bool changed = b->dirty;
  for (Slot& S : b->slots) {
    if (!S.active) continue;
    BBox& B = S.B;
    double m = B.dx[0];
    for (int d = 1; d < ND; ++d) m = std::min(m, B.dx[d]);
    const double cm = p->courant * m;
    changed = changed || cm != B.cmdx;
    B.cmdx = cm;
    if (!b->pool) {
      for (int f = 0; f < 6; ++f) {
        changed = changed || B.bc.t[f] != p->bc_types[f];
        B.bc.t[f] = p->bc_types[f];
      }
      changed = changed || B.limiter != p->limiter;
      B.limiter = p->limiter;
    }
  }
  int rc;
  if (changed && (rc = build_tiles(b))) return rc;
  if (b->active.empty()) {
    b->last_nsteps = 0;
    return RHY_OK;
  }
  const int lim = b->pool ? 0 : p->limiter;
  const int64_t need = (int64_t)nsteps * (int64_t)b->slots.size();
  if (need > (int64_t)b->dtlog_steps * (int64_t)b->slots.size()) {
    BHIP_TRY(hipStreamSynchronize(b->stream));
    if (b->dtlog) (void)hipFree(b->dtlog);
    b->dtlog = nullptr;
    if (dev_alloc(&b->dtlog, sizeof(double) * (size_t)need) != hipSuccess)
      return bfail(RHY_ERR_OOM, "dt log allocation failed");
    b->dtlog_steps = nsteps;
  }
  // steps with a given dt need no CFL partials; their updates rewrite them
  if ((rc = prologue<ND, R>(b, k, !(mode == 1 && nsteps > 0)))) return rc;
  if (mode == 2) {
    hipLaunchKernelGGL(batch_dt_kernel, dim3((unsigned)b->active.size()), dim3(256), 0,
                       b->stream, b->dboxes, b->dactive, b->partial, b->dt, (double*)nullptr,
                       b->step, 0, (int32_t)b->slots.size());
    BHIP_TRY(hipGetLastError());
    return RHY_OK;
  }
  if (mode == 1) {
    for (int32_t s = 0; s < nsteps; ++s)
      enqueue_step<ND, R, PW, P>(b, k, p->solver_type, lim, fixed_dt);
    BHIP_TRY(hipGetLastError());
    b->last_nsteps = nsteps;
    return RHY_OK;
  }
  if (!use_graph) {
    for (int32_t s = 0; s < nsteps; ++s) enqueue_step<ND, R, PW, P>(b, k, p->solver_type, lim);
    BHIP_TRY(hipGetLastError());
    b->last_nsteps = nsteps;
    return RHY_OK;
  }
  // the captured step depends on the parameters and the tile table
  uint64_t key = 1469598103934665603ull;
  auto mix = [&key](const void* data, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) key = (key ^ c[i]) * 1099511628211ull;
  };
  mix(p, sizeof(*p));
  mix(&b->dtlog_steps, sizeof(b->dtlog_steps));
  const int32_t d_every = defer_every_setting();
  const int64_t d_cap = defer_cap_setting();
  mix(&d_every, sizeof(d_every));
  mix(&d_cap, sizeof(d_cap));
  if (!b->graph || b->graph_key != key) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    hipGraph_t g = nullptr;
    BHIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step<ND, R, PW, P>(b, k, p->solver_type, lim);
    BHIP_TRY(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(RHY_ERR_HIP, "hipGraphInstantiate failed");
    b->graph_key = key;
  }
  for (int32_t s = 0; s < nsteps; ++s) BHIP_TRY(hipGraphLaunch(b->graph, b->stream));
  b->last_nsteps = nsteps;
  return RHY_OK;
}

rhy_batch* new_batch(int32_t ndim, int32_t ncmp, int32_t precision, int64_t arena_elems,
                     int32_t max_slots, int* rc) {
  *rc = rhy_init(0);
  if (*rc) {
    bfail(*rc, rhy_last_error());
    return nullptr;
  }
  if (ndim < 1 || ndim > 3 || max_slots < 1) {
    *rc = bfail(RHY_ERR_ARG, "bad batch arguments");
    return nullptr;
  }
  if (ncmp < ndim + 2 || ncmp > RHY_MAX_CMP) {
    *rc = bfail(RHY_ERR_ARG, "ncmp out of range");
    return nullptr;
  }
  if (precision != RHY_FP64 && precision != RHY_MIXED) {
    *rc = bfail(RHY_ERR_ARG, "unsupported precision");
    return nullptr;
  }
  rhy_batch* b = new rhy_batch();
  b->ndim = ndim;
  b->ncmp = ncmp;
  b->precision = precision;
  b->esz = precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  b->slots.resize(max_slots);
  b->arena_elems = arena_elems;
  if (arena_elems > 0) b->free_blocks[0] = arena_elems;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      (arena_elems > 0 && dev_alloc(&b->arena, b->esz * (size_t)arena_elems) != hipSuccess) ||
      dev_alloc(&b->dboxes, sizeof(BBox) * max_slots) != hipSuccess ||
      dev_alloc(&b->dactive, sizeof(int32_t) * max_slots) != hipSuccess ||
      dev_alloc(&b->dt, sizeof(double) * max_slots) != hipSuccess ||
      dev_alloc(&b->step, sizeof(int32_t)) != hipSuccess ||
      dev_alloc(&b->inflow, sizeof(double) * 6 * RHY_MAX_CMP) != hipSuccess) {
    rhy_batch_destroy(b);
    *rc = bfail(RHY_ERR_OOM, "batch allocation failed");
    return nullptr;
  }
  // deferred-face list: about one entry per 64 arena elements, within
  // [4096, 2^20] (the fixup recomputes every face past that)
  b->defer_cap = (uint32_t)std::min<int64_t>(std::max<int64_t>(arena_elems / 64, 4096), 1 << 20);
  if (dev_alloc(&b->defer_count, sizeof(uint32_t)) != hipSuccess ||
      dev_alloc(&b->defer_items, sizeof(uint64_t) * (size_t)b->defer_cap) != hipSuccess ||
      hipMemsetAsync(b->defer_count, 0, sizeof(uint32_t), b->stream) != hipSuccess) {
    rhy_batch_destroy(b);
    *rc = bfail(RHY_ERR_OOM, "deferred-face list allocation failed");
    return nullptr;
  }
  *rc = RHY_OK;
  return b;
}

template <int ND>
void launch_paint(rhy_batch* b, const Geom& G, const PaintIC& ic, int64_t off) {
  const unsigned grid = cdiv(G.plane, 256);
  const int64_t xo = off + (int64_t)(ND + 2) * G.plane;  // the non-hydro planes
  if (b->precision == RHY_MIXED)
    hipLaunchKernelGGL((paint_kernel<ND, float>), dim3(grid), dim3(256), 0, b->stream, G, ic,
                       b->ncmp, static_cast<float*>(b->arena) + off,
                       static_cast<float*>(b->arena) + xo);
  else
    hipLaunchKernelGGL((paint_kernel<ND, double>), dim3(grid), dim3(256), 0, b->stream, G, ic,
                       b->ncmp, static_cast<double*>(b->arena) + off,
                       static_cast<double*>(b->arena) + xo);
}

// Place a box: its state from host `cells`, painted from `ic`, or zero.
int add_box(rhy_batch* b, const int32_t* dims, const double* dx, const double* cells,
            const PaintIC* ic, int32_t limiter, const int32_t* bc_types, int32_t* slot_out) {
  Geom G{};
  int rc;
  if ((rc = box_geom(b->ndim, dims, G))) return rc;
  for (int d = 0; d < b->ndim; ++d)
    if (!(dx[d] > 0.0)) return bfail(RHY_ERR_ARG, "dx must be positive on active axes");
  int slot = -1;
  for (int i = 0; i < (int)b->slots.size(); ++i)
    if (!b->slots[i].active) {
      slot = i;
      break;
    }
  if (slot < 0) return bfail(RHY_ERR_OOM, "no free slot");
  const int64_t len = box_elems(G, b->ndim, b->ncmp);
  const int64_t off = arena_alloc(b, len);
  if (off < 0) return bfail(RHY_ERR_OOM, "box does not fit in the arena");
  // the state first; the slot becomes active only once it is written
  hipError_t e = hipSuccess;
  if (cells) {
    const size_t n = (size_t)G.plane * b->ncmp;
    if (b->precision == RHY_MIXED) {
      std::vector<float> tmp(n);
      for (size_t i = 0; i < n; ++i) tmp[i] = (float)cells[i];
      e = hipMemcpyAsync(static_cast<float*>(b->arena) + off, tmp.data(), sizeof(float) * n,
                         hipMemcpyHostToDevice, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    } else {
      e = hipMemcpyAsync(static_cast<double*>(b->arena) + off, cells, sizeof(double) * n,
                         hipMemcpyHostToDevice, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    }
  } else if (ic) {
    if (b->ndim == 1) launch_paint<1>(b, G, *ic, off);
    else if (b->ndim == 2) launch_paint<2>(b, G, *ic, off);
    else launch_paint<3>(b, G, *ic, off);
    e = hipGetLastError();
  } else {
    e = hipMemsetAsync(static_cast<char*>(b->arena) + off * b->esz, 0,
                       b->esz * (size_t)G.plane * b->ncmp, b->stream);
  }
  if (e != hipSuccess) {
    arena_free(b, off, len);
    return bfail(RHY_ERR_HIP, std::string("writing the box state: ") + hipGetErrorString(e));
  }
  Slot& S = b->slots[slot];
  S.active = true;
  S.off = off;
  S.len = len;
  S.B = BBox{};
  S.B.G = G;
  S.B.cell_off = off;
  S.B.flux_off = off + round_up(G.plane * b->ncmp);
  for (int d = 0; d < 3; ++d) S.B.dx[d] = dx[d];
  S.B.cmdx = -1.0;
  S.B.limiter = limiter;
  for (int f = 0; f < 6; ++f) S.B.bc.t[f] = bc_types ? bc_types[f] : 1;
  S.needs_bc = true;
  b->dirty = true;
  *slot_out = slot;
  return RHY_OK;
}

}  // namespace

extern "C" {

const char* rhy_batch_last_error(void) { return b_err.c_str(); }

int rhy_batch_create(int32_t ndim, int32_t nboxes, const int32_t* dims, const double* dx,
                     int32_t ncmp, int32_t precision, rhy_batch_t** out) {
  int rc;
  if (!out) return bfail(RHY_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (nboxes < 1 || !dims || !dx) return bfail(RHY_ERR_ARG, "bad batch arguments");
  int64_t total = 0;
  for (int i = 0; i < nboxes; ++i) {
    Geom G{};
    if ((rc = box_geom(ndim, dims + 3 * i, G))) return rc;
    total += box_elems(G, ndim, ncmp);
  }
  rhy_batch* b = new_batch(ndim, ncmp, precision, total, nboxes, &rc);
  if (!b) return rc;
  for (int i = 0; i < nboxes; ++i) {
    int32_t slot;
    if ((rc = add_box(b, dims + 3 * i, dx + 3 * i, nullptr, nullptr, 2, nullptr, &slot))) {
      rhy_batch_destroy(b);
      return rc;
    }
  }
  if ((rc = build_tiles(b))) {
    rhy_batch_destroy(b);
    return rc;
  }
  *out = b;
  return RHY_OK;
}

int rhy_pool_create(int32_t ndim, int32_t ncmp, int32_t precision, int64_t budget_bytes,
                    int32_t max_boxes, rhy_batch_t** out) {
  int rc;
  if (!out) return bfail(RHY_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (budget_bytes <= 0) return bfail(RHY_ERR_ARG, "budget must be positive");
  const size_t esz = precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  rhy_batch* b = new_batch(ndim, ncmp, precision, budget_bytes / (int64_t)esz / kAlign * kAlign,
                           max_boxes, &rc);
  if (!b) return rc;
  b->pool = true;
  *out = b;
  return RHY_OK;
}

int rhy_pool_box_bytes(int32_t ndim, const int32_t dims[3], int32_t ncmp, int32_t precision,
                       int64_t* bytes) {
  int rc;
  if (!dims || !bytes) return bfail(RHY_ERR_ARG, "NULL argument");
  Geom G{};
  if ((rc = box_geom(ndim, dims, G))) return rc;
  *bytes = box_elems(G, ndim, ncmp) * (precision == RHY_MIXED ? 4 : 8);
  return RHY_OK;
}

int rhy_pool_add(rhy_batch_t* b, const int32_t dims[3], const double dx[3], const double* cells,
                 int32_t limiter, const int32_t bc_types[6], int32_t* slot) {
  if (!b || !dims || !dx || !cells || !slot) return bfail(RHY_ERR_ARG, "NULL argument");
  if (!b->pool) return bfail(RHY_ERR_ARG, "not a pool");
  if (limiter < 1 || limiter > 4) return bfail(RHY_ERR_ARG, "unknown slope limiter");
  return add_box(b, dims, dx, cells, nullptr, limiter, bc_types, slot);
}

int rhy_pool_add_ic(rhy_batch_t* b, const int32_t dims[3], const double dx[3], const rhy_ic_t* ic,
                    int32_t limiter, const int32_t bc_types[6], int32_t* slot) {
  if (!b || !dims || !dx || !ic || !slot) return bfail(RHY_ERR_ARG, "NULL argument");
  if (!b->pool) return bfail(RHY_ERR_ARG, "not a pool");
  if (limiter < 1 || limiter > 4) return bfail(RHY_ERR_ARG, "unknown slope limiter");
  if (ic->kind != RHY_IC_TWO_STATE && ic->kind != RHY_IC_BLAST)
    return bfail(RHY_ERR_ARG, "unknown initial condition kind");
  if (ic->kind == RHY_IC_TWO_STATE && (ic->axis < 0 || ic->axis >= b->ndim))
    return bfail(RHY_ERR_ARG, "two-state axis out of range");
  if (b->ncmp < b->ndim + 2) return bfail(RHY_ERR_ARG, "too few components");
  const PaintIC pc =
      make_paint(ic->kind, ic->axis, ic->gamma, ic->rho_l, ic->p_l, ic->rho_r, ic->p_r,
                 ic->radius, ic->jitter, ic->p0, ic->energy, ic->seed, b->ndim, dims);
  return add_box(b, dims, dx, nullptr, &pc, limiter, bc_types, slot);
}

int rhy_pool_remove(rhy_batch_t* b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].active)
    return bfail(RHY_ERR_ARG, "bad slot");
  BHIP_TRY(hipStreamSynchronize(b->stream));
  arena_free(b, b->slots[slot].off, b->slots[slot].len);
  b->slots[slot].active = false;
  b->dirty = true;
  return RHY_OK;
}

int rhy_pool_free_bytes(rhy_batch_t* b, int64_t* free_bytes, int64_t* largest_bytes) {
  if (!b || !free_bytes || !largest_bytes) return bfail(RHY_ERR_ARG, "NULL argument");
  int64_t tot = 0, big = 0;
  for (auto& kv : b->free_blocks) {
    tot += kv.second;
    big = std::max(big, kv.second);
  }
  *free_bytes = tot * (int64_t)b->esz;
  *largest_bytes = big * (int64_t)b->esz;
  return RHY_OK;
}

int rhy_batch_destroy(rhy_batch_t* b) {
  if (!b) return RHY_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (void* p : {(void*)b->arena, (void*)b->dboxes, (void*)b->dactive, (void*)b->dsegs,
                  (void*)b->dtile_seg, (void*)b->partial, (void*)b->dt, (void*)b->dtlog, (void*)b->step,
                  (void*)b->inflow, (void*)b->diag, (void*)b->defer_count,
                  (void*)b->defer_items, (void*)b->snap, (void*)b->amr_items,
                  (void*)b->amr_reg, (void*)b->reg_items})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RHY_OK;
}

int rhy_batch_bytes(rhy_batch_t* b, int64_t* bytes) {
  if (!b || !bytes) return bfail(RHY_ERR_ARG, "NULL argument");
  *bytes = (int64_t)b->esz * b->arena_elems;
  return RHY_OK;
}

static int slot_ok(rhy_batch_t* b, int32_t slot) {
  return b && slot >= 0 && slot < (int32_t)b->slots.size() && b->slots[slot].active;
}

int rhy_batch_upload(rhy_batch_t* b, int32_t box, const double* cells) {
  if (!slot_ok(b, box) || !cells) return bfail(RHY_ERR_ARG, "bad upload");
  const BBox& B = b->slots[box].B;
  b->slots[box].needs_bc = true;
  b->partials_valid = false;
  const size_t n = (size_t)B.G.plane * b->ncmp;
  if (b->precision == RHY_MIXED) {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)cells[i];
    BHIP_TRY(hipMemcpyAsync(static_cast<float*>(b->arena) + B.cell_off, tmp.data(),
                            sizeof(float) * n, hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    return RHY_OK;
  }
  BHIP_TRY(hipMemcpyAsync(static_cast<double*>(b->arena) + B.cell_off, cells,
                          sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

// Snapshot payload of the whole level: the active boxes in slot order, each
// its components' interior cells (i fastest) one component after another, as
// FP32 -- a level's data:datatype=0 dataset (rhyme_chombo_write_level_data.f90:41-71);
// unpack is the restart reader's half (rhyme_initial_condition_load_rhyme.f90:57-76).
// (The slots are walked directly: the active list is only rebuilt by a run.)
static int64_t snapshot_elems(const rhy_batch_t* b) {
  int64_t n = 0;
  for (const Slot& S : b->slots) {
    if (!S.active) continue;
    const Geom& G = S.B.G;
    n += (int64_t)G.n[0] * G.n[1] * G.n[2] * b->ncmp;
  }
  return n;
}

static int batch_snapshot(rhy_batch_t* b, float* out, const float* in) {
  if (!b || (!out && !in)) return bfail(RHY_ERR_ARG, "NULL argument");
  const int64_t total = snapshot_elems(b);
  if (total == 0) return RHY_OK;
  if (total > b->snap_elems) {
    BHIP_TRY(hipStreamSynchronize(b->stream));
    if (b->snap) (void)hipFree(b->snap);
    b->snap = nullptr;
    b->snap_elems = 0;
    if (dev_alloc(&b->snap, sizeof(float) * (size_t)total) != hipSuccess) {
      b->snap = nullptr;
      return bfail(RHY_ERR_OOM, "snapshot staging allocation failed");
    }
    b->snap_elems = total;
  }
  if (in)
    BHIP_TRY(hipMemcpyAsync(b->snap, in, sizeof(float) * (size_t)total, hipMemcpyHostToDevice,
                            b->stream));
  int64_t ofs = 0;
  for (Slot& S : b->slots) {
    if (!S.active) continue;
    const Geom& G = S.B.G;
    const int64_t blen = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const dim3 grid = snapshot_grid(G);
    const int nv = b->ndim + 2;
    const int64_t co = S.B.cell_off, xo = S.B.cell_off + (int64_t)nv * G.plane;
    if (in) {
      S.needs_bc = true;
      if (b->precision == RHY_MIXED)
        hipLaunchKernelGGL(snapshot_unpack_kernel<float>, grid, dim3(256), 0, b->stream, G, b->ncmp,
                           nv, static_cast<float*>(b->arena) + co,
                           static_cast<float*>(b->arena) + xo, b->snap + ofs);
      else
        hipLaunchKernelGGL(snapshot_unpack_kernel<double>, grid, dim3(256), 0, b->stream, G, b->ncmp,
                           nv, static_cast<double*>(b->arena) + co,
                           static_cast<double*>(b->arena) + xo, b->snap + ofs);
    } else {
      if (b->precision == RHY_MIXED)
        hipLaunchKernelGGL(snapshot_pack_kernel<float>, grid, dim3(256), 0, b->stream, G, b->ncmp,
                           nv, static_cast<const float*>(b->arena) + co,
                           static_cast<const float*>(b->arena) + xo, b->snap + ofs);
      else
        hipLaunchKernelGGL(snapshot_pack_kernel<double>, grid, dim3(256), 0, b->stream, G, b->ncmp,
                           nv, static_cast<const double*>(b->arena) + co,
                           static_cast<const double*>(b->arena) + xo, b->snap + ofs);
    }
    ofs += blen * b->ncmp;
  }
  BHIP_TRY(hipGetLastError());
  if (in) b->partials_valid = false;
  if (out)
    BHIP_TRY(hipMemcpyAsync(out, b->snap, sizeof(float) * (size_t)total, hipMemcpyDeviceToHost,
                            b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_batch_snapshot_length(rhy_batch_t* b, int64_t* n) {
  if (!b || !n) return bfail(RHY_ERR_ARG, "NULL argument");
  *n = snapshot_elems(b);
  return RHY_OK;
}

int rhy_batch_snapshot_pack(rhy_batch_t* b, float* data) { return batch_snapshot(b, data, nullptr); }

int rhy_batch_snapshot_unpack(rhy_batch_t* b, const float* data) {
  return batch_snapshot(b, nullptr, data);
}

int rhy_batch_download(rhy_batch_t* b, int32_t box, double* cells) {
  if (!slot_ok(b, box) || !cells) return bfail(RHY_ERR_ARG, "bad download");
  const BBox& B = b->slots[box].B;
  const size_t n = (size_t)B.G.plane * b->ncmp;
  if (b->precision == RHY_MIXED) {
    std::vector<float> tmp(n);
    BHIP_TRY(hipMemcpyAsync(tmp.data(), static_cast<float*>(b->arena) + B.cell_off,
                            sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < n; ++i) cells[i] = (double)tmp[i];
    return RHY_OK;
  }
  BHIP_TRY(hipMemcpyAsync(cells, static_cast<double*>(b->arena) + B.cell_off,
                          sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_batch_set_tiles(rhy_batch_t* b, int32_t block, int32_t nf_x, int32_t nf_yz,
                        int32_t grid_cap) {
  if (!b) return bfail(RHY_ERR_ARG, "batch is NULL");
  if (block < 64 || block > 256 || block % 64 || nf_x < 1 || nf_x > 64 || nf_yz < 1 ||
      nf_yz > 64 || grid_cap < 0)
    return bfail(RHY_ERR_ARG, "tile shape out of range");
  BHIP_TRY(hipStreamSynchronize(b->stream));
  b->block = block;
  b->nf_x = nf_x;
  b->nf_yz = nf_yz;
  b->grid_cap = grid_cap;
  return build_tiles(b);
}

int rhy_batch_run_async(rhy_batch_t* b, const rhy_params_t* p, int32_t nsteps, int32_t use_graph) {
  if (!b || !p) return bfail(RHY_ERR_ARG, "NULL argument");
  if (p->ndim != b->ndim) return bfail(RHY_ERR_ARG, "params.ndim does not match the batch");
  if (!b->pool && (p->limiter < 1 || p->limiter > 4))
    return bfail(RHY_ERR_ARG, "unknown slope limiter");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return bfail(RHY_ERR_SOLVER, "Unknown solver type");
  if (nsteps < 0) return bfail(RHY_ERR_ARG, "nsteps < 0");
  const bool mixed = b->precision == RHY_MIXED;
  switch (b->ndim) {
    case 1:
      return mixed ? run_t<1, float, MixedFlux, MixedMath>(b, p, nsteps, use_graph)
                   : run_t<1, double, PowF64Flux, PowF64Fix>(b, p, nsteps, use_graph);
    case 2:
      return mixed ? run_t<2, float, MixedFlux, MixedMath>(b, p, nsteps, use_graph)
                   : run_t<2, double, PowF64Flux, PowF64Fix>(b, p, nsteps, use_graph);
    default:
      return mixed ? run_t<3, float, MixedFlux, MixedMath>(b, p, nsteps, use_graph)
                   : run_t<3, double, PowF64Flux, PowF64Fix>(b, p, nsteps, use_graph);
  }
}

static int run_checks(rhy_batch_t* b, const rhy_params_t* p) {
  if (!b || !p) return bfail(RHY_ERR_ARG, "NULL argument");
  if (p->ndim != b->ndim) return bfail(RHY_ERR_ARG, "params.ndim does not match the batch");
  if (!b->pool && (p->limiter < 1 || p->limiter > 4))
    return bfail(RHY_ERR_ARG, "unknown slope limiter");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return bfail(RHY_ERR_SOLVER, "Unknown solver type");
  return RHY_OK;
}

static int run_mode(rhy_batch_t* b, const rhy_params_t* p, int32_t nsteps, int mode, double dt) {
  const bool mixed = b->precision == RHY_MIXED;
  switch (b->ndim) {
    case 1:
      return mixed ? run_t<1, float, MixedFlux, MixedMath>(b, p, nsteps, 0, mode, dt)
                   : run_t<1, double, PowF64Flux, PowF64Fix>(b, p, nsteps, 0, mode, dt);
    case 2:
      return mixed ? run_t<2, float, MixedFlux, MixedMath>(b, p, nsteps, 0, mode, dt)
                   : run_t<2, double, PowF64Flux, PowF64Fix>(b, p, nsteps, 0, mode, dt);
    default:
      return mixed ? run_t<3, float, MixedFlux, MixedMath>(b, p, nsteps, 0, mode, dt)
                   : run_t<3, double, PowF64Flux, PowF64Fix>(b, p, nsteps, 0, mode, dt);
  }
}

int rhy_batch_run_dt_async(rhy_batch_t* b, const rhy_params_t* p, double dt, int32_t nsteps) {
  int rc;
  if ((rc = run_checks(b, p))) return rc;
  if (!(dt > 0.0) || nsteps < 0) return bfail(RHY_ERR_ARG, "dt must be positive, nsteps >= 0");
  return run_mode(b, p, nsteps, 1, dt);
}

int rhy_batch_cfl_dts(rhy_batch_t* b, const rhy_params_t* p, double* dts) {
  int rc;
  if ((rc = run_checks(b, p))) return rc;
  if (!dts) return bfail(RHY_ERR_ARG, "dts is NULL");
  if ((rc = run_mode(b, p, 0, 2, 0.0))) return rc;
  BHIP_TRY(hipMemcpyAsync(dts, b->dt, sizeof(double) * b->slots.size(), hipMemcpyDeviceToHost,
                          b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

// Coarse parents of every fine box (ghosts included) must be interior
// coarse cells: 2 <= lo[d] and lo[d] + n_f[d] / 2 <= n_c[d], n_f[d] even.
// A box of a fixed batch as the coarse side of a hierarchy call (a level
// between the coarsest and the finest): its state and flux planes in the
// batch's arena, after the batch's queued work.
static int slot_view(rhy_batch_t* b, int32_t slot, BoxView* v) {
  if (!b) return bfail(RHY_ERR_ARG, "NULL argument");
  if (b->pool) return bfail(RHY_ERR_ARG, "the coarse level is a fixed batch, not a pool");
  if (slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].active)
    return bfail(RHY_ERR_ARG, "no such coarse box");
  BHIP_TRY(hipStreamSynchronize(b->stream));
  const BBox& B = b->slots[slot].B;
  v->cells = static_cast<char*>(b->arena) + b->esz * (size_t)B.cell_off;
  v->flux = static_cast<char*>(b->arena) + b->esz * (size_t)B.flux_off;
  v->prev = nullptr;
  v->k = Consts{};
  v->step_dt = 0.0;
  v->step_dx[0] = v->step_dx[1] = v->step_dx[2] = 1.0;
  v->geom = B.G;
  v->ndim = b->ndim;
  v->ncmp = b->ncmp;
  v->precision = b->precision;
  return RHY_OK;
}

static int amr_check(rhy_batch_t* f, const BoxView& c, const int32_t* lo) {
  if (!f || !lo) return bfail(RHY_ERR_ARG, "NULL argument");
  if (f->pool) return bfail(RHY_ERR_ARG, "the fine level is a fixed batch, not a pool");
  if (c.ndim != f->ndim || c.precision != f->precision)
    return bfail(RHY_ERR_ARG, "coarse box and fine batch differ in ndim or precision");
  for (int i = 0; i < (int)f->slots.size(); ++i) {
    const Slot& S = f->slots[i];
    if (!S.active) return bfail(RHY_ERR_ARG, "every fine box must be active");
    for (int d = 0; d < f->ndim; ++d) {
      const int nf = S.B.G.n[d], l = lo[3 * i + d];
      if (nf % 2) return bfail(RHY_ERR_SHAPE, "fine extents must be even (ratio 2)");
      if (l < 2 || l + nf / 2 > c.geom.n[d])
        return bfail(RHY_ERR_SHAPE, "a fine box or its ghosts leave the coarse interior");
    }
  }
  return RHY_OK;
}

// The host items to the device (one copy, in stream order).
static int amr_upload(rhy_batch_t* f) {
  const int64_t n = (int64_t)f->amr_host.size();
  if (n > f->amr_cap) {
    BHIP_TRY(hipStreamSynchronize(f->stream));
    if (f->amr_items) (void)hipFree(f->amr_items);
    f->amr_items = nullptr;
    f->amr_cap = 0;
    if (dev_alloc(&f->amr_items, sizeof(AmrItem) * (size_t)n) != hipSuccess) {
      f->amr_items = nullptr;
      return bfail(RHY_ERR_OOM, "hierarchy work items allocation failed");
    }
    f->amr_cap = n;
  }
  if (n)
    BHIP_TRY(hipMemcpyAsync(f->amr_items, f->amr_host.data(), sizeof(AmrItem) * (size_t)n,
                            hipMemcpyHostToDevice, f->stream));
  return RHY_OK;
}

static int amr_prolong_v(const BoxView& vc, rhy_batch_t* f, const int32_t* lo, int32_t half) {
  int rc;
  if ((rc = amr_check(f, vc, lo))) return rc;
  const int nd = f->ndim, nb = (int)f->slots.size();
  f->amr_host.clear();
  // items [0, nb): one per fine box, its face ghost cells
  int64_t nprol = 0;
  for (int i = 0; i < nb; ++i) {
    const BBox& B = f->slots[i].B;
    AmrItem it{};
    it.first = nprol;
    it.dst_off = B.cell_off;
    it.GD = B.G;
    for (int d = 0; d < 3; ++d) it.lo[d] = lo[3 * i + d];
    f->amr_host.push_back(it);
    for (int a = 0; a < nd; ++a) {
      int64_t m = 4;
      for (int d = 0; d < 3; ++d)
        if (d != a) m *= B.G.n[d];
      nprol += m;
    }
  }
  // items [nb, ...): face ghosts inside a sibling's interior take the
  // sibling's cells (same level, same time).  In fine-level indices box i's
  // cell f lies at 2 (lo_i - 1) + f; interiors never overlap, so each ghost
  // has at most one sibling source; the copies run after the coarse fill.
  int64_t ncopy = 0;
  for (int i = 0; i < nb; ++i) {
    const BBox& D = f->slots[i].B;
    for (int a = 0; a < nd; ++a)
      for (int side = 0; side < 2; ++side)
        for (int j = 0; j < nb; ++j) {
          if (j == i) continue;
          const BBox& S = f->slots[j].B;
          int g0[3], g1[3];  // the face slab in fine-level indices, intersected with j
          bool empty = false;
          for (int d = 0; d < 3; ++d) {
            const int bi = d < nd ? 2 * (lo[3 * i + d] - 1) : 0;
            const int bj = d < nd ? 2 * (lo[3 * j + d] - 1) : 0;
            int s0 = bi + 1, s1 = bi + D.G.n[d];
            if (d == a) {
              s0 = side ? bi + D.G.n[d] + 1 : bi - 1;
              s1 = s0 + 1;
            }
            g0[d] = std::max(s0, bj + 1);
            g1[d] = std::min(s1, bj + S.G.n[d]);
            empty = empty || g0[d] > g1[d];
          }
          if (empty) continue;
          AmrItem it{};
          it.first = ncopy;
          it.dst_off = D.cell_off;
          it.src_off = S.cell_off;
          it.GD = D.G;
          it.GS = S.G;
          int64_t vol = 1;
          for (int d = 0; d < 3; ++d) {
            const int bi = d < nd ? 2 * (lo[3 * i + d] - 1) : 0;
            const int bj = d < nd ? 2 * (lo[3 * j + d] - 1) : 0;
            it.lo[d] = g0[d] - bi;
            it.slo[d] = g0[d] - bj;
            it.ext[d] = g1[d] - g0[d] + 1;
            // dst within its ghost layers, src within its interior
            if (it.lo[d] < 1 - D.G.g[d] || it.lo[d] + it.ext[d] - 1 > D.G.n[d] + D.G.g[d] ||
                it.slo[d] < 1 || it.slo[d] + it.ext[d] - 1 > S.G.n[d])
              return bfail(RHY_ERR_SHAPE, "sibling ghost region outside a box");
            vol *= it.ext[d];
          }
          f->amr_host.push_back(it);
          ncopy += vol;
        }
  }
  if ((rc = amr_upload(f))) return rc;
  const int ncitems = (int)f->amr_host.size() - nb;
  const AmrItem* items = f->amr_items;
#define RHY_AMR_FILL(ND, R)                                                                    \
  do {                                                                                         \
    R* arena = static_cast<R*>(f->arena);                                                      \
    hipLaunchKernelGGL((amr_prolong_kernel<ND, R>), dim3(cdiv(nprol, 256)), dim3(256), 0,     \
                       f->stream, items, nb, nprol, arena, vc.geom,                            \
                       static_cast<const R*>(vc.cells), half ? 1 : 0);                         \
    if (ncopy)                                                                                 \
      hipLaunchKernelGGL((amr_copy_kernel<ND, R>), dim3(cdiv(ncopy, 256)), dim3(256), 0,      \
                         f->stream, items + nb, ncitems, ncopy, arena);                        \
  } while (0)
  if (nb) {
    if (f->precision == RHY_MIXED) {
      if (nd == 1) RHY_AMR_FILL(1, float);
      else if (nd == 2) RHY_AMR_FILL(2, float);
      else RHY_AMR_FILL(3, float);
    } else {
      if (nd == 1) RHY_AMR_FILL(1, double);
      else if (nd == 2) RHY_AMR_FILL(2, double);
      else RHY_AMR_FILL(3, double);
    }
  }
#undef RHY_AMR_FILL
  BHIP_TRY(hipGetLastError());
  BHIP_TRY(hipStreamSynchronize(f->stream));
  return RHY_OK;
}

int rhy_amr_prolong(rhy_box_t* coarse, rhy_batch_t* f, const int32_t* lo, int32_t half) {
  BoxView vc;
  int rc;
  if (!coarse) return bfail(RHY_ERR_ARG, "NULL argument");
  if ((rc = box_view(coarse, &vc)))
    return bfail(rc, "coarse box: " + std::string(rhy_last_error()));
  return amr_prolong_v(vc, f, lo, half);
}

int rhy_amr_prolong_slot(rhy_batch_t* coarse, int32_t slot, rhy_batch_t* f, const int32_t* lo,
                         int32_t half) {
  BoxView vc;
  int rc;
  if (coarse == f) return bfail(RHY_ERR_ARG, "the coarse and fine levels are one batch");
  if ((rc = slot_view(coarse, slot, &vc))) return rc;
  return amr_prolong_v(vc, f, lo, half);
}

static int amr_restrict_v(const BoxView& vc, rhy_batch_t* f, const int32_t* lo) {
  int rc;
  if ((rc = amr_check(f, vc, lo))) return rc;
  const int nd = f->ndim, nb = (int)f->slots.size();
  f->amr_host.clear();
  int64_t total = 0;
  for (int i = 0; i < nb; ++i) {
    const BBox& B = f->slots[i].B;
    AmrItem it{};
    it.first = total;
    it.dst_off = B.cell_off;
    it.GD = B.G;
    for (int d = 0; d < 3; ++d) it.lo[d] = lo[3 * i + d];
    f->amr_host.push_back(it);
    int64_t n = 1;
    for (int d = 0; d < nd; ++d) n *= B.G.n[d] / 2;
    total += n;
  }
  if ((rc = amr_upload(f))) return rc;
#define RHY_RESTRICT(ND, R)                                                                      \
  hipLaunchKernelGGL((amr_restrict_kernel<ND, R>), dim3(cdiv(total, 256)), dim3(256), 0,        \
                     f->stream, f->amr_items, nb, total, static_cast<const R*>(f->arena),        \
                     vc.geom, static_cast<R*>(vc.cells))
  if (total) {
    if (f->precision == RHY_MIXED) {
      if (nd == 1) RHY_RESTRICT(1, float);
      else if (nd == 2) RHY_RESTRICT(2, float);
      else RHY_RESTRICT(3, float);
    } else {
      if (nd == 1) RHY_RESTRICT(1, double);
      else if (nd == 2) RHY_RESTRICT(2, double);
      else RHY_RESTRICT(3, double);
    }
  }
#undef RHY_RESTRICT
  BHIP_TRY(hipGetLastError());
  BHIP_TRY(hipStreamSynchronize(f->stream));
  return RHY_OK;
}

int rhy_amr_restrict(rhy_box_t* coarse, rhy_batch_t* f, const int32_t* lo) {
  BoxView vc;
  int rc;
  if (!coarse) return bfail(RHY_ERR_ARG, "NULL argument");
  if ((rc = box_view(coarse, &vc)))
    return bfail(rc, "coarse box: " + std::string(rhy_last_error()));
  return amr_restrict_v(vc, f, lo);
}

int rhy_amr_restrict_slot(rhy_batch_t* coarse, int32_t slot, rhy_batch_t* f, const int32_t* lo) {
  BoxView vc;
  int rc;
  if (coarse == f) return bfail(RHY_ERR_ARG, "the coarse and fine levels are one batch");
  if ((rc = slot_view(coarse, slot, &vc))) return rc;
  return amr_restrict_v(vc, f, lo);
}

static int amr_regrid_v(const BoxView& vc, rhy_batch_t* from, const int32_t* from_lo,
                        rhy_batch_t* to, const int32_t* to_lo) {
  int rc;
  if (!from || !to || !from_lo || !to_lo) return bfail(RHY_ERR_ARG, "NULL argument");
  if (from == to) return bfail(RHY_ERR_ARG, "the old and new fine levels are one batch");
  if ((rc = amr_check(from, vc, from_lo))) return rc;
  if ((rc = amr_check(to, vc, to_lo))) return rc;
  BHIP_TRY(hipStreamSynchronize(from->stream));
  const int nd = to->ndim, nto = (int)to->slots.size(), nfrom = (int)from->slots.size();
  to->amr_host.clear();
  int64_t total = 0;
  for (int i = 0; i < nto; ++i) {
    const BBox& B = to->slots[i].B;
    AmrItem it{};
    it.first = total;
    it.dst_off = B.cell_off;
    it.GD = B.G;
    for (int d = 0; d < 3; ++d) it.lo[d] = to_lo[3 * i + d];
    to->amr_host.push_back(it);
    int64_t m = 1;
    for (int d = 0; d < nd; ++d) m *= B.G.n[d] + 4;
    total += m;
  }
  for (int j = 0; j < nfrom; ++j) {
    const BBox& B = from->slots[j].B;
    AmrItem it{};
    it.first = total;
    it.dst_off = B.cell_off;
    it.GD = B.G;
    for (int d = 0; d < 3; ++d) it.lo[d] = from_lo[3 * j + d];
    to->amr_host.push_back(it);
  }
  if ((rc = amr_upload(to))) return rc;
  if (total > 0) {
#define RHY_REGRID(ND, R)                                                                       \
  hipLaunchKernelGGL((amr_regrid_kernel<ND, R>), dim3(cdiv(total, 256)), dim3(256), 0,         \
                     to->stream, to->amr_items, nto, nfrom, total, static_cast<R*>(to->arena), \
                     static_cast<const R*>(from->arena), vc.geom, static_cast<const R*>(vc.cells))
    if (to->precision == RHY_MIXED) {
      if (nd == 1) RHY_REGRID(1, float);
      else if (nd == 2) RHY_REGRID(2, float);
      else RHY_REGRID(3, float);
    } else {
      if (nd == 1) RHY_REGRID(1, double);
      else if (nd == 2) RHY_REGRID(2, double);
      else RHY_REGRID(3, double);
    }
#undef RHY_REGRID
  }
  BHIP_TRY(hipGetLastError());
  BHIP_TRY(hipStreamSynchronize(to->stream));
  return RHY_OK;
}

int rhy_amr_regrid(rhy_box_t* coarse, rhy_batch_t* from, const int32_t* from_lo,
                   rhy_batch_t* to, const int32_t* to_lo) {
  BoxView vc;
  int rc;
  if (!coarse) return bfail(RHY_ERR_ARG, "NULL argument");
  if ((rc = box_view(coarse, &vc))) return bfail(rc, "coarse box: " + std::string(rhy_last_error()));
  return amr_regrid_v(vc, from, from_lo, to, to_lo);
}

int rhy_amr_regrid_slot(rhy_batch_t* coarse, int32_t slot, rhy_batch_t* from,
                        const int32_t* from_lo, rhy_batch_t* to, const int32_t* to_lo) {
  BoxView vc;
  int rc;
  if (coarse == from || coarse == to)
    return bfail(RHY_ERR_ARG, "the coarse level and a fine level are one batch");
  if ((rc = slot_view(coarse, slot, &vc))) return rc;
  return amr_regrid_v(vc, from, from_lo, to, to_lo);
}

// The register items of every fine box, grouped by (axis, side) pass in
// pass order a = 0..ndim-1, low side then high side; pass_begin[p] and
// pass_total[p] locate pass p's items and threads.  Items number their
// threads across all passes (pass p's start after those of passes < p), so
// one launch can take every pass.  Returns the register length in doubles.
static int64_t reflux_items(rhy_batch_t* f, const int32_t* lo, int pass_begin[7],
                            int64_t pass_total[6]) {
  const int nd = f->ndim, nb = (int)f->slots.size();
  f->amr_host.clear();
  int64_t reg = 0, first = 0;
  for (int a = 0; a < nd; ++a)
    for (int s = 0; s < 2; ++s) {
      const int pass = 2 * a + s;
      pass_begin[pass] = (int)f->amr_host.size();
      const int64_t pass_first = first;
      for (int i = 0; i < nb; ++i) {
        const BBox& B = f->slots[i].B;
        AmrItem it{};
        it.first = first;
        it.dst_off = reg;
        it.src_off = B.flux_off;
        it.GD = B.G;
        it.slo[0] = a;
        it.slo[1] = s;
        for (int d = 0; d < 3; ++d) it.lo[d] = lo[3 * i + d];
        int64_t m = 1;
        for (int d = 0; d < nd; ++d)
          if (d != a) m *= B.G.n[d] / 2;
        f->amr_host.push_back(it);
        first += m;
        reg += m * (nd + 2);
      }
      pass_total[pass] = first - pass_first;
    }
  pass_begin[2 * nd] = (int)f->amr_host.size();
  return reg;
}

// The register/reflux work items for this placement on the device (built
// and uploaded when the placement or the boxes change, synchronously, so
// the calls themselves only enqueue kernels); returns the register length.
static int reflux_prepare(rhy_batch_t* f, const int32_t* lo, int64_t* len) {
  std::vector<int64_t> key;
  for (int i = 0; i < (int)f->slots.size(); ++i)
    for (int d = 0; d < 3; ++d) {
      key.push_back(lo[3 * i + d]);
      key.push_back(f->slots[i].B.G.n[d]);
    }
  if (key == f->reg_key && f->reg_items) {
    *len = f->amr_reg_len;
    return RHY_OK;
  }
  BHIP_TRY(hipStreamSynchronize(f->stream));
  const int64_t n = reflux_items(f, lo, f->reg_pb, f->reg_pt);
  const int64_t ni = (int64_t)f->amr_host.size();
  if (ni > f->reg_items_cap) {
    if (f->reg_items) (void)hipFree(f->reg_items);
    f->reg_items = nullptr;
    f->reg_items_cap = 0;
    if (dev_alloc(&f->reg_items, sizeof(AmrItem) * (size_t)std::max<int64_t>(ni, 1)) !=
        hipSuccess) {
      f->reg_items = nullptr;
      return bfail(RHY_ERR_OOM, "flux register items allocation failed");
    }
    f->reg_items_cap = ni;
  }
  if (ni)
    BHIP_TRY(hipMemcpy(f->reg_items, f->amr_host.data(), sizeof(AmrItem) * (size_t)ni,
                       hipMemcpyHostToDevice));
  if (n > f->amr_reg_cap) {
    if (f->amr_reg) (void)hipFree(f->amr_reg);
    f->amr_reg = nullptr;
    f->amr_reg_cap = 0;
    if (dev_alloc(&f->amr_reg, sizeof(double) * (size_t)std::max<int64_t>(n, 1)) !=
        hipSuccess) {
      f->amr_reg = nullptr;
      f->reg_key.clear();
      return bfail(RHY_ERR_OOM, "flux register allocation failed");
    }
    f->amr_reg_cap = n;
  }
  f->amr_reg_len = n;
  f->reg_key = key;
  *len = n;
  return RHY_OK;
}

int rhy_amr_flux_register(rhy_batch_t* f, const int32_t* lo, int32_t reset) {
  if (!f || !lo) return bfail(RHY_ERR_ARG, "NULL argument");
  if (f->pool) return bfail(RHY_ERR_ARG, "the fine level is a fixed batch, not a pool");
  for (int i = 0; i < (int)f->slots.size(); ++i)
    for (int d = 0; d < f->ndim; ++d)
      if (f->slots[i].B.G.n[d] % 2) return bfail(RHY_ERR_SHAPE, "fine extents must be even");
  const std::vector<int64_t> before = f->reg_key;
  int64_t len = 0;
  int rc;
  if ((rc = reflux_prepare(f, lo, &len))) return rc;
  if (!reset && f->reg_key != before)
    return bfail(RHY_ERR_ARG, "flux register not started (reset = 1) for these boxes");
  // one launch over the items of every pass (each thread writes its own
  // register entry), enqueued on the batch's stream behind the substep; no
  // wait, so the coarse step on the box's stream overlaps the fine substep
  const int nd = f->ndim;
  const int ni = f->reg_pb[2 * nd];
  int64_t total = 0;
  for (int pass = 0; pass < 2 * nd; ++pass) total += f->reg_pt[pass];
  if (!total || !ni) return RHY_OK;
#define RHY_REG(ND, R)                                                                         \
  hipLaunchKernelGGL((amr_register_kernel<ND, R>), dim3(cdiv(total, 256)), dim3(256), 0,      \
                     f->stream, f->reg_items, ni, total, static_cast<const R*>(f->arena),     \
                     f->amr_reg, reset ? 1 : 0)
  if (f->precision == RHY_MIXED) {
    if (nd == 1) RHY_REG(1, float);
    else if (nd == 2) RHY_REG(2, float);
    else RHY_REG(3, float);
  } else {
    if (nd == 1) RHY_REG(1, double);
    else if (nd == 2) RHY_REG(2, double);
    else RHY_REG(3, double);
  }
#undef RHY_REG
  BHIP_TRY(hipGetLastError());
  return RHY_OK;
}

static int amr_reflux_v(const BoxView& vc, rhy_batch_t* f, const int32_t* lo, double dt,
                        const double dx[3]) {
  int rc;
  if (!dx) return bfail(RHY_ERR_ARG, "NULL argument");
  if ((rc = amr_check(f, vc, lo))) return rc;
  const std::vector<int64_t> before = f->reg_key;
  int64_t len = 0;
  if ((rc = reflux_prepare(f, lo, &len))) return rc;
  if (f->reg_key != before || !f->amr_reg)
    return bfail(RHY_ERR_ARG, "flux register not filled for these boxes");
  // one launch per (axis, side) pass: faces of two passes can share their
  // outside coarse cell (a corner between two fine boxes)
  const int nd = f->ndim;
  int64_t base = 0;
  for (int pass = 0; pass < 2 * nd; base += f->reg_pt[pass], ++pass) {
    const int ni = f->reg_pb[pass + 1] - f->reg_pb[pass];
    if (!f->reg_pt[pass] || !ni) continue;
    const AmrItem* items = f->reg_items + f->reg_pb[pass];
    const double dtdx = dt / dx[pass / 2];
    const int a = pass / 2;
#define RHY_REFLUX(ND, R)                                                                       \
  hipLaunchKernelGGL((amr_reflux_kernel<ND, R>), dim3(cdiv(f->reg_pt[pass], 256)), dim3(256), 0, \
                     f->stream, items, ni, f->reg_pt[pass], base, f->amr_reg, vc.geom,         \
                     static_cast<R*>(vc.cells), static_cast<const R*>(vc.flux), dtdx)
#define RHY_REFLUX_S(ND, A, R, P)                                                               \
  hipLaunchKernelGGL((amr_reflux_kernel<ND, R, A, P>), dim3(cdiv(f->reg_pt[pass], 256)),       \
                     dim3(256), 0, f->stream, items, ni, f->reg_pt[pass], base, f->amr_reg,    \
                     vc.geom, static_cast<R*>(vc.cells), static_cast<const R*>(vc.flux), dtdx, \
                     vc.k, static_cast<const R*>(vc.prev), vc.step_dt, vc.step_dx[a])
    if (vc.prev && a > 0) {  // y/z coarse fluxes of an accumulating coarse step
      if (f->precision == RHY_MIXED) {
        if (nd == 2) RHY_REFLUX_S(2, 1, float, MixedMath);
        else if (a == 1) RHY_REFLUX_S(3, 1, float, MixedMath);
        else RHY_REFLUX_S(3, 2, float, MixedMath);
      } else {
        if (nd == 2) RHY_REFLUX_S(2, 1, double, PowF64Fix);
        else if (a == 1) RHY_REFLUX_S(3, 1, double, PowF64Fix);
        else RHY_REFLUX_S(3, 2, double, PowF64Fix);
      }
    } else if (f->precision == RHY_MIXED) {
      if (nd == 1) RHY_REFLUX(1, float);
      else if (nd == 2) RHY_REFLUX(2, float);
      else RHY_REFLUX(3, float);
    } else {
      if (nd == 1) RHY_REFLUX(1, double);
      else if (nd == 2) RHY_REFLUX(2, double);
      else RHY_REFLUX(3, double);
    }
#undef RHY_REFLUX_S
#undef RHY_REFLUX
  }
  BHIP_TRY(hipGetLastError());
  BHIP_TRY(hipStreamSynchronize(f->stream));
  return RHY_OK;
}

int rhy_amr_reflux(rhy_box_t* coarse, rhy_batch_t* f, const int32_t* lo, double dt,
                   const double dx[3]) {
  BoxView vc;
  int rc;
  if (!coarse) return bfail(RHY_ERR_ARG, "NULL argument");
  if ((rc = box_view(coarse, &vc)))
    return bfail(rc, "coarse box: " + std::string(rhy_last_error()));
  return amr_reflux_v(vc, f, lo, dt, dx);
}

int rhy_amr_reflux_slot(rhy_batch_t* coarse, int32_t slot, rhy_batch_t* f, const int32_t* lo,
                        double dt, const double dx[3]) {
  BoxView vc;
  int rc;
  if (coarse == f) return bfail(RHY_ERR_ARG, "the coarse and fine levels are one batch");
  if ((rc = slot_view(coarse, slot, &vc))) return rc;
  return amr_reflux_v(vc, f, lo, dt, dx);
}

int rhy_batch_sync(rhy_batch_t* b) {
  if (!b) return bfail(RHY_ERR_ARG, "batch is NULL");
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_batch_last_deferred(rhy_batch_t* b, int64_t* n) {
  if (!b || !n) return bfail(RHY_ERR_ARG, "NULL argument");
  uint32_t c = 0;
  BHIP_TRY(hipStreamSynchronize(b->stream));
  BHIP_TRY(hipMemcpy(&c, b->defer_count, sizeof(c), hipMemcpyDeviceToHost));
  *n = c;
  return RHY_OK;
}

int rhy_batch_diagnostics(rhy_batch_t* b, int32_t box, rhy_diag_t* out) {
  if (!slot_ok(b, box) || !out) return bfail(RHY_ERR_ARG, "bad diagnostics request");
  constexpr int kBlocks = 256;
  if (!b->diag && dev_alloc(&b->diag, sizeof(DiagPartial) * kBlocks) != hipSuccess)
    return bfail(RHY_ERR_OOM, "diagnostics buffer allocation failed");
  const BBox& B = b->slots[box].B;
  const Geom& G = B.G;
  const unsigned nb = std::min<unsigned>(cdiv(G.plane, 256), kBlocks);
  const bool mixed = b->precision == RHY_MIXED;
#define RHY_BDIAG(ND)                                                                        \
  do {                                                                                       \
    if (mixed)                                                                               \
      hipLaunchKernelGGL((diag_kernel<ND, float>), dim3(nb), dim3(256), 0, b->stream, G,      \
                         b->ncmp, static_cast<const float*>(b->arena) + B.cell_off,           \
                         static_cast<const float*>(b->arena) + B.cell_off + (ND + 2) * G.plane, \
                         b->diag);                                                           \
    else                                                                                     \
      hipLaunchKernelGGL((diag_kernel<ND, double>), dim3(nb), dim3(256), 0, b->stream, G,     \
                         b->ncmp, static_cast<const double*>(b->arena) + B.cell_off,          \
                         static_cast<const double*>(b->arena) + B.cell_off + (ND + 2) * G.plane, \
                         b->diag);                                                           \
  } while (0)
  if (b->ndim == 1) RHY_BDIAG(1);
  else if (b->ndim == 2) RHY_BDIAG(2);
  else RHY_BDIAG(3);
#undef RHY_BDIAG
  BHIP_TRY(hipGetLastError());
  std::vector<DiagPartial> parts(nb);
  BHIP_TRY(hipMemcpyAsync(parts.data(), b->diag, sizeof(DiagPartial) * nb,
                          hipMemcpyDeviceToHost, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  diag_finish(parts, G, out);
  return RHY_OK;
}

int rhy_batch_last_dts(rhy_batch_t* b, int32_t n, double* dts) {
  if (!b || !dts || n < 0 || n > b->last_nsteps) return bfail(RHY_ERR_ARG, "bad dts request");
  if (n == 0) return RHY_OK;
  BHIP_TRY(hipMemcpyAsync(dts, b->dtlog, sizeof(double) * (size_t)n * b->slots.size(),
                          hipMemcpyDeviceToHost, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

}  // extern "C"
