// rhy_batch.hip -- many independent boxes of uneven shapes advanced
// together: one launch per phase (dt, fluxes of every axis of every box,
// update of every box) over a tile table, with persistent workgroups that
// stride over the tiles, and the step captured once in a hipGraph and
// replayed.  Each box keeps its own ghost fill (the configured faces) and
// its own CFL dt, i.e. every box evolves exactly as it would alone
// (rhy_run on that box), which is what the tests check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

// per-box descriptor, device resident
struct BBox {
  Geom G;
  int64_t cell_off;  // element offset of the box's planes in the cell arena
  int64_t flux_off;  // element offset of its flux planes
  double dx[3];
  double cmdx;       // courant * minval(dx)
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
  int32_t pad;
};

template <int ND, int LIM, typename R, class PW>
__global__ __launch_bounds__(256) void batch_flux_kernel(Consts k, const BBox* __restrict__ boxes,
                                                         const BTile* __restrict__ tiles,
                                                         int32_t ntiles, const R* __restrict__ cells,
                                                         R* __restrict__ flux,
                                                         const double* __restrict__ dt,
                                                         int32_t nf_x, int32_t nf_yz) {
  for (int32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const BTile T = tiles[ti];
    if ((int32_t)threadIdx.x >= T.count) continue;
    const BBox& B = boxes[T.box];
    const int64_t t = T.base + threadIdx.x;
    const R* c = cells + B.cell_off;
    R* F = flux + B.flux_off + (int64_t)T.kind * (ND + 2) * B.G.plane;
    const double d = dt[T.box];
    if (T.kind == 0) flux_walk<ND, 0, LIM, R, PW>(k, B.G, c, F, d, B.dx[0], nf_x, t);
    if constexpr (ND > 1)
      if (T.kind == 1) flux_walk<ND, 1, LIM, R, PW>(k, B.G, c, F, d, B.dx[1], nf_yz, t);
    if constexpr (ND > 2)
      if (T.kind == 2) flux_walk<ND, 2, LIM, R, PW>(k, B.G, c, F, d, B.dx[2], nf_yz, t);
  }
}

// Update (memory_intensive or cpu_intensive order) with the fused ghost
// images and CFL partial of every box; tile ti writes partial[ti - first].
template <int ND, int ORDER, typename R>
__global__ __launch_bounds__(256) void batch_update_kernel(Consts k, const BBox* __restrict__ boxes,
                                                           const BTile* __restrict__ tiles,
                                                           int32_t first, int32_t ntiles,
                                                           R* __restrict__ cells,
                                                           const R* __restrict__ flux,
                                                           const double* __restrict__ dt,
                                                           BcTypes bc, double* __restrict__ partial) {
  constexpr int NV = ND + 2;
  for (int32_t ti = first + blockIdx.x; ti < first + ntiles; ti += gridDim.x) {
    const BTile T = tiles[ti];
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
      R* c = cells + B.cell_off;
      const R* F0 = flux + B.flux_off;
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
      write_ghost_images<ND, R>(G, bc, c, i, j, kk, u);
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
                                                        const BTile* __restrict__ tiles,
                                                        int32_t first, int32_t ntiles,
                                                        const R* __restrict__ cells,
                                                        double* __restrict__ partial) {
  for (int32_t ti = first + blockIdx.x; ti < first + ntiles; ti += gridDim.x) {
    const BTile T = tiles[ti];
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
      load_state<ND, R>(cells + B.cell_off, G.plane, goff(G, i, j, kk), ur);
      double u[ND + 2];
#pragma unroll
      for (int c = 0; c < ND + 2; ++c) u[c] = (double)ur[c];
      v = cfl_speed<ND>(k, u);
    }
    block_max_to_index(v, partial, ti - first);
    __syncthreads();
  }
}

// dt of every box (block b reduces box b's partials); logs it at
// dtlog[*step * nboxes + b] when the step counter is below the log size.
__global__ __launch_bounds__(256) void batch_dt_kernel(const BBox* __restrict__ boxes,
                                                       const double* __restrict__ partial,
                                                       double* __restrict__ dt,
                                                       double* __restrict__ dtlog,
                                                       const int32_t* __restrict__ step,
                                                       int32_t log_steps, int32_t nboxes) {
  __shared__ double wave_max[4];
  const int b = blockIdx.x;
  const BBox& B = boxes[b];
  double v = 0.0;
  for (int i = B.part_begin + threadIdx.x; i < B.part_end; i += blockDim.x)
    v = partial[i] > v ? partial[i] : v;
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
    if (dtlog && s < log_steps) dtlog[(int64_t)s * nboxes + b] = d;
  }
}

__global__ void batch_step_counter_kernel(int32_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = *step + 1;
}

__global__ void batch_set_counter_kernel(int32_t* step, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}

unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct rhy_batch {
  int32_t ndim = 0, ncmp = 0, precision = RHY_FP64, nboxes = 0;
  size_t esz = sizeof(double);
  std::vector<BBox> hboxes;
  std::vector<BTile> htiles;
  int32_t nflux_tiles = 0, nupd_tiles = 0;
  int32_t block = 256, nf_x = 4, nf_yz = 8, grid_cap = 0;
  hipStream_t stream = nullptr;
  void* cells = nullptr;
  void* flux = nullptr;
  BBox* dboxes = nullptr;
  BTile* dtiles = nullptr;
  double* partial = nullptr;
  double* dt = nullptr;
  double* dtlog = nullptr;
  int32_t dtlog_steps = 0;
  int32_t* step = nullptr;
  int32_t last_nsteps = 0;
  double* inflow = nullptr;
  hipGraphExec_t graph = nullptr;
  uint64_t graph_key = 0;
  int64_t cells_elems = 0, flux_elems = 0;
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
  return k;
}

// Rebuild the tile table for the current tile shape; box b's update tiles
// are contiguous so its CFL partials are the range [part_begin, part_end).
int build_tiles(rhy_batch* b) {
  b->htiles.clear();
  const int nd = b->ndim;
  for (int a = 0; a < nd; ++a)
    for (int i = 0; i < b->nboxes; ++i) {
      const Geom& G = b->hboxes[i].G;
      const int64_t n = flux_threads(G, a, a == 0 ? b->nf_x : b->nf_yz);
      for (int64_t base = 0; base < n; base += b->block)
        b->htiles.push_back({i, a, base, (int32_t)std::min<int64_t>(b->block, n - base), 0});
    }
  b->nflux_tiles = (int32_t)b->htiles.size();
  int32_t part = 0;
  for (int i = 0; i < b->nboxes; ++i) {
    const Geom& G = b->hboxes[i].G;
    const int64_t n = (int64_t)G.n[0] * G.n[1] * G.n[2];
    b->hboxes[i].part_begin = part;
    for (int64_t base = 0; base < n; base += b->block) {
      b->htiles.push_back({i, 3, base, (int32_t)std::min<int64_t>(b->block, n - base), 0});
      ++part;
    }
    b->hboxes[i].part_end = part;
  }
  b->nupd_tiles = part;
  if (b->dtiles) hipFree(b->dtiles);
  if (b->partial) hipFree(b->partial);
  b->dtiles = nullptr;
  b->partial = nullptr;
  if (hipMalloc(&b->dtiles, sizeof(BTile) * b->htiles.size()) != hipSuccess ||
      hipMalloc(&b->partial, sizeof(double) * std::max(1, b->nupd_tiles)) != hipSuccess)
    return bfail(RHY_ERR_OOM, "tile table allocation failed");
  BHIP_TRY(hipMemcpy(b->dtiles, b->htiles.data(), sizeof(BTile) * b->htiles.size(),
                     hipMemcpyHostToDevice));
  BHIP_TRY(hipMemcpy(b->dboxes, b->hboxes.data(), sizeof(BBox) * b->nboxes,
                     hipMemcpyHostToDevice));
  if (b->graph) {
    hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return RHY_OK;
}

unsigned persistent_grid(rhy_batch* b, int32_t ntiles) {
  unsigned cap = b->grid_cap > 0 ? (unsigned)b->grid_cap : 256u * 16u;
  return std::max(1u, std::min(cap, (unsigned)ntiles));
}

template <int ND, typename R, class PW>
void launch_flux_phase(rhy_batch* b, const Consts& k) {
  const unsigned g = persistent_grid(b, b->nflux_tiles);
  const R* c = static_cast<const R*>(b->cells);
  R* f = static_cast<R*>(b->flux);
  switch (k.limiter) {
    case 1:
      hipLaunchKernelGGL((batch_flux_kernel<ND, 1, R, PW>), dim3(g), dim3(b->block), 0, b->stream,
                         k, b->dboxes, b->dtiles, b->nflux_tiles, c, f, b->dt, b->nf_x, b->nf_yz);
      break;
    case 2:
      hipLaunchKernelGGL((batch_flux_kernel<ND, 2, R, PW>), dim3(g), dim3(b->block), 0, b->stream,
                         k, b->dboxes, b->dtiles, b->nflux_tiles, c, f, b->dt, b->nf_x, b->nf_yz);
      break;
    case 3:
      hipLaunchKernelGGL((batch_flux_kernel<ND, 3, R, PW>), dim3(g), dim3(b->block), 0, b->stream,
                         k, b->dboxes, b->dtiles, b->nflux_tiles, c, f, b->dt, b->nf_x, b->nf_yz);
      break;
    default:
      hipLaunchKernelGGL((batch_flux_kernel<ND, 4, R, PW>), dim3(g), dim3(b->block), 0, b->stream,
                         k, b->dboxes, b->dtiles, b->nflux_tiles, c, f, b->dt, b->nf_x, b->nf_yz);
      break;
  }
}

template <int ND, typename R>
void launch_update_phase(rhy_batch* b, const Consts& k, int order, BcTypes bc) {
  const unsigned g = persistent_grid(b, b->nupd_tiles);
  R* c = static_cast<R*>(b->cells);
  const R* f = static_cast<const R*>(b->flux);
  if (order == RHY_MEMORY_INTENSIVE)
    hipLaunchKernelGGL((batch_update_kernel<ND, 2, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                       b->dboxes, b->dtiles, b->nflux_tiles, b->nupd_tiles, c, f, b->dt, bc,
                       b->partial);
  else
    hipLaunchKernelGGL((batch_update_kernel<ND, 1, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                       b->dboxes, b->dtiles, b->nflux_tiles, b->nupd_tiles, c, f, b->dt, bc,
                       b->partial);
}

template <int ND, typename R, class PW>
void enqueue_step(rhy_batch* b, const Consts& k, int order, BcTypes bc) {
  hipLaunchKernelGGL(batch_dt_kernel, dim3(b->nboxes), dim3(256), 0, b->stream, b->dboxes,
                     b->partial, b->dt, b->dtlog, b->step, b->dtlog_steps, b->nboxes);
  launch_flux_phase<ND, R, PW>(b, k);
  launch_update_phase<ND, R>(b, k, order, bc);
  hipLaunchKernelGGL(batch_step_counter_kernel, dim3(1), dim3(64), 0, b->stream, b->step);
}

template <int ND, typename R>
int prologue(rhy_batch* b, const rhy_params_t* p, const Consts& k) {
  // ghost fill of every box, then CFL partials of the current state
  BHIP_TRY(hipMemcpyAsync(b->inflow, p->cons_inflows, sizeof(double) * 6 * RHY_MAX_CMP,
                          hipMemcpyHostToDevice, b->stream));
  BcTypes bc{};
  for (int f = 0; f < 6; ++f) bc.t[f] = p->bc_types[f];
  R* c = static_cast<R*>(b->cells);
  for (int i = 0; i < b->nboxes; ++i) {
    const Geom& G = b->hboxes[i].G;
    int64_t maxface = 1;
    for (int a = 0; a < ND; ++a) {
      int64_t m = 1;
      for (int d = 0; d < 3; ++d)
        if (d != a) m *= G.n[d];
      maxface = std::max(maxface, m);
    }
    hipLaunchKernelGGL((bc_kernel<ND, R>), dim3(cdiv(maxface, 256), 2 * ND), dim3(256), 0,
                       b->stream, G, bc, b->inflow, b->ncmp, c + b->hboxes[i].cell_off, -1);
  }
  const unsigned g = persistent_grid(b, b->nupd_tiles);
  hipLaunchKernelGGL((batch_cfl_kernel<ND, R>), dim3(g), dim3(b->block), 0, b->stream, k,
                     b->dboxes, b->dtiles, b->nflux_tiles, b->nupd_tiles,
                     static_cast<const R*>(b->cells), b->partial);
  hipLaunchKernelGGL(batch_set_counter_kernel, dim3(1), dim3(64), 0, b->stream, b->step, 0);
  BHIP_TRY(hipGetLastError());
  return RHY_OK;
}

template <int ND, typename R, class PW>
int run_t(rhy_batch* b, const rhy_params_t* p, int32_t nsteps, int32_t use_graph) {
  const Consts k = consts_of(p);
  BcTypes bc{};
  for (int f = 0; f < 6; ++f) bc.t[f] = p->bc_types[f];
  // courant * minval(dx) per box
  bool changed = false;
  for (auto& B : b->hboxes) {
    double m = B.dx[0];
    for (int d = 1; d < ND; ++d) m = std::min(m, B.dx[d]);
    const double cm = p->courant * m;
    changed = changed || cm != B.cmdx;
    B.cmdx = cm;
  }
  if (changed) {
    BHIP_TRY(hipMemcpy(b->dboxes, b->hboxes.data(), sizeof(BBox) * b->nboxes,
                       hipMemcpyHostToDevice));
  }
  if (nsteps > b->dtlog_steps) {
    if (b->dtlog) hipFree(b->dtlog);
    b->dtlog = nullptr;
    if (hipMalloc(&b->dtlog, sizeof(double) * (size_t)nsteps * b->nboxes) != hipSuccess)
      return bfail(RHY_ERR_OOM, "dt log allocation failed");
    b->dtlog_steps = nsteps;
  }
  int rc = prologue<ND, R>(b, p, k);
  if (rc) return rc;
  if (!use_graph) {
    for (int32_t s = 0; s < nsteps; ++s) enqueue_step<ND, R, PW>(b, k, p->solver_type, bc);
    BHIP_TRY(hipGetLastError());
    b->last_nsteps = nsteps;
    return RHY_OK;
  }
  // the captured step depends on the parameters and the tile shape
  uint64_t key = 1469598103934665603ull;
  auto mix = [&key](const void* data, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) key = (key ^ c[i]) * 1099511628211ull;
  };
  mix(p, sizeof(*p));
  mix(&b->block, sizeof(b->block));
  mix(&b->nf_x, sizeof(b->nf_x));
  mix(&b->nf_yz, sizeof(b->nf_yz));
  mix(&b->grid_cap, sizeof(b->grid_cap));
  if (!b->graph || b->graph_key != key) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    hipGraph_t g = nullptr;
    BHIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step<ND, R, PW>(b, k, p->solver_type, bc);
    BHIP_TRY(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(RHY_ERR_HIP, "hipGraphInstantiate failed");
    b->graph_key = key;
  }
  for (int32_t s = 0; s < nsteps; ++s) BHIP_TRY(hipGraphLaunch(b->graph, b->stream));
  b->last_nsteps = nsteps;
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
  if ((rc = rhy_init(0))) return bfail(rc, rhy_last_error());
  if (ndim < 1 || ndim > 3 || nboxes < 1 || !dims || !dx)
    return bfail(RHY_ERR_ARG, "bad batch arguments");
  if (ncmp < ndim + 2 || ncmp > RHY_MAX_CMP) return bfail(RHY_ERR_ARG, "ncmp out of range");
  if (precision != RHY_FP64 && precision != RHY_MIXED)
    return bfail(RHY_ERR_ARG, "unsupported precision");
  rhy_batch* b = new rhy_batch();
  b->ndim = ndim;
  b->ncmp = ncmp;
  b->precision = precision;
  b->nboxes = nboxes;
  b->esz = precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  int64_t coff = 0, foff = 0;
  for (int i = 0; i < nboxes; ++i) {
    BBox B{};
    int64_t st = 1;
    for (int d = 0; d < 3; ++d) {
      const int n = dims[3 * i + d];
      if (d < ndim && n < 2) {
        delete b;
        return bfail(RHY_ERR_SHAPE, "active extents must be >= 2");
      }
      if (d >= ndim && n != 1) {
        delete b;
        return bfail(RHY_ERR_SHAPE, "inactive extents must be 1");
      }
      B.G.n[d] = n;
      B.G.g[d] = d < ndim ? 2 : 0;
      B.G.s[d] = st;
      st *= (int64_t)(n + 2 * B.G.g[d]);
      B.dx[d] = dx[3 * i + d];
    }
    B.G.plane = st;
    B.cell_off = coff;
    B.flux_off = foff;
    B.cmdx = -1.0;
    coff += st * ncmp;
    foff += st * (ndim + 2) * ndim;
    b->hboxes.push_back(B);
  }
  b->cells_elems = coff;
  b->flux_elems = foff;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&b->cells, b->esz * (size_t)coff) != hipSuccess ||
      hipMalloc(&b->flux, b->esz * (size_t)foff) != hipSuccess ||
      hipMalloc(&b->dboxes, sizeof(BBox) * nboxes) != hipSuccess ||
      hipMalloc(&b->dt, sizeof(double) * nboxes) != hipSuccess ||
      hipMalloc(&b->step, sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&b->inflow, sizeof(double) * 6 * RHY_MAX_CMP) != hipSuccess) {
    rhy_batch_destroy(b);
    return bfail(RHY_ERR_OOM, "batch allocation failed");
  }
  if (hipMemsetAsync(b->cells, 0, b->esz * (size_t)coff, b->stream) != hipSuccess) {
    rhy_batch_destroy(b);
    return bfail(RHY_ERR_HIP, "initial clear failed");
  }
  if ((rc = build_tiles(b))) {
    rhy_batch_destroy(b);
    return rc;
  }
  *out = b;
  return RHY_OK;
}

int rhy_batch_destroy(rhy_batch_t* b) {
  if (!b) return RHY_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->graph) hipGraphExecDestroy(b->graph);
  for (void* p : {(void*)b->cells, (void*)b->flux, (void*)b->dboxes, (void*)b->dtiles,
                  (void*)b->partial, (void*)b->dt, (void*)b->dtlog, (void*)b->step,
                  (void*)b->inflow})
    if (p) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return RHY_OK;
}

int rhy_batch_bytes(rhy_batch_t* b, int64_t* bytes) {
  if (!b || !bytes) return bfail(RHY_ERR_ARG, "NULL argument");
  *bytes = (int64_t)b->esz * (b->cells_elems + b->flux_elems);
  return RHY_OK;
}

int rhy_batch_upload(rhy_batch_t* b, int32_t box, const double* cells) {
  if (!b || !cells || box < 0 || box >= b->nboxes) return bfail(RHY_ERR_ARG, "bad upload");
  const BBox& B = b->hboxes[box];
  const size_t n = (size_t)B.G.plane * b->ncmp;
  if (b->precision == RHY_MIXED) {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)cells[i];
    BHIP_TRY(hipMemcpyAsync(static_cast<float*>(b->cells) + B.cell_off, tmp.data(),
                            sizeof(float) * n, hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    return RHY_OK;
  }
  BHIP_TRY(hipMemcpyAsync(static_cast<double*>(b->cells) + B.cell_off, cells,
                          sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_batch_download(rhy_batch_t* b, int32_t box, double* cells) {
  if (!b || !cells || box < 0 || box >= b->nboxes) return bfail(RHY_ERR_ARG, "bad download");
  const BBox& B = b->hboxes[box];
  const size_t n = (size_t)B.G.plane * b->ncmp;
  if (b->precision == RHY_MIXED) {
    std::vector<float> tmp(n);
    BHIP_TRY(hipMemcpyAsync(tmp.data(), static_cast<float*>(b->cells) + B.cell_off,
                            sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < n; ++i) cells[i] = (double)tmp[i];
    return RHY_OK;
  }
  BHIP_TRY(hipMemcpyAsync(cells, static_cast<double*>(b->cells) + B.cell_off,
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
  if (p->limiter < 1 || p->limiter > 4) return bfail(RHY_ERR_ARG, "unknown slope limiter");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return bfail(RHY_ERR_SOLVER, "Unknown solver type");
  if (nsteps < 0) return bfail(RHY_ERR_ARG, "nsteps < 0");
  const bool mixed = b->precision == RHY_MIXED;
  switch (b->ndim) {
    case 1:
      return mixed ? run_t<1, float, PowFast>(b, p, nsteps, use_graph)
                   : run_t<1, double, PowLib>(b, p, nsteps, use_graph);
    case 2:
      return mixed ? run_t<2, float, PowFast>(b, p, nsteps, use_graph)
                   : run_t<2, double, PowLib>(b, p, nsteps, use_graph);
    default:
      return mixed ? run_t<3, float, PowFast>(b, p, nsteps, use_graph)
                   : run_t<3, double, PowLib>(b, p, nsteps, use_graph);
  }
}

int rhy_batch_sync(rhy_batch_t* b) {
  if (!b) return bfail(RHY_ERR_ARG, "batch is NULL");
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_batch_last_dts(rhy_batch_t* b, int32_t n, double* dts) {
  if (!b || !dts || n < 0 || n > b->last_nsteps) return bfail(RHY_ERR_ARG, "bad dts request");
  if (n == 0) return RHY_OK;
  BHIP_TRY(hipMemcpyAsync(dts, b->dtlog, sizeof(double) * (size_t)n * b->nboxes,
                          hipMemcpyDeviceToHost, b->stream));
  BHIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

}  // extern "C"
