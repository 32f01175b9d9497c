// rhy_api.hip -- the C ABI of include/rhyme_hip.h: device selection, the
// device-resident box (state, flux scratch, CFL slots, dt log, stream),
// host mirror transfers, and the launch sequences of the module procedures.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rhyme_hip.h"
#include "rhy_kernels.hpp"

using namespace rhy;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(RHY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

std::once_flag g_init_once;
int g_init_rc = RHY_ERR_NODEV;
int g_device = -1;

// Scratch of rhy_mh_solve_noalloc (flux planes, deferred-face list, dt):
// one copy per device, held for the life of the process.  Its device
// address on each device is taken once, at rhy_init, with that device
// current; host code uses that address, never the symbol.
constexpr size_t kNoAllocScratch = (size_t)RHY_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

// Round up to a 256-byte boundary (parts of one device buffer).
size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

constexpr int kThreads = 256;
static_assert(kThreads <= kWalkLanes, "the walks' LDS ring has one slot column per thread");
// workgroups of the fixup kernel (grid-stride over the deferred faces)
constexpr int kFixupBlocks = 128;
constexpr int kNumKernelIds = 6;  // flux/step, update, bc, cfl, dt, deferred-face fixup

struct TimedLaunch {
  int kidx;
  hipEvent_t a, b;
};

}  // namespace

struct rhy_box {
  int32_t ndim = 0;
  int32_t dims[3] = {1, 1, 1};
  int32_t ncmp = 0;
  int32_t precision = RHY_FP64;
  Geom geom{};
  hipStream_t stream = nullptr;
  // State: two buffers of the ndim+2 hydro planes (the step reads one and
  // writes the other, then they swap) and one set of the other ncmp-ndim-2
  // component planes, in one block [hydro 0 | hydro 1 | others]; `cells`
  // is the current hydro buffer.  Double or float.
  void* block = nullptr;
  void* state[2] = {nullptr, nullptr};
  void* extra = nullptr;
  int32_t cur = 0;
  void* cells = nullptr;
  void* flux = nullptr;           // ndim * (ndim+2) planes, same type
  size_t esz = sizeof(double);    // bytes per stored value
  double* partial = nullptr;      // per-block CFL maxima of the last reduction
  int32_t partial_n = 0;          // partials of the stand-alone CFL pass
  int32_t acc_blocks = 0;         // partials of the last accumulating walk
  int32_t nf_x = 8, acc_L = 16;   // walk lengths for this shape (choose_walks)
  int32_t partial_cap = 0;
  double* scalars = nullptr;      // [0] max_u, [16..16+6*RHY_MAX_CMP) inflow states
  double inflow_copied[6 * RHY_MAX_CMP] = {};  // what the device inflow slots hold
  bool inflow_valid = false;
  double* dt = nullptr;           // current dt
  double* dt_log = nullptr;       // per-step dt of the last run
  void* diag = nullptr;           // workgroup partials of rhy_box_diagnostics
  uint32_t* defer_count = nullptr;  // faces the flux kernels left to the fixup:
                                    // [0] or per axis [0..ndim)
  uint64_t* defer_items = nullptr;
  uint32_t defer_cap = 0;           // the whole list (one list for every axis)
  uint32_t defer_cap_axis[3] = {0, 0, 0};  // per-axis lists carved from it
  int64_t defer_off_axis[3] = {0, 0, 0};
  float* snap = nullptr;          // staging of the FP32 snapshot payload (lazy)
  int32_t dt_log_cap = 0;
  int32_t last_nsteps = 0;
  bool coarse_walked = false;  // last step: rhy_amr_coarse_step (pre-step state kept)
  Consts coarse_k{};           // its constants, dt and dx
  double coarse_dt = 0.0, coarse_dx[3] = {1.0, 1.0, 1.0};
  int32_t prof_every = 0;  // 0: no timing, 1: every launch, N: the steps s % N == 0
  bool profiling = false;   // launches are being timed now
  std::vector<TimedLaunch> pending;
  int64_t launches[kNumKernelIds] = {0, 0, 0, 0, 0, 0};
  double total_ms[kNumKernelIds] = {0, 0, 0, 0, 0, 0};
};

namespace rhy {
namespace {
std::atomic<int32_t> g_defer_every{0};
std::atomic<int64_t> g_defer_cap{-1};
}  // namespace
int32_t defer_every_setting() { return g_defer_every.load(); }
int64_t defer_cap_setting() { return g_defer_cap.load(); }
}  // namespace rhy

namespace {

Geom make_geom(int nd, const int32_t* dims) {
  Geom G{};
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

int check_params(const rhy_params_t* p, int nd) {
  if (!p) return fail(RHY_ERR_ARG, "params is NULL");
  if (p->ndim != nd) return fail(RHY_ERR_ARG, "params.ndim does not match the box");
  if (p->limiter < 1 || p->limiter > 4) return fail(RHY_ERR_ARG, "unknown slope limiter");
  return RHY_OK;
}

Consts make_consts(const rhy_params_t* p) {
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

BcTypes make_bc(const rhy_params_t* p) {
  BcTypes b{};
  for (int f = 0; f < 6; ++f) b.t[f] = p->bc_types[f];
  return b;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// A launch is timed only if both events could be created and recorded;
// otherwise it is left out of the statistics.
void timed_begin(rhy_box* b, int kidx, TimedLaunch& tl) {
  tl.kidx = kidx;
  tl.a = tl.b = nullptr;
  if (!b->profiling) return;
  if (hipEventCreate(&tl.a) != hipSuccess || hipEventCreate(&tl.b) != hipSuccess ||
      hipEventRecord(tl.a, b->stream) != hipSuccess) {
    if (tl.a) (void)hipEventDestroy(tl.a);
    if (tl.b) (void)hipEventDestroy(tl.b);
    tl.a = tl.b = nullptr;
  }
}

void timed_end(rhy_box* b, TimedLaunch& tl) {
  if (!b->profiling || !tl.a) return;
  if (hipEventRecord(tl.b, b->stream) != hipSuccess) {
    (void)hipEventDestroy(tl.a);
    (void)hipEventDestroy(tl.b);
    return;
  }
  b->pending.push_back(tl);
}

void collect_timings(rhy_box* b) {
  for (auto& tl : b->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, tl.a, tl.b) == hipSuccess) {
      b->launches[tl.kidx] += 1;
      b->total_ms[tl.kidx] += ms;
    }
    (void)hipEventDestroy(tl.a);
    (void)hipEventDestroy(tl.b);
  }
  b->pending.clear();
}

// Steps of 64 row cells per wave along x (the wave-cooperative walk) and
// faces walked per thread along y and z (lanes consecutive in x).
#ifndef RHY_NF_X
#define RHY_NF_X 8
#endif
#ifndef RHY_NF_YZ
#define RHY_NF_YZ 8
#endif
constexpr int kNfX = RHY_NF_X;
constexpr int kNfYZ = RHY_NF_YZ;

// Cells walked per thread by the accumulating y/z walks of rhy_run.
#ifndef RHY_ACC_L
#define RHY_ACC_L 16
#endif
constexpr int kAccL = RHY_ACC_L;

// The walk lengths of a box: kNfX steps per x wave and kAccL cells per y/z
// chunk, halved (down to 2 and 4) while a walk would have fewer waves than
// fill the chip a few times over (256 CUs x 4 SIMDs x 3 waves = 3072 at a
// time).  Longer walks repeat less work (a warm-up half step per x wave,
// one extra face per chunk); more waves hide more latency.  Measured on
// the 2D 1024^2 spheres: 8 / 16 -> 2 / 4 took 3.2e9 to 4.9e9 zone-updates/s
// (FP64); at 3D 256^3 the longest lengths already give 33k and 16k waves.
constexpr int64_t kMinWavesX = 8192, kMinWavesAcc = 4096;

// Faces of axis a a step solves (faces 0..n_a along a, interior across).
int64_t axis_faces(const Geom& G, int a) {
  int64_t n = (int64_t)G.n[a] + 1;
  for (int d = 0; d < 3; ++d)
    if (d != a) n *= G.n[d];
  return n;
}


DeferList defer_list(rhy_box* b) {
  DeferList dl;
  dl.count = b->defer_count;
  dl.items = b->defer_items;
  const int64_t cap = defer_cap_setting();
  dl.cap = cap >= 0 ? (uint32_t)std::min<int64_t>(cap, b->defer_cap) : b->defer_cap;
  return dl;
}

template <int ND, int A, int LIM, typename R, class PW>
void launch_flux_axis(rhy_box* b, const Consts& k, const double* dx, const DeferList& dl) {
  const Geom& G = b->geom;
  const int nf = A == 0 ? b->nf_x : kNfYZ;
  const int64_t nthreads = flux_threads(G, A, nf);
  R* F = static_cast<R*>(b->flux) + (int64_t)A * (ND + 2) * G.plane;
  TimedLaunch tl;
  timed_begin(b, 0, tl);
  hipLaunchKernelGGL((flux_kernel<ND, A, LIM, R, PW>), dim3(grid_for(nthreads)),
                     dim3(kThreads), 0, b->stream, k, G, static_cast<const R*>(b->cells), F, b->dt,
                     dx[A], nf, dl);
  timed_end(b, tl);
}

template <int ND, int LIM, typename R, class PW>
void launch_flux_lim(rhy_box* b, const Consts& k, const double* dx, const DeferList& dl) {
  launch_flux_axis<ND, 0, LIM, R, PW>(b, k, dx, dl);
  if constexpr (ND > 1) launch_flux_axis<ND, 1, LIM, R, PW>(b, k, dx, dl);
  if constexpr (ND > 2) launch_flux_axis<ND, 2, LIM, R, PW>(b, k, dx, dl);
}

// The flux kernels run with policy PW (Deferring<P> unless built with libm
// pow) and list the faces that need P's slow paths; the fixup kernel then
// recomputes those with P.  The list is emptied before the first axis.
template <int ND, typename R, class PW, class P>
void launch_flux_r(rhy_box* b, const Consts& k, const double* dx) {
  const DeferList dl = defer_list(b);
  TimedLaunch tr;
  timed_begin(b, 5, tr);
  hipLaunchKernelGGL(step_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->defer_count, 4,
                     (double*)nullptr);
  timed_end(b, tr);
  switch (k.limiter) {
    case 1: launch_flux_lim<ND, 1, R, PW>(b, k, dx, dl); break;
    case 2: launch_flux_lim<ND, 2, R, PW>(b, k, dx, dl); break;
    case 3: launch_flux_lim<ND, 3, R, PW>(b, k, dx, dl); break;
    default: launch_flux_lim<ND, 4, R, PW>(b, k, dx, dl); break;
  }
  if constexpr (PW::kDefer) {
    TimedLaunch tf;
    timed_begin(b, 5, tf);
    hipLaunchKernelGGL((flux_fixup_kernel<ND, R, P>), dim3(kFixupBlocks), dim3(kThreads), 0,
                       b->stream, k, b->geom, static_cast<const R*>(b->cells),
                       static_cast<R*>(b->flux), b->dt, dx[0], dx[1], dx[2], dl, -1);
    timed_end(b, tf);
  }
}

template <int ND>
int launch_flux(rhy_box* b, const Consts& k, const double* dx) {
  if (b->precision == RHY_MIXED)
    launch_flux_r<ND, float, MixedFlux, MixedMath>(b, k, dx);
  else
    launch_flux_r<ND, double, PowF64Flux, PowF64Fix>(b, k, dx);
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

template <int ND, int ORDER, bool FUSED, typename R>
int launch_update_t(rhy_box* b, const Consts& k, const double* dx, BcTypes bc) {
  const Geom& G = b->geom;
  const int64_t P = G.plane;
  const int nv = ND + 2;
  int64_t n = (int64_t)G.n[0] * G.n[1] * G.n[2];
  const R* F0 = static_cast<const R*>(b->flux);
  const R* F1 = ND > 1 ? F0 + (int64_t)nv * P : nullptr;
  const R* F2 = ND > 2 ? F0 + (int64_t)2 * nv * P : nullptr;
  TimedLaunch tl;
  timed_begin(b, 1, tl);
  hipLaunchKernelGGL((update_kernel<ND, ORDER, FUSED, R>), dim3(grid_for(n)), dim3(kThreads), 0,
                     b->stream, k, G, static_cast<R*>(b->cells), F0, F1, F2, b->dt, dx[0], dx[1],
                     dx[2], bc, b->partial);
  timed_end(b, tl);
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

template <int ND, typename R>
int launch_update_r(rhy_box* b, const Consts& k, const double* dx, int order, bool fused,
                    BcTypes bc) {
  if (order == RHY_MEMORY_INTENSIVE)
    return fused ? launch_update_t<ND, 2, true, R>(b, k, dx, bc)
                 : launch_update_t<ND, 2, false, R>(b, k, dx, bc);
  return fused ? launch_update_t<ND, 1, true, R>(b, k, dx, bc)
               : launch_update_t<ND, 1, false, R>(b, k, dx, bc);
}

template <int ND>
int launch_update(rhy_box* b, const Consts& k, const double* dx, int order, bool fused,
                  BcTypes bc) {
  if (b->precision == RHY_MIXED) return launch_update_r<ND, float>(b, k, dx, order, fused, bc);
  return launch_update_r<ND, double>(b, k, dx, order, fused, bc);
}

// Ghost fill of the current state (hydro: of another hydro buffer instead).
template <int ND>
int launch_bc(rhy_box* b, const rhy_params_t* p, int face_only, void* hydro = nullptr) {
  const Geom& G = b->geom;
  int64_t maxface = 1;
  for (int a = 0; a < ND; ++a) {
    int64_t m = 1;
    for (int d = 0; d < 3; ++d)
      if (d != a) m *= G.n[d];
    maxface = m > maxface ? m : maxface;
  }
  // inflow states travel in device memory; they are copied only when they
  // change, after the box's queued work (which may still read them) is done,
  // so nothing refers to the caller's parameters once this call returns
  double* dinf = b->scalars + 16;
  if (!b->inflow_valid ||
      std::memcmp(b->inflow_copied, p->cons_inflows, sizeof(b->inflow_copied)) != 0) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::memcpy(b->inflow_copied, p->cons_inflows, sizeof(b->inflow_copied));
    HIP_TRY(hipMemcpy(dinf, b->inflow_copied, sizeof(b->inflow_copied), hipMemcpyHostToDevice));
    b->inflow_valid = true;
  }
  TimedLaunch tl;
  timed_begin(b, 2, tl);
  dim3 grid(grid_for(maxface), face_only >= 0 ? 1 : 2 * ND);
  if (b->precision == RHY_MIXED)
    hipLaunchKernelGGL((bc_kernel<ND, float>), grid, dim3(kThreads), 0, b->stream, G, make_bc(p),
                       dinf, b->ncmp, static_cast<float*>(hydro ? hydro : b->cells),
                       static_cast<float*>(b->extra), face_only);
  else
    hipLaunchKernelGGL((bc_kernel<ND, double>), grid, dim3(kThreads), 0, b->stream, G, make_bc(p),
                       dinf, b->ncmp, static_cast<double*>(hydro ? hydro : b->cells),
                       static_cast<double*>(b->extra), face_only);
  timed_end(b, tl);
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

template <int ND>
int launch_cfl(rhy_box* b, const Consts& k) {
  const Geom& G = b->geom;
  int64_t n = (int64_t)G.n[0] * G.n[1] * G.n[2];
  TimedLaunch tl;
  timed_begin(b, 3, tl);
  if (b->precision == RHY_MIXED)
    hipLaunchKernelGGL((cfl_kernel<ND, float>), dim3(grid_for(n)), dim3(kThreads), 0, b->stream,
                       k, G, static_cast<const float*>(b->cells), b->partial);
  else
    hipLaunchKernelGGL((cfl_kernel<ND, double>), dim3(grid_for(n)), dim3(kThreads), 0, b->stream,
                       k, G, static_cast<const double*>(b->cells), b->partial);
  timed_end(b, tl);
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

double min_dx(int nd, const double* dx) {
  double m = dx[0];
  for (int d = 1; d < nd; ++d) m = dx[d] < m ? dx[d] : m;
  return m;
}

int ensure_dt_log(rhy_box* b, int32_t n) {
  if (n <= b->dt_log_cap) return RHY_OK;
  if (b->dt_log) (void)hipFree(b->dt_log);
  b->dt_log = nullptr;
  b->dt_log_cap = 0;
  if (dev_alloc(&b->dt_log, sizeof(double) * (size_t)n) != hipSuccess)
    return fail(RHY_ERR_OOM, "dt log allocation failed");
  b->dt_log_cap = n;
  return RHY_OK;
}

// dt and dt/dx, (.5 dt)/dx per axis of the current step (dt_kernel writes
// them with dt): scalars[8..14]
double* step_consts(rhy_box* b) { return b->scalars + 8; }

// The deferred-face list of axis a in the accumulating step.
DeferList defer_list_axis(rhy_box* b, int a) {
  DeferList dl;
  dl.count = b->defer_count + a;
  dl.items = b->defer_items + b->defer_off_axis[a];
  const int64_t cap = defer_cap_setting();
  dl.cap = cap >= 0 ? (uint32_t)std::min<int64_t>(cap, b->defer_cap_axis[a])
                    : b->defer_cap_axis[a];
  return dl;
}

// One accumulating walk (and its fixup) along A of an ND box: reads the
// current state, writes Q (kAccMid) or the other state buffer (kAccLast).
template <int ND, int A, int LIM, typename R, class PW, class P, int MODE>
void launch_acc_axis(rhy_box* b, const Consts& k, const double* dx,
                     int order) {
  const Geom& G = b->geom;
  const int nv = ND + 2;
  const R* cur = static_cast<const R*>(b->cells);
  const R* fx = static_cast<const R*>(b->flux);
  R* q = static_cast<R*>(b->flux) + (int64_t)nv * G.plane;
  R* other = static_cast<R*>(b->state[1 - b->cur]);
  const R* acc = MODE == kAccLast && ND == 3 ? q : nullptr;
  R* out = MODE == kAccMid ? q : other;
  double* slot = b->partial + b->acc_blocks;
  const DeferList dl = defer_list_axis(b, A);
  TimedLaunch tl;
  timed_begin(b, 0, tl);
  hipLaunchKernelGGL((acc_walk_kernel<ND, A, LIM, R, PW, MODE>),
                     dim3(grid_for(acc_threads(G, A, b->acc_L))), dim3(kThreads), 0, b->stream, k,
                     G, cur, fx, acc, out, step_consts(b), b->acc_L, order, dl, b->partial);
  timed_end(b, tl);
  if constexpr (PW::kDefer) {
    TimedLaunch tf;
    timed_begin(b, 5, tf);
    hipLaunchKernelGGL((acc_fixup_kernel<ND, A, R, P, MODE>), dim3(kFixupBlocks), dim3(kThreads),
                       0, b->stream, k, G, cur, fx, acc, out, b->dt, dx[0], dx[A], order, dl,
                       slot);
    timed_end(b, tf);
  }
}

// One step of a 2D/3D box without the flux-plane update pass: the x walk's
// fluxes into the first flux planes (and their fixup), then the
// accumulating walks; the new state lands in the other state buffer.
template <int ND, int LIM, typename R, class PW, class P>
void launch_acc_step_lim(rhy_box* b, const Consts& k, const double* dx,
                         int order) {
  static_assert(ND > 1, "1D boxes step through the flux planes");
  // the deferred-face counters and the fixup's partial slot were reset by
  // the step's dt_kernel (rhy_run) or set_step_kernel (rhy_amr_coarse_step)
  const DeferList dlx = defer_list_axis(b, 0);
  launch_flux_axis<ND, 0, LIM, R, PW>(b, k, dx, dlx);
  if constexpr (PW::kDefer) {
    TimedLaunch tf;
    timed_begin(b, 5, tf);
    hipLaunchKernelGGL((flux_fixup_kernel<ND, R, P>), dim3(kFixupBlocks), dim3(kThreads), 0,
                       b->stream, k, b->geom, static_cast<const R*>(b->cells),
                       static_cast<R*>(b->flux), b->dt, dx[0], dx[1], dx[2], dlx, 0);
    timed_end(b, tf);
  }
  if constexpr (ND == 3) {
    launch_acc_axis<ND, 1, LIM, R, PW, P, kAccMid>(b, k, dx, order);
    launch_acc_axis<ND, 2, LIM, R, PW, P, kAccLast>(b, k, dx, order);
  } else {
    launch_acc_axis<ND, 1, LIM, R, PW, P, kAccLast>(b, k, dx, order);
  }
  b->cur = 1 - b->cur;
  b->cells = b->state[b->cur];
}

template <int ND, typename R, class PW, class P>
void launch_acc_step_r(rhy_box* b, const Consts& k, const double* dx,
                       int order) {
  switch (k.limiter) {
    case 1: launch_acc_step_lim<ND, 1, R, PW, P>(b, k, dx, order); break;
    case 2: launch_acc_step_lim<ND, 2, R, PW, P>(b, k, dx, order); break;
    case 3: launch_acc_step_lim<ND, 3, R, PW, P>(b, k, dx, order); break;
    default: launch_acc_step_lim<ND, 4, R, PW, P>(b, k, dx, order); break;
  }
}

template <int ND>
int launch_acc_step(rhy_box* b, const Consts& k, const double* dx,
                    int order) {
  if (b->precision == RHY_MIXED)
    launch_acc_step_r<ND, float, MixedFlux, MixedMath>(b, k, dx, order);
  else
    launch_acc_step_r<ND, double, PowF64Flux, PowF64Fix>(b, k, dx, order);
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

// The main loop's steps (src/main.f90:109-136) on the device.  1D boxes:
// fluxes into planes, then the update with the fused epilogue (ghost
// images and CFL partials of the new state).  2D/3D boxes: the
// accumulating walks, whose last one writes the new state's interior and
// CFL partials into the other state buffer (the buffers swap every step),
// then the ghost fill of the new state.  Writing the ghost images from the
// last walk instead costs that walk more than the separate fill takes.
template <int ND>
int run_steps(rhy_box* b, const rhy_params_t* p, const double* dx, int32_t nsteps) {
  b->coarse_walked = false;
  const Consts k = make_consts(p);
  const BcTypes bc = make_bc(p);
  int rc;
  if ((rc = ensure_dt_log(b, nsteps > 0 ? nsteps : 1))) return rc;
  // prologue: CFL of the current state and the ghost fill it is stepped with
  if ((rc = launch_cfl<ND>(b, k))) return rc;
  if ((rc = launch_bc<ND>(b, p, -1))) return rc;
  const double cmdx = p->courant * min_dx(ND, dx);
  for (int32_t s = 0; s < nsteps; ++s) {
    if (b->prof_every > 1) b->profiling = s % b->prof_every == 0;
    // partials: the CFL pass's before the first step, the last walk's and
    // the fixup's slot after
    const int32_t np = ND > 1 && s > 0 ? b->acc_blocks + 1 : b->partial_n;
    TimedLaunch tl;
    timed_begin(b, 4, tl);
    hipLaunchKernelGGL(dt_kernel, dim3(1), dim3(1024), 0, b->stream, b->partial, np, cmdx, b->dt,
                       b->dt_log + s, (double*)nullptr, dx[0], dx[1], dx[2], step_consts(b),
                       ND > 1 ? b->defer_count : (uint32_t*)nullptr,
                       ND > 1 ? b->partial + b->acc_blocks : (double*)nullptr);
    timed_end(b, tl);
    if constexpr (ND > 1) {
      if ((rc = launch_acc_step<ND>(b, k, dx, p->solver_type))) return rc;
      if ((rc = launch_bc<ND>(b, p, -1))) return rc;  // the new state's ghosts
    } else {
      if ((rc = launch_flux<ND>(b, k, dx))) return rc;
      if ((rc = launch_update<ND>(b, k, dx, p->solver_type, true, bc))) return rc;
    }
  }
  b->profiling = b->prof_every == 1;
  b->last_nsteps = nsteps;
  HIP_TRY(hipGetLastError());
  return RHY_OK;
}

int check_box(rhy_box* b) {
  if (!b) return fail(RHY_ERR_ARG, "box is NULL");
  if (!b->cells) return fail(RHY_ERR_ARG, "box has no device storage");
  return RHY_OK;
}

void free_box_memory(rhy_box* b) {
  if (b->block) (void)hipFree(b->block);
  if (b->flux) (void)hipFree(b->flux);
  if (b->partial) (void)hipFree(b->partial);
  if (b->defer_count) (void)hipFree(b->defer_count);
  if (b->defer_items) (void)hipFree(b->defer_items);
  if (b->snap) (void)hipFree(b->snap);
  b->snap = nullptr;
  b->block = b->state[0] = b->state[1] = b->extra = b->cells = nullptr;
  b->cur = 0;
  b->flux = nullptr;
  b->partial = nullptr;
  b->defer_count = nullptr;
  b->defer_items = nullptr;
  b->defer_cap = 0;
}

int alloc_box_memory(rhy_box* b) {
  b->geom = make_geom(b->ndim, b->dims);
  const int64_t P = b->geom.plane;
  const int nv = b->ndim + 2;
  b->esz = b->precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  const size_t planes = (size_t)(2 * nv + (b->ncmp - nv));
  if (dev_alloc(&b->block, b->esz * (size_t)P * planes) != hipSuccess) {
    free_box_memory(b);
    return fail(RHY_ERR_OOM, "state allocation failed");
  }
  unsigned char* base = static_cast<unsigned char*>(b->block);
  b->state[0] = base;
  b->state[1] = base + b->esz * (size_t)P * nv;
  b->extra = base + b->esz * (size_t)P * 2 * nv;
  b->cur = 0;
  b->cells = b->state[0];
  if (dev_alloc(&b->flux, b->esz * (size_t)P * nv * b->ndim) != hipSuccess) {
    free_box_memory(b);
    return fail(RHY_ERR_OOM, "flux scratch allocation failed");
  }
  const int64_t ncell = (int64_t)b->geom.n[0] * b->geom.n[1] * b->geom.n[2];
  b->partial_n = (int32_t)grid_for(ncell);
  b->nf_x = kNfX;
  while (b->nf_x > 2 && flux_threads(b->geom, 0, b->nf_x) / 64 < kMinWavesX) b->nf_x /= 2;
  b->acc_L = kAccL;
  auto acc_waves = [&](int L) {
    int64_t w = (int64_t)1 << 62;
    for (int a = 1; a < b->ndim; ++a) w = std::min(w, acc_threads(b->geom, a, L) / 64);
    return w;
  };
  while (b->ndim > 1 && b->acc_L > 4 && acc_waves(b->acc_L) < kMinWavesAcc) b->acc_L /= 2;
  b->acc_blocks =
      b->ndim > 1 ? (int32_t)grid_for(acc_threads(b->geom, b->ndim - 1, b->acc_L)) : 0;
  // one more slot: the CFL speeds of cells the fixup recomputed
  b->partial_cap = std::max(b->partial_n, b->acc_blocks) + 1;
  if (dev_alloc(&b->partial, sizeof(double) * (size_t)b->partial_cap) != hipSuccess) {
    free_box_memory(b);
    return fail(RHY_ERR_OOM, "reduction scratch allocation failed");
  }
  // deferred-face lists: per axis an eighth of the axis' faces, within
  // [4096, 2^22] entries (the fixup recomputes the whole axis past that);
  // the flux-plane path uses them as one list for every axis
  int64_t total = 0;
  for (int a = 0; a < b->ndim; ++a) {
    const int64_t cap = std::min<int64_t>(std::max<int64_t>(axis_faces(b->geom, a) / 8, 4096),
                                          1 << 22);
    b->defer_cap_axis[a] = (uint32_t)cap;
    b->defer_off_axis[a] = total;
    total += cap;
  }
  b->defer_cap = (uint32_t)std::min<int64_t>(total, 0xffffffffll);
  if (dev_alloc(&b->defer_count, sizeof(uint32_t) * 4) != hipSuccess ||
      dev_alloc(&b->defer_items, sizeof(uint64_t) * (size_t)total) != hipSuccess) {
    free_box_memory(b);
    return fail(RHY_ERR_OOM, "deferred-face list allocation failed");
  }
  // ghosts start at 0 (rhyme_samr_bc_init.f90:61-87)
  if (hipMemsetAsync(b->defer_count, 0, sizeof(uint32_t) * 4, b->stream) != hipSuccess ||
      hipMemsetAsync(b->block, 0, b->esz * (size_t)P * planes, b->stream) != hipSuccess) {
    free_box_memory(b);
    return fail(RHY_ERR_HIP, "clearing the new box storage failed");
  }
  return RHY_OK;
}

int check_dims(int32_t ndim, const int32_t* dims) {
  if (ndim < 1 || ndim > 3) return fail(RHY_ERR_SHAPE, "ndim must be 1, 2 or 3");
  if (!dims) return fail(RHY_ERR_ARG, "dims is NULL");
  for (int d = 0; d < 3; ++d) {
    if (d < ndim && dims[d] < 2) return fail(RHY_ERR_SHAPE, "active extents must be >= 2");
    if (d >= ndim && dims[d] != 1) return fail(RHY_ERR_SHAPE, "inactive extents must be 1");
  }
  int64_t cells = 1;
  for (int d = 0; d < 3; ++d) cells *= (int64_t)dims[d] + (d < ndim ? 4 : 0);
  if (cells >= ((int64_t)1 << 31)) return fail(RHY_ERR_SHAPE, "box exceeds 2^31 cells");
  return RHY_OK;
}

}  // namespace

extern "C" {

int rhy_init(int32_t device) {
  std::call_once(g_init_once, [device]() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
      g_init_rc = RHY_ERR_NODEV;
      g_err = "no HIP device";
      return;
    }
    int d = device >= 0 && device < n ? device : 0;
    if (hipSetDevice(d) != hipSuccess) {
      g_init_rc = RHY_ERR_NODEV;
      g_err = "hipSetDevice failed";
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) != hipSuccess) {
      g_init_rc = RHY_ERR_NODEV;
      g_err = "hipGetDeviceProperties failed";
      return;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      g_init_rc = RHY_ERR_NODEV;
      g_err = std::string("device is ") + prop.gcnArchName + ", library is built for gfx950";
      return;
    }
    g_device = d;
    g_init_rc = RHY_OK;
    void* ptr = nullptr;
    if (d < kMaxDevices && hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess)
      g_scratch[d] = ptr;
  });
  if (g_init_rc != RHY_OK && g_err.empty()) g_err = "device initialisation failed";
  return g_init_rc;
}

const char* rhy_last_error(void) { return g_err.c_str(); }

int32_t rhy_version(void) { return 2; }

int32_t rhy_build_options(void) {
  return (RHY_FP64_POW_FAST ? RHY_OPT_POW_FAST : 0) |
         (RHY_CONTACT_EXACT ? RHY_OPT_CONTACT_EXACT : 0) |
         (RHY_LDS_WALK ? RHY_OPT_LDS_WALK : 0) |
         (RHY_FP64_FAST_RIEMANN ? RHY_OPT_FAST_RIEMANN : 0);
}

int rhy_box_create(int32_t ndim, const int32_t dims[3], int32_t ncmp, int32_t precision,
                   rhy_box_t** out) {
  int rc;
  if (!out) return fail(RHY_ERR_ARG, "out is NULL");
  *out = nullptr;
  if ((rc = rhy_init(0))) return rc;
  if ((rc = check_dims(ndim, dims))) return rc;
  if (ncmp < ndim + 2 || ncmp > RHY_MAX_CMP) return fail(RHY_ERR_ARG, "ncmp out of range");
  if (precision != RHY_FP64 && precision != RHY_MIXED)
    return fail(RHY_ERR_ARG, "unsupported precision");
  rhy_box* b = new rhy_box();
  b->ndim = ndim;
  for (int d = 0; d < 3; ++d) b->dims[d] = dims[d];
  b->ncmp = ncmp;
  b->precision = precision;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(RHY_ERR_HIP, "stream creation failed");
  }
  if (dev_alloc(&b->scalars, sizeof(double) * (16 + 6 * RHY_MAX_CMP)) != hipSuccess ||
      dev_alloc(&b->dt, sizeof(double)) != hipSuccess) {
    rhy_box_destroy(b);
    return fail(RHY_ERR_OOM, "scalar allocation failed");
  }
  if ((rc = alloc_box_memory(b))) {
    rhy_box_destroy(b);
    return rc;
  }
  *out = b;
  return RHY_OK;
}

int rhy_box_resize(rhy_box_t* b, const int32_t dims[3]) {
  int rc;
  if ((rc = check_box(b))) return rc;
  b->coarse_walked = false;
  if ((rc = check_dims(b->ndim, dims))) return rc;
  bool same = true;
  for (int d = 0; d < 3; ++d) same = same && dims[d] == b->dims[d];
  if (same) return RHY_OK;
  HIP_TRY(hipStreamSynchronize(b->stream));
  // storage of the new shape first: if it cannot be had, the box keeps its
  // old shape and storage and stays usable
  rhy_box t;
  t.ndim = b->ndim;
  for (int d = 0; d < 3; ++d) t.dims[d] = dims[d];
  t.ncmp = b->ncmp;
  t.precision = b->precision;
  t.stream = b->stream;
  if ((rc = alloc_box_memory(&t))) return rc;
  free_box_memory(b);
  for (int d = 0; d < 3; ++d) b->dims[d] = dims[d];
  b->geom = t.geom;
  b->esz = t.esz;
  b->block = t.block;
  b->state[0] = t.state[0];
  b->state[1] = t.state[1];
  b->extra = t.extra;
  b->cur = t.cur;
  b->cells = t.cells;
  b->flux = t.flux;
  b->partial = t.partial;
  b->partial_n = t.partial_n;
  b->acc_blocks = t.acc_blocks;
  b->nf_x = t.nf_x;
  b->acc_L = t.acc_L;
  b->partial_cap = t.partial_cap;
  b->defer_count = t.defer_count;
  b->defer_items = t.defer_items;
  b->defer_cap = t.defer_cap;
  for (int a = 0; a < 3; ++a) {
    b->defer_cap_axis[a] = t.defer_cap_axis[a];
    b->defer_off_axis[a] = t.defer_off_axis[a];
  }
  b->last_nsteps = 0;
  return RHY_OK;
}

int rhy_box_destroy(rhy_box_t* b) {
  if (!b) return RHY_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  collect_timings(b);
  free_box_memory(b);
  if (b->scalars) (void)hipFree(b->scalars);
  if (b->dt) (void)hipFree(b->dt);
  if (b->dt_log) (void)hipFree(b->dt_log);
  if (b->diag) (void)hipFree(b->diag);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RHY_OK;
}

}  // extern "C"

namespace {

// Host array (ncmp planes, hydro first) to or from the box's current hydro
// buffer and its other-component planes.
int box_copy(rhy_box* b, const double* in, double* out) {
  const size_t P = (size_t)b->geom.plane;
  const size_t nv = (size_t)(b->ndim + 2);
  const size_t n = P * b->ncmp;
  const size_t nh = P * nv;
  const hipMemcpyKind kind = in ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  auto copy = [&](void* host, size_t esz) -> int {
    unsigned char* h = static_cast<unsigned char*>(host);
    void* dh = b->cells;
    void* dx = b->extra;
    if (in) {
      HIP_TRY(hipMemcpyAsync(dh, h, esz * nh, kind, b->stream));
      if (n > nh) HIP_TRY(hipMemcpyAsync(dx, h + esz * nh, esz * (n - nh), kind, b->stream));
    } else {
      HIP_TRY(hipMemcpyAsync(h, dh, esz * nh, kind, b->stream));
      if (n > nh) HIP_TRY(hipMemcpyAsync(h + esz * nh, dx, esz * (n - nh), kind, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return RHY_OK;
  };
  if (b->precision == RHY_MIXED) {
    std::vector<float> tmp(n);
    if (in) {
      for (size_t i = 0; i < n; ++i) tmp[i] = (float)in[i];
      return copy(tmp.data(), sizeof(float));
    }
    int rc = copy(tmp.data(), sizeof(float));
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return RHY_OK;
  }
  return copy(in ? const_cast<double*>(in) : out, sizeof(double));
}

}  // namespace

extern "C" {

int rhy_box_upload(rhy_box_t* b, const double* cells) {
  int rc;
  if ((rc = check_box(b))) return rc;
  b->coarse_walked = false;
  if (!cells) return fail(RHY_ERR_ARG, "cells is NULL");
  return box_copy(b, cells, nullptr);
}

int rhy_box_download(rhy_box_t* b, double* cells) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (!cells) return fail(RHY_ERR_ARG, "cells is NULL");
  return box_copy(b, nullptr, cells);
}

int rhy_box_sync(rhy_box_t* b) {
  int rc;
  if ((rc = check_box(b))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  collect_timings(b);
  return RHY_OK;
}

int rhy_bc_set_boundaries(rhy_box_t* b, const rhy_params_t* p) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if ((rc = check_params(p, b->ndim))) return rc;
  if (b->ndim == 1) rc = launch_bc<1>(b, p, -1);
  else if (b->ndim == 2) rc = launch_bc<2>(b, p, -1);
  else rc = launch_bc<3>(b, p, -1);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_debug_defer(int32_t every, int64_t cap) {
  if (every < 0) return fail(RHY_ERR_ARG, "every must be >= 0");
  g_defer_every.store(every);
  g_defer_cap.store(cap < 0 ? -1 : cap);
  return RHY_OK;
}

int rhy_box_last_deferred(rhy_box_t* b, int64_t* n) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (!n) return fail(RHY_ERR_ARG, "n is NULL");
  uint32_t c[4] = {0, 0, 0, 0};
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(c, b->defer_count, sizeof(c), hipMemcpyDeviceToHost));
  *n = (int64_t)c[0] + c[1] + c[2] + c[3];
  return RHY_OK;
}

// FP32 snapshot payload of the interior (pack) or from it (unpack), through a
// device staging buffer that lives as long as the box's storage.
static int box_snapshot(rhy_box_t* b, float* out, const float* in) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (!out && !in) return fail(RHY_ERR_ARG, "data is NULL");
  const int64_t blen = (int64_t)b->geom.n[0] * b->geom.n[1] * b->geom.n[2];
  const size_t bytes = sizeof(float) * (size_t)blen * b->ncmp;
  if (!b->snap && dev_alloc(&b->snap, bytes) != hipSuccess) {
    b->snap = nullptr;
    return fail(RHY_ERR_OOM, "snapshot staging allocation failed");
  }
  const dim3 grid = snapshot_grid(b->geom);
  if (in) {
    HIP_TRY(hipMemcpyAsync(b->snap, in, bytes, hipMemcpyHostToDevice, b->stream));
    if (b->precision == RHY_MIXED)
      hipLaunchKernelGGL(snapshot_unpack_kernel<float>, grid, dim3(kThreads), 0, b->stream,
                         b->geom, b->ncmp, b->ndim + 2, static_cast<float*>(b->cells),
                         static_cast<float*>(b->extra), b->snap);
    else
      hipLaunchKernelGGL(snapshot_unpack_kernel<double>, grid, dim3(kThreads), 0, b->stream,
                         b->geom, b->ncmp, b->ndim + 2, static_cast<double*>(b->cells),
                         static_cast<double*>(b->extra), b->snap);
    HIP_TRY(hipGetLastError());
  } else {
    if (b->precision == RHY_MIXED)
      hipLaunchKernelGGL(snapshot_pack_kernel<float>, grid, dim3(kThreads), 0, b->stream,
                         b->geom, b->ncmp, b->ndim + 2, static_cast<const float*>(b->cells),
                         static_cast<const float*>(b->extra), b->snap);
    else
      hipLaunchKernelGGL(snapshot_pack_kernel<double>, grid, dim3(kThreads), 0, b->stream,
                         b->geom, b->ncmp, b->ndim + 2, static_cast<const double*>(b->cells),
                         static_cast<const double*>(b->extra), b->snap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, b->snap, bytes, hipMemcpyDeviceToHost, b->stream));
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_box_snapshot_pack(rhy_box_t* b, float* data) {
  if (!data) return fail(RHY_ERR_ARG, "data is NULL");
  return box_snapshot(b, data, nullptr);
}

int rhy_box_snapshot_unpack(rhy_box_t* b, const float* data) {
  if (!data) return fail(RHY_ERR_ARG, "data is NULL");
  const int rc = box_snapshot(b, nullptr, data);
  if (rc == RHY_OK) b->coarse_walked = false;
  return rc;
}

constexpr int kDiagBlocks = 256;

int rhy_box_diagnostics(rhy_box_t* b, rhy_diag_t* out) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (!out) return fail(RHY_ERR_ARG, "out is NULL");
  if (!b->diag && dev_alloc(&b->diag, sizeof(DiagPartial) * kDiagBlocks) != hipSuccess)
    return fail(RHY_ERR_OOM, "diagnostics buffer allocation failed");
  const Geom& G = b->geom;
  const unsigned nb = std::min<unsigned>(grid_for(G.plane), kDiagBlocks);
  DiagPartial* dp = static_cast<DiagPartial*>(b->diag);
  const bool mixed = b->precision == RHY_MIXED;
#define RHY_DIAG(ND)                                                                          \
  do {                                                                                        \
    if (mixed)                                                                                \
      hipLaunchKernelGGL((diag_kernel<ND, float>), dim3(nb), dim3(kThreads), 0, b->stream, G, \
                         b->ncmp, static_cast<const float*>(b->cells),                       \
                         static_cast<const float*>(b->extra), dp);                            \
    else                                                                                      \
      hipLaunchKernelGGL((diag_kernel<ND, double>), dim3(nb), dim3(kThreads), 0, b->stream,   \
                         G, b->ncmp, static_cast<const double*>(b->cells),                    \
                         static_cast<const double*>(b->extra), dp);                           \
  } while (0)
  if (b->ndim == 1) RHY_DIAG(1);
  else if (b->ndim == 2) RHY_DIAG(2);
  else RHY_DIAG(3);
#undef RHY_DIAG
  HIP_TRY(hipGetLastError());
  std::vector<DiagPartial> parts(nb);
  HIP_TRY(hipMemcpyAsync(parts.data(), dp, sizeof(DiagPartial) * nb, hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  diag_finish(parts, G, out);
  return RHY_OK;
}

int rhy_box_paint(rhy_box_t* b, const rhy_ic_t* ic) {
  int rc;
  if ((rc = check_box(b))) return rc;
  b->coarse_walked = false;
  if (!ic || (ic->kind != RHY_IC_TWO_STATE && ic->kind != RHY_IC_BLAST))
    return fail(RHY_ERR_ARG, "unknown initial condition kind");
  if (ic->kind == RHY_IC_TWO_STATE && (ic->axis < 0 || ic->axis >= b->ndim))
    return fail(RHY_ERR_ARG, "two-state axis out of range");
  if (b->ncmp < b->ndim + 2) return fail(RHY_ERR_ARG, "box has too few components");
  const PaintIC pc =
      make_paint(ic->kind, ic->axis, ic->gamma, ic->rho_l, ic->p_l, ic->rho_r, ic->p_r,
                 ic->radius, ic->jitter, ic->p0, ic->energy, ic->seed, b->ndim, b->dims);
  const Geom& G = b->geom;
  const dim3 grid(grid_for(G.plane));
  const bool mixed = b->precision == RHY_MIXED;
#define RHY_PAINT(ND)                                                                         \
  do {                                                                                        \
    if (mixed)                                                                                \
      hipLaunchKernelGGL((paint_kernel<ND, float>), grid, dim3(kThreads), 0, b->stream, G, pc, \
                         b->ncmp, static_cast<float*>(b->cells),                              \
                         static_cast<float*>(b->extra));                                      \
    else                                                                                      \
      hipLaunchKernelGGL((paint_kernel<ND, double>), grid, dim3(kThreads), 0, b->stream, G,   \
                         pc, b->ncmp, static_cast<double*>(b->cells),                         \
                         static_cast<double*>(b->extra));                                     \
  } while (0)
  if (b->ndim == 1) RHY_PAINT(1);
  else if (b->ndim == 2) RHY_PAINT(2);
  else RHY_PAINT(3);
#undef RHY_PAINT
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_box_draw(rhy_box_t* b, const rhy_params_t* p, const rhy_drawing_t* d) {
  int rc;
  if ((rc = check_box(b))) return rc;
  b->coarse_walked = false;
  if (!p || !d) return fail(RHY_ERR_ARG, "params or drawing is NULL");
  if (b->ncmp < b->ndim + 2) return fail(RHY_ERR_ARG, "box has too few components");
  if (d->nshapes < 0 || d->nshapes > RHY_DRAW_MAX_SHAPES || (d->nshapes && !d->shapes))
    return fail(RHY_ERR_ARG, "shape count out of range (0.." +
                                 std::to_string(RHY_DRAW_MAX_SHAPES) + ") or shapes NULL");
  if (d->nharmonics < 0 || d->nharmonics > RHY_DRAW_MAX_HARMONICS ||
      (d->nharmonics && !d->harmonics))
    return fail(RHY_ERR_ARG, "harmonic count out of range (0.." +
                                 std::to_string(RHY_DRAW_MAX_HARMONICS) +
                                 ") or harmonics NULL");
  if (!(p->gamma > 1.0)) return fail(RHY_ERR_ARG, "gamma must exceed 1");
  DrawCanvas cv{};
  const double cw[5] = {d->rho, d->v[0], d->v[1], d->v[2], d->p};
  // the colour of a box of fewer axes keeps its own velocity slots: (rho,
  // v[0..ndim), p) packed into w[0..ndim+1], p last at w[4]
  auto pack = [&](const double* src, double* w) {
    for (int c = 0; c < 5; ++c) w[c] = 0.0;
    w[0] = src[0];
    for (int a = 0; a < b->ndim; ++a) w[1 + a] = src[1 + a];
    w[4] = src[4];
  };
  pack(cw, cv.w);
  cv.gm1 = p->gamma - 1.0;
  cv.nshapes = d->nshapes;
  cv.nharmonics = d->nharmonics;
  std::vector<DrawShape> hs((size_t)d->nshapes);
  for (int s = 0; s < d->nshapes; ++s) {
    const rhy_shape_t& S = d->shapes[s];
    if (S.kind != RHY_DRAW_SPHERE && S.kind != RHY_DRAW_CUBOID)
      return fail(RHY_ERR_ARG, "unknown shape kind " + std::to_string(S.kind));
    DrawShape& T = hs[s];
    T.kind = S.kind;
    for (int a = 0; a < 3; ++a) {
      T.lo[a] = S.lo[a];
      T.hi[a] = S.hi[a];
      T.c[a] = S.centre[a];
    }
    T.r2 = S.radius * S.radius;
    const double sw[5] = {S.rho, S.v[0], S.v[1], S.v[2], S.p};
    pack(sw, T.w);
  }
  std::vector<DrawHarmonic> hh((size_t)d->nharmonics);
  int64_t nprof = 0;
  for (int m = 0; m < d->nharmonics; ++m) {
    const rhy_harmonic_t& H = d->harmonics[m];
    if (H.axis < 0 || H.axis >= b->ndim || !H.profile)
      return fail(RHY_ERR_ARG, "harmonic axis out of range or profile NULL");
    hh[m].axis = H.axis;
    hh[m].off = nprof;
    const double hw[5] = {H.rho, H.v[0], H.v[1], H.v[2], H.p};
    pack(hw, hh[m].w);
    nprof += b->dims[H.axis];
  }
  std::vector<double> prof((size_t)std::max<int64_t>(nprof, 1));
  for (int m = 0; m < d->nharmonics; ++m)
    std::memcpy(prof.data() + hh[m].off, d->harmonics[m].profile,
                sizeof(double) * (size_t)b->dims[d->harmonics[m].axis]);
  // the descriptors and profiles, in one short-lived device buffer
  const size_t o1 = align256(sizeof(DrawShape) * std::max<size_t>(hs.size(), 1));
  const size_t o2 = o1 + align256(sizeof(DrawHarmonic) * std::max<size_t>(hh.size(), 1));
  const size_t bytes = o2 + sizeof(double) * prof.size();
  unsigned char* dev = nullptr;
  HIP_TRY(dev_alloc(&dev, bytes));
  std::vector<unsigned char> host(bytes, 0);
  std::memcpy(host.data(), hs.data(), sizeof(DrawShape) * hs.size());
  std::memcpy(host.data() + o1, hh.data(), sizeof(DrawHarmonic) * hh.size());
  std::memcpy(host.data() + o2, prof.data(), sizeof(double) * prof.size());
  hipError_t e = hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) {
    const Geom& G = b->geom;
    const dim3 grid(grid_for(G.plane));
    const bool mixed = b->precision == RHY_MIXED;
    const DrawShape* ds = reinterpret_cast<const DrawShape*>(dev);
    const DrawHarmonic* dh = reinterpret_cast<const DrawHarmonic*>(dev + o1);
    const double* dp = reinterpret_cast<const double*>(dev + o2);
#define RHY_DRAW(ND)                                                                           \
  do {                                                                                         \
    if (mixed)                                                                                 \
      hipLaunchKernelGGL((draw_kernel<ND, float>), grid, dim3(kThreads), 0, b->stream, G, cv,  \
                         ds, dh, dp, b->ncmp, static_cast<float*>(b->cells),                   \
                         static_cast<float*>(b->extra));                                       \
    else                                                                                       \
      hipLaunchKernelGGL((draw_kernel<ND, double>), grid, dim3(kThreads), 0, b->stream, G, cv, \
                         ds, dh, dp, b->ncmp, static_cast<double*>(b->cells),                  \
                         static_cast<double*>(b->extra));                                      \
  } while (0)
    if (b->ndim == 1) RHY_DRAW(1);
    else if (b->ndim == 2) RHY_DRAW(2);
    else RHY_DRAW(3);
#undef RHY_DRAW
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  }
  (void)hipFree(dev);
  if (e != hipSuccess)
    return fail(RHY_ERR_HIP, std::string("rhy_box_draw: ") + hipGetErrorString(e));
  return RHY_OK;
}

int rhy_bc_set_face(rhy_box_t* b, const rhy_params_t* p, int32_t face) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if ((rc = check_params(p, b->ndim))) return rc;
  if (face < 0 || face >= 6) return fail(RHY_ERR_ARG, "face out of range");
  if (face >= 2 * b->ndim) return RHY_OK;  // no such face at this NDIM
  if (b->ndim == 1) rc = launch_bc<1>(b, p, face);
  else if (b->ndim == 2) rc = launch_bc<2>(b, p, face);
  else rc = launch_bc<3>(b, p, face);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_cfl_time_step(rhy_box_t* b, const rhy_params_t* p, const double dx[3], double* dt) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if ((rc = check_params(p, b->ndim))) return rc;
  if (!dx || !dt) return fail(RHY_ERR_ARG, "dx/dt is NULL");
  const Consts k = make_consts(p);
  if (b->ndim == 1) rc = launch_cfl<1>(b, k);
  else if (b->ndim == 2) rc = launch_cfl<2>(b, k);
  else rc = launch_cfl<3>(b, k);
  if (rc) return rc;
  hipLaunchKernelGGL(dt_kernel, dim3(1), dim3(1024), 0, b->stream, b->partial, b->partial_n, 0.0,
                     (double*)nullptr, (double*)nullptr, b->scalars);
  HIP_TRY(hipGetLastError());
  double maxu = 0.0;
  HIP_TRY(hipMemcpyAsync(&maxu, b->scalars, sizeof(maxu), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  *dt = p->courant * min_dx(b->ndim, dx) / maxu;
  return RHY_OK;
}

}  // extern "C"

namespace {

// The launches of one MUSCL-Hancock solve of a box, enqueued on its stream.
int mh_solve_enqueue(rhy_box* b, const rhy_params_t* p, const double* dx, double dt) {
  int rc;
  b->coarse_walked = false;
  if ((rc = check_params(p, b->ndim))) return rc;
  if (!dx) return fail(RHY_ERR_ARG, "dx is NULL");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return fail(RHY_ERR_SOLVER, "Unknown solver type");
  const Consts k = make_consts(p);
  hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(64), 0, b->stream, b->dt, dt);
  const BcTypes bc = make_bc(p);
  if (b->ndim == 1) {
    if ((rc = launch_flux<1>(b, k, dx))) return rc;
    rc = launch_update<1>(b, k, dx, p->solver_type, false, bc);
  } else if (b->ndim == 2) {
    if ((rc = launch_flux<2>(b, k, dx))) return rc;
    rc = launch_update<2>(b, k, dx, p->solver_type, false, bc);
  } else {
    if ((rc = launch_flux<3>(b, k, dx))) return rc;
    rc = launch_update<3>(b, k, dx, p->solver_type, false, bc);
  }
  return rc;
}

}  // namespace

extern "C" {

int rhy_mh_solve(rhy_box_t* b, const rhy_params_t* p, const double dx[3], double dt) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if ((rc = mh_solve_enqueue(b, p, dx, dt))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_amr_coarse_step(rhy_box_t* b, const rhy_params_t* p, const double dx[3], double dt) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (b->ndim == 1) {  // 1D boxes step through the flux planes anyway
    if ((rc = launch_bc<1>(b, p, -1))) return rc;
    if ((rc = mh_solve_enqueue(b, p, dx, dt))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return RHY_OK;
  }
  if ((rc = check_params(p, b->ndim))) return rc;
  if (!dx) return fail(RHY_ERR_ARG, "dx is NULL");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return fail(RHY_ERR_SOLVER, "Unknown solver type");
  const Consts k = make_consts(p);
  if ((rc = b->ndim == 2 ? launch_bc<2>(b, p, -1) : launch_bc<3>(b, p, -1))) return rc;
  hipLaunchKernelGGL(set_step_kernel, dim3(1), dim3(64), 0, b->stream, b->dt, step_consts(b),
                     dt, dx[0], b->ndim > 1 ? dx[1] : 1.0, b->ndim > 2 ? dx[2] : 1.0,
                     b->defer_count, b->partial + b->acc_blocks);
  rc = b->ndim == 2 ? launch_acc_step<2>(b, k, dx, p->solver_type)
                    : launch_acc_step<3>(b, k, dx, p->solver_type);
  if (rc) return rc;
  b->coarse_walked = true;  // after launch_acc_step: the state it stepped from is
  b->coarse_k = k;          // state[1 - cur], its x fluxes in the first flux planes
  b->coarse_dt = dt;
  for (int d = 0; d < 3; ++d) b->coarse_dx[d] = d < b->ndim ? dx[d] : 1.0;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

}  // extern "C"

namespace {

// Scratch bytes rhy_mh_solve_noalloc needs for a box: its flux planes, its
// deferred-face list, the list's counter and dt, each part 256-byte aligned.
int noalloc_scratch_bytes(int32_t ndim, const int32_t dims[3], int32_t precision,
                          int64_t* bytes) {
  int rc;
  if (!bytes) return fail(RHY_ERR_ARG, "bytes is NULL");
  if ((rc = check_dims(ndim, dims))) return rc;
  if (precision != RHY_FP64 && precision != RHY_MIXED)
    return fail(RHY_ERR_ARG, "unsupported precision");
  const Geom G = make_geom(ndim, dims);
  const size_t esz = precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
  int64_t faces = 0;
  for (int a = 0; a < ndim; ++a) faces += ncell / G.n[a] * (G.n[a] + 1);
  const int64_t cap = std::min<int64_t>(std::max<int64_t>(faces / 8, 4096), 1 << 22);
  *bytes = (int64_t)(align256(esz * (size_t)G.plane * (ndim + 2) * ndim) +
                     align256(sizeof(uint64_t) * (size_t)cap) + 2 * 256);
  return RHY_OK;
}

}  // namespace

extern "C" {

int rhy_mh_solve_noalloc(const rhy_params_t* p, int32_t ndim, const int32_t dims[3],
                         int32_t ncmp, int32_t precision, const double dx[3], double dt,
                         void* cells, void* stream) {
  int rc;
  if ((rc = rhy_init(0))) return rc;
  if (!cells) return fail(RHY_ERR_ARG, "cells is NULL");
  if (ncmp < ndim + 2 || ncmp > RHY_MAX_CMP) return fail(RHY_ERR_ARG, "ncmp out of range");
  int64_t need = 0;
  if ((rc = noalloc_scratch_bytes(ndim, dims, precision, &need))) return rc;
  if (need > (int64_t)kNoAllocScratch)
    return fail(RHY_ERR_SHAPE, "box needs " + std::to_string(need) +
                                   " bytes of scratch; the no-allocation entry holds " +
                                   std::to_string(kNoAllocScratch));
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev])
    return fail(RHY_ERR_HIP, "no scratch address for the current device (rhy_init selects it)");
  // a box view of the caller's state, its scratch carved from `scratch`:
  // flux planes, then the deferred-face list, its counter and dt
  unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
  rhy_box b;
  b.ndim = ndim;
  for (int d = 0; d < 3; ++d) b.dims[d] = dims[d];
  b.ncmp = ncmp;
  b.precision = precision;
  b.geom = make_geom(ndim, dims);
  b.stream = static_cast<hipStream_t>(stream);
  b.esz = precision == RHY_MIXED ? sizeof(float) : sizeof(double);
  b.cells = cells;
  b.extra = static_cast<unsigned char*>(cells) + b.esz * (size_t)b.geom.plane * (ndim + 2);
  const size_t flux_bytes = align256(b.esz * (size_t)b.geom.plane * (ndim + 2) * ndim);
  b.flux = base;
  const int64_t ncell = (int64_t)b.geom.n[0] * b.geom.n[1] * b.geom.n[2];
  int64_t faces = 0;
  for (int a = 0; a < ndim; ++a) faces += ncell / b.geom.n[a] * (b.geom.n[a] + 1);
  b.defer_cap = (uint32_t)std::min<int64_t>(std::max<int64_t>(faces / 8, 4096), 1 << 22);
  b.defer_items = reinterpret_cast<uint64_t*>(base + flux_bytes);
  const size_t list_bytes = align256(sizeof(uint64_t) * (size_t)b.defer_cap);
  b.defer_count = reinterpret_cast<uint32_t*>(base + flux_bytes + list_bytes);
  b.dt = reinterpret_cast<double*>(base + flux_bytes + list_bytes + 256);
  b.profiling = false;
  rc = mh_solve_enqueue(&b, p, dx, dt);
  b.cells = b.flux = nullptr;  // nothing here is the view's to free
  return rc;
}

int rhy_run_async(rhy_box_t* b, const rhy_params_t* p, const double dx[3], int32_t nsteps) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if ((rc = check_params(p, b->ndim))) return rc;
  if (!dx) return fail(RHY_ERR_ARG, "dx is NULL");
  if (nsteps < 0) return fail(RHY_ERR_ARG, "nsteps < 0");
  if (p->solver_type != RHY_CPU_INTENSIVE && p->solver_type != RHY_MEMORY_INTENSIVE)
    return fail(RHY_ERR_SOLVER, "Unknown solver type");
  if (b->ndim == 1) return run_steps<1>(b, p, dx, nsteps);
  if (b->ndim == 2) return run_steps<2>(b, p, dx, nsteps);
  return run_steps<3>(b, p, dx, nsteps);
}

int rhy_run(rhy_box_t* b, const rhy_params_t* p, const double dx[3], int32_t nsteps,
            double* dts) {
  int rc = rhy_run_async(b, p, dx, nsteps);
  if (rc) return rc;
  if ((rc = rhy_box_sync(b))) return rc;
  if (dts && nsteps > 0) return rhy_box_last_dts(b, nsteps, dts);
  return RHY_OK;
}

int rhy_box_last_dts(rhy_box_t* b, int32_t n, double* dts) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (!dts || n < 0 || n > b->last_nsteps) return fail(RHY_ERR_ARG, "bad dts request");
  if (n == 0) return RHY_OK;
  HIP_TRY(hipMemcpyAsync(dts, b->dt_log, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return RHY_OK;
}

int rhy_box_set_profiling(rhy_box_t* b, int32_t on) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (on < 0) return fail(RHY_ERR_ARG, "profiling period must be >= 0");
  HIP_TRY(hipStreamSynchronize(b->stream));
  collect_timings(b);
  b->prof_every = on;
  b->profiling = on == 1;
  for (int i = 0; i < kNumKernelIds; ++i) {
    b->launches[i] = 0;
    b->total_ms[i] = 0.0;
  }
  return RHY_OK;
}

int rhy_box_kernel_stats(rhy_box_t* b, int32_t kidx, int64_t* launches, double* total_ms) {
  int rc;
  if ((rc = check_box(b))) return rc;
  if (kidx < 0 || kidx >= kNumKernelIds || !launches || !total_ms)
    return fail(RHY_ERR_ARG, "bad kernel stats request");
  HIP_TRY(hipStreamSynchronize(b->stream));
  collect_timings(b);
  *launches = b->launches[kidx];
  *total_ms = b->total_ms[kidx];
  return RHY_OK;
}

// ------------------------------------------------------------ unit batches

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int to_dev(DevBuf& d, const void* h, size_t bytes) {
  if (dev_alloc(&d.p, bytes ? bytes : 8) != hipSuccess) return fail(RHY_ERR_OOM, "batch alloc");
  if (h && bytes) HIP_TRY(hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice));
  return RHY_OK;
}

// axis ids of a unit batch must name an axis of the problem (1..ndim)
int check_axes(const int32_t* axis, int64_t n, int nd) {
  if (!axis) return fail(RHY_ERR_ARG, "axis is NULL");
  for (int64_t q = 0; q < n; ++q)
    if (axis[q] < 1 || axis[q] > nd) return fail(RHY_ERR_ARG, "axis out of range 1..ndim");
  return RHY_OK;
}

int batch_prelude(const rhy_params_t* p, int64_t n) {
  int rc;
  if ((rc = rhy_init(0))) return rc;
  if (!p) return fail(RHY_ERR_ARG, "params is NULL");
  if (p->ndim < 1 || p->ndim > 3) return fail(RHY_ERR_ARG, "ndim");
  if (n < 0) return fail(RHY_ERR_ARG, "n < 0");
  return RHY_OK;
}

}  // namespace

// dx/dt arrays (or nullptr: x/t = 0) for riemann_batch_kernel
static int riemann_solve_common(const rhy_params_t* p, int64_t n, const double* l,
                                const double* r, const int32_t* axis, const double* dx,
                                const double* dt, double* u) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if (!l || !r || !u || (dx && !dt)) return fail(RHY_ERR_ARG, "array is NULL");
  if ((rc = check_axes(axis, n, p->ndim))) return rc;
  const int nv = p->ndim + 2;
  const size_t vb = sizeof(double) * (size_t)n * nv;
  DevBuf dl, dr, da, du, ddx, ddt;
  if ((rc = to_dev(dl, l, vb)) || (rc = to_dev(dr, r, vb)) ||
      (rc = to_dev(da, axis, sizeof(int32_t) * (size_t)n)) || (rc = to_dev(du, nullptr, vb)))
    return rc;
  if (dx && ((rc = to_dev(ddx, dx, sizeof(double) * (size_t)n)) ||
             (rc = to_dev(ddt, dt, sizeof(double) * (size_t)n))))
    return rc;
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  const double* L = (const double*)dl.p;
  const double* R = (const double*)dr.p;
  const int32_t* A = (const int32_t*)da.p;
  const double* DX = (const double*)ddx.p;
  const double* DT = (const double*)ddt.p;
  double* U = (double*)du.p;
  if (p->ndim == 1)
    hipLaunchKernelGGL(riemann_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n, L, R, A, DX,
                       DT, U);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(riemann_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n, L, R, A, DX,
                       DT, U);
  else
    hipLaunchKernelGGL(riemann_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n, L, R, A, DX,
                       DT, U);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(u, du.p, vb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_riemann_batch(const rhy_params_t* p, int64_t n, const double* l, const double* r,
                      const int32_t* axis, double* u) {
  return riemann_solve_common(p, n, l, r, axis, nullptr, nullptr, u);
}

int rhy_riemann_solve_batch(const rhy_params_t* p, int64_t n, const double* l, const double* r,
                            const int32_t* axis, const double* dx, const double* dt, double* u) {
  if (n > 0 && (!dx || !dt)) return fail(RHY_ERR_ARG, "dx/dt is NULL");
  return riemann_solve_common(p, n, l, r, axis, dx, dt, u);
}

int rhy_riemann_star_batch(const rhy_params_t* p, int64_t n, rhy_rp_solution_t* sol,
                           const int32_t* axis) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if (!sol) return fail(RHY_ERR_ARG, "array is NULL");
  if ((rc = check_axes(axis, n, p->ndim))) return rc;
  const size_t sb = sizeof(rhy_rp_solution_t) * (size_t)n;
  DevBuf ds, da;
  if ((rc = to_dev(ds, sol, sb)) || (rc = to_dev(da, axis, sizeof(int32_t) * (size_t)n)))
    return rc;
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  rhy_rp_solution_t* S = (rhy_rp_solution_t*)ds.p;
  const int32_t* A = (const int32_t*)da.p;
  if (p->ndim == 1)
    hipLaunchKernelGGL(rp_star_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(rp_star_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A);
  else
    hipLaunchKernelGGL(rp_star_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(sol, ds.p, sb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

// sampling at dx/dt (dx != nullptr) or one region per record at dxdt
static int rp_sample_common(const rhy_params_t* p, int64_t n, const rhy_rp_solution_t* sol,
                            const int32_t* axis, const double* dx, const double* dt,
                            const int32_t* region, const double* dxdt, double* u) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if (!sol || !u) return fail(RHY_ERR_ARG, "array is NULL");
  if ((rc = check_axes(axis, n, p->ndim))) return rc;
  const int nv = p->ndim + 2;
  const size_t vb = sizeof(double) * (size_t)n * nv;
  const size_t nb = sizeof(double) * (size_t)n;
  DevBuf ds, da, du, d1, d2, dreg;
  if ((rc = to_dev(ds, sol, sizeof(rhy_rp_solution_t) * (size_t)n)) ||
      (rc = to_dev(da, axis, sizeof(int32_t) * (size_t)n)) || (rc = to_dev(du, nullptr, vb)))
    return rc;
  if (region) {
    if ((rc = to_dev(dreg, region, sizeof(int32_t) * (size_t)n)) || (rc = to_dev(d1, dxdt, nb)))
      return rc;
  } else if ((rc = to_dev(d1, dx, nb)) || (rc = to_dev(d2, dt, nb))) {
    return rc;
  }
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  const rhy_rp_solution_t* S = (const rhy_rp_solution_t*)ds.p;
  const int32_t* A = (const int32_t*)da.p;
  const double* DX = region ? nullptr : (const double*)d1.p;
  const double* DT = region ? nullptr : (const double*)d2.p;
  const int32_t* REG = region ? (const int32_t*)dreg.p : nullptr;
  const double* DXDT = region ? (const double*)d1.p : nullptr;
  double* U = (double*)du.p;
  if (p->ndim == 1)
    hipLaunchKernelGGL(rp_sample_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A, DX,
                       DT, REG, DXDT, U);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(rp_sample_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A, DX,
                       DT, REG, DXDT, U);
  else
    hipLaunchKernelGGL(rp_sample_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n, S, A, DX,
                       DT, REG, DXDT, U);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(u, du.p, vb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_riemann_sampling_batch(const rhy_params_t* p, int64_t n, const rhy_rp_solution_t* sol,
                               const int32_t* axis, const double* dx, const double* dt,
                               double* u) {
  if (n > 0 && (!dx || !dt)) return fail(RHY_ERR_ARG, "dx/dt is NULL");
  return rp_sample_common(p, n, sol, axis, dx, dt, nullptr, nullptr, u);
}

int rhy_riemann_region_batch(const rhy_params_t* p, int64_t n, const rhy_rp_solution_t* sol,
                             const int32_t* region, const int32_t* axis, const double* dxdt,
                             double* u) {
  if (n > 0 && (!region || !dxdt)) return fail(RHY_ERR_ARG, "region/dxdt is NULL");
  for (int64_t q = 0; q < n; ++q)
    if (region[q] < RHY_W_L || region[q] > RHY_W_RFAN)
      return fail(RHY_ERR_ARG, "unknown region id");
  return rp_sample_common(p, n, sol, axis, nullptr, nullptr, region, dxdt, u);
}

static int irs_common(const rhy_params_t* p, int64_t n, const double* l, const double* r,
                      double* p_star, double* l_f, double* r_f, double* guesses) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if (!l || !r) return fail(RHY_ERR_ARG, "array is NULL");
  const size_t rb = sizeof(double) * 4 * (size_t)n;
  const size_t nb = sizeof(double) * (size_t)n;
  DevBuf dl, dr, dp, dlf, drf, dg;
  if ((rc = to_dev(dl, l, rb)) || (rc = to_dev(dr, r, rb))) return rc;
  if (guesses) {
    if ((rc = to_dev(dg, nullptr, 6 * nb))) return rc;
  } else if ((rc = to_dev(dp, nullptr, nb)) || (rc = to_dev(dlf, nullptr, nb)) ||
             (rc = to_dev(drf, nullptr, nb))) {
    return rc;
  }
  const Consts k = make_consts(p);
  hipLaunchKernelGGL(irs_batch_kernel, dim3(grid_for(n)), dim3(kThreads), 0, 0, k, n,
                     (const double*)dl.p, (const double*)dr.p, (double*)dp.p, (double*)dlf.p,
                     (double*)drf.p, (double*)dg.p);
  HIP_TRY(hipGetLastError());
  if (guesses) {
    HIP_TRY(hipMemcpy(guesses, dg.p, 6 * nb, hipMemcpyDeviceToHost));
    return RHY_OK;
  }
  HIP_TRY(hipMemcpy(p_star, dp.p, nb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(l_f, dlf.p, nb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(r_f, drf.p, nb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_irs_exec_batch(const rhy_params_t* p, int64_t n, const double* l, const double* r,
                       double* p_star, double* l_f, double* r_f) {
  if (n > 0 && (!p_star || !l_f || !r_f)) return fail(RHY_ERR_ARG, "output is NULL");
  return irs_common(p, n, l, r, p_star, l_f, r_f, nullptr);
}

int rhy_irs_guess_batch(const rhy_params_t* p, int64_t n, const double* l, const double* r,
                        double* guesses) {
  if (n > 0 && !guesses) return fail(RHY_ERR_ARG, "guesses is NULL");
  return irs_common(p, n, l, r, nullptr, nullptr, nullptr, guesses);
}

int rhy_pow_batch(int32_t policy, int64_t n, const double* x, const double* y, double* out) {
  int rc;
  if ((rc = rhy_init(0))) return rc;
  if (policy < RHY_POW_PATH || policy > RHY_POW_MIXED) return fail(RHY_ERR_ARG, "unknown policy");
  if (n < 0) return fail(RHY_ERR_ARG, "n < 0");
  if (n == 0) return RHY_OK;
  if (!x || !y || !out) return fail(RHY_ERR_ARG, "array is NULL");
  const size_t nb = sizeof(double) * (size_t)n;
  DevBuf dx, dy, dout;
  if ((rc = to_dev(dx, x, nb)) || (rc = to_dev(dy, y, nb)) || (rc = to_dev(dout, nullptr, nb)))
    return rc;
  hipLaunchKernelGGL(pow_batch_kernel, dim3(grid_for(n)), dim3(kThreads), 0, 0, policy, n,
                     (const double*)dx.p, (const double*)dy.p, (double*)dout.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dout.p, nb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_limiter_batch(const rhy_params_t* p, int64_t n, const double* ul, const double* u,
                      const double* ur, double* delta) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (p->limiter < 1 || p->limiter > 4) return fail(RHY_ERR_ARG, "unknown slope limiter");
  if (n == 0) return RHY_OK;
  const int nv = p->ndim + 2;
  const size_t vb = sizeof(double) * (size_t)n * nv;
  DevBuf a, b, c, d;
  if ((rc = to_dev(a, ul, vb)) || (rc = to_dev(b, u, vb)) || (rc = to_dev(c, ur, vb)) ||
      (rc = to_dev(d, nullptr, vb)))
    return rc;
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  if (p->ndim == 1)
    hipLaunchKernelGGL(limiter_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)a.p, (const double*)b.p, (const double*)c.p, (double*)d.p);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(limiter_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)a.p, (const double*)b.p, (const double*)c.p, (double*)d.p);
  else
    hipLaunchKernelGGL(limiter_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)a.p, (const double*)b.p, (const double*)c.p, (double*)d.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(delta, d.p, vb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_half_step_batch(const rhy_params_t* p, int64_t n, const double* u, const double* delta,
                        const int32_t* axis, double dx, double dt, double* l, double* r) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if ((rc = check_axes(axis, n, p->ndim))) return rc;
  const int nv = p->ndim + 2;
  const size_t vb = sizeof(double) * (size_t)n * nv;
  DevBuf du, dd, da, dl, dr;
  if ((rc = to_dev(du, u, vb)) || (rc = to_dev(dd, delta, vb)) ||
      (rc = to_dev(da, axis, sizeof(int32_t) * (size_t)n)) || (rc = to_dev(dl, nullptr, vb)) ||
      (rc = to_dev(dr, nullptr, vb)))
    return rc;
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  if (p->ndim == 1)
    hipLaunchKernelGGL(half_step_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const double*)dd.p, (const int32_t*)da.p, dx, dt,
                       (double*)dl.p, (double*)dr.p);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(half_step_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const double*)dd.p, (const int32_t*)da.p, dx, dt,
                       (double*)dl.p, (double*)dr.p);
  else
    hipLaunchKernelGGL(half_step_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const double*)dd.p, (const int32_t*)da.p, dx, dt,
                       (double*)dl.p, (double*)dr.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(l, dl.p, vb, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(r, dr.p, vb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

int rhy_flux_batch(const rhy_params_t* p, int64_t n, const double* u, const int32_t* axis,
                   double* f) {
  int rc;
  if ((rc = batch_prelude(p, n))) return rc;
  if (n == 0) return RHY_OK;
  if ((rc = check_axes(axis, n, p->ndim))) return rc;
  const int nv = p->ndim + 2;
  const size_t vb = sizeof(double) * (size_t)n * nv;
  DevBuf du, da, df;
  if ((rc = to_dev(du, u, vb)) || (rc = to_dev(da, axis, sizeof(int32_t) * (size_t)n)) ||
      (rc = to_dev(df, nullptr, vb)))
    return rc;
  const Consts k = make_consts(p);
  const unsigned g = grid_for(n);
  if (p->ndim == 1)
    hipLaunchKernelGGL(flux_batch_kernel<1>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const int32_t*)da.p, (double*)df.p);
  else if (p->ndim == 2)
    hipLaunchKernelGGL(flux_batch_kernel<2>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const int32_t*)da.p, (double*)df.p);
  else
    hipLaunchKernelGGL(flux_batch_kernel<3>, dim3(g), dim3(kThreads), 0, 0, k, n,
                       (const double*)du.p, (const int32_t*)da.p, (double*)df.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(f, df.p, vb, hipMemcpyDeviceToHost));
  return RHY_OK;
}

}  // extern "C"

namespace rhy {

// The device state of a box, for the two-level hierarchy's kernels (which
// live with the batches): waits for the box's stream first.
int box_view(rhy_box_t* b, BoxView* v) {
  int rc;
  if ((rc = check_box(b))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  v->cells = b->cells;
  v->flux = b->flux;
  v->prev = b->coarse_walked ? b->state[1 - b->cur] : nullptr;
  v->k = b->coarse_k;
  v->step_dt = b->coarse_dt;
  for (int d = 0; d < 3; ++d) v->step_dx[d] = b->coarse_dx[d];
  v->geom = b->geom;
  v->ndim = b->ndim;
  v->ncmp = b->ncmp;
  v->precision = b->precision;
  return RHY_OK;
}

}  // namespace rhy
