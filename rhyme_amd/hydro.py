"""Host-side mirror of the reference's hot-path module interfaces.

The reference is Fortran; its hot path is reached through these module
procedures (file:line relative to the reference's src/):

  rhyme_cfl_time_step(c, samr)                 rhyme_cfl/src/rhyme_cfl.f90:19-23
  rhyme_samr_bc_set_boundaries(bc, samr)       rhyme_samr_bc/src/rhyme_samr_bc.f90:38-41
  rhyme_samr_bc_set_<face>_boundary(bc, box)   rhyme_samr_bc/src/rhyme_samr_bc.f90:43-71
  rhyme_muscl_hancock_solve(mh, box, dx, dt, rp, sl, ws, logger)
                                               rhyme_muscl_hancock/src/rhyme_muscl_hancock.f90:47-56
  rhyme_riemann_problem_solve(rp, l, r, dx, dt, axis, u)
                                               rhyme_riemann_problem/src/rhyme_riemann_problem.f90:70-76
  rhyme_slope_limiter_run(sl, ul, u, ur, delta) rhyme_slope_limiter/src/rhyme_slope_limiter.f90:18-23
  rhyme_muscl_hancock_half_step_extrapolation(u, delta, axis, dx, dt, l, r)
                                               rhyme_muscl_hancock/src/rhyme_muscl_hancock.f90:58-63

This module keeps those names, argument meanings and defaults, and routes
every call to the HIP library through the C ABI (rhyme_amd._lib).  Box
state lives on the device between calls; `SamrBox.cells` is the host mirror
in the reference's layout, synchronised only by `upload()`/`download()`.
Errors are raised (the reference logs through logger%err and continues; a
raised RhyError is the Python form of that).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import RhyParams, check, dptr, iptr, dx3, dims3

# --- enumerations, same integers as the reference ------------------------
# slid (rhyme_slope_limiter/src/rhyme_slope_limiter.f90:6-8)
VAN_LEER, MINMOD, VAN_ALBADA, SUPERBEE = 1, 2, 3, 4
LIMITERS = {"van_leer": VAN_LEER, "minmod": MINMOD, "van_albada": VAN_ALBADA,
            "superbee": SUPERBEE}
# mhid (rhyme_muscl_hancock/src/rhyme_muscl_hancock.f90:10-15)
CPU_INTENSIVE, MEMORY_INTENSIVE = 1, 2
# bcid (rhyme_samr_bc/src/rhyme_samr_bc.f90:9-22)
REFLECTIVE, OUTFLOW, PERIODIC, INFLOW = 1, 2, 3, 4
# thid / ig_gamma (rhyme_thermo_base/src/rhyme_thermo_base.f90:8-22)
MONATOMIC, DIATOMIC, POLYATOMIC = 1, 2, 3
IG_GAMMA = {MONATOMIC: 5.0 / 3.0, DIATOMIC: 7.0 / 5.0, POLYATOMIC: 4.0 / 3.0}

# kB / amu in SI code units: the FP32 literals of rhyme_units.f90:25,29
# promoted to double, divided in double (rhyme_thermo_base.f90:130).
KB_AMU_SI = float(np.float32(1.38064852e-23)) / float(np.float32(1.6605e-27))

FP64, MIXED = 0, 1


@dataclass
class SlopeLimiter:
    """slope_limiter_t (rhyme_slope_limiter.f90:12-15)."""
    type: int = MINMOD
    w: float = 0.0


@dataclass
class Irs:
    """irs_t (rhyme_irs/src/rhyme_irs.f90:16-19)."""
    n_iteration: int = 10000
    tolerance: float = 1e-8


@dataclass
class RiemannProblem:
    """riemann_problem_t (rhyme_riemann_problem.f90:24-31); only irs%* and
    w_vacuum are read by the solver (…_star.f90:16-20)."""
    w_vacuum_rho: float = 0.0
    w_vacuum_p: float = 0.0
    tolerance: float = 1e-8
    n_iteration: int = 1000
    irs: Irs = field(default_factory=Irs)


@dataclass
class Cfl:
    """cfl_t (rhyme_cfl/src/rhyme_cfl.f90:7-9)."""
    courant_number: float = 0.8


@dataclass
class SamrBc:
    """samr_bc_t (rhyme_samr_bc.f90:24-28): per face left, right, bottom,
    top, back, front."""
    types: List[int] = field(default_factory=lambda: [REFLECTIVE] * 6)
    cons_inflows: Optional[np.ndarray] = None  # (6, ncmp)


@dataclass
class MusclHancock:
    """muscl_hancock_t (rhyme_muscl_hancock.f90:17-19)."""
    solver_type: int = MEMORY_INTENSIVE


@dataclass
class Thermo:
    """thermo_base_t + the module globals it sets (rhyme_thermo_base.f90:118-134)."""
    state_of_matter: int = MONATOMIC
    kb_amu: float = KB_AMU_SI

    @property
    def gamma(self):
        return IG_GAMMA[self.state_of_matter]


def ncmp_for(ndim, nspe=3):
    """NCMP = 1 + NDIM + 1 + 1 + NSPE (cmake/CalcNumberOfComponents.cmake:1-10)."""
    return 1 + ndim + 1 + 1 + nspe


def make_params(ndim, thermo=None, rp=None, sl=None, mh=None, cfl=None, bc=None,
                ncmp=None):
    """Pack the reference's configuration objects into rhy_params_t."""
    thermo = thermo or Thermo()
    rp = rp or RiemannProblem()
    sl = sl or SlopeLimiter()
    mh = mh or MusclHancock()
    cfl = cfl or Cfl()
    bc = bc or SamrBc()
    p = RhyParams()
    p.ndim = ndim
    p.limiter = sl.type
    p.solver_type = mh.solver_type
    p.n_iteration = rp.irs.n_iteration
    p.gamma = thermo.gamma
    p.kb_amu = thermo.kb_amu
    p.omega = sl.w
    p.tolerance = rp.irs.tolerance
    p.w_vacuum_rho = rp.w_vacuum_rho
    p.w_vacuum_p = rp.w_vacuum_p
    p.courant = cfl.courant_number
    for f in range(6):
        p.bc_types[f] = int(bc.types[f])
    p.ncmp = ncmp if ncmp is not None else ncmp_for(ndim)
    if bc.cons_inflows is not None:
        inf = np.asarray(bc.cons_inflows, dtype=np.float64)
        for f in range(6):
            for c in range(min(_lib.MAX_CMP, inf.shape[1])):
                p.cons_inflows[f][c] = inf[f, c]
    return p


def box_shape(ndim, dims, ncmp):
    """Host array shape of samr_box_t%cells in C order: (ncmp, nz+4, ny+4, nx+4)
    for 3D (Fortran cells(i,j,k,c) with i fastest)."""
    ext = [d + 4 for d in dims[:ndim]]
    return tuple([ncmp] + ext[::-1])


class SamrBox:
    """A box (samr_box_t, rhyme_samr.f90:34-42) with its device copy.

    `cells` is the host mirror, shape box_shape(ndim, dims, ncmp); the device
    copy is authoritative between upload() and download()."""

    def __init__(self, ndim, dims, ncmp=None, precision=FP64, cells=None):
        _lib.load()
        self.ndim = ndim
        self.dims = list(dims)[:ndim]
        self.ncmp = ncmp if ncmp is not None else ncmp_for(ndim)
        self.precision = precision
        shape = box_shape(ndim, self.dims, self.ncmp)
        if cells is None:
            cells = np.zeros(shape, dtype=np.float64)
        assert cells.shape == shape and cells.dtype == np.float64
        self.cells = np.ascontiguousarray(cells)
        h = C.c_void_p()
        check(_lib.load().rhy_box_create(ndim, iptr(dims3(self.dims)), self.ncmp,
                                         precision, C.byref(h)), "rhy_box_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            _lib.load().rhy_box_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self):
        check(_lib.load().rhy_box_upload(self._h, dptr(self.cells)), "rhy_box_upload")

    def download(self):
        check(_lib.load().rhy_box_download(self._h, dptr(self.cells)), "rhy_box_download")
        return self.cells

    def snapshot(self):
        """FP32 snapshot payload of the box (rhy_box_snapshot_pack): every
        component's interior cells, i fastest, one component after another."""
        n = self.ncmp * int(np.prod(self.dims))
        out = np.empty(n, dtype=np.float32)
        check(_lib.load().rhy_box_snapshot_pack(self._h, _lib.fptr(out)),
              "rhy_box_snapshot_pack")
        return out

    def load_snapshot(self, data):
        """The interior from an FP32 snapshot payload (rhy_box_snapshot_unpack);
        ghosts are left as they are."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.size != self.ncmp * int(np.prod(self.dims)):
            raise ValueError("payload holds %d values, the box %d" %
                             (data.size, self.ncmp * int(np.prod(self.dims))))
        check(_lib.load().rhy_box_snapshot_unpack(self._h, _lib.fptr(data)),
              "rhy_box_snapshot_unpack")

    def interior(self, var=None):
        sl = tuple(slice(2, 2 + n) for n in self.dims[::-1])
        if var is None:
            return self.cells[(slice(0, self.ndim + 2),) + sl]
        return self.cells[(var,) + sl]

    def sync(self):
        check(_lib.load().rhy_box_sync(self._h), "rhy_box_sync")

    def diagnostics(self):
        """rhy_box_diagnostics of the device state, as a dict: NaN cells per
        component, negative rho / e_tot counts, their minima and 1-based
        positions, interior sums of rho and e_tot (compensated), and the
        reference's masked totals (positive values over the ghosted array)."""
        d = _lib.RhyDiag()
        check(_lib.load().rhy_box_diagnostics(self._h, C.byref(d)), "rhy_box_diagnostics")
        return d.as_dict(self.ncmp, self.ndim)

    def last_deferred(self):
        """Faces the last flux phase left to the fixup kernel (the slow
        paths of the Riemann solve, recomputed in the same step)."""
        n = C.c_int64(0)
        check(_lib.load().rhy_box_last_deferred(self._h, C.byref(n)), "rhy_box_last_deferred")
        return n.value

    def paint(self, ic):
        """Paint the device state from `ic` (a dict of rhy_ic_t fields); the
        host mirror is not touched (download() to see it)."""
        c_ic = _lib.make_ic(ic)
        check(_lib.load().rhy_box_paint(self._h, C.byref(c_ic)), "rhy_box_paint")

    def draw(self, params, drawing):
        """Paint the device state from a drawing (rhyme_drawing_init: uniform
        canvas, sharp shapes in order, harmonic perturbations; the dict
        layout of ics.draw); gamma from params.  The host mirror is not
        touched (download() to see it)."""
        d, keep = _lib.make_drawing(drawing)
        check(_lib.load().rhy_box_draw(self._h, C.byref(params), C.byref(d)), "rhy_box_draw")
        del keep

    # -- timing helpers used by bench.py
    def set_profiling(self, on):
        """on: False/0 off, True/1 every launch, N > 1 every launch of the
        steps s % N == 0 of each run (rhy_box_set_profiling)."""
        check(_lib.load().rhy_box_set_profiling(self._h, int(on)), "rhy_box_set_profiling")

    def kernel_stats(self, kidx):
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        check(_lib.load().rhy_box_kernel_stats(self._h, kidx, C.byref(n), C.byref(ms)),
              "rhy_box_kernel_stats")
        return n.value, ms.value


# --- module procedures ------------------------------------------------------

def rhyme_samr_bc_set_boundaries(params, box):
    check(_lib.load().rhy_bc_set_boundaries(box.handle, C.byref(params)),
          "rhy_bc_set_boundaries")


def rhyme_samr_bc_set_face(params, box, face):
    check(_lib.load().rhy_bc_set_face(box.handle, C.byref(params), face), "rhy_bc_set_face")


def rhyme_cfl_time_step(params, box, dx):
    dt = C.c_double(0.0)
    d = dx3(dx)
    check(_lib.load().rhy_cfl_time_step(box.handle, C.byref(params), dptr(d), C.byref(dt)),
          "rhy_cfl_time_step")
    return dt.value


def rhyme_muscl_hancock_solve(params, box, dx, dt):
    d = dx3(dx)
    check(_lib.load().rhy_mh_solve(box.handle, C.byref(params), dptr(d), float(dt)),
          "rhy_mh_solve")


def run(params, box, dx, nsteps):
    """nsteps of the main loop (src/main.f90:109-136) on the device; returns dts."""
    d = dx3(dx)
    dts = np.zeros(max(nsteps, 1))
    check(_lib.load().rhy_run(box.handle, C.byref(params), dptr(d), nsteps, dptr(dts)),
          "rhy_run")
    return dts[:nsteps]


def run_async(params, box, dx, nsteps):
    d = dx3(dx)
    check(_lib.load().rhy_run_async(box.handle, C.byref(params), dptr(d), nsteps),
          "rhy_run_async")


def last_dts(box, n):
    dts = np.zeros(max(n, 1))
    check(_lib.load().rhy_box_last_dts(box.handle, n, dptr(dts)), "rhy_box_last_dts")
    return dts[:n]


def riemann_batch(params, l, r, axis):
    l = np.ascontiguousarray(l, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    u = np.zeros_like(l)
    check(_lib.load().rhy_riemann_batch(C.byref(params), len(axis), dptr(l), dptr(r),
                                        iptr(axis), dptr(u)), "rhy_riemann_batch")
    return u


def riemann_solve_batch(params, l, r, axis, dx, dt):
    """rhyme_riemann_problem_solve on records: record q sampled at
    x/t = dx[q] / dt[q] (scalars broadcast)."""
    l = np.ascontiguousarray(l, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    n = len(axis)
    dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, dtype=np.float64), (n,)))
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
    u = np.zeros_like(l)
    check(_lib.load().rhy_riemann_solve_batch(C.byref(params), n, dptr(l), dptr(r), iptr(axis),
                                              dptr(dx), dptr(dt), dptr(u)),
          "rhy_riemann_solve_batch")
    return u


def rhyme_riemann_problem_solve(params, l, r, dx, dt, axis):
    """rhyme_riemann_problem_solve(rp, l, r, dx, dt, axis, u): the conserved
    state at x/t = dx/dt (dx = 0: the Godunov state the MUSCL-Hancock loops
    use)."""
    return riemann_solve_batch(params, np.asarray(l)[None], np.asarray(r)[None],
                               np.array([axis], dtype=np.int32), dx, dt)[0]


def rp_solutions(sides):
    """An array of rhy_rp_solution_t from a list of dicts with the sides'
    primitive states: l_rho, l_v (sequence), l_p, l_cs and the same for r."""
    arr = (_lib.RhyRpSolution * len(sides))()
    for q, sd in enumerate(sides):
        s = arr[q]
        for side in ("l", "r"):
            setattr(s, side + "_rho", float(sd[side + "_rho"]))
            v = list(sd[side + "_v"])
            for d in range(len(v)):
                getattr(s, side + "_v")[d] = float(v[d])
            setattr(s, side + "_p", float(sd[side + "_p"]))
            setattr(s, side + "_cs", float(sd[side + "_cs"]))
    return arr


def riemann_star_batch(params, sols, axis):
    """rhyme_riemann_problem_star on every record of `sols` (an array of
    RhyRpSolution, updated in place)."""
    axis = np.ascontiguousarray(np.broadcast_to(np.asarray(axis, dtype=np.int32), (len(sols),)))
    check(_lib.load().rhy_riemann_star_batch(C.byref(params), len(sols), sols, iptr(axis)),
          "rhy_riemann_star_batch")
    return sols


def riemann_sampling_batch(params, sols, axis, dx, dt):
    """rhyme_riemann_problem_sampling of record q at x/t = dx[q] / dt[q]."""
    n = len(sols)
    nv = params.ndim + 2
    axis = np.ascontiguousarray(np.broadcast_to(np.asarray(axis, dtype=np.int32), (n,)))
    dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, dtype=np.float64), (n,)))
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
    u = np.zeros((n, nv))
    check(_lib.load().rhy_riemann_sampling_batch(C.byref(params), n, sols, iptr(axis), dptr(dx),
                                                 dptr(dt), dptr(u)),
          "rhy_riemann_sampling_batch")
    return u


def riemann_region_batch(params, sols, region, axis, dxdt):
    """The state of region id region[q] (_lib.W_*) of record q; the fans at
    x/t = dxdt[q]."""
    n = len(sols)
    nv = params.ndim + 2
    region = np.ascontiguousarray(np.broadcast_to(np.asarray(region, dtype=np.int32), (n,)))
    axis = np.ascontiguousarray(np.broadcast_to(np.asarray(axis, dtype=np.int32), (n,)))
    dxdt = np.ascontiguousarray(np.broadcast_to(np.asarray(dxdt, dtype=np.float64), (n,)))
    u = np.zeros((n, nv))
    check(_lib.load().rhy_riemann_region_batch(C.byref(params), n, sols, iptr(region),
                                               iptr(axis), dptr(dxdt), dptr(u)),
          "rhy_riemann_region_batch")
    return u


def irs_exec_batch(params, l, r):
    """rhyme_irs_exec on records (rho, v(axis), p, cs): returns p*, l_f, r_f."""
    l = np.ascontiguousarray(l, dtype=np.float64).reshape(-1, 4)
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 4)
    n = l.shape[0]
    ps, lf, rf = np.zeros(n), np.zeros(n), np.zeros(n)
    check(_lib.load().rhy_irs_exec_batch(C.byref(params), n, dptr(l), dptr(r), dptr(ps),
                                         dptr(lf), dptr(rf)), "rhy_irs_exec_batch")
    return ps, lf, rf


def irs_guess_batch(params, l, r):
    """rhyme_irs_guess_p_star on records (rho, v(axis), p, cs): [n, 6]."""
    l = np.ascontiguousarray(l, dtype=np.float64).reshape(-1, 4)
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 4)
    g = np.zeros((l.shape[0], 6))
    check(_lib.load().rhy_irs_guess_batch(C.byref(params), l.shape[0], dptr(l), dptr(r),
                                          dptr(g)), "rhy_irs_guess_batch")
    return g


def pow_batch(policy, x, y):
    """x**y on the device under a library policy (_lib.POW_PATH, POW_LIBM,
    POW_MIXED)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, dtype=np.float64), x.shape))
    out = np.zeros_like(x)
    check(_lib.load().rhy_pow_batch(policy, x.size, dptr(x), dptr(y), dptr(out)),
          "rhy_pow_batch")
    return out


def build_options():
    """The compile-time options of the loaded library (_lib.OPT_* bits)."""
    return int(_lib.load().rhy_build_options())


def limiter_batch(params, ul, u, ur):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ul, u, ur)]
    d = np.zeros_like(a[1])
    check(_lib.load().rhy_limiter_batch(C.byref(params), a[1].shape[0], dptr(a[0]),
                                        dptr(a[1]), dptr(a[2]), dptr(d)),
          "rhy_limiter_batch")
    return d


def rhyme_slope_limiter_run(params, ul, u, ur):
    return limiter_batch(params, np.asarray(ul)[None], np.asarray(u)[None],
                         np.asarray(ur)[None])[0]


def half_step_batch(params, u, delta, axis, dx, dt):
    u = np.ascontiguousarray(u, dtype=np.float64)
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    l = np.zeros_like(u)
    r = np.zeros_like(u)
    check(_lib.load().rhy_half_step_batch(C.byref(params), u.shape[0], dptr(u), dptr(delta),
                                          iptr(axis), float(dx), float(dt), dptr(l), dptr(r)),
          "rhy_half_step_batch")
    return l, r


def rhyme_muscl_hancock_half_step_extrapolation(params, u, delta, axis, dx, dt):
    l, r = half_step_batch(params, np.asarray(u)[None], np.asarray(delta)[None],
                           np.array([axis], dtype=np.int32), dx, dt)
    return l[0], r[0]


def flux_batch(params, u, axis):
    u = np.ascontiguousarray(u, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    f = np.zeros_like(u)
    check(_lib.load().rhy_flux_batch(C.byref(params), u.shape[0], dptr(u), iptr(axis),
                                     dptr(f)), "rhy_flux_batch")
    return f
