"""Two-level SAMR hierarchy on the device (refinement ratio 2, subcycled).

The reference has no coarse-fine coupling: its main loop advances every box
of every level alone with the level's dt (src/main.f90:118-136).  Here a
coarse box and a fixed batch of fine boxes are coupled as follows, per
coarse step:

1. dt0 = min(coarse CFL dt, 2 * min fine CFL dt), each with its level's dx.
2. The coarse box gets its ghost fill and one MUSCL-Hancock step with dt0
   (rhy_amr_coarse_step: the accumulating walks, bit for bit the flux-plane
   step rhy_mh_solve; the reflux forms the coarse y/z interface fluxes
   from the pre-step state the walks leave behind).
3. Two fine substeps of dt0 / 2, each after the fine boxes' face ghosts are
   set from their coarse parents, piecewise constant in space: U_old before
   the first, 0.5 * U_old + 0.5 * U_new before the second; ghosts inside
   another fine box's interior are copied from that box instead.
4. Refluxing (on by default): after each substep the fine fluxes through
   every fine box's boundary are summed per coarse face into a flux
   register (rhy_amr_flux_register); after the second, every coarse cell
   next to a fine box has its coarse face flux replaced by the fine fluxes'
   mean (rhy_amr_reflux), so mass and energy are conserved to rounding.
5. Every coarse cell under a fine box takes the mean of its 2^ndim children.

Every kernel is the library's (the box and batch steps, rhy_amr_prolong,
rhy_amr_flux_register, rhy_amr_reflux, rhy_amr_restrict); the orchestration
is on the host.  The
ghosts for the first substep are set before the coarse step, so no copy of
U_old is kept (the second fill interpolates in place), and the first fine
substep runs on the batch's stream while the coarse step runs on the box's.
oracle/amr.py states the same step on the CPU.
"""
import ctypes as C

import numpy as np

from . import hydro
from .batch import SamrBatch, _L, _check

EXTERNAL = 0  # face type of the fine boxes: ghosts set by the hierarchy


def fine_from_coarse(ndim, coarse_cells, lo, fine_dims):
    """Host array of a fine box (ghosts included) whose every cell holds its
    coarse parent's state: a fine level that starts as the coarse one."""
    idx = []
    for d in reversed(range(ndim)):
        f = np.arange(-1, fine_dims[d] + 3)
        idx.append(lo[d] + (f + 1) // 2 - 1 + 1)  # array index of the parent cell
    return np.ascontiguousarray(coarse_cells[(slice(None),) + np.ix_(*idx)])


def fine_params(params):
    """The fine level's parameters: params with every face type EXTERNAL."""
    q = type(params).from_buffer_copy(params)
    for f in range(6):
        q.bc_types[f] = EXTERNAL
    return q


class AmrHierarchy:
    """A coarse box (base_dims cells of base_dx) and fine boxes (fine_dims[i]
    cells of base_dx / 2) placed at fine_lo[i], the coarse cell (1-based) that
    holds fine cells 1 and 2 along each axis."""

    def __init__(self, ndim, base_dims, base_dx, fine_lo, fine_dims, precision=hydro.FP64,
                 ncmp=None, refluxing=True, coarse_walks=True, level2=()):
        self.ndim = ndim
        self.refluxing = refluxing
        # the coarse step through rhy_amr_coarse_step (accumulating walks), or
        # (False) through rhyme_samr_bc_set_boundaries + rhyme_muscl_hancock_solve
        self.coarse_walks = coarse_walks
        self.base_dims = list(base_dims)[:ndim]
        self.dx0 = [float(x) for x in list(base_dx)[:ndim]]
        self.dx1 = [x / 2.0 for x in self.dx0]
        self.fine_dims = [list(d)[:ndim] for d in fine_dims]
        self.fine_lo = [list(l)[:ndim] for l in fine_lo]
        self._check_placement()
        self.ncmp = ncmp if ncmp is not None else hydro.ncmp_for(ndim)
        self.precision = precision
        self.coarse = hydro.SamrBox(ndim, self.base_dims, self.ncmp, precision)
        self.fine = SamrBatch(ndim, self.fine_dims, [self.dx1] * len(self.fine_dims),
                              ncmp=self.ncmp, precision=precision)
        lo = np.ones((len(self.fine_dims), 3), dtype=np.int32)
        for i, l in enumerate(self.fine_lo):
            lo[i, :ndim] = l
        self._lo = np.ascontiguousarray(lo.ravel())
        # a third level: per level-1 parent box, a batch of its children
        # (dx0 / 4), lo relative to the parent; (parent, lo, dims, batch, lo3)
        self.kids = []
        self.dx2 = [x / 4.0 for x in self.dx0]
        for spec in level2:
            par = int(spec["parent"])
            kl = [list(l)[:ndim] for l in spec["lo"]]
            kd = [list(d)[:ndim] for d in spec["dims"]]
            if not 0 <= par < len(self.fine_dims):
                raise ValueError("no such level-1 parent box")
            self._check_placement(kd, kl, self.fine_dims[par])
            b = SamrBatch(ndim, kd, [self.dx2] * len(kd), ncmp=self.ncmp, precision=precision)
            l3 = np.ones((len(kd), 3), dtype=np.int32)
            for i, l in enumerate(kl):
                l3[i, :ndim] = l
            self.kids.append((par, kl, kd, b, np.ascontiguousarray(l3.ravel())))

    def _check_placement(self, fine_dims=None, fine_lo=None, parent_dims=None):
        nd = self.ndim
        fine_dims = self.fine_dims if fine_dims is None else fine_dims
        fine_lo = self.fine_lo if fine_lo is None else fine_lo
        parent_dims = self.base_dims if parent_dims is None else parent_dims
        for n_f, l in zip(fine_dims, fine_lo):
            for d in range(nd):
                if n_f[d] % 2:
                    raise ValueError("fine extents must be even (ratio 2)")
                if l[d] < 2 or l[d] + n_f[d] // 2 > parent_dims[d]:
                    raise ValueError("a fine box or its ghosts leave the coarse interior")
        cover = [[(l[d], l[d] + n_f[d] // 2 - 1) for d in range(nd)]
                 for n_f, l in zip(fine_dims, fine_lo)]
        for a in range(len(cover)):
            for b in range(a + 1, len(cover)):
                if all(cover[a][d][0] <= cover[b][d][1] and cover[b][d][0] <= cover[a][d][1]
                       for d in range(nd)):
                    raise ValueError("fine boxes %d and %d cover the same coarse cells" % (a, b))

    def close(self):
        for k in self.kids:
            k[3].close()
        self.fine.close()
        self.coarse.close()

    def upload_kids(self, cells):
        """Host arrays of the level-2 boxes, cells[j][i] for box i of the
        j-th level2 entry."""
        for (_, _, _, b, _), cs in zip(self.kids, cells):
            for i, c in enumerate(cs):
                b.upload(i, c)

    def download_kids(self):
        return [[b.download(i).copy() for i in range(b.n)] for (_, _, _, b, _) in self.kids]

    def regrid(self, fine_lo, fine_dims):
        """Move the fine level to new boxes (fine_dims[i] cells at
        fine_lo[i]): a new fine cell takes the old fine cell at its position
        when an old box held it, else its coarse parent's value
        (rhy_amr_regrid; oracle/amr.py regrid).  Call it between coarse
        steps, when the coarse cells under the old boxes hold their
        children's mean (a coarse step ends with that restriction)."""
        if self.kids:
            raise ValueError("regrid moves the finest level; this hierarchy has level-2 boxes")
        nd = self.ndim
        fine_dims = [list(d)[:nd] for d in fine_dims]
        fine_lo = [list(l)[:nd] for l in fine_lo]
        self._check_placement(fine_dims, fine_lo)
        new = SamrBatch(nd, fine_dims, [self.dx1] * len(fine_dims), ncmp=self.ncmp,
                        precision=self.precision)
        lo = np.ones((len(fine_dims), 3), dtype=np.int32)
        for i, l in enumerate(fine_lo):
            lo[i, :nd] = l
        lo = np.ascontiguousarray(lo.ravel())
        try:
            _check(_L().rhy_amr_regrid(self.coarse.handle, self.fine._h,
                                       self._lo.ctypes.data_as(C.POINTER(C.c_int32)), new._h,
                                       lo.ctypes.data_as(C.POINTER(C.c_int32))),
                   "rhy_amr_regrid")
        except Exception:
            new.close()
            raise
        self.fine.close()
        self.fine, self.fine_dims, self.fine_lo, self._lo = new, fine_dims, fine_lo, lo

    def upload(self, coarse_cells, fine_cells):
        self.coarse.cells[...] = coarse_cells
        self.coarse.upload()
        for i, c in enumerate(fine_cells):
            self.fine.upload(i, c)

    def download(self):
        return self.coarse.download().copy(), [self.fine.download(i).copy()
                                                for i in range(len(self.fine_dims))]

    def zones_per_step(self):
        """Zone-updates of one coarse step: the coarse cells once, the fine
        cells twice (two substeps), level-2 cells four times."""
        return (int(np.prod(self.base_dims)) + 2 * sum(int(np.prod(d)) for d in self.fine_dims)
                + 4 * sum(int(np.prod(d)) for k in self.kids for d in k[2]))

    def _prolong(self, half):
        _check(_L().rhy_amr_prolong(self.coarse.handle, self.fine._h,
                                    self._lo.ctypes.data_as(C.POINTER(C.c_int32)), half),
               "rhy_amr_prolong")

    def _restrict(self):
        _check(_L().rhy_amr_restrict(self.coarse.handle, self.fine._h,
                                     self._lo.ctypes.data_as(C.POINTER(C.c_int32))),
               "rhy_amr_restrict")

    def _register(self, reset):
        _check(_L().rhy_amr_flux_register(self.fine._h,
                                          self._lo.ctypes.data_as(C.POINTER(C.c_int32)),
                                          1 if reset else 0), "rhy_amr_flux_register")

    def _reflux(self, dt0):
        dx = (C.c_double * 3)(*(self.dx0 + [0.0] * (3 - self.ndim)))
        _check(_L().rhy_amr_reflux(self.coarse.handle, self.fine._h,
                                   self._lo.ctypes.data_as(C.POINTER(C.c_int32)), dt0, dx),
               "rhy_amr_reflux")

    def _kids_prolong0(self):
        """Level-2 face ghosts from their parents' U_old: before the level-1
        substep (and its sibling fill)."""
        p32 = C.POINTER(C.c_int32)
        for par, _, _, b, l3 in self.kids:
            _check(_L().rhy_amr_prolong_slot(self.fine._h, par, b._h, l3.ctypes.data_as(p32), 0),
                   "rhy_amr_prolong_slot")

    def _kids_cycle(self, fp, dt1):
        """Level 2 after a level-1 substep of dt1 (its ghosts set from U_old by
        _kids_prolong0 before that substep): for each parent, two substeps of
        its children, the second after the time-interpolated prolongation,
        then their reflux and restriction into the parent (oracle/amr.py
        step_levels)."""
        p32 = C.POINTER(C.c_int32)
        dx1 = (C.c_double * 3)(*(self.dx1 + [1.0] * (3 - self.ndim)))
        for par, _, _, b, l3 in self.kids:
            lo = l3.ctypes.data_as(p32)
            for half in (0, 1):
                if half:
                    _check(_L().rhy_amr_prolong_slot(self.fine._h, par, b._h, lo, 1),
                           "rhy_amr_prolong_slot")
                b.run_dt_async(fp, 0.5 * dt1, 1)
                if self.refluxing:
                    _check(_L().rhy_amr_flux_register(b._h, lo, 1 if half == 0 else 0),
                           "rhy_amr_flux_register")
            if self.refluxing:
                _check(_L().rhy_amr_reflux_slot(self.fine._h, par, b._h, lo, dt1, dx1),
                       "rhy_amr_reflux_slot")
            _check(_L().rhy_amr_restrict_slot(self.fine._h, par, b._h, lo),
                   "rhy_amr_restrict_slot")

    def step(self, params):
        """One coarse step; returns dt0."""
        fp = fine_params(params)
        dt0 = hydro.rhyme_cfl_time_step(params, self.coarse, self.dx0)
        dt0 = min(dt0, 2.0 * float(self.fine.cfl_dts(fp).min()))
        for k in self.kids:
            dt0 = min(dt0, 4.0 * float(k[3].cfl_dts(fp).min()))
        self._prolong(0)                              # U_old: before the coarse step
        if self.kids:
            self._kids_prolong0()                     # level 2 from level-1 U_old
        self.fine.run_dt_async(fp, 0.5 * dt0, 1)      # overlaps the coarse step
        if self.refluxing:
            self._register(True)                      # the batch's stream, after substep 1
        if self.coarse_walks:
            dx = (C.c_double * 3)(*(self.dx0 + [1.0] * (3 - self.ndim)))
            _check(_L().rhy_amr_coarse_step(self.coarse.handle, C.byref(params), dx, dt0),
                   "rhy_amr_coarse_step")
        else:
            hydro.rhyme_samr_bc_set_boundaries(params, self.coarse)
            hydro.rhyme_muscl_hancock_solve(params, self.coarse, self.dx0, dt0)
        if self.kids:
            self._kids_cycle(fp, 0.5 * dt0)           # level 2 under substep 1
        self._prolong(1)                              # after substep 1 (stream order)
        if self.kids:
            self._kids_prolong0()
        self.fine.run_dt_async(fp, 0.5 * dt0, 1)
        if self.refluxing:
            self._register(False)
        if self.kids:
            self._kids_cycle(fp, 0.5 * dt0)           # level 2 under substep 2
        if self.refluxing:
            self._reflux(dt0)                         # waits for the coarse box's stream
        self._restrict()                              # waits for the batch's stream
        return dt0

    def run(self, params, nsteps):
        return np.array([self.step(params) for _ in range(nsteps)])
