"""Many independent boxes of uneven shapes on one device (the batched form).

`SamrBatch` wraps rhy_batch_t (include/rhyme_hip.h): one device arena for
every box's state and flux scratch, one launch per phase over all boxes,
the step captured in a hipGraph.  Each box keeps its own ghost fill and its
own CFL dt, so a box evolves exactly as `hydro.run` would advance it alone.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr
from .hydro import FP64, box_shape, ncmp_for

_BATCH_SIGS = {
    "rhy_batch_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                   C.POINTER(C.c_void_p)]),
    "rhy_batch_destroy": (C.c_int, [C.c_void_p]),
    "rhy_batch_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "rhy_batch_upload": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "rhy_batch_download": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "rhy_batch_set_tiles": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32]),
    "rhy_batch_run_async": (C.c_int, [C.c_void_p, C.POINTER(_lib.RhyParams), C.c_int32,
                                      C.c_int32]),
    "rhy_batch_sync": (C.c_int, [C.c_void_p]),
    "rhy_batch_last_dts": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "rhy_batch_last_error": (C.c_char_p, []),
    "rhy_batch_diagnostics": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(_lib.RhyDiag)]),
    "rhy_batch_last_deferred": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "rhy_batch_cfl_dts": (C.c_int, [C.c_void_p, C.POINTER(_lib.RhyParams),
                                    C.POINTER(C.c_double)]),
    "rhy_batch_run_dt_async": (C.c_int, [C.c_void_p, C.POINTER(_lib.RhyParams), C.c_double,
                                         C.c_int32]),
    "rhy_amr_prolong": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_int32]),
    "rhy_amr_restrict": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]),
    "rhy_amr_flux_register": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]),
    "rhy_amr_reflux": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_double,
                                 C.POINTER(C.c_double)]),
    "rhy_amr_coarse_step": (C.c_int, [C.c_void_p, C.POINTER(_lib.RhyParams),
                                      C.POINTER(C.c_double), C.c_double]),
    "rhy_amr_regrid": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_void_p,
                                 C.POINTER(C.c_int32)]),
    "rhy_amr_prolong_slot": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p,
                                       C.POINTER(C.c_int32), C.c_int32]),
    "rhy_amr_restrict_slot": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p,
                                        C.POINTER(C.c_int32)]),
    "rhy_amr_reflux_slot": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32),
                                      C.c_double, C.POINTER(C.c_double)]),
    "rhy_amr_regrid_slot": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32),
                                      C.c_void_p, C.POINTER(C.c_int32)]),
    "rhy_batch_snapshot_length": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "rhy_batch_snapshot_pack": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "rhy_batch_snapshot_unpack": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "rhy_pool_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int32,
                                  C.POINTER(C.c_void_p)]),
    "rhy_pool_box_bytes": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                     C.POINTER(C.c_int64)]),
    "rhy_pool_add": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32)]),
    "rhy_pool_add_ic": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                  C.POINTER(_lib.RhyIC), C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
    "rhy_pool_remove": (C.c_int, [C.c_void_p, C.c_int32]),
    "rhy_pool_free_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]),
}


def _L():
    L = _lib.load()
    if not getattr(L, "_batch_sigs", False):
        for name, (res, args) in _BATCH_SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._batch_sigs = True
    return L


def _check(rc, what):
    if rc != _lib.RHY_OK:
        L = _L()
        msg = L.rhy_batch_last_error() or L.rhy_last_error()
        raise _lib.RhyError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


class SamrBatch:
    """Boxes (ndim, dims[i], dx[i]) advanced together; host arrays per box in
    the reference layout (hydro.box_shape)."""

    def __init__(self, ndim, dims, dx, ncmp=None, precision=FP64):
        self.ndim = ndim
        self.ncmp = ncmp if ncmp is not None else ncmp_for(ndim)
        self.precision = precision
        self.dims = [list(d)[:ndim] for d in dims]
        self.n = len(self.dims)
        d3 = np.ones((self.n, 3), dtype=np.int32)
        x3 = np.ones((self.n, 3), dtype=np.float64)
        for i, d in enumerate(self.dims):
            d3[i, :ndim] = d
            x3[i, :ndim] = list(dx[i])[:ndim]
        self.dx = x3
        h = C.c_void_p()
        _check(_L().rhy_batch_create(ndim, self.n, iptr(d3), dptr(x3), self.ncmp, precision,
                                     C.byref(h)), "rhy_batch_create")
        self._h = h
        self.tiles = (256, 8, 8, 0)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _L().rhy_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i):
        return box_shape(self.ndim, self.dims[i], self.ncmp)

    def upload(self, i, cells):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        assert cells.shape == self.shape(i)
        _check(_L().rhy_batch_upload(self._h, i, dptr(cells)), "rhy_batch_upload")

    def download(self, i):
        out = np.empty(self.shape(i), dtype=np.float64)
        _check(_L().rhy_batch_download(self._h, i, dptr(out)), "rhy_batch_download")
        return out

    def nbytes(self):
        b = C.c_int64(0)
        _check(_L().rhy_batch_bytes(self._h, C.byref(b)), "rhy_batch_bytes")
        return b.value

    def set_tiles(self, block=256, nf_x=8, nf_yz=8, grid_cap=0):
        _check(_L().rhy_batch_set_tiles(self._h, block, nf_x, nf_yz, grid_cap),
               "rhy_batch_set_tiles")
        self.tiles = (block, nf_x, nf_yz, grid_cap)

    def run_async(self, params, nsteps, graph=True):
        _check(_L().rhy_batch_run_async(self._h, C.byref(params), nsteps, 1 if graph else 0),
               "rhy_batch_run_async")

    def sync(self):
        _check(_L().rhy_batch_sync(self._h), "rhy_batch_sync")

    def run(self, params, nsteps, graph=True):
        self.run_async(params, nsteps, graph)
        self.sync()
        return self.last_dts(nsteps)

    def last_dts(self, n):
        out = np.zeros((max(n, 1), self.n))
        _check(_L().rhy_batch_last_dts(self._h, n, dptr(out)), "rhy_batch_last_dts")
        return out[:n]

    def zones(self):
        return int(sum(int(np.prod(d)) for d in self.dims))

    def diagnostics(self, i):
        d = _lib.RhyDiag()
        _check(_L().rhy_batch_diagnostics(self._h, i, C.byref(d)), "rhy_batch_diagnostics")
        return d.as_dict(self.ncmp, self.ndim)

    def last_deferred(self):
        """Faces the last flux phase left to the fixup kernel."""
        n = C.c_int64(0)
        _check(_L().rhy_batch_last_deferred(self._h, C.byref(n)), "rhy_batch_last_deferred")
        return n.value

    def snapshot(self):
        """FP32 snapshot payload of the level (rhy_batch_snapshot_pack): the
        boxes in order, each as SamrBox.snapshot lays it out."""
        n = C.c_int64(0)
        _check(_L().rhy_batch_snapshot_length(self._h, C.byref(n)), "rhy_batch_snapshot_length")
        out = np.empty(n.value, dtype=np.float32)
        _check(_L().rhy_batch_snapshot_pack(self._h, _lib.fptr(out)), "rhy_batch_snapshot_pack")
        return out

    def load_snapshot(self, data):
        """Every box's interior from a level payload (rhy_batch_snapshot_unpack)."""
        n = C.c_int64(0)
        _check(_L().rhy_batch_snapshot_length(self._h, C.byref(n)), "rhy_batch_snapshot_length")
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.size != n.value:
            raise ValueError("payload holds %d values, the level %d" % (data.size, n.value))
        _check(_L().rhy_batch_snapshot_unpack(self._h, _lib.fptr(data)),
               "rhy_batch_snapshot_unpack")

    def cfl_dts(self, params):
        """The CFL dt of every box for its current state (no step)."""
        out = np.zeros(self.n)
        _check(_L().rhy_batch_cfl_dts(self._h, C.byref(params), dptr(out)), "rhy_batch_cfl_dts")
        return out

    def run_dt_async(self, params, dt, nsteps):
        """nsteps steps with the given dt for every box (enqueue only)."""
        _check(_L().rhy_batch_run_dt_async(self._h, C.byref(params), float(dt), nsteps),
               "rhy_batch_run_dt_async")


def box_bytes(ndim, dims, ncmp=None, precision=FP64):
    """Arena bytes one box takes in a pool (state planes + flux scratch)."""
    d3 = np.ones(3, dtype=np.int32)
    d3[:ndim] = list(dims)[:ndim]
    b = C.c_int64(0)
    _check(_L().rhy_pool_box_bytes(ndim, iptr(d3), ncmp if ncmp is not None else ncmp_for(ndim),
                                   precision, C.byref(b)), "rhy_pool_box_bytes")
    return b.value


class PoolFull(Exception):
    """The box does not fit in the pool's arena (or every slot is taken)."""


class SamrPool:
    """A batch whose boxes come and go: a device arena of `budget_bytes`,
    boxes placed first-fit with their own limiter and face types, advanced
    together by `run_async`.  Boxes are addressed by slot."""

    def __init__(self, ndim, budget_bytes, max_boxes=4096, ncmp=None, precision=FP64):
        self.ndim = ndim
        self.ncmp = ncmp if ncmp is not None else ncmp_for(ndim)
        self.precision = precision
        self.max_boxes = max_boxes
        self.budget = int(budget_bytes)
        h = C.c_void_p()
        _check(_L().rhy_pool_create(ndim, self.ncmp, precision, self.budget, max_boxes,
                                    C.byref(h)), "rhy_pool_create")
        self._h = h
        self.boxes = {}  # slot -> dims
        self.tiles = (256, 8, 8, 0)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _L().rhy_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def box_bytes(self, dims):
        return box_bytes(self.ndim, dims, self.ncmp, self.precision)

    def add(self, dims, dx, cells, limiter, bc_types=None):
        """Place a box; returns its slot, raises PoolFull when it does not fit."""
        d3 = np.ones(3, dtype=np.int32)
        x3 = np.ones(3, dtype=np.float64)
        d3[:self.ndim] = list(dims)[:self.ndim]
        x3[:self.ndim] = list(dx)[:self.ndim]
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        assert cells.shape == box_shape(self.ndim, list(dims)[:self.ndim], self.ncmp)
        bc = None
        if bc_types is not None:
            bc = (C.c_int32 * 6)(*[int(t) for t in bc_types])
        slot = C.c_int32(-1)
        rc = _L().rhy_pool_add(self._h, iptr(d3), dptr(x3), dptr(cells), int(limiter),
                               C.cast(bc, C.POINTER(C.c_int32)) if bc is not None else None,
                               C.byref(slot))
        if rc == _lib.RHY_ERR_OOM:
            raise PoolFull(_L().rhy_batch_last_error().decode())
        _check(rc, "rhy_pool_add")
        self.boxes[slot.value] = list(dims)[:self.ndim]
        return slot.value

    def add_ic(self, dims, dx, ic, limiter, bc_types=None):
        """Place a box whose state is painted on the device from `ic` (a dict
        of rhy_ic_t fields, see ics.ic_two_state / ics.ic_blast); returns its
        slot, raises PoolFull when it does not fit."""
        d3 = np.ones(3, dtype=np.int32)
        x3 = np.ones(3, dtype=np.float64)
        d3[:self.ndim] = list(dims)[:self.ndim]
        x3[:self.ndim] = list(dx)[:self.ndim]
        bc = None
        if bc_types is not None:
            bc = (C.c_int32 * 6)(*[int(t) for t in bc_types])
        c_ic = _lib.make_ic(ic)
        slot = C.c_int32(-1)
        rc = _L().rhy_pool_add_ic(self._h, iptr(d3), dptr(x3), C.byref(c_ic), int(limiter),
                                  C.cast(bc, C.POINTER(C.c_int32)) if bc is not None else None,
                                  C.byref(slot))
        if rc == _lib.RHY_ERR_OOM:
            raise PoolFull(_L().rhy_batch_last_error().decode())
        _check(rc, "rhy_pool_add_ic")
        self.boxes[slot.value] = list(dims)[:self.ndim]
        return slot.value

    def remove(self, slot):
        _check(_L().rhy_pool_remove(self._h, slot), "rhy_pool_remove")
        del self.boxes[slot]

    def free_bytes(self):
        tot, big = C.c_int64(0), C.c_int64(0)
        _check(_L().rhy_pool_free_bytes(self._h, C.byref(tot), C.byref(big)),
               "rhy_pool_free_bytes")
        return tot.value, big.value

    def download(self, slot):
        out = np.empty(box_shape(self.ndim, self.boxes[slot], self.ncmp), dtype=np.float64)
        _check(_L().rhy_batch_download(self._h, slot, dptr(out)), "rhy_batch_download")
        return out

    def upload(self, slot, cells):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        assert cells.shape == box_shape(self.ndim, self.boxes[slot], self.ncmp)
        _check(_L().rhy_batch_upload(self._h, slot, dptr(cells)), "rhy_batch_upload")

    snapshot = SamrBatch.snapshot            # the resident boxes in slot order
    load_snapshot = SamrBatch.load_snapshot

    def set_tiles(self, block=256, nf_x=8, nf_yz=8, grid_cap=0):
        _check(_L().rhy_batch_set_tiles(self._h, block, nf_x, nf_yz, grid_cap),
               "rhy_batch_set_tiles")
        self.tiles = (block, nf_x, nf_yz, grid_cap)

    def run_async(self, params, nsteps, graph=True):
        _check(_L().rhy_batch_run_async(self._h, C.byref(params), nsteps, 1 if graph else 0),
               "rhy_batch_run_async")

    def sync(self):
        _check(_L().rhy_batch_sync(self._h), "rhy_batch_sync")

    def last_dts(self, n):
        """[n][max_boxes] dts of the last run, by slot."""
        out = np.zeros((max(n, 1), self.max_boxes))
        _check(_L().rhy_batch_last_dts(self._h, n, dptr(out)), "rhy_batch_last_dts")
        return out[:n]

    def zones(self):
        return int(sum(int(np.prod(d)) for d in self.boxes.values()))

    def diagnostics(self, slot):
        """rhy_diag_t of the box in `slot`, as a dict."""
        d = _lib.RhyDiag()
        _check(_L().rhy_batch_diagnostics(self._h, slot, C.byref(d)), "rhy_batch_diagnostics")
        return d.as_dict(self.ncmp, self.ndim)

    def last_deferred(self):
        """Faces the last flux phase left to the fixup kernel."""
        n = C.c_int64(0)
        _check(_L().rhy_batch_last_deferred(self._h, C.byref(n)), "rhy_batch_last_deferred")
        return n.value
