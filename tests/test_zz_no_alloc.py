"""rhy_mh_solve_noalloc: the MUSCL-Hancock solve on a caller-owned device
state, its scratch in the library's static device array, no allocation.
Checked through the C ABI against rhy_mh_solve (the entry point it is a
variant of) bit for bit, on two calls in a row with the second box smaller,
FP64 and mixed, both solver types; and the size limit is refused cleanly.

The caller's device memory comes from the HIP runtime the library itself
links (hipMalloc / hipMemcpy through ctypes): torch ships a HIP runtime of
its own, which cannot initialise the device in a process where the library's
runtime already has."""
import ctypes as C

import numpy as np
import pytest

from rhyme_amd import _lib, hydro, ics

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    _lib.load()  # the library has loaded its runtime; this is the same one
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArray:
    """Device memory of the same size as a host array."""

    def __init__(self, host):
        self.h = hip()
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.h.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        self.put(host)

    def put(self, host):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr.value:
            self.h.hipFree(self.ptr)
        self.ptr = C.c_void_p()


def say(*a):
    print(*a, flush=True)


def params(ndim, solver, limiter):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), mh=hydro.MusclHancock(solver_type=solver),
        cfl=hydro.Cfl(0.4), bc=hydro.SamrBc(types=[1, 2, 3, 3, 2, 1]))


def with_ghosts(ndim, dims, p, seed):
    """A state whose ghost cells hold a fill of its faces (the solve reads
    the ghosts as they are)."""
    import oracle
    cells = ics.random_smooth(ndim, dims, seed=seed, gamma=p.gamma)
    oracle.bc_set_boundaries(oracle.params_from(p), dims, cells)
    return cells


def reference_solve(ndim, dims, cells, p, dx, dt, precision):
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    hydro.rhyme_muscl_hancock_solve(p, box, dx, dt)
    out = box.download().copy()
    box.close()
    return out


def noalloc_solve(ndim, dims, cells, p, dx, dt, precision):
    dtype = np.float64 if precision == hydro.FP64 else np.float32
    dev = DevArray(cells.astype(dtype))
    say("  noalloc", ndim, dims, "precision", precision, "solver", p.solver_type)
    try:
        rc = _lib.load().rhy_mh_solve_noalloc(
            C.byref(p), ndim, _lib.iptr(_lib.dims3(dims)), cells.shape[0], precision,
            _lib.dptr(_lib.dx3(dx)), float(dt), dev.ptr, None)
        _lib.check(rc, "rhy_mh_solve_noalloc")
        out = dev.get(dtype, cells.shape).astype(np.float64)
    finally:
        dev.free()
    return out


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("solver", [hydro.MEMORY_INTENSIVE, hydro.CPU_INTENSIVE])
def test_noalloc_matches_mh_solve_twice(precision, solver):
    cases = [(3, [40, 36, 44], hydro.SUPERBEE, 11), (3, [12, 10, 14], hydro.VAN_LEER, 12),
             (2, [64, 48], hydro.MINMOD, 13), (1, [96], hydro.VAN_ALBADA, 14)]
    for ndim, dims, lim, seed in cases:
        p = params(ndim, solver, lim)
        cells = with_ghosts(ndim, dims, p, seed)
        dx = [1.0 / 48, 1.0 / 40, 1.0 / 36][:ndim]
        dt = 2e-3
        want = reference_solve(ndim, dims, cells, p, dx, dt, precision)
        got = noalloc_solve(ndim, dims, cells, p, dx, dt, precision)
        assert np.array_equal(got, want), (ndim, dims)


def test_noalloc_deferred_faces_match():
    """The deferred-face list lives in the static array too: force every
    7th face through the fixup and compare with rhy_mh_solve under the same
    hook, then a smaller box with the normal path."""
    p = params(3, hydro.MEMORY_INTENSIVE, hydro.SUPERBEE)
    for dims, every in (([24, 20, 28], 7), ([10, 12, 9], 0)):
        cells = with_ghosts(3, dims, p, 21)
        dx = [1.0 / 32] * 3
        with _lib.debug_defer(every):
            want = reference_solve(3, dims, cells, p, dx, 1e-3, hydro.FP64)
            got = noalloc_solve(3, dims, cells, p, dx, 1e-3, hydro.FP64)
        assert np.array_equal(got, want), dims


def test_noalloc_refuses_boxes_past_its_scratch():
    """A box whose scratch would exceed the array is refused before any work
    (RHY_ERR_SHAPE), and the caller's state is untouched; the largest FP64 3D
    box that fits (77^3) runs."""
    p = params(3, hydro.MEMORY_INTENSIVE, hydro.MINMOD)
    zeros = np.zeros(9 * 82 ** 3)
    dev = DevArray(zeros)
    try:
        rc = _lib.load().rhy_mh_solve_noalloc(
            C.byref(p), 3, _lib.iptr(_lib.dims3([78, 78, 78])), 9, hydro.FP64,
            _lib.dptr(_lib.dx3([1.0 / 78] * 3)), 1e-3, dev.ptr, None)
        assert rc == _lib.RHY_ERR_SHAPE
        assert "scratch" in _lib.load().rhy_last_error().decode()
        assert np.all(dev.get(np.float64, zeros.shape) == 0.0)
    finally:
        dev.free()
    cells = with_ghosts(3, [77, 77, 77], p, 5)
    want = reference_solve(3, [77] * 3, cells, p, [1.0 / 77] * 3, 1e-3, hydro.FP64)
    got = noalloc_solve(3, [77] * 3, cells, p, [1.0 / 77] * 3, 1e-3, hydro.FP64)
    assert np.array_equal(got, want)
