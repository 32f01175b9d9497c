"""The main-loop step of 2D/3D boxes (rhy_run) walks y and z accumulating the
update as it goes and writes the new state into a second buffer, with no
flux planes for y/z and no update pass; rhy_mh_solve keeps the flux planes
and the update kernel.  Both compute the reference's update expressions
operation for operation, so:

* N steps of rhy_run equal N steps of rhyme_cfl_time_step +
  rhyme_samr_bc_set_boundaries + rhyme_muscl_hancock_solve bit for bit
  (interior and dts), for every face type, both solver types, FP64 and
  mixed, and box extents that are and are not multiples of the walk's chunk;
* the ghost faces rhy_run leaves are the fill of its final state;
* a resized box (rhy_box_resize) steps like a new box of that shape, and its
  device storage survives a resize the device cannot hold.
"""
import numpy as np
import pytest

import oracle
from rhyme_amd import _lib, hydro, ics

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def params(ndim, limiter, solver, bc, courant=0.3, inflows=None):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), mh=hydro.MusclHancock(solver_type=solver),
        cfl=hydro.Cfl(courant), bc=hydro.SamrBc(types=bc, cons_inflows=inflows))


def interior(ndim, a):
    sl = tuple(slice(2, -2) for _ in range(ndim))
    return a[(slice(0, ndim + 2),) + sl]


def faces_of(ndim, a):
    """The face ghost layers of a ghosted array (edges and corners left out)."""
    out = []
    nv = ndim + 2
    for ax in range(ndim):
        for side in (0, 1):
            idx = [slice(2, -2)] * ndim
            idx[ndim - 1 - ax] = slice(0, 2) if side == 0 else slice(-2, None)
            out.append(a[(slice(0, nv),) + tuple(idx)])
    return out


def manual_loop(ndim, dims, cells, p, dx, nsteps, precision):
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    dts = []
    for _ in range(nsteps):
        dt = hydro.rhyme_cfl_time_step(p, box, dx)
        hydro.rhyme_samr_bc_set_boundaries(p, box)
        hydro.rhyme_muscl_hancock_solve(p, box, dx, dt)
        dts.append(dt)
    out = box.download().copy()
    box.close()
    return out, np.array(dts)


def run_loop(ndim, dims, cells, p, dx, nsteps, precision):
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    dts = hydro.run(p, box, dx, nsteps)
    out = box.download().copy()
    box.close()
    return out, dts


CASES = [
    (3, [20, 18, 22], [1, 2, 3, 3, 2, 1], hydro.SUPERBEE),
    (3, [16, 32, 33], [3, 3, 1, 1, 2, 2], hydro.VAN_LEER),
    (3, [9, 7, 2], [2, 2, 2, 2, 3, 3], hydro.MINMOD),
    (2, [37, 29], [1, 3, 3, 1, 2, 2], hydro.VAN_ALBADA),
    (2, [64, 17], [4, 2, 3, 3, 4, 4], hydro.MINMOD),
]


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("solver", [hydro.MEMORY_INTENSIVE, hydro.CPU_INTENSIVE])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_run_equals_flux_plane_loop(case, solver, precision):
    ndim, dims, bc, lim = CASES[case]
    say("run vs flux-plane loop", ndim, dims, bc, "solver", solver, "precision", precision)
    ncmp = hydro.ncmp_for(ndim)
    inflows = None
    if 4 in bc:
        inflows = np.zeros((6, ncmp))
        inflows[:, 0] = 1.1
        inflows[:, 1] = 0.05
        inflows[:, ndim + 1] = 2.0
        inflows[:, ndim + 2:] = 0.25
    p = params(ndim, lim, solver, bc, inflows=inflows)
    cells = ics.random_smooth(ndim, dims, seed=60 + case)
    dx = [1.0 / 32, 1.0 / 30, 1.0 / 28][:ndim]
    n = 7
    a, adts = run_loop(ndim, dims, cells, p, dx, n, precision)
    b, bdts = manual_loop(ndim, dims, cells, p, dx, n, precision)
    np.testing.assert_array_equal(adts, bdts)
    np.testing.assert_array_equal(interior(ndim, a), interior(ndim, b))
    # rhy_run leaves the ghost faces of its final state filled
    filled = a.copy()
    oracle.bc_set_boundaries(oracle.params_from(p), dims, filled)
    if precision == hydro.MIXED:  # the device holds the inflow states in FP32
        filled = filled.astype(np.float32).astype(np.float64)
    for got, want in zip(faces_of(ndim, a), faces_of(ndim, filled)):
        np.testing.assert_array_equal(got, want)
    # the other components are carried through untouched
    np.testing.assert_array_equal(a[ndim + 2:], b[ndim + 2:])


def test_run_continues_across_calls():
    """Three rhy_run calls of 2, 3 and 4 steps equal one call of 9 (the state
    buffers swap every step, an odd number of times per call here)."""
    p = params(3, hydro.SUPERBEE, hydro.MEMORY_INTENSIVE, [1] * 6)
    dims = [24, 20, 18]
    cells = ics.blast_3d(dims=dims, seed=9)
    dx = [1.0 / 24] * 3
    one, d1 = run_loop(3, dims, cells, p, dx, 9, hydro.FP64)
    box = hydro.SamrBox(3, dims, cells=cells.copy())
    box.upload()
    d = np.concatenate([hydro.run(p, box, dx, k) for k in (2, 3, 4)])
    out = box.download()
    box.close()
    np.testing.assert_array_equal(d, d1)
    np.testing.assert_array_equal(interior(3, out), interior(3, one))


def test_sampled_kernel_timing():
    """rhy_box_set_profiling(N): events around the launches of steps 0, N,
    2N, ... of each rhy_run call and nothing else; N = 1 every launch; the
    state is the same whatever the timing."""
    p = params(3, hydro.SUPERBEE, hydro.MEMORY_INTENSIVE, [1] * 6)
    dims = [24, 20, 18]
    cells = ics.blast_3d(dims=dims, seed=4)
    dx = [1.0 / 24] * 3
    ref, dref = run_loop(3, dims, cells, p, dx, 9, hydro.FP64)
    box = hydro.SamrBox(3, dims, cells=cells.copy())
    box.upload()
    box.set_profiling(3)
    d = [hydro.run(p, box, dx, 7)]           # steps 0, 3, 6 timed
    say("sampled", [box.kernel_stats(k) for k in range(6)])
    assert box.kernel_stats(4)[0] == 3       # one dt launch per step
    assert box.kernel_stats(0)[0] == 9       # x walk, y walk, z walk
    assert box.kernel_stats(2)[0] == 3       # the new state's ghost fill
    assert box.kernel_stats(3)[0] == 0       # the prologue's CFL pass: untimed
    assert all(box.kernel_stats(k)[1] > 0.0 for k in (0, 4))
    box.set_profiling(1)                     # every launch, prologue included
    d.append(hydro.run(p, box, dx, 2))
    say("every", [box.kernel_stats(k) for k in range(6)])
    assert box.kernel_stats(4)[0] == 2
    assert box.kernel_stats(0)[0] == 6
    assert box.kernel_stats(2)[0] == 3       # the prologue's fill and 2 steps
    assert box.kernel_stats(3)[0] == 1
    box.set_profiling(0)
    assert box.kernel_stats(4)[0] == 0
    with pytest.raises(_lib.RhyError):
        box.set_profiling(-1)
    out = box.download()
    box.close()
    np.testing.assert_array_equal(np.concatenate(d), dref)
    np.testing.assert_array_equal(interior(3, out), interior(3, ref))


def test_resize_then_step_matches_oracle():
    """rhy_box_resize (the device side of rhyme_mh_workspace_check's
    reallocation on a shape change, rhyme_mh_workspace_check.f90:51-59):
    step a box, resize it, upload a state of the new shape, step again; the
    result equals the oracle on the new shape."""
    p = params(3, hydro.MINMOD, hydro.MEMORY_INTENSIVE, [2] * 6)
    dims0, dims1 = [16, 12, 10], [22, 9, 14]
    c0 = ics.random_smooth(3, dims0, seed=3)
    c1 = ics.random_smooth(3, dims1, seed=4)
    box = hydro.SamrBox(3, dims0, cells=c0.copy())
    box.upload()
    hydro.run(p, box, [1.0 / 16] * 3, 3)
    _lib.check(_lib.load().rhy_box_resize(box.handle, _lib.iptr(_lib.dims3(dims1))),
               "rhy_box_resize")
    box.dims = list(dims1)
    box.cells = c1.copy()
    box.upload()
    dts = hydro.run(p, box, [1.0 / 22] * 3, 5)
    out = box.download()
    box.close()
    ref = c1.copy()
    rdts = oracle.run(oracle.params_from(p), dims1, ref, [1.0 / 22] * 3, 5)
    for c in range(5):
        e = np.abs(interior(3, out)[c] - interior(3, ref)[c]).sum() / \
            np.abs(interior(3, ref)[c]).sum()
        assert e <= 1e-12, (c, e)
    assert np.max(np.abs(dts - rdts) / rdts) <= 1e-12


def test_resize_refused_keeps_box_usable():
    """A resize the device cannot hold returns an error and leaves the box
    at its old shape and storage: it still downloads and steps."""
    p = params(3, hydro.MINMOD, hydro.MEMORY_INTENSIVE, [2] * 6)
    dims = [12, 10, 8]
    cells = ics.random_smooth(3, dims, seed=5)
    box = hydro.SamrBox(3, dims, cells=cells.copy())
    box.upload()
    big = [1400, 1400, 1400]  # 1404^3 cells: past the 2^31 limit
    rc = _lib.load().rhy_box_resize(box.handle, _lib.iptr(_lib.dims3(big)))
    assert rc == _lib.RHY_ERR_SHAPE
    # 1288^3 cells (under 2^31): state and flux planes together need ~495 GB
    huge = [1284, 1284, 1284]
    rc = _lib.load().rhy_box_resize(box.handle, _lib.iptr(_lib.dims3(huge)))
    assert rc in (_lib.RHY_ERR_SHAPE, _lib.RHY_ERR_OOM)
    out = box.download().copy()
    np.testing.assert_array_equal(out, cells)
    dts = hydro.run(p, box, [1.0 / 12] * 3, 2)
    assert np.all(np.isfinite(dts)) and np.all(dts > 0)
    box.close()
