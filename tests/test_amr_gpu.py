"""Two-level hierarchy on the device (rhyme_amd.amr) against its CPU mirror
(oracle/amr.py): ratio-2 prolongation of the fine boxes' face ghosts (linear
in time over the two substeps), subcycled fine steps, refluxing of the
coarse cells next to the fine boxes, restriction.  The
reference has no coarse-fine coupling (src/main.f90:118-136 advances every
box alone), so oracle/amr.py defines the semantics; the kernels under it are
the single-box and batch steps whose parity the other tests establish."""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import amr as oamr
from rhyme_amd import _lib, amr, hydro, ics
from rhyme_amd.batch import SamrBatch, _L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def params(ndim, limiter, bc, courant=0.3):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), cfl=hydro.Cfl(courant),
        bc=hydro.SamrBc(types=[bc] * 6))


def inject(ndim, coarse, lo, fdims, seed):
    """A fine box (ghosts included) filled from its coarse parents, every
    conserved variable of a cell scaled by one factor in [0.97, 1.03] (same
    velocity, pressure scaled) so that the levels differ."""
    ncmp = coarse.shape[0]
    fine = np.zeros(hydro.box_shape(ndim, fdims, ncmp))
    idx = []
    for d in reversed(range(ndim)):
        f = np.arange(-1, fdims[d] + 3)
        idx.append(lo[d] + (f + 1) // 2 - 1 + 1)     # array index of the parent
    fine[...] = coarse[(slice(None),) + np.ix_(*idx)]
    s = 1.0 + 0.03 * np.random.default_rng(seed).uniform(-1, 1, size=fine.shape[1:])
    fine[:ndim + 2] *= s
    return fine


def rel_l1(a, b, ndim):
    inner = (slice(None),) + (slice(2, -2),) * ndim
    a, b = a[inner], b[inner]
    return [float(np.abs(a[c] - b[c]).sum() / max(np.abs(b[c]).sum(), 1e-300))
            for c in range(ndim + 2)]


CASES = {
    "1d": (1, [96], [[24], [16]], [[10], [60]], hydro.VAN_LEER, hydro.OUTFLOW, 12),
    "2d": (2, [40, 32], [[24, 16], [8, 12]], [[4, 6], [24, 18]], hydro.SUPERBEE,
           hydro.PERIODIC, 8),
    "3d": (3, [20, 18, 16], [[12, 8, 10], [6, 6, 6]], [[3, 4, 3], [12, 11, 10]],
           hydro.MINMOD, hydro.OUTFLOW, 4),
    # fine boxes that touch: ghosts from siblings across the shared faces
    "2d_adjacent": (2, [32, 32], [[12, 16], [8, 16], [12, 8]], [[6, 8], [12, 8], [8, 16]],
                    hydro.VAN_ALBADA, hydro.PERIODIC, 6),
    "3d_adjacent": (3, [20, 20, 20], [[8, 8, 8]] * 4, [[5, 5, 5], [9, 5, 5], [5, 9, 5],
                                                       [9, 9, 9]], hydro.VAN_LEER,
                    hydro.OUTFLOW, 4),
}


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("case", sorted(CASES))
def test_hierarchy_matches_oracle(case, precision):
    ndim, cdims, fdims, lo, lim, bc, nsteps = CASES[case]
    say("hierarchy", case, "precision", precision, "coarse", cdims, "fine", fdims, "lo", lo)
    coarse = ics.random_smooth(ndim, cdims, seed=7 + ndim)
    fines = [inject(ndim, coarse, l, d, 100 + i) for i, (l, d) in enumerate(zip(lo, fdims))]
    p = params(ndim, lim, bc)
    dx0 = [1.0 / cdims[0]] * ndim
    h = amr.AmrHierarchy(ndim, cdims, dx0, lo, fdims, precision=precision)
    h.upload(coarse, fines)
    dts = h.run(p, nsteps)
    dc, dfs = h.download()
    h.close()
    rc, rfs = coarse.copy(), [f.copy() for f in fines]
    pc = oracle.params_from(p)
    pf = oamr.fine_params(pc)
    rdts = np.array([oamr.step(pc, pf, cdims, rc, dx0, rfs, lo) for _ in range(nsteps)])
    tol = 1e-12 if precision == hydro.FP64 else 1e-5
    e_dt = float(np.max(np.abs(dts - rdts) / rdts))
    say("  dt rel err", e_dt)
    assert e_dt <= tol
    e = rel_l1(dc, rc, ndim)
    say("  coarse rel L1", e)
    assert max(e) <= tol
    for i, (a, b) in enumerate(zip(dfs, rfs)):
        e = rel_l1(a, b, ndim)
        say("  fine", i, "rel L1", e)
        assert max(e) <= tol
    # restriction ran last: covered coarse cells are the mean of the device's fine cells
    again = dc.copy()
    for f, l in zip(dfs, lo):
        oamr.restrict(ndim, f, l, again)
    if precision == hydro.FP64:
        assert np.array_equal(again, dc)
    else:
        # FP32 sums of 2^ndim children: a few ulp of the variable's scale
        for c in range(ndim + 2):
            np.testing.assert_allclose(again[c], dc[c], rtol=0,
                                       atol=1e-6 * np.abs(dc[c]).max())


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_split_fine_box_identical(precision):
    """A 3D fine box cut in two along x, and each half cut again along y:
    the device hierarchy gives the whole box's states and dts bit for bit
    (the shared faces' ghosts come from the siblings)."""
    ndim, cdims = 3, [24, 20, 16]
    whole, lo = [16, 12, 8], [5, 4, 4]
    coarse = ics.random_smooth(ndim, cdims, seed=31)
    fw = inject(ndim, coarse, lo, whole, 77)
    parts, plo = [], []
    for x0 in (0, 8):
        for y0 in (0, 6):
            parts.append(np.ascontiguousarray(fw[:, :, y0:y0 + 10, x0:x0 + 12]))
            plo.append([lo[0] + x0 // 2, lo[1] + y0 // 2, lo[2]])
    p = params(ndim, hydro.SUPERBEE, hydro.OUTFLOW)
    dx0 = [1.0 / cdims[0]] * ndim

    def go(fdims, lo_list, fines):
        h = amr.AmrHierarchy(ndim, cdims, dx0, lo_list, fdims, precision=precision)
        h.upload(coarse, fines)
        dts = h.run(p, 5)
        out = h.download()
        h.close()
        return dts, out
    d1, (c1, f1) = go([whole], [lo], [fw])
    d2, (c2, f2) = go([[8, 6, 8]] * 4, plo, parts)
    say("split box, precision", precision, "dts", d1[:2])
    np.testing.assert_array_equal(d1, d2)
    inner = (slice(0, ndim + 2),) + (slice(2, -2),) * ndim
    np.testing.assert_array_equal(c1[inner], c2[inner])
    k = 0
    for x0 in (0, 8):
        for y0 in (0, 6):
            np.testing.assert_array_equal(
                f1[0][:ndim + 2, 2:-2, 2 + y0:8 + y0, 2 + x0:10 + x0], f2[k][inner])
            k += 1


def test_hierarchy_blast_fp64():
    """A 2D blast whose shock crosses both fine boxes (strong shocks, faces
    that defer to the full solver)."""
    ndim, cdims = 2, [48, 48]
    fdims, lo = [[20, 24], [16, 16]], [[14, 12], [30, 30]]
    coarse = ics.blast(ndim, cdims, seed=3)
    fines = [inject(ndim, coarse, l, d, 200 + i) for i, (l, d) in enumerate(zip(lo, fdims))]
    p = params(ndim, hydro.VAN_ALBADA, hydro.OUTFLOW)
    dx0 = [1.0 / 48] * 2
    h = amr.AmrHierarchy(ndim, cdims, dx0, lo, fdims)
    h.upload(coarse, fines)
    dts = h.run(p, 10)
    dc, dfs = h.download()
    h.close()
    rc, rfs = coarse.copy(), [f.copy() for f in fines]
    pc = oracle.params_from(p)
    rdts = np.array([oamr.step(pc, oamr.fine_params(pc), cdims, rc, dx0, rfs, lo)
                     for _ in range(10)])
    assert np.max(np.abs(dts - rdts) / rdts) <= 1e-12
    assert max(rel_l1(dc, rc, ndim)) <= 1e-12
    for a, b in zip(dfs, rfs):
        e = rel_l1(a, b, ndim)
        say("blast fine rel L1", e)
        assert max(e) <= 1e-12


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_uniform_state_stays_uniform_large(precision):
    """At a full-size coarse box (128^3 with four 64^3 fine boxes) a uniform
    state is a fixed point of every kernel: both levels stay bit-identical."""
    ndim, cdims = 3, [128, 128, 128]
    fdims = [[64, 64, 64]] * 4
    lo = [[4, 4, 4], [70, 4, 4], [4, 70, 40], [70, 70, 90]]
    ncmp = hydro.ncmp_for(ndim)

    def uniform(dims):
        # rho 1, u 0.25, p 1 (gamma 5/3): E = 1.5 + 0.03125, exact in FP32 too
        c = np.zeros(hydro.box_shape(ndim, dims, ncmp))
        c[0] = 1.0
        c[1] = 0.25
        c[ndim + 1] = 1.53125
        return c
    h = amr.AmrHierarchy(ndim, cdims, [1.0 / 128] * 3, lo, fdims, precision=precision)
    h.upload(uniform(cdims), [uniform(d) for d in fdims])
    dts = h.run(params(ndim, hydro.SUPERBEE, hydro.PERIODIC), 3)
    dc, dfs = h.download()
    h.close()
    say("uniform 128^3 + 4 x 64^3, precision", precision, "dts", dts)
    inner = (slice(0, ndim + 2),) + (slice(2, -2),) * ndim
    assert np.array_equal(dc[inner], uniform(cdims)[inner])
    for f, d in zip(dfs, fdims):
        assert np.array_equal(f[inner], uniform(d)[inner])
    assert np.all(dts == dts[0])


def test_batch_cfl_dts_match_single_boxes():
    boxes = ics.amr_boxes(seed=11, nboxes=5, base=24, lo=6, hi=20)
    p = params(3, hydro.MINMOD, hydro.OUTFLOW)
    b = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes])
    for i, (_, _, c) in enumerate(boxes):
        b.upload(i, c)
    dts = b.cfl_dts(p)
    for i, (dims, dx, c) in enumerate(boxes):
        box = hydro.SamrBox(3, dims, cells=c.copy())
        box.upload()
        assert dts[i] == hydro.rhyme_cfl_time_step(p, box, dx)
        box.close()
    b.close()


def test_bad_placement_rejected_by_library():
    ndim, cdims = 2, [16, 16]
    h = amr.AmrHierarchy(ndim, cdims, [1.0 / 16] * 2, [[2, 2]], [[8, 8]])
    L = _L()
    for bad in ([1, 2, 1], [2, 13, 1]):
        lo = np.array(bad, dtype=np.int32)
        rc = L.rhy_amr_prolong(h.coarse.handle, h.fine._h,
                               lo.ctypes.data_as(C.POINTER(C.c_int32)), 0)
        assert rc == _lib.RHY_ERR_SHAPE
        rc = L.rhy_amr_restrict(h.coarse.handle, h.fine._h,
                                lo.ctypes.data_as(C.POINTER(C.c_int32)))
        assert rc == _lib.RHY_ERR_SHAPE
    h.close()


def total(ndim, cells, c):
    import math
    return math.fsum(cells[(c,) + (slice(2, -2),) * ndim].ravel())


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim", [2, 3])
def test_refluxing_conserves_on_device(ndim, precision):
    """Periodic coarse box, adjacent and separate fine boxes: with refluxing
    the device hierarchy keeps total mass and energy (the coarse level after
    restriction) to rounding in FP64 and to FP32 rounding in the mixed form,
    and matches the oracle's hierarchy; without refluxing both move."""
    cases = {
        2: ([32, 24], [[8, 8], [8, 8], [6, 4]], [[5, 6], [9, 6], [22, 15]]),
        3: ([16, 16, 12], [[8, 6, 4], [4, 6, 4], [4, 4, 4]], [[3, 4, 4], [7, 4, 4], [11, 11, 7]]),
    }
    cdims, fdims, lo = cases[ndim]
    p = params(ndim, hydro.MINMOD, hydro.PERIODIC, courant=0.2)
    dx0 = [1.0 / cdims[0]] * ndim
    coarse = ics.random_smooth(ndim, cdims, seed=40 + ndim, vamp=0.8)
    fines = [amr.fine_from_coarse(ndim, coarse, l, d) for l, d in zip(lo, fdims)]
    for f, l in zip(fines, lo):
        oamr.restrict(ndim, f, l, coarse)
    drift = {}
    for refl in (True, False):
        say("reflux conservation", ndim, "precision", precision, "refluxing", refl)
        h = amr.AmrHierarchy(ndim, cdims, dx0, lo, fdims, precision=precision, refluxing=refl)
        h.upload(coarse, fines)
        dc0, _ = h.download()
        m0, e0 = total(ndim, dc0, 0), total(ndim, dc0, ndim + 1)
        dts = h.run(p, 4)
        dc, dfs = h.download()
        h.close()
        drift[refl] = max(abs(total(ndim, dc, 0) - m0) / abs(m0),
                          abs(total(ndim, dc, ndim + 1) - e0) / abs(e0))
        rc, rfs = coarse.copy(), [f.copy() for f in fines]
        pc = oracle.params_from(p)
        rdts = [oamr.step(pc, oamr.fine_params(pc), cdims, rc, dx0, rfs, lo, refluxing=refl)
                for _ in range(4)]
        tol = 1e-12 if precision == hydro.FP64 else 1e-5
        assert np.max(np.abs(dts - rdts) / rdts) <= tol
        assert max(rel_l1(dc, rc, ndim)) <= tol
        for a, b in zip(dfs, rfs):
            assert max(rel_l1(a, b, ndim)) <= tol
    say("  drift with / without refluxing", drift[True], drift[False])
    assert drift[True] <= (1e-14 if precision == hydro.FP64 else 1e-6)
    assert drift[False] > 10 * drift[True]


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("case", ["2d", "3d", "2d_adjacent", "3d_adjacent"])
def test_coarse_walks_equal_flux_plane_coarse_step(case, precision):
    """The coarse step through rhy_amr_coarse_step (accumulating walks; the
    reflux forms the coarse y/z interface fluxes from the pre-step state)
    gives the hierarchy the flux-plane coarse step (ghost fill +
    rhy_mh_solve, fluxes from the planes) gives, bit for bit: dts, coarse
    and fine interiors, with refluxing on."""
    ndim, cdims, fdims, lo, lim, bc, nsteps = CASES[case]
    say("coarse walks", case, "precision", precision)
    coarse = ics.random_smooth(ndim, cdims, seed=17 + ndim, vamp=0.6)
    fines = [inject(ndim, coarse, l, d, 200 + i) for i, (l, d) in enumerate(zip(lo, fdims))]
    p = params(ndim, lim, bc)
    dx0 = [1.0 / cdims[0]] * ndim
    out = {}
    for walks in (True, False):
        h = amr.AmrHierarchy(ndim, cdims, dx0, lo, fdims, precision=precision,
                             coarse_walks=walks)
        h.upload(coarse, fines)
        dts = h.run(p, nsteps)
        dc, dfs = h.download()
        h.close()
        out[walks] = (dts, dc, dfs)
    inner = (slice(None),) + (slice(2, -2),) * ndim
    np.testing.assert_array_equal(out[True][0], out[False][0])
    np.testing.assert_array_equal(out[True][1][inner], out[False][1][inner])
    for a, b in zip(out[True][2], out[False][2]):
        np.testing.assert_array_equal(a[inner], b[inner])


def test_coarse_step_equals_fill_and_solve():
    """rhy_amr_coarse_step on a lone box equals rhy_bc_set_boundaries +
    rhy_mh_solve bit for bit (interior), both solver types."""
    for solver in (hydro.MEMORY_INTENSIVE, hydro.CPU_INTENSIVE):
        dims = [18, 14, 12]
        cells = ics.random_smooth(3, dims, seed=5, vamp=0.5)
        p = hydro.make_params(
            3, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
            sl=hydro.SlopeLimiter(type=hydro.VAN_LEER), mh=hydro.MusclHancock(solver_type=solver),
            cfl=hydro.Cfl(0.3), bc=hydro.SamrBc(types=[hydro.REFLECTIVE] * 6))
        dx = [1.0 / 18] * 3
        a = hydro.SamrBox(3, dims, cells=cells.copy())
        a.upload()
        dt = hydro.rhyme_cfl_time_step(p, a, dx)
        rc = _L().rhy_amr_coarse_step(a.handle, C.byref(p), (C.c_double * 3)(*dx), dt)
        assert rc == 0
        ua = a.download().copy()
        a.close()
        b = hydro.SamrBox(3, dims, cells=cells.copy())
        b.upload()
        hydro.rhyme_samr_bc_set_boundaries(p, b)
        hydro.rhyme_muscl_hancock_solve(p, b, dx, dt)
        ub = b.download().copy()
        b.close()
        inner = (slice(None),) + (slice(2, -2),) * 3
        np.testing.assert_array_equal(ua[inner], ub[inner])


REGRID = {
    2: ([24, 20], [[8, 6], [6, 4]], [[3, 3], [12, 10]], [[10, 8], [4, 6], [6, 4]],
        [[4, 4], [10, 9], [17, 14]]),
    3: ([14, 12, 12], [[6, 6, 4], [4, 4, 4]], [[3, 3, 3], [8, 7, 7]], [[8, 4, 6], [4, 6, 4]],
        [[4, 3, 4], [9, 6, 8]]),
}


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim", [2, 3])
def test_regrid_matches_oracle(ndim, precision):
    """Two coarse steps, a regrid to shifted, grown and new fine boxes, two
    more steps: the device's regrid equals oracle/amr.py's regrid of the
    device's own arrays bit for bit (every cell of the new boxes, ghosts
    included; the coarse box untouched), and the steps after it match the
    oracle hierarchy started from those arrays."""
    cdims, odims, olo, ndims, nlo = REGRID[ndim]
    say("regrid", ndim, "precision", precision, "old", odims, olo, "new", ndims, nlo)
    coarse = ics.random_smooth(ndim, cdims, seed=80 + ndim, vamp=0.6)
    fines = [inject(ndim, coarse, l, d, 300 + i) for i, (l, d) in enumerate(zip(olo, odims))]
    p = params(ndim, hydro.MINMOD, hydro.OUTFLOW)
    dx0 = [1.0 / cdims[0]] * ndim
    h = amr.AmrHierarchy(ndim, cdims, dx0, olo, odims, precision=precision)
    h.upload(coarse, fines)
    h.run(p, 2)
    dc, dfs = h.download()
    h.regrid(nlo, ndims)
    dc2, dfs2 = h.download()
    assert np.array_equal(dc2, dc)
    want = oamr.regrid(ndim, dc, dfs, olo, ndims, nlo)
    for a, b in zip(dfs2, want):
        assert np.array_equal(a, b)
    dts = h.run(p, 2)
    ec, efs = h.download()
    h.close()
    pc = oracle.params_from(p)
    pf = oamr.fine_params(pc)
    rc, rfs = dc2.copy(), [w.copy() for w in want]
    rdts = np.array([oamr.step(pc, pf, cdims, rc, dx0, rfs, nlo) for _ in range(2)])
    tol = 1e-12 if precision == hydro.FP64 else 1e-5
    assert np.max(np.abs(dts - rdts) / rdts) <= tol
    assert max(rel_l1(ec, rc, ndim)) <= tol
    for a, b in zip(efs, rfs):
        assert max(rel_l1(a, b, ndim)) <= tol


LEVELS3 = {
    2: ([24, 20], [[12, 10], [6, 6]], [[4, 4], [15, 10]], [[8, 6], [4, 4]], [[2, 2], [7, 6]]),
    3: ([14, 12, 12], [[12, 8, 8]], [[3, 3, 3]], [[8, 6, 4]], [[2, 2, 3]]),
}


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim", [2, 3])
def test_three_levels_match_oracle(ndim, precision):
    """A third level (boxes inside level-1 box 0, dx0 / 4, four substeps
    per coarse step) through the slot calls (rhy_amr_*_slot) against
    oracle/amr.py step_levels: dts, coarse, level-1 and level-2 interiors,
    refluxing at both interfaces."""
    cdims, d1, lo1, d2, lo2 = LEVELS3[ndim]
    say("three levels", ndim, "precision", precision)
    coarse = ics.random_smooth(ndim, cdims, seed=110 + ndim, vamp=0.6)
    f1 = [inject(ndim, coarse, l, d, 400 + i) for i, (l, d) in enumerate(zip(lo1, d1))]
    f2 = [inject(ndim, f1[0], l, d, 500 + i) for i, (l, d) in enumerate(zip(lo2, d2))]
    p = params(ndim, hydro.VAN_LEER, hydro.OUTFLOW)
    dx0 = [1.0 / cdims[0]] * ndim
    h = amr.AmrHierarchy(ndim, cdims, dx0, lo1, d1, precision=precision,
                         level2=[{"parent": 0, "lo": lo2, "dims": d2}])
    h.upload(coarse, f1)
    h.upload_kids([f2])
    dts = h.run(p, 3)
    dc, dfs = h.download()
    (dks,) = h.download_kids()
    h.close()
    pc = oracle.params_from(p)
    rc, r1, r2 = coarse.copy(), [f.copy() for f in f1], [f.copy() for f in f2]
    levels = [{"boxes": r1, "lo": lo1, "parent": [None] * len(r1)},
              {"boxes": r2, "lo": lo2, "parent": [0] * len(r2)}]
    rdts = np.array([oamr.step_levels(pc, oamr.fine_params(pc), cdims, rc, dx0, levels)
                     for _ in range(3)])
    tol = 1e-12 if precision == hydro.FP64 else 1e-5
    e_dt = float(np.max(np.abs(dts - rdts) / rdts))
    say("  dt rel err", e_dt)
    assert e_dt <= tol
    assert max(rel_l1(dc, rc, ndim)) <= tol
    for a, b in zip(dfs + dks, r1 + r2):
        e = rel_l1(a, b, ndim)
        say("  box rel L1", e)
        assert max(e) <= tol
