"""The two-level hierarchy's CPU mirror (oracle/amr.py): placement rules,
prolongation and restriction against explicit loops over the cells, and
whole hierarchy steps that must keep a uniform state uniform, bit for bit,
on both levels."""
import numpy as np
import pytest

import oracle
from oracle import amr as oamr
from rhyme_amd import hydro, ics


def mk(ndim, bc=hydro.PERIODIC):
    return oracle.params_from(hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=hydro.MINMOD), cfl=hydro.Cfl(0.2),
        bc=hydro.SamrBc(types=[bc] * 6)))


def test_placement_rules():
    oamr.check_placement(2, [16, 16], [[8, 6]], [[2, 3]])
    with pytest.raises(ValueError):
        oamr.check_placement(2, [16, 16], [[7, 6]], [[2, 3]])          # odd extent
    with pytest.raises(ValueError):
        oamr.check_placement(2, [16, 16], [[8, 6]], [[1, 3]])          # ghost outside
    with pytest.raises(ValueError):
        oamr.check_placement(2, [16, 16], [[8, 6]], [[13, 3]])         # 13 + 4 > 16
    with pytest.raises(ValueError):
        oamr.check_placement(2, [16, 16], [[8, 6], [4, 4]], [[2, 3], [5, 4]])  # overlap
    oamr.check_placement(2, [16, 16], [[8, 6], [4, 4]], [[2, 3], [6, 4]])


def _fidx(f):
    return f + 1      # array index of Fortran index f (2 ghost layers)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_prolong_against_loops(ndim):
    g = np.random.default_rng(ndim)
    cdims = [10, 9, 8][:ndim]
    fdims = [6, 4, 8][:ndim]
    lo = [3, 2, 2][:ndim]
    ncmp = hydro.ncmp_for(ndim)
    cnew = g.uniform(1, 2, size=hydro.box_shape(ndim, cdims, ncmp))
    cold = g.uniform(1, 2, size=cnew.shape)
    for half in (0, 1):
        fine = np.zeros(hydro.box_shape(ndim, fdims, ncmp))
        oamr.prolong(ndim, fine, lo, cnew, cold, half)
        # every face ghost cell, one at a time
        for a in range(ndim):
            for gidx in (-1, 0, fdims[a] + 1, fdims[a] + 2):
                others = [range(1, fdims[d] + 1) if d != a else [gidx] for d in range(ndim)]
                for f in np.ndindex(*[len(o) for o in others]):
                    fi = [list(others[d])[f[d]] for d in range(ndim)]
                    par = [lo[d] + (fi[d] + 1) // 2 - 1 for d in range(ndim)]
                    fa = tuple(_fidx(fi[d]) for d in reversed(range(ndim)))
                    ca = tuple(_fidx(par[d]) for d in reversed(range(ndim)))
                    for c in range(ndim + 2):
                        want = cold[c][ca] if not half else 0.5 * cold[c][ca] + 0.5 * cnew[c][ca]
                        assert fine[c][fa] == want
        # interior and edge/corner ghosts untouched
        inner = (slice(None),) + (slice(2, -2),) * ndim
        assert not fine[inner].any()


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_restrict_against_loops(ndim):
    g = np.random.default_rng(10 + ndim)
    cdims = [10, 9, 8][:ndim]
    fdims = [6, 4, 8][:ndim]
    lo = [3, 2, 2][:ndim]
    ncmp = hydro.ncmp_for(ndim)
    fine = g.uniform(1, 2, size=hydro.box_shape(ndim, fdims, ncmp))
    coarse = g.uniform(1, 2, size=hydro.box_shape(ndim, cdims, ncmp))
    before = coarse.copy()
    oamr.restrict(ndim, fine, lo, coarse)
    covered = np.zeros(coarse.shape[1:], dtype=bool)
    for cell in np.ndindex(*[fdims[d] // 2 for d in range(ndim)]):
        cc = [lo[d] + cell[d] for d in range(ndim)]
        ca = tuple(_fidx(cc[d]) for d in reversed(range(ndim)))
        covered[ca] = True
        for c in range(ndim + 2):
            s = 0.0
            for z in range(2 if ndim > 2 else 1):
                for y in range(2 if ndim > 1 else 1):
                    for x in range(2):
                        off = (x, y, z)
                        fi = [2 * cell[d] + 1 + off[d] for d in range(ndim)]
                        s = s + fine[c][tuple(_fidx(fi[d]) for d in reversed(range(ndim)))]
            assert coarse[c][ca] == s * [0.5, 0.25, 0.125][ndim - 1]
    assert np.array_equal(coarse[:, ~covered], before[:, ~covered])
    assert np.array_equal(coarse[ndim + 2:], before[ndim + 2:])


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_uniform_state_stays_uniform(ndim):
    cdims = [16, 12, 10][:ndim]
    fdims = [[8, 6, 4][:ndim], [4, 4, 4][:ndim]]
    lo = [[2, 2, 2][:ndim], [9, 7, 6][:ndim]]
    oamr.check_placement(ndim, cdims, fdims, lo)
    ncmp = hydro.ncmp_for(ndim)
    gamma = 5.0 / 3.0

    def uniform(dims):
        c = np.zeros(hydro.box_shape(ndim, dims, ncmp))
        c[0] = 1.0
        c[ndim + 1] = 1.0 / (gamma - 1.0)
        return c
    coarse = uniform(cdims)
    fines = [uniform(d) for d in fdims]
    pc = mk(ndim)
    pf = oamr.fine_params(pc)
    dx0 = [1.0 / cdims[0]] * ndim
    for _ in range(3):
        dt = oamr.step(pc, pf, cdims, coarse, dx0, fines, lo)
        assert dt > 0
    inner = (slice(0, ndim + 2),) + (slice(2, -2),) * ndim
    assert np.array_equal(coarse[inner], uniform(cdims)[inner])
    for f, d in zip(fines, fdims):
        assert np.array_equal(f[inner], uniform(d)[inner])


def test_fill_siblings_against_loops():
    """2D, three fine boxes: two share a face along x, the third touches the
    second along y.  Every face ghost cell inside another box's interior
    takes that cell; every other cell is left alone."""
    ndim = 2
    fdims = [[8, 6], [6, 6], [4, 4]]
    lo = [[3, 3], [7, 3], [8, 6]]
    oamr.check_placement(ndim, [16, 16], fdims, lo)
    ncmp = hydro.ncmp_for(ndim)
    g = np.random.default_rng(5)
    fines = [g.uniform(1, 2, size=hydro.box_shape(ndim, d, ncmp)) for d in fdims]
    before = [f.copy() for f in fines]
    oamr.fill_siblings(ndim, fines, lo)

    def owner(gx, gy, skip):
        for j, (d, l) in enumerate(zip(fdims, lo)):
            fx, fy = gx - 2 * (l[0] - 1), gy - 2 * (l[1] - 1)
            if j != skip and 1 <= fx <= d[0] and 1 <= fy <= d[1]:
                return j, fx, fy
        return None
    copied = 0
    for i, (d, l) in enumerate(zip(fdims, lo)):
        for fy in range(-1, d[1] + 3):
            for fx in range(-1, d[0] + 3):
                face = (1 <= fx <= d[0]) != (1 <= fy <= d[1])   # face ghost, not edge
                o = owner(2 * (l[0] - 1) + fx, 2 * (l[1] - 1) + fy, i) if face else None
                for c in range(ncmp):
                    got = fines[i][c][fy + 1, fx + 1]
                    if o is not None and c < ndim + 2:
                        j, sx, sy = o
                        assert got == before[j][c][sy + 1, sx + 1]
                        copied += c == 0
                    else:
                        assert got == before[i][c][fy + 1, fx + 1]
    # box 0 <-> box 1 share a 6-cell x face, box 1 <-> box 2 a 4-cell y face
    assert copied == 2 * (2 * 6) + 2 * (2 * 4)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_split_fine_box_identical(ndim):
    """A fine box cut in two along x gives the same hierarchy step, bit for
    bit, as the whole box: the shared face's ghosts come from the sibling."""
    cdims = [24, 16, 12][:ndim]
    whole, lo = [16, 12, 8][:ndim], [5, 3, 3][:ndim]
    left = [8] + whole[1:]
    right = [8] + whole[1:]
    lo_r = [lo[0] + 4] + lo[1:]
    p = mk(ndim, hydro.OUTFLOW)
    pf = oamr.fine_params(p)
    dx0 = [1.0 / cdims[0]] * ndim
    coarse = ics.random_smooth(ndim, cdims, seed=20 + ndim)
    fw = np.zeros(hydro.box_shape(ndim, whole, coarse.shape[0]))
    idx = [lo[d] + (np.arange(-1, whole[d] + 3) + 1) // 2 for d in reversed(range(ndim))]
    fw[...] = coarse[(slice(None),) + np.ix_(*idx)]
    fw[:ndim + 2] *= 1.0 + 0.02 * np.random.default_rng(1).uniform(-1, 1, fw.shape[1:])
    fl = np.ascontiguousarray(fw[..., :left[0] + 4])
    fr = np.ascontiguousarray(fw[..., right[0]:])
    c1, c2 = coarse.copy(), coarse.copy()
    fines2 = [fl, fr]
    for _ in range(3):
        d1 = oamr.step(p, pf, cdims, c1, dx0, [fw], [lo])
        d2 = oamr.step(p, pf, cdims, c2, dx0, fines2, [lo, lo_r])
        assert d1 == d2
    inner = (slice(0, ndim + 2),) + (slice(2, -2),) * ndim
    assert np.array_equal(c1[inner], c2[inner])
    inner_w = (slice(0, ndim + 2),) + (slice(2, -2),) * (ndim - 1)
    assert np.array_equal(fw[inner_w + (slice(2, 2 + left[0]),)], fl[inner_w + (slice(2, -2),)])
    assert np.array_equal(fw[inner_w + (slice(2 + left[0], -2),)], fr[inner_w + (slice(2, -2),)])
