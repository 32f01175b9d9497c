"""CPU mirror of the two-level hierarchy (refinement ratio 2).

TEST INFRASTRUCTURE ONLY, like the rest of oracle/.  The reference has no
coarse-fine coupling: its main loop advances every box of every level alone
with the level's dt (src/main.f90:118-136).  The hierarchy's semantics are
therefore defined here, and the device hierarchy (rhyme_amd.amr) must
reproduce this step bit for bit in FP64:

1. dt0 = min(CFL dt of the coarse box, 2 * min over fine boxes of their CFL
   dt), each CFL dt the reference's (rhyme_cfl_time_step.f90:3-60) with its
   level's dx.
2. The coarse state U_old is kept; the coarse box gets its ghost fill and one
   MUSCL-Hancock step with dt0 (rhyme_muscl_hancock_solve with the coarse
   faces' boundary types).
3. Two fine substeps of dt0 / 2.  Before each, every fine box's face ghosts
   (2 layers) take the value of the coarse cell that contains them: U_old
   before the first substep, 0.5 * U_old + 0.5 * U_new before the second
   (linear in time).  Then the face ghosts that lie in another fine box's
   interior take that box's cells instead (same level and time).  Fine boxes
   have no boundary fill of their own.
4. Refluxing (when on): for every coarse face on the boundary of a fine box,
   the fluxes of its 2^(ndim-1) fine faces are summed over both substeps
   (register: S_1 + S_2, S_k the pairwise sum (F_00 + F_10) + (F_01 + F_11)
   of substep k's fine faces, first index on the lower transverse axis, so
   that equal fluxes sum exactly), and the coarse cell outside the box takes
   U += dt0/dx0_a (F_c - reg 2^-ndim) on the box's low side along a and
   U += dt0/dx0_a (reg 2^-ndim - F_c) on its high side, F_c the coarse
   step's flux on that face; passes in order a = 0.., low side then high
   side, each over the boxes in order.  The fine fluxes through a box's
   boundary then replace the coarse ones, and mass and energy are conserved
   to rounding (the reference has no coarse-fine coupling, so this step has
   no reference counterpart either).
5. Every coarse cell under a fine box takes the mean of its 2^ndim children
   (summed with i fastest, then times 2^-ndim).

Placement: lo[d] is the coarse cell (1-based) that holds fine cells 1 and 2
along d; fine extents are even; every fine box with its ghost layers lies in
the coarse interior (2 <= lo[d], lo[d] + n_f[d] / 2 <= n_c[d]).
"""
import numpy as np

from . import bc_set_boundaries, cfl_time_step, mh_solve, mh_solve_fluxes


def _axis(ndim, d):
    """Array axis of Fortran index d (arrays are (ncmp, [z,] [y,] x))."""
    return ndim - d


def _fine_dims(cells, ndim):
    return [cells.shape[_axis(ndim, d)] - 4 for d in range(ndim)]


def check_placement(ndim, coarse_dims, fine_dims, lo):
    for n_f, l in zip(fine_dims, lo):
        for d in range(ndim):
            if n_f[d] % 2:
                raise ValueError("fine extents must be even (ratio 2)")
            if l[d] < 2 or l[d] + n_f[d] // 2 > coarse_dims[d]:
                raise ValueError("a fine box or its ghosts leave the coarse interior")
    # interiors must not overlap (each coarse cell restricted from one box)
    boxes = [[(l[d], l[d] + n_f[d] // 2 - 1) for d in range(ndim)] for n_f, l in zip(fine_dims, lo)]
    for a in range(len(boxes)):
        for b in range(a + 1, len(boxes)):
            if all(boxes[a][d][0] <= boxes[b][d][1] and boxes[b][d][0] <= boxes[a][d][1]
                   for d in range(ndim)):
                raise ValueError("fine boxes %d and %d cover the same coarse cells" % (a, b))


def prolong(ndim, fine, lo, cnew, cold, half):
    """Face ghosts of one fine box from its coarse parents (see step 3)."""
    nv = ndim + 2
    n_f = _fine_dims(fine, ndim)
    for a in range(ndim):
        for side in (0, 1):
            for layer in (0, 1):
                g = n_f[a] + 1 + layer if side else -layer
                fidx = [None] * ndim
                cidx = [None] * ndim
                for d in range(ndim):
                    f = np.array([g]) if d == a else np.arange(1, n_f[d] + 1)
                    par = lo[d] + (f + 1) // 2 - 1
                    fidx[d] = f + 1        # array index of Fortran index f
                    cidx[d] = par + 1
                # numpy axes are reversed (x last)
                fsel = np.ix_(*[fidx[d] for d in reversed(range(ndim))])
                csel = np.ix_(*[cidx[d] for d in reversed(range(ndim))])
                for c in range(nv):
                    o = cold[c][csel]
                    fine[c][fsel] = 0.5 * o + 0.5 * cnew[c][csel] if half else o


def fill_siblings(ndim, fines, lo):
    """Face ghosts of each fine box that lie in another fine box's interior
    take that box's cells (see step 3).  Box i's cell f lies at fine-level
    index 2 (lo_i - 1) + f; interiors do not overlap."""
    nv = ndim + 2
    dims = [_fine_dims(f, ndim) for f in fines]
    for i, fi in enumerate(fines):
        bi = [2 * (lo[i][d] - 1) for d in range(ndim)]
        for a in range(ndim):
            for side in (0, 1):
                for j, fj in enumerate(fines):
                    if j == i:
                        continue
                    bj = [2 * (lo[j][d] - 1) for d in range(ndim)]
                    g0, g1 = [], []
                    for d in range(ndim):
                        s0, s1 = bi[d] + 1, bi[d] + dims[i][d]
                        if d == a:
                            s0 = bi[d] + dims[i][d] + 1 if side else bi[d] - 1
                            s1 = s0 + 1
                        g0.append(max(s0, bj[d] + 1))
                        g1.append(min(s1, bj[d] + dims[j][d]))
                    if any(g0[d] > g1[d] for d in range(ndim)):
                        continue
                    # array index of Fortran index f is f + 1; numpy axes reversed
                    dsel = tuple(slice(g0[d] - bi[d] + 1, g1[d] - bi[d] + 2)
                                 for d in reversed(range(ndim)))
                    ssel = tuple(slice(g0[d] - bj[d] + 1, g1[d] - bj[d] + 2)
                                 for d in reversed(range(ndim)))
                    for c in range(nv):
                        fi[c][dsel] = fj[c][ssel]


def restrict(ndim, fine, lo, coarse):
    """Coarse cells under one fine box from their children (see step 4)."""
    nv = ndim + 2
    n_f = _fine_dims(fine, ndim)
    m = [n // 2 for n in n_f]
    w = {1: 0.5, 2: 0.25, 3: 0.125}[ndim]
    csel = tuple(slice(lo[d] + 1, lo[d] + 1 + m[d]) for d in reversed(range(ndim)))
    for c in range(nv):
        s = np.zeros([m[d] for d in reversed(range(ndim))])
        for z in range(2 if ndim > 2 else 1):
            for y in range(2 if ndim > 1 else 1):
                for x in range(2):
                    off = (x, y, z)
                    fsel = tuple(slice(2 + off[d], 2 + off[d] + 2 * m[d], 2)
                                 for d in reversed(range(ndim)))
                    s = s + fine[c][fsel]
        coarse[c][csel] = s * w


def regrid(ndim, coarse, fines, lo, new_dims, new_lo):
    """The fine level moved to new boxes (new_dims[i] cells at new_lo[i]):
    every cell of a new box, both ghost layers included, takes the old fine
    cell at the same fine-level position (cell f of a box at lo lies at
    2 (lo - 1) + f) when an old box's interior holds it, else the value of
    its coarse parent (cell lo + (f + 1) // 2 - 1), hydro components only;
    the others are 0.  Returns the new fine arrays."""
    nv = ndim + 2
    out = []
    for n_f, l in zip(new_dims, new_lo):
        shape = (coarse.shape[0],) + tuple(n_f[d] + 4 for d in reversed(range(ndim)))
        new = np.zeros(shape)
        f = [np.arange(-1, n_f[d] + 3) for d in range(ndim)]
        parent = [l[d] + (f[d] + 1) // 2 - 1 + 1 for d in range(ndim)]  # array indices
        new[:nv] = coarse[(slice(0, nv),) + np.ix_(*[parent[d] for d in reversed(range(ndim))])]
        gf = [2 * (l[d] - 1) + f[d] for d in range(ndim)]
        for old, lb in zip(fines, lo):
            m = _fine_dims(old, ndim)
            sel_new, sel_old = [], []
            for d in range(ndim):
                sc = gf[d] - 2 * (lb[d] - 1)            # the old box's cell index
                hit = np.nonzero((sc >= 1) & (sc <= m[d]))[0]
                if hit.size == 0:
                    break
                sel_new.append(slice(hit[0], hit[-1] + 1))
                sel_old.append(slice(sc[hit[0]] + 1, sc[hit[-1]] + 2))
            else:
                new[(slice(0, nv),) + tuple(reversed(sel_new))] = \
                    old[(slice(0, nv),) + tuple(reversed(sel_old))]
        out.append(new)
    return out


def _transverse(ndim, a):
    return [d for d in range(ndim) if d != a]


def register(ndim, fr, lo, reg, box, reset):
    """Add one fine box's boundary fluxes fr (mh_solve_fluxes of its
    substep) to the register entries reg[(box, a, side)] (see step 4)."""
    nv = ndim + 2
    n_f = [fr.shape[2 + ndim - 1 - d] - 4 for d in range(ndim)]
    for a in range(ndim):
        tr = _transverse(ndim, a)
        for side in (0, 1):
            fa = n_f[a] if side else 0

            def face(c, x, y):
                sel = [None] * ndim
                sel[a] = slice(fa + 1, fa + 2)
                for k, d in enumerate(tr):
                    o = x if k == 0 else y
                    sel[d] = slice(2 + o, 2 + o + n_f[d], 2)
                return fr[a][c][tuple(sel[d] for d in reversed(range(ndim)))]
            key = (box, a, side)
            out = []
            for c in range(nv):
                if ndim == 1:
                    s = face(c, 0, 0)
                elif ndim == 2:
                    s = face(c, 0, 0) + face(c, 1, 0)
                else:
                    s = (face(c, 0, 0) + face(c, 1, 0)) + (face(c, 0, 1) + face(c, 1, 1))
                out.append(s.copy() if reset else reg[key][c] + s)
            reg[key] = out


def reflux(ndim, coarse, cflux, fine_dims, lo, reg, dt, dx):
    """The coarse cells next to the fine boxes take the flux corrections
    (see step 4)."""
    nv = ndim + 2
    w = {1: 0.5, 2: 0.25, 3: 0.125}[ndim]
    for a in range(ndim):
        dtdx = dt / dx[a]
        for side in (0, 1):
            for i, (n_f, l) in enumerate(zip(fine_dims, lo)):
                out = l[a] + n_f[a] // 2 if side else l[a] - 1
                low = out - 1 if side else out
                so, sl = [None] * ndim, [None] * ndim
                for d in range(ndim):
                    if d == a:
                        so[d] = slice(out + 1, out + 2)
                        sl[d] = slice(low + 1, low + 2)
                    else:
                        so[d] = sl[d] = slice(l[d] + 1, l[d] + 1 + n_f[d] // 2)
                so = tuple(so[d] for d in reversed(range(ndim)))
                sl = tuple(sl[d] for d in reversed(range(ndim)))
                for c in range(nv):
                    favg = reg[(i, a, side)][c] * w
                    fc = cflux[a][c][sl]
                    dd = dtdx * (favg - fc) if side else dtdx * (fc - favg)
                    coarse[c][so] = coarse[c][so] + dd


def fine_params(p):
    """The fine level's parameters: the coarse ones with every face type 0
    (ghosts set by the hierarchy)."""
    q = type(p).from_buffer_copy(p)
    for f in range(6):
        q.bc_types[f] = 0
    return q


def step(pc, pf, coarse_dims, coarse, dx0, fines, lo, refluxing=True):
    """One coarse step of the hierarchy, in place; returns dt0.  pc / pf are
    oracle parameters of the coarse and fine levels, fines the fine boxes'
    host arrays, lo their placements."""
    ndim = len(coarse_dims)
    dx1 = [x / 2.0 for x in dx0]
    fdims = [_fine_dims(f, ndim) for f in fines]
    dt0 = cfl_time_step(pc, coarse_dims, coarse, dx0)
    for fd, f in zip(fdims, fines):
        dt0 = min(dt0, 2.0 * cfl_time_step(pf, fd, f, dx1))
    old = coarse.copy()
    bc_set_boundaries(pc, coarse_dims, coarse)
    cflux = mh_solve_fluxes(pc, coarse_dims, coarse, dx0, dt0) if refluxing else \
        mh_solve(pc, coarse_dims, coarse, dx0, dt0)
    reg = {}
    for half in (0, 1):
        for f, l in zip(fines, lo):
            prolong(ndim, f, l, coarse, old, half)
        fill_siblings(ndim, fines, lo)
        for i, (fd, f) in enumerate(zip(fdims, fines)):
            if refluxing:
                register(ndim, mh_solve_fluxes(pf, fd, f, dx1, 0.5 * dt0), lo[i], reg, i,
                         half == 0)
            else:
                mh_solve(pf, fd, f, dx1, 0.5 * dt0)
    if refluxing:
        reflux(ndim, coarse, cflux, fdims, lo, reg, dt0, dx0)
    for f, l in zip(fines, lo):
        restrict(ndim, f, l, coarse)
    return dt0


def _kids_cycle(ndim, pf, parent, parent_old, parent_flux, dt, dx, dxs, lev, levels, kids,
                refluxing):
    """Two substeps of dt / 2 of the boxes `kids` of levels[lev] (children of
    `parent`, whose step of dt from parent_old gave parent_flux), each followed
    by its own children's cycle; then their reflux and restriction into the
    parent.  Siblings are the children of one parent."""
    fines = [levels[lev]["boxes"][i] for i in kids]
    lo = [levels[lev]["lo"][i] for i in kids]
    fdims = [_fine_dims(f, ndim) for f in fines]
    reg = {}
    for half in (0, 1):
        for f, l in zip(fines, lo):
            prolong(ndim, f, l, parent, parent_old, half)
        fill_siblings(ndim, fines, lo)
        for j, (fd, f) in enumerate(zip(fdims, fines)):
            f_old = f.copy()
            fl = mh_solve_fluxes(pf, fd, f, dxs[lev + 1], 0.5 * dt)
            if refluxing:
                register(ndim, fl, lo[j], reg, j, half == 0)
            if lev + 1 < len(levels):
                grand = [g for g, par in enumerate(levels[lev + 1]["parent"]) if par == kids[j]]
                if grand:
                    _kids_cycle(ndim, pf, f, f_old, fl, 0.5 * dt, dxs[lev + 1], dxs, lev + 1,
                                levels, grand, refluxing)
    if refluxing:
        reflux(ndim, parent, parent_flux, fdims, lo, reg, dt, dx)
    for f, l in zip(fines, lo):
        restrict(ndim, f, l, parent)


def step_levels(pc, pf, coarse_dims, coarse, dx0, levels, refluxing=True):
    """One coarse step of a hierarchy of 1 + len(levels) levels, in place;
    returns dt0.  levels[k] = {"boxes": arrays, "lo": placements, "parent":
    index of each box's parent in levels[k - 1] (level 1: the coarse box,
    None)}, level k + 1's boxes at dx0 / 2^(k + 1), each inside one parent
    box (lo relative to it, the placement rules above).  dt0 = min over
    levels of 2^k times the level's CFL dt; the coarse box takes its ghost
    fill and one step of dt0; every box of level k + 1 then takes two
    substeps of half its parent's dt, each after the prolongation from its
    parent (U_old, then 0.5 U_old + 0.5 U_new) and the sibling fill among
    the parent's children, and followed by its own children's cycle; then
    the children's reflux into the parent (from the parent step's fluxes)
    and restriction.  With one fine level this is step() bit for bit."""
    ndim = len(coarse_dims)
    dxs = [list(dx0)] + [[x / 2.0 ** (k + 1) for x in dx0] for k in range(len(levels))]
    dt0 = cfl_time_step(pc, coarse_dims, coarse, dx0)
    for k, lev in enumerate(levels):
        for f in lev["boxes"]:
            dt0 = min(dt0, 2.0 ** (k + 1) * cfl_time_step(pf, _fine_dims(f, ndim), f, dxs[k + 1]))
    old = coarse.copy()
    bc_set_boundaries(pc, coarse_dims, coarse)
    cflux = mh_solve_fluxes(pc, coarse_dims, coarse, dx0, dt0)
    if levels and levels[0]["boxes"]:
        _kids_cycle(ndim, pf, coarse, old, cflux, dt0, dxs[0], dxs, 0, levels,
                    list(range(len(levels[0]["boxes"]))), refluxing)
    return dt0
