"""Seeded initial conditions of the benchmark workloads, in the reference's
box layout (host arrays of box_shape(ndim, dims, ncmp), ghosts zero as after
rhyme_samr_bc_init, rhyme_samr_bc/src/rhyme_samr_bc_init.f90:61-87).

The generator is numpy's counter-based Philox; its arrays are the *inputs*
to both the device path and the CPU oracle, so the choice of generator cannot
affect parity.  The non-hydro components (temp, ntr_frac_*) are left zero:
the hot path never reads them.
"""
import numpy as np

from .hydro import box_shape, ncmp_for


def rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def prim_to_cons(gamma, rho, v, p):
    """conv_prim_vars_to_cons (rhyme_ideal_gas_primitive_vars_to_conserved.f90:9-11)
    on arrays: v has shape (ndim, ...)."""
    s = np.zeros_like(rho)
    for d in range(v.shape[0]):
        s = s + v[d] * v[d]
    out = [rho]
    for d in range(v.shape[0]):
        out.append(rho * v[d])
    out.append(.5 * rho * s + p / (gamma - 1.0))
    return out


def _fill(ndim, dims, ncmp, gamma, rho, v, p):
    cells = np.zeros(box_shape(ndim, dims, ncmp), dtype=np.float64)
    cons = prim_to_cons(gamma, rho, v, p)
    sl = tuple(slice(2, 2 + n) for n in list(dims)[::-1])
    for c, arr in enumerate(cons):
        cells[(c,) + sl] = arr
    return cells


def cell_centres(dims):
    """x_d = (i_d - 1/2) / n_d on [0, 1] per axis, arrays in C order (z, y, x)."""
    axes = [(np.arange(n) + 0.5) / n for n in dims]
    grids = np.meshgrid(*axes[::-1], indexing="ij")
    return grids[::-1]  # x, y, z


def two_state_1d(n=1024, seed=0, gamma=5.0 / 3.0, ncmp=None):
    """1D two-state workload: discontinuity at x = 0.5 (between cells n/2 and
    n/2+1), u = 0, rho and p on each side drawn from U[0.1, 1]."""
    g = rng(seed)
    rl, pl, rr, pr = g.uniform(0.1, 1.0, size=4)
    ncmp = ncmp or ncmp_for(1)
    rho = np.where(np.arange(n) < n // 2, rl, rr).astype(np.float64)
    p = np.where(np.arange(n) < n // 2, pl, pr).astype(np.float64)
    v = np.zeros((1, n))
    return _fill(1, [n], ncmp, gamma, rho, v, p), dict(rho_l=rl, p_l=pl, rho_r=rr, p_r=pr)


SPHERE, CUBOID = 1, 2


def harmonic_profile(n, amp, lam):
    """A cos(kx) of a harmonic perturbation at the n cell centres x = q - 1/2
    of an axis, kx = 2 pi / lambda * x as rhyme_drawing_apply_perturbations.f90
    :115 evaluates it (pi = 4 atan(1))."""
    x = np.arange(1, n + 1, dtype=np.float64) - 0.5
    return amp * np.cos(2.0 * np.pi / lam * x)


def draw(drawing, ndim, dims, gamma=5.0 / 3.0, ncmp=None):
    """Host array of the state rhy_box_draw paints (rhyme_drawing_init.f90
    :19-78 for a uniform canvas, sharp shapes and harmonic perturbations):

    * canvas: every cell gets conv_prim_to_cons of (rho, v, p);
    * shapes in order, later over earlier: a sphere takes the cells whose
      centre x_d = q_d - 1/2 (1-based q) satisfies sum_d (x_d - c_d)^2 <=
      radius^2 (axis order), a cuboid the cells with lo <= q <= hi;
    * then every cell gets conv_prim_to_cons(h) added to its hydro
      components, h = 0 + sum over the perturbations of profile * (rho, v, p)
      (rhyme_drawing_apply_perturbations.f90:51-60, 115, 184).

    drawing: dict(rho, v, p, shapes=[dict(kind, centre, radius | lo, hi, rho,
    v, p)], harmonics=[dict(axis, profile, rho, v, p)])."""
    ncmp = ncmp or ncmp_for(ndim)
    dims = list(dims)[:ndim]
    shape = dims[::-1]
    pos = [(np.arange(1, n + 1, dtype=np.float64) - 0.5) for n in dims]
    grids = np.meshgrid(*pos[::-1], indexing="ij")[::-1]  # x, y, z arrays
    qs = np.meshgrid(*[np.arange(1, n + 1) for n in dims[::-1]], indexing="ij")[::-1]

    def vel(c):
        v = list(c["v"]) + [0.0] * 3
        return v[:ndim]
    rho = np.full(shape, float(drawing["rho"]))
    v = [np.full(shape, x) for x in vel(drawing)]
    p = np.full(shape, float(drawing["p"]))
    for sh in drawing.get("shapes", []):
        if sh["kind"] == SPHERE:
            r2 = None
            for d in range(ndim):
                x = grids[d] - sh["centre"][d]
                r2 = x * x if r2 is None else r2 + x * x
            m = r2 <= sh["radius"] * sh["radius"]
        else:
            m = np.ones(shape, dtype=bool)
            for d in range(ndim):
                m &= (qs[d] >= sh["lo"][d]) & (qs[d] <= sh["hi"][d])
        rho[m] = sh["rho"]
        for d, x in enumerate(vel(sh)):
            v[d][m] = x
        p[m] = sh["p"]
    cons = prim_to_cons(gamma, rho, np.array(v).reshape([ndim] + shape), p)
    harms = drawing.get("harmonics", [])
    if harms:
        h = [np.zeros(shape) for _ in range(ndim + 2)]  # rho, v[ndim], p
        for hm in harms:
            prof = np.asarray(hm["profile"], dtype=np.float64)
            bshape = [1] * ndim
            bshape[ndim - 1 - hm["axis"]] = dims[hm["axis"]]
            a = prof.reshape(bshape)
            base = [hm["rho"]] + vel(hm) + [hm["p"]]
            for c in range(ndim + 2):
                h[c] = h[c] + a * base[c]
        up = prim_to_cons(gamma, h[0], np.array(h[1:ndim + 1]).reshape([ndim] + shape), h[-1])
        cons = [c + u for c, u in zip(cons, up)]
    cells = np.zeros(box_shape(ndim, dims, ncmp), dtype=np.float64)
    sl = tuple(slice(2, 2 + n) for n in dims[::-1])
    for c, arr in enumerate(cons):
        cells[(c,) + sl] = arr
    return cells


def spheres_2d_drawing(n=1024, seed=0, nspheres=16, amp=0.05, lam=32.0):
    """The 2D spheres workload as a drawing: canvas rho = p = 1 at rest;
    nspheres sharp spheres at seeded centres U[0, n)^2 (cell units), radius
    U[8, 32] cells, rho and p U[0.1, 10], later over earlier; then a harmonic
    perturbation of p along x, amplitude amp, wavelength lam cells (the
    reference's order: shapes, then perturbations)."""
    g = rng(seed)
    shapes = []
    for _ in range(nspheres):
        cx, cy = g.uniform(0.0, n, size=2)
        rad = g.uniform(8, 32)
        r_s, p_s = g.uniform(0.1, 10.0, size=2)
        shapes.append(dict(kind=SPHERE, centre=[float(cx), float(cy), 0.0], radius=float(rad),
                           rho=float(r_s), v=[0.0, 0.0], p=float(p_s)))
    harm = dict(axis=0, profile=harmonic_profile(n, amp, lam), rho=0.0, v=[0.0, 0.0], p=1.0)
    return dict(rho=1.0, v=[0.0, 0.0], p=1.0, shapes=shapes, harmonics=[harm])


def spheres_2d(n=1024, seed=0, gamma=5.0 / 3.0, nspheres=16, amp=0.05, lam=32.0, ncmp=None):
    """2D spheres workload (host array of spheres_2d_drawing)."""
    return draw(spheres_2d_drawing(n, seed, nspheres, amp, lam), 2, [n, n], gamma, ncmp)


def blast_3d(n=256, seed=0, gamma=5.0 / 3.0, radius=3.0, jitter=0.01, p0=1e-5, dims=None,
             ncmp=None):
    """3D point blast: rho = 1 + jitter*U[-1, 1], p = p0, at rest; unit energy
    deposited as pressure in the cells whose centre lies within `radius`
    cells of the box centre: p = (gamma-1) * 1 / (N_in * dx^3)."""
    dims = list(dims) if dims is not None else [n, n, n]
    g = rng(seed)
    ncmp = ncmp or ncmp_for(3)
    nz, ny, nx = dims[2], dims[1], dims[0]
    rho = 1.0 + jitter * g.uniform(-1.0, 1.0, size=(nz, ny, nx))
    k = np.arange(nz)[:, None, None] + 0.5 - nz / 2.0
    j = np.arange(ny)[None, :, None] + 0.5 - ny / 2.0
    i = np.arange(nx)[None, None, :] + 0.5 - nx / 2.0
    inside = (i * i + j * j + k * k) <= radius * radius
    n_in = int(inside.sum())
    dx = 1.0 / max(dims)
    p = np.full((nz, ny, nx), p0)
    p[inside] = (gamma - 1.0) * 1.0 / (n_in * dx ** 3)
    v = np.zeros((3, nz, ny, nx))
    return _fill(3, dims, ncmp, gamma, rho, v, p)


def blast_2d(dims, seed=0, gamma=5.0 / 3.0, radius=3.0, jitter=0.01, p0=1e-5, ncmp=None):
    """The 2D analogue of blast_3d: unit energy as pressure in the cells whose
    centre lies within `radius` cells of the box centre, p = (gamma-1) /
    (N_in * dx^2), over rho = 1 + jitter*U[-1, 1], p = p0, at rest."""
    g = rng(seed)
    ncmp = ncmp or ncmp_for(2)
    ny, nx = dims[1], dims[0]
    rho = 1.0 + jitter * g.uniform(-1.0, 1.0, size=(ny, nx))
    j = np.arange(ny)[:, None] + 0.5 - ny / 2.0
    i = np.arange(nx)[None, :] + 0.5 - nx / 2.0
    inside = (i * i + j * j) <= radius * radius
    n_in = int(inside.sum())
    dx = 1.0 / max(dims)
    p = np.full((ny, nx), p0)
    p[inside] = (gamma - 1.0) * 1.0 / (n_in * dx ** 2)
    v = np.zeros((2, ny, nx))
    return _fill(2, dims, ncmp, gamma, rho, v, p)


def blast(ndim, dims, seed=0, gamma=5.0 / 3.0, ncmp=None):
    if ndim == 3:
        return blast_3d(dims=dims, seed=seed, gamma=gamma, ncmp=ncmp)
    if ndim == 2:
        return blast_2d(dims, seed=seed, gamma=gamma, ncmp=ncmp)
    raise ValueError("blast needs ndim 2 or 3")


def random_two_state(ndim, dims, seed=0, gamma=5.0 / 3.0, ncmp=None, vmax=0.0):
    """Seeded two-state problem along a random axis: rho, p ~ U[0.1, 1] per side,
    optional velocities U[-vmax, vmax] per side."""
    g = rng(seed)
    ncmp = ncmp or ncmp_for(ndim)
    axis = int(g.integers(0, ndim))
    rl, pl, rr, pr = g.uniform(0.1, 1.0, size=4)
    vl = g.uniform(-vmax, vmax, size=ndim) if vmax else np.zeros(ndim)
    vr = g.uniform(-vmax, vmax, size=ndim) if vmax else np.zeros(ndim)
    shape = list(dims)[::-1]
    idx = np.indices(shape)[::-1][axis]  # index along `axis`
    left = idx < dims[axis] // 2
    rho = np.where(left, rl, rr).astype(np.float64)
    p = np.where(left, pl, pr).astype(np.float64)
    v = np.stack([np.where(left, vl[d], vr[d]) for d in range(ndim)]).astype(np.float64)
    return _fill(ndim, dims, ncmp, gamma, rho, v, p)


def random_smooth(ndim, dims, seed=0, gamma=5.0 / 3.0, ncmp=None, amp=0.3, vamp=0.5):
    """Seeded smooth-plus-noise state exercising every branch of the limiters
    and both shock and rarefaction waves (used by the parity tests)."""
    g = rng(seed)
    ncmp = ncmp or ncmp_for(ndim)
    shape = list(dims)[::-1]
    rho = 1.0 + amp * g.uniform(-1, 1, size=shape)
    p = 1.0 + amp * g.uniform(-1, 1, size=shape)
    v = vamp * g.uniform(-1, 1, size=[ndim] + shape)
    return _fill(ndim, dims, ncmp, gamma, rho, v, p)


def amr_boxes(seed=0, nboxes=200, base=128, lo=8, hi=64, gamma=5.0 / 3.0, ndim=3):
    """Batched-AMR workload: one base box of base^ndim cells (dx0 = 1/base)
    plus `nboxes` refined boxes (dx = dx0/2) whose per-axis extents are drawn
    uniformly from [lo, hi]; each refined box gets a seeded two-state or
    blast initial condition, the base box a blast.  Returns a list of
    (dims, dx, cells)."""
    g = rng(seed)
    out = []
    dx0 = 1.0 / base
    if base:
        dims = [base] * ndim
        if ndim == 3:
            cells = blast_3d(dims=dims, seed=int(g.integers(1 << 30)), gamma=gamma)
        else:
            cells = random_two_state(ndim, dims, seed=int(g.integers(1 << 30)), gamma=gamma)
        out.append((dims, [dx0] * ndim, cells))
    for _ in range(nboxes):
        dims = [int(x) for x in g.integers(lo, hi + 1, size=ndim)]
        s = int(g.integers(1 << 30))
        if ndim == 3 and g.uniform() < 0.5:
            cells = blast_3d(dims=dims, seed=s, gamma=gamma)
        else:
            cells = random_two_state(ndim, dims, seed=s, gamma=gamma)
        out.append((dims, [dx0 / 2.0] * ndim, cells))
    return out


# --- initial conditions painted on the device (rhy_ic_t) --------------------
# Descriptors of the seeded states that rhy_box_paint / rhy_pool_add_ic
# write on the device, and `paint`, the same arithmetic on the host, so that
# a painted state can be checked value for value and a job can be run from
# either.

IC_TWO_STATE, IC_BLAST = 1, 2
_MASK64 = (1 << 64) - 1


def splitmix64(seed, counter):
    """SplitMix64 (Steele, Lea, Flood 2014; constants of Vigna's
    splitmix64.c): draw number `counter` (0-based, uint64 array) of stream
    `seed`, in numpy uint64 arithmetic (wrapping mod 2^64)."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK64) + (c + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def ic_two_state(ndim, seed, gamma=5.0 / 3.0):
    """The two-state descriptor of random_two_state (same draws: axis, then
    rho_l, p_l, rho_r, p_r ~ U[0.1, 1]); painting it gives random_two_state's
    array."""
    g = rng(seed)
    axis = int(g.integers(0, ndim))
    rl, pl, rr, pr = (float(x) for x in g.uniform(0.1, 1.0, size=4))
    return dict(kind=IC_TWO_STATE, axis=axis, gamma=gamma, rho_l=rl, p_l=pl, rho_r=rr, p_r=pr)


def ic_blast(seed, gamma=5.0 / 3.0, radius=3.0, jitter=0.01, p0=1e-5, energy=1.0):
    """A blast over a jittered medium: rho = 1 + jitter*(-1 + 2u) with u the
    SplitMix64 draw of the cell's C-order index, p = p0, and `energy`
    deposited as pressure in the cells within `radius` cells of the centre."""
    return dict(kind=IC_BLAST, axis=0, gamma=gamma, radius=radius, jitter=jitter, p0=p0,
                energy=energy, seed=int(seed) & _MASK64)


def _centre_offsets(n):
    return (np.arange(n) + 0.5) - n / 2.0


def paint(ic, ndim, dims, ncmp=None):
    """Host array of the state rhy_box_paint writes for descriptor `ic`."""
    ncmp = ncmp or ncmp_for(ndim)
    dims = list(dims)[:ndim]
    shape = dims[::-1]  # C order (z, y, x)
    gamma = ic["gamma"]
    if ic["kind"] == IC_TWO_STATE:
        axis = ic["axis"]
        idx = np.indices(shape)[::-1][axis]
        left = idx < dims[axis] // 2
        rho = np.where(left, ic["rho_l"], ic["rho_r"]).astype(np.float64)
        p = np.where(left, ic["p_l"], ic["p_r"]).astype(np.float64)
    elif ic["kind"] == IC_BLAST:
        n = int(np.prod(dims))
        u = (splitmix64(ic["seed"], np.arange(n, dtype=np.uint64)) >> np.uint64(11)).astype(
            np.float64) * 2.0 ** -53
        rho = (1.0 + ic["jitter"] * (-1.0 + 2.0 * u)).reshape(shape)
        offs = [_centre_offsets(d) for d in dims]
        grids = np.meshgrid(*offs[::-1], indexing="ij")[::-1]  # x, y, z arrays
        rr = grids[0] * grids[0]
        for d in range(1, ndim):
            rr = rr + grids[d] * grids[d]
        r2 = ic["radius"] * ic["radius"]
        inside = rr <= r2
        n_in = int(inside.sum())
        dx = 1.0 / max(dims)
        p = np.full(shape, ic["p0"])
        p[inside] = (gamma - 1.0) * ic["energy"] / (n_in * dx ** ndim)
    else:
        raise ValueError("unknown initial condition kind %r" % ic["kind"])
    v = np.zeros([ndim] + shape)
    return _fill(ndim, dims, ncmp, gamma, rho, v, p)
