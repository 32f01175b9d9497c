"""The library's documented build options, each built from the same sources
(rhyme_amd/build.py VARIANTS) and run through the C ABI against the oracle:

  powlib       RHY_FP64_POW_FAST=0: device libm pow in the FP64 path
  nocontact    RHY_CONTACT_EXACT=0: the Newton iteration on contacts too
  noldswalk    RHY_LDS_WALK=0: the y/z flux walks re-read their cells
  fastriemann  RHY_FP64_FAST_RIEMANN=1: the mixed form's FP64 Riemann policy

Each one passes the FP64 trajectory subset of the parity tests at the same
1e-12 relative L1, and the options that must not change any result
(noldswalk: where the walks stage their cells) are bit-identical to the
default build, single boxes and batches, FP64 and mixed.
"""
import os

import numpy as np
import pytest

import oracle
from rhyme_amd import _lib, build, hydro, ics
from rhyme_amd.batch import SamrBatch

pytestmark = pytest.mark.gpu

REL_L1 = 1e-12
EXPECT = {
    "powlib": _lib.OPT_CONTACT_EXACT | _lib.OPT_LDS_WALK,
    "nocontact": _lib.OPT_POW_FAST | _lib.OPT_LDS_WALK,
    "noldswalk": _lib.OPT_POW_FAST | _lib.OPT_CONTACT_EXACT,
    "fastriemann": _lib.OPT_POW_FAST | _lib.OPT_CONTACT_EXACT | _lib.OPT_LDS_WALK |
                   _lib.OPT_FAST_RIEMANN,
}


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)


def interior(ndim, a):
    sl = tuple(slice(2, -2) for _ in range(ndim))
    return a[(slice(0, ndim + 2),) + sl]


def params(ndim, limiter, bc, courant=0.8, som=hydro.MONATOMIC):
    return hydro.make_params(
        ndim, thermo=hydro.Thermo(state_of_matter=som),
        rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), cfl=hydro.Cfl(courant),
        bc=hydro.SamrBc(types=bc))


# (name, ndim, dims, state, params, dx, steps): the trajectory subset
def cases():
    n3 = 24
    return [
        ("1D two-state", 1, [1024], ics.two_state_1d(1024, seed=1)[0],
         params(1, hydro.VAN_LEER, [hydro.OUTFLOW] * 6), [1.0 / 1024], 100),
        ("2D spheres periodic", 2, [64, 64], ics.spheres_2d(64, seed=2, nspheres=4),
         params(2, hydro.MINMOD, [hydro.PERIODIC] * 6, courant=0.2), [1.0 / 64] * 2, 20),
        ("3D blast", 3, [n3] * 3, ics.blast_3d(dims=[n3] * 3, seed=3),
         params(3, hydro.SUPERBEE, [hydro.REFLECTIVE] * 6, courant=0.2), [1.0 / n3] * 3, 12),
        ("3D smooth polyatomic", 3, [12, 10, 14],
         ics.random_smooth(3, [12, 10, 14], seed=5, gamma=4.0 / 3.0),
         params(3, hydro.VAN_ALBADA, [1, 2, 3, 3, 2, 1], courant=0.4, som=hydro.POLYATOMIC),
         [1.0 / 16] * 3, 10),
    ]


def run_box(ndim, dims, cells, p, dx, steps, precision=hydro.FP64):
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    dts = hydro.run(p, box, dx, steps)
    out = box.download().copy()
    box.close()
    return out, dts


ORACLE_CACHE = {}


def oracle_run(name, ndim, dims, cells, p, dx, steps):
    if name not in ORACLE_CACHE:
        ref = cells.copy()
        dts = oracle.run(oracle.params_from(p), dims, ref, dx, steps)
        ORACLE_CACHE[name] = (ref, dts)
    return ORACLE_CACHE[name]


@pytest.mark.parametrize("variant", sorted(build.VARIANTS))
def test_variant_trajectories_against_oracle(variant):
    path = build.variant_path(variant)
    assert os.path.exists(path), "variant library missing: run __graft_entry__.build()"
    for name, ndim, dims, cells, p, dx, steps in cases():
        ref, rdts = oracle_run(name, ndim, dims, cells, p, dx, steps)
        with _lib.using(path):
            assert hydro.build_options() == EXPECT[variant]
            out, dts = run_box(ndim, dims, cells, p, dx, steps)
        a, b = interior(ndim, out), interior(ndim, ref)
        errs = [rel_l1(a[c], b[c]) for c in range(ndim + 2)]
        say(variant, name, "max rel L1", max(errs))
        assert max(errs) <= REL_L1
        assert np.max(np.abs(dts - rdts) / rdts) <= 1e-12


def solve_box(ndim, dims, cells, p, dx, precision):
    """One rhy_mh_solve (the flux-plane walks) after a ghost fill."""
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    hydro.rhyme_samr_bc_set_boundaries(p, box)
    dt = hydro.rhyme_cfl_time_step(p, box, dx)
    hydro.rhyme_muscl_hancock_solve(p, box, dx, dt)
    out = box.download().copy()
    box.close()
    return out


def test_noldswalk_bit_identical_to_default():
    """Where the y/z walks stage their cells (LDS ring or re-reads) changes no
    bit: single boxes (rhy_run and rhy_mh_solve, FP64 and mixed) and a batch
    of uneven boxes."""
    path = build.variant_path("noldswalk")
    for name, ndim, dims, cells, p, dx, steps in cases()[1:3]:
        for prec in (hydro.FP64, hydro.MIXED):
            base, bd = run_box(ndim, dims, cells, p, dx, steps, prec)
            bs = solve_box(ndim, dims, cells, p, dx, prec)
            with _lib.using(path):
                var, vd = run_box(ndim, dims, cells, p, dx, steps, prec)
                vs = solve_box(ndim, dims, cells, p, dx, prec)
            say("noldswalk", name, "precision", prec)
            assert np.array_equal(base, var) and np.array_equal(bd, vd)
            assert np.array_equal(bs, vs)
    boxes = ics.amr_boxes(seed=4, nboxes=12, base=24, lo=8, hi=20)
    p = hydro.make_params(3, rp=hydro.RiemannProblem(irs=hydro.Irs(100, 1e-6)),
                          sl=hydro.SlopeLimiter(type=hydro.VAN_ALBADA), cfl=hydro.Cfl(0.2),
                          bc=hydro.SamrBc(types=[hydro.OUTFLOW] * 6))

    def batch_run():
        out = []
        for prec in (hydro.FP64, hydro.MIXED):
            b = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=prec)
            for i, (_, _, c) in enumerate(boxes):
                b.upload(i, c)
            b.run(p, 5, graph=False)
            out.append([b.download(i).copy() for i in range(b.n)])
            b.close()
        return out

    base = batch_run()
    with _lib.using(path):
        var = batch_run()
    for pb, pv in zip(base, var):
        for x, y in zip(pb, pv):
            assert np.array_equal(x, y)


def test_powlib_riemann_closer_to_libm_host():
    """With the device's libm pow the Riemann states of random records stay
    within the same 1e-10 of the oracle (glibc pow)."""
    g = np.random.default_rng(9)
    p = hydro.make_params(3, rp=hydro.RiemannProblem(irs=hydro.Irs(100, 1e-6)))
    n = 5000
    rho = np.exp(g.uniform(-2, 2, (2, n)))
    pr = np.exp(g.uniform(-2, 2, (2, n)))
    v = g.uniform(-2, 2, (2, 3, n))
    l = np.stack(ics.prim_to_cons(p.gamma, rho[0], v[0], pr[0]), 1)
    r = np.stack(ics.prim_to_cons(p.gamma, rho[1], v[1], pr[1]), 1)
    axis = g.integers(1, 4, size=n).astype(np.int32)
    ref = oracle.rp_solve_batch(oracle.params_from(p), l, r, axis, 1.0)
    fracs = {}
    for name, path in (("default", None), ("powlib", build.variant_path("powlib"))):
        if path is None:
            u = hydro.riemann_batch(p, l, r, axis)
        else:
            with _lib.using(path):
                u = hydro.riemann_batch(p, l, r, axis)
        rel = np.abs(u - ref) / np.maximum(np.abs(ref), 1e-300)
        fracs[name] = float(np.mean(u == ref))
        say(name, "max rel", rel.max(), "bit-identical fraction", fracs[name])
        assert rel.max() < 1e-10
