"""Deferred faces.  The flux kernels leave the rare slow paths of the
Riemann solve (libm pow outside its fast domain, the Newton restarts from
guesses 2-6) to a fixup kernel that recomputes those faces with the full
solver in the same step.

* A host-side test hook (rhy_debug_defer) also sends every k-th face down
  that path, or caps the list so that it overflows and the fixup recomputes
  every face; results must be bit-identical to the normal run (single boxes
  in 1D/2D/3D, batches; FP64 and mixed).
* Strong collisions with large density contrasts make the Newton iteration
  from guess 1 overshoot to p* < 0 (rhyme_irs_exec.f90:24-42 then restarts
  from the next guess): such a box defers faces by itself and must still
  match the oracle, and give the same state inside a batch as alone.
"""
import numpy as np
import pytest

import oracle
from rhyme_amd import _lib, hydro, ics
from rhyme_amd.batch import SamrBatch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def params(ndim, limiter, bc, courant=0.2):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), cfl=hydro.Cfl(courant),
        bc=hydro.SamrBc(types=[bc] * 6))


def run_box(ndim, dims, cells, p, nsteps, precision):
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    dts = hydro.run(p, box, [1.0 / dims[0]] * ndim, nsteps)
    n = box.last_deferred()
    out = box.download().copy()
    box.close()
    return out, dts, n


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim,dims", [(1, [64]), (2, [40, 36]), (3, [20, 18, 22])])
def test_forced_deferral_bit_identical(ndim, dims, precision):
    cells = ics.random_smooth(ndim, dims, seed=40 + ndim)
    p = params(ndim, hydro.VAN_LEER, hydro.OUTFLOW)
    ref, rdts, n0 = run_box(ndim, dims, cells, p, 6, precision)
    say("ndim", ndim, "precision", precision, "deferred normally", n0)
    with _lib.debug_defer(7):
        out, dts, n1 = run_box(ndim, dims, cells, p, 6, precision)
    say("  every 7th face:", n1)
    assert n1 > n0
    np.testing.assert_array_equal(dts, rdts)
    np.testing.assert_array_equal(out, ref)
    with _lib.debug_defer(3, cap=16):
        out, dts, n2 = run_box(ndim, dims, cells, p, 6, precision)
    say("  every 3rd face, list of 16:", n2)
    assert n2 > 16
    np.testing.assert_array_equal(dts, rdts)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_forced_deferral_batch_bit_identical(precision):
    boxes = ics.amr_boxes(seed=5, nboxes=9, base=24, lo=6, hi=20)
    p = params(3, hydro.VAN_ALBADA, hydro.OUTFLOW)

    def go():
        b = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=precision)
        for i, (_, _, c) in enumerate(boxes):
            b.upload(i, c)
        dts = b.run(p, 6, graph=True)
        n = b.last_deferred()
        outs = [b.download(i).copy() for i in range(len(boxes))]
        b.close()
        return dts, outs, n

    rdts, ref, n0 = go()
    for every, cap in ((5, -1), (2, 8)):
        with _lib.debug_defer(every, cap):
            dts, outs, n = go()
        say("batch precision", precision, "every", every, "cap", cap, "deferred", n, "normally", n0)
        assert n > n0
        np.testing.assert_array_equal(dts, rdts)
        for a, b in zip(outs, ref):
            np.testing.assert_array_equal(a, b)


# A collision with a 1:830 density contrast: Newton from the two-shock guess
# overshoots to p* < 0 on the interface, and the solve restarts from guess 2.
COLLISION = dict(rho_l=0.0190126, u_l=5.35396, p_l=0.34856,
                 rho_r=15.7988, u_r=-4.9701, p_r=0.0333912)


def collision_1d(n, gamma=5.0 / 3.0):
    c = COLLISION
    left = np.arange(n) < n // 2
    rho = np.where(left, c["rho_l"], c["rho_r"])
    u = np.where(left, c["u_l"], c["u_r"])[None, :]
    p = np.where(left, c["p_l"], c["p_r"])
    return ics._fill(1, [n], hydro.ncmp_for(1), gamma, rho, u, p)


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_natural_deferral_matches_oracle(precision):
    n = 64
    cells = collision_1d(n)
    p = params(1, hydro.MINMOD, hydro.OUTFLOW)
    _, _, n1 = run_box(1, [n], cells, p, 1, precision)
    say("collision: faces deferred on the first step", n1)
    assert n1 >= 1
    out, dts, _ = run_box(1, [n], cells, p, 20, precision)
    ref = cells.copy()
    rdts = oracle.run(oracle.params_from(p), [n], ref, [1.0 / n], 20)
    tol = 1e-12 if precision == hydro.FP64 else 1e-5
    for c in range(3):
        a, b = out[c, 2:-2], ref[c, 2:-2]
        e = np.abs(a - b).sum() / np.abs(b).sum()
        say("  var", c, "rel L1", e)
        assert e <= tol
    assert np.max(np.abs(dts - rdts) / rdts) <= tol


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_natural_deferral_in_batch_identical_to_alone(precision):
    """Boxes whose faces defer by themselves, advanced in one batch: the
    batch's fixup kernel gives every box the state and dts of the box run
    alone (whose fixup is the single-box kernel)."""
    boxes = [([64], [1.0 / 64], collision_1d(64)), ([48], [1.0 / 48], collision_1d(48)),
             ([40], [1.0 / 40], ics.random_smooth(1, [40], seed=3))]
    p = params(1, hydro.MINMOD, hydro.OUTFLOW)
    b = SamrBatch(1, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=precision)
    for i, (_, _, c) in enumerate(boxes):
        b.upload(i, c)
    b.run(p, 1, graph=True)
    n1 = b.last_deferred()
    say("batch precision", precision, "faces deferred on the first step", n1)
    assert n1 >= 2
    for i, (_, _, c) in enumerate(boxes):
        b.upload(i, c)
    dts = b.run(p, 12, graph=True)
    for i, (dims, dx, c) in enumerate(boxes):
        box = hydro.SamrBox(1, dims, cells=c.copy(), precision=precision)
        box.upload()
        adts = hydro.run(p, box, dx, 12)
        np.testing.assert_array_equal(dts[:, i], adts)
        np.testing.assert_array_equal(b.download(i), box.download())
        box.close()
    b.close()
