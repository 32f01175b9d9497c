"""The batched form (rhyme_amd.batch): many uneven boxes advanced by one
launch per phase, each box with its own ghost fill and dt, checked box by
box against the CPU oracle run on that box alone; hipGraph replays against
eager runs; every tile shape giving bit-identical results."""
import numpy as np
import pytest

import oracle
from rhyme_amd import autotune, hydro, ics
from rhyme_amd.batch import SamrBatch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def params(ndim, limiter=hydro.VAN_ALBADA, solver=hydro.MEMORY_INTENSIVE, courant=0.2):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(100, 1e-6)),
        sl=hydro.SlopeLimiter(type=limiter), mh=hydro.MusclHancock(solver_type=solver),
        cfl=hydro.Cfl(courant), bc=hydro.SamrBc(types=[hydro.OUTFLOW] * 6))


def small_set(ndim, seed, n=9):
    return ics.amr_boxes(seed=seed, nboxes=n, base=24 if ndim == 3 else 48, lo=6, hi=20,
                         ndim=ndim)


def inner(ndim, a):
    return a[(slice(0, ndim + 2),) + tuple(slice(2, -2) for _ in range(ndim))]


def rel_l1(a, b):
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)


def make_batch(ndim, boxes, precision=hydro.FP64):
    b = SamrBatch(ndim, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=precision)
    for i, (_, _, c) in enumerate(boxes):
        b.upload(i, c)
    return b


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("solver", [hydro.CPU_INTENSIVE, hydro.MEMORY_INTENSIVE])
def test_batch_boxes_match_oracle(ndim, solver):
    say("batch vs oracle ndim", ndim, "solver", solver)
    boxes = small_set(ndim, 11 + ndim)
    p = params(ndim, solver=solver)
    b = make_batch(ndim, boxes)
    nsteps = 8
    dts = b.run(p, nsteps, graph=True)
    op = oracle.params_from(p)
    for i, (dims, dx, cells) in enumerate(boxes):
        ref = cells.copy()
        rdts = oracle.run(op, dims, ref, dx, nsteps)
        out = b.download(i)
        for c in range(ndim + 2):
            e = rel_l1(inner(ndim, out)[c], inner(ndim, ref)[c])
            assert e <= 1e-12, (i, dims, c, e)
        assert np.max(np.abs(dts[:, i] - rdts) / rdts) <= 1e-12


def test_batch_graph_replays_match_eager():
    """The captured step replayed (with downloads/uploads and syncs between
    replays) gives exactly the eager launches' states after every step."""
    say("batch graph replays")
    boxes = small_set(3, 21)
    p = params(3)
    eager = make_batch(3, boxes)
    graph = make_batch(3, boxes)
    for step in range(4):
        eager.run(p, 1, graph=False)
        graph.run(p, 1, graph=True)
        for i in range(len(boxes)):
            a = graph.download(i)
            assert np.array_equal(a, eager.download(i)), (step, i)
            graph.upload(i, a)  # eager work between replays


def test_batch_tile_shapes_bit_identical():
    say("batch tile shapes")
    boxes = small_set(3, 31, n=6)
    p = params(3, limiter=hydro.SUPERBEE)
    ref = None
    for cand in autotune.CANDIDATES:
        b = make_batch(3, boxes)
        b.set_tiles(*cand)
        b.run(p, 3, graph=True)
        states = [b.download(i) for i in range(len(boxes))]
        if ref is None:
            ref = states
        else:
            for i in range(len(boxes)):
                assert np.array_equal(states[i], ref[i]), (cand, i)
        b.close()


def test_batch_mixed_vs_oracle():
    say("batch mixed")
    boxes = small_set(3, 41)
    p = params(3)
    b = make_batch(3, boxes, precision=hydro.MIXED)
    b.run(p, 8, graph=True)
    op = oracle.params_from(p)
    for i, (dims, dx, cells) in enumerate(boxes):
        ref = cells.copy()
        oracle.run(op, dims, ref, dx, 8)
        out = b.download(i)
        for c in range(5):
            assert rel_l1(inner(3, out)[c], inner(3, ref)[c]) <= 1e-5, (i, c)


def test_autotune_picks_and_caches(tmp_path, monkeypatch):
    say("autotune")
    monkeypatch.setenv("RHYME_AMD_TILE_CACHE", str(tmp_path / "tiles.json"))
    boxes = small_set(3, 51, n=5)
    p = params(3)
    b = make_batch(3, boxes)
    best, timings = autotune.tune(b, p, [c for _, _, c in boxes], steps=2,
                                  candidates=autotune.CANDIDATES[:3])
    assert tuple(best) in autotune.CANDIDATES[:3] and len(timings) == 3
    assert autotune.lookup(b, p)["tiles"] == list(best)
    # the tuned batch still gives the oracle's answer
    b.run(p, 3, graph=True)
    dims, dx, cells = boxes[1]
    ref = cells.copy()
    oracle.run(oracle.params_from(p), dims, ref, dx, 3)
    assert rel_l1(inner(3, b.download(1))[0], inner(3, ref)[0]) <= 1e-12
