"""Snapshot payloads on the device (rhy_*_snapshot_pack / _unpack) against
the restatement of the reference's Chombo level writer and restart reader
(oracle/chombo.py): bit for bit, FP64 and mixed, single boxes, batches,
pools, and the 256^3 blast after a few steps."""
import numpy as np
import pytest

from oracle import chombo as och
from rhyme_amd import chombo, hydro, ics
from rhyme_amd.batch import SamrBatch, SamrPool, box_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def params(ndim, bc=hydro.OUTFLOW):
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=100, tolerance=1e-6)),
        sl=hydro.SlopeLimiter(type=hydro.VAN_LEER), cfl=hydro.Cfl(0.3),
        bc=hydro.SamrBc(types=[bc] * 6))


def stored(cells, precision):
    """What the device holds after an upload: FP32-rounded in the mixed form."""
    return cells.astype(np.float32).astype(np.float64) if precision == hydro.MIXED else cells


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim,dims", [(1, [37]), (2, [19, 12]), (3, [9, 7, 11])])
def test_box_pack_unpack(ndim, dims, precision):
    say("box snapshot", ndim, dims, "precision", precision)
    ncmp = hydro.ncmp_for(ndim)
    cells = ics.random_smooth(ndim, dims, seed=ndim)
    cells[ncmp - 1] = np.random.default_rng(9).uniform(0, 1, cells.shape[1:])
    box = hydro.SamrBox(ndim, dims, cells=cells.copy(), precision=precision)
    box.upload()
    data = box.snapshot()
    want = och.level_data(ndim, [cells], ncmp)
    assert data.dtype == np.float32 and data.size == want.size
    np.testing.assert_array_equal(data.view(np.uint32), want.view(np.uint32))
    # the reader: interiors from a payload, ghosts left alone
    g = np.random.default_rng(1)
    payload = g.uniform(-2, 2, size=data.size).astype(np.float32)
    box.load_snapshot(payload)
    out = box.download().copy()
    inner = (slice(None),) + (slice(2, -2),) * ndim
    np.testing.assert_array_equal(out[inner], och.load_level(ndim, ncmp, [dims], payload)[0])
    mask = np.ones(out.shape[1:], dtype=bool)
    mask[inner[1:]] = False
    np.testing.assert_array_equal(out[:, mask], stored(cells, precision)[:, mask])
    np.testing.assert_array_equal(box.snapshot(), payload)
    with pytest.raises(ValueError):
        box.load_snapshot(payload[:-1])
    box.close()


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
def test_level_pack_batch_and_pool(precision):
    boxes = ics.amr_boxes(seed=4, nboxes=7, base=20, lo=4, hi=14)
    ncmp = hydro.ncmp_for(3)
    b = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=precision)
    for i, (_, _, c) in enumerate(boxes):
        b.upload(i, c)
    p = params(3)
    b.run(p, 3)
    evolved = [b.download(i).copy() for i in range(len(boxes))]
    data = b.snapshot()
    np.testing.assert_array_equal(data, och.level_data(3, evolved, ncmp))
    # restart into a fresh batch: same payload, and the same next steps as the original
    b2 = SamrBatch(3, [d for d, _, _ in boxes], [x for _, x, _ in boxes], precision=precision)
    for i in range(len(boxes)):
        b2.upload(i, np.zeros_like(evolved[i]))
    b2.load_snapshot(data)
    np.testing.assert_array_equal(b2.snapshot(), data)
    if precision == hydro.MIXED:       # the payload is the mixed state itself
        d1, d2 = b.run(p, 2), b2.run(p, 2)
        np.testing.assert_array_equal(d1, d2)
        for i in range(len(boxes)):
            np.testing.assert_array_equal(b.download(i), b2.download(i))
    b.close()
    b2.close()
    # a pool: resident boxes in slot order, a removed box leaves the level
    budget = sum(box_bytes(3, d, precision=precision) for d, _, _ in boxes) + (1 << 22)
    pool = SamrPool(3, budget, precision=precision)
    slots = [pool.add(d, x, c, hydro.MINMOD) for d, x, c in boxes]
    pool.remove(slots[2])
    kept = [c for i, (_, _, c) in enumerate(boxes) if i != 2]
    want = och.level_data(3, [stored(c, precision) for c in kept], ncmp)
    np.testing.assert_array_equal(pool.snapshot(), want)
    pool.close()


def test_blast256_snapshot_round_trip():
    """Full size: the 256^3 blast after 5 FP64 steps; the payload equals the
    FP32 conversion of the downloaded state, and a restart from it packs to
    the same payload."""
    n = 256
    box = hydro.SamrBox(3, [n] * 3, cells=ics.blast_3d(n=n, seed=0))
    box.upload()
    p = hydro.make_params(3, sl=hydro.SlopeLimiter(type=hydro.SUPERBEE), cfl=hydro.Cfl(0.2),
                          bc=hydro.SamrBc(types=[hydro.REFLECTIVE] * 6))
    hydro.run(p, box, [1.0 / n] * 3, 5)
    data = box.snapshot()
    state = box.download().copy()
    np.testing.assert_array_equal(data, och.level_data(3, [state], box.ncmp))
    box2 = hydro.SamrBox(3, [n] * 3)
    box2.upload()
    box2.load_snapshot(data)
    np.testing.assert_array_equal(box2.snapshot(), data)
    table = chombo.boxes_table(3, [[1, 1, 1]], [[n] * 3])
    (left, dims, part), = chombo.split_level(3, table, data)
    assert left == [1, 1, 1] and dims == [n] * 3 and part.size == data.size
    say("256^3 payload", data.nbytes / 2 ** 20, "MiB")
    box.close()
    box2.close()
