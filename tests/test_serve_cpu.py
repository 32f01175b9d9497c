"""Host logic of the serving loop, without a device: the workload generator
(sizes, halves, arrivals, determinism), first-come-first-served admission
under the byte budget, backfill admission and its bound on how long a
blocked job is passed, quanta that never overshoot a job's step count, and
the latency report.  The device pools are replaced by a stand-in that keeps
the same byte accounting (rhy_pool_box_bytes, a host-only call)."""
import numpy as np
import pytest

from rhyme_amd import batch, hydro, ics, serve


class FakePool:
    instances = []

    def __init__(self, ndim, budget_bytes, max_boxes=4096, ncmp=None, precision=0):
        self.ndim, self.budget, self.precision = ndim, budget_bytes, precision
        self.ncmp = ncmp or hydro.ncmp_for(ndim)
        self.boxes, self.used, self.runs = {}, 0, []
        FakePool.instances.append(self)

    def add(self, dims, dx, cells, limiter, bc_types=None):
        nb = batch.box_bytes(self.ndim, dims, self.ncmp, self.precision)
        if self.used + nb > self.budget:
            raise batch.PoolFull("full")
        slot = next(i for i in range(10 ** 6) if i not in self.boxes)
        self.boxes[slot] = (list(dims), nb)
        self.used += nb
        return slot

    def add_ic(self, dims, dx, ic, limiter, bc_types=None):
        return self.add(dims, dx, None, limiter, bc_types)

    def remove(self, slot):
        self.used -= self.boxes.pop(slot)[1]

    def download(self, slot):
        return np.zeros(hydro.box_shape(self.ndim, self.boxes[slot][0], self.ncmp))

    def run_async(self, params, nsteps, graph=True):
        self.runs.append((nsteps, sorted(self.boxes)))

    def sync(self):
        pass

    def set_tiles(self, *a):
        pass

    def last_dts(self, n):
        return np.zeros((n, 4096))

    def close(self):
        pass


@pytest.fixture
def fake_pools(monkeypatch):
    FakePool.instances = []
    monkeypatch.setattr(serve, "SamrPool", FakePool)
    return FakePool


def test_workload_shape_and_determinism():
    jobs = serve.serving_jobs(njobs=1000, rate=50.0, seed=7)
    again = serve.serving_jobs(njobs=1000, rate=50.0, seed=7)
    assert [(j.dims, j.limiter, j.seed, j.arrival) for j in jobs] == \
           [(j.dims, j.limiter, j.seed, j.arrival) for j in again]
    nd = np.array([j.ndim for j in jobs])
    assert (nd == 2).sum() == 500 and (nd == 3).sum() == 500
    n2 = np.array([j.dims[0] for j in jobs if j.ndim == 2])
    n3 = np.array([j.dims[0] for j in jobs if j.ndim == 3])
    assert n2.min() >= 64 and n2.max() <= 512 and n3.min() >= 32 and n3.max() <= 128
    # log-uniform: the median size is near the geometric mean of the range
    assert 150 < np.median(n2) < 210 and 55 < np.median(n3) < 75
    arr = np.array([j.arrival for j in jobs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert abs(arr[-1] / 999 - 1 / 50.0) < 0.2 / 50.0   # mean gap 1/rate
    assert set(j.limiter for j in jobs) == {1, 2, 3, 4}
    assert {j.kind for j in jobs} == {"blast", "two_state"}
    for j in jobs:
        assert j.bc_types == [hydro.REFLECTIVE if j.kind == "blast" else hydro.OUTFLOW] * 6
        assert j.dx == [1.0 / j.dims[0]] * j.ndim


def test_make_cells_shapes():
    for j in serve.serving_jobs(njobs=8, seed=1, n2d=(8, 16), n3d=(6, 10)):
        c = j.make_cells()
        assert c.shape == hydro.box_shape(j.ndim, j.dims, hydro.ncmp_for(j.ndim))
        assert np.all(c[0][(slice(2, -2),) * j.ndim] > 0)


def test_fifo_admission_under_budget(fake_pools):
    jobs = serve.serving_jobs(njobs=30, rate=1e9, seed=2, steps=7, n2d=(8, 32), n3d=(6, 16))
    sizes = [batch.box_bytes(j.ndim, j.dims) for j in jobs]
    budget = max(sizes) * 2
    srv = serve.HydroServer(budget, quantum=3, keep_results=False, admission="fifo")
    srv.serve(jobs)
    assert srv.peak_bytes <= budget and srv.deferrals > 0
    # admission order is arrival order
    order = sorted(jobs, key=lambda j: (j.admitted, j.jid))
    assert [j.jid for j in order] == sorted(j.jid for j in jobs)
    for j in jobs:
        assert j.done == j.steps == 7
        assert j.arrival <= j.admitted <= j.finished
    # no quantum ran past the smallest remaining step count of its pool
    for pool in FakePool.instances:
        for q, _ in pool.runs:
            assert 1 <= q <= 3


def test_backfill_passes_a_blocked_head_for_a_bounded_time(fake_pools):
    """Backfill admits arrived jobs that fit past a head job that does not,
    only while the head has been blocked for less than max_bypass_s (on the
    server's clock); every job still completes, within the budget."""
    def mk(jid, nd, n):
        return serve.Job(jid=jid, ndim=nd, dims=[n] * nd, dx=[1.0 / n] * nd,
                         limiter=hydro.MINMOD, bc_types=[hydro.OUTFLOW] * 6, steps=6,
                         arrival=0.0, kind="blast", seed=jid, ic=ics.ic_blast(jid))
    jobs = [mk(0, 3, 24), mk(1, 3, 24)] + [mk(2 + i, 2, 12) for i in range(40)]
    nb = [batch.box_bytes(j.ndim, j.dims) for j in jobs]
    budget = nb[0] + 3 * nb[2]
    t = [0.0]

    def clock():  # every reading of the clock advances it by 10 ms
        t[0] += 0.01
        return t[0]
    window = 0.3
    srv = serve.HydroServer(budget, quantum=3, keep_results=False, max_bypass_s=window)
    srv.serve(jobs, clock=clock, sleep=lambda s: None)
    head = jobs[1]
    assert head.blocked_since is not None and srv.bypassed > 0
    passed = [j for j in jobs[2:] if j.admitted < head.admitted]
    assert passed and all(j.admitted < head.blocked_since + window + 0.05 for j in passed)
    assert any(j.admitted > head.admitted for j in jobs[2:])  # the rest waited for it
    assert srv.peak_bytes <= budget
    assert all(j.done == j.steps for j in jobs)
    # FIFO on the same stream passes nothing
    t[0] = 0.0
    for j in jobs:
        j.admitted = j.finished = j.blocked_since = None
        j.done = 0
    fifo = serve.HydroServer(budget, quantum=3, keep_results=False, admission="fifo")
    fifo.serve(jobs, clock=clock, sleep=lambda s: None)
    assert fifo.bypassed == 0
    order = sorted(jobs, key=lambda j: (j.admitted, j.jid))
    assert [j.jid for j in order] == list(range(len(jobs)))


def test_quanta_bounded_when_nothing_waits(fake_pools):
    jobs = serve.serving_jobs(njobs=6, rate=1e9, seed=4, steps=10, n2d=(8, 16), n3d=(6, 8))
    srv = serve.HydroServer(1 << 30, quantum=3, keep_results=False)
    srv.serve(jobs)
    assert srv.deferrals == 0
    for pool in FakePool.instances:
        for q, _ in pool.runs:
            assert 1 <= q <= 3
    assert all(j.done == 10 for j in jobs)


def test_job_larger_than_budget_is_an_error(fake_pools):
    jobs = serve.serving_jobs(njobs=2, rate=1e9, seed=2, steps=3, n2d=(64, 64), n3d=(32, 32))
    srv = serve.HydroServer(1 << 16, quantum=3, keep_results=False)
    with pytest.raises(ValueError):
        srv.serve(jobs)


def test_latency_report_fields():
    jobs = serve.serving_jobs(njobs=4, rate=10.0, seed=0, steps=2, n2d=(8, 8), n3d=(8, 8))
    for i, j in enumerate(jobs):
        j.admitted = j.arrival + 0.01
        j.finished = j.arrival + 0.1 * (i + 1)
        j.done = j.steps
    rep = serve.latency_report(jobs)
    assert rep["jobs"] == 4 and rep["failed"] == 0
    assert np.isclose(rep["p50_ms"], 250.0) and rep["max_ms"] == pytest.approx(400.0)
    assert rep["zone_updates"] == sum(2 * 8 ** j.ndim for j in jobs)
    span = max(j.finished for j in jobs) - jobs[0].arrival
    assert rep["zone_updates_per_s"] == pytest.approx(rep["zone_updates"] / span)


def _splitmix_reference(seed, counter):
    """SplitMix64 in Python integers, independent of numpy's uint64 wrap."""
    m = (1 << 64) - 1
    z = (seed + (counter + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_splitmix64_mirror():
    for seed in (0, 1, 12345, (1 << 64) - 1, 0x9E3779B97F4A7C15):
        got = ics.splitmix64(seed, np.arange(50, dtype=np.uint64))
        assert [int(v) for v in got] == [_splitmix_reference(seed, c) for c in range(50)]
    # Vigna's splitmix64.c from state 0 gives 0xE220A8397B1DCDAF first
    assert int(ics.splitmix64(0, np.array([0], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("ndim,dims", [(1, [31]), (2, [12, 17]), (3, [9, 8, 10])])
def test_paint_two_state_equals_seeded_two_state(ndim, dims):
    for seed in range(4):
        np.testing.assert_array_equal(ics.paint(ics.ic_two_state(ndim, seed), ndim, dims),
                                      ics.random_two_state(ndim, dims, seed=seed))


def test_paint_blast_properties():
    dims = [20, 18, 16]
    ic = ics.ic_blast(99, radius=3.0, jitter=0.01, p0=1e-5, energy=1.0)
    c = ics.paint(ic, 3, dims)
    inner = (slice(2, -2),) * 3
    rho = c[0][inner]
    assert np.all(rho >= 0.99) and np.all(rho < 1.01) and abs(rho.mean() - 1.0) < 1e-3
    assert np.all(c[1][inner] == 0) and np.all(c[3][inner] == 0)
    p = c[4][inner] * (2.0 / 3.0)        # E = p / (gamma - 1) at rest
    hot = p > 1e-3
    # cells within 3 cells of the centre (10, 9, 8) -> offsets q + .5 - n/2
    k, j, i = np.indices(rho.shape)
    rr = (i + .5 - 10) ** 2 + (j + .5 - 9) ** 2 + (k + .5 - 8) ** 2
    assert np.array_equal(hot, rr <= 9.0)
    # unit energy deposited: sum(p_in * dx^3) / (gamma-1) == 1 with dx = 1/20
    assert abs((p[hot] * (1 / 20.0) ** 3).sum() / (2.0 / 3.0) - 1.0) < 1e-12
    ghosts = c.copy()
    ghosts[(slice(None),) + inner] = 0
    assert not ghosts.any()


def test_failed_jobs_are_reported_apart():
    jobs = serve.serving_jobs(njobs=4, rate=10.0, seed=0, steps=6, n2d=(8, 8), n3d=(8, 8))
    for i, j in enumerate(jobs):
        j.admitted = j.arrival
        j.finished = j.arrival + 0.1 * (i + 1)
        j.done = j.steps if i else 2
    jobs[0].failed = "NaN in components [0]"
    rep = serve.latency_report(jobs)
    assert rep["failed"] == 1
    assert rep["p50_ms"] == pytest.approx(300.0)     # of the three completed jobs
    assert rep["zone_updates"] == 2 * 8 ** jobs[0].ndim + sum(6 * 8 ** j.ndim for j in jobs[1:])


def test_unhealthy_rules():
    ok = dict(nan_count=[0] * 9, neg_rho=0, neg_e_tot=0)
    assert serve.HydroServer.unhealthy(ok) is None
    assert "NaN" in serve.HydroServer.unhealthy(dict(ok, nan_count=[0, 0, 3] + [0] * 6))
    assert "density" in serve.HydroServer.unhealthy(dict(ok, neg_rho=1))
    assert "energy" in serve.HydroServer.unhealthy(dict(ok, neg_e_tot=2))


class _FakeBatch:
    """Records what autotune.tune does to a batch (no device)."""

    def __init__(self, dims):
        self.ndim, self.precision, self.dims = 3, 0, dims
        self.uploads, self.tiles, self.runs = {}, None, 0

    def set_tiles(self, *t):
        self.tiles = tuple(t)

    def upload(self, i, c):
        self.uploads[i] = c

    def run(self, params, n, graph=True):
        self.runs += n
        self.uploads.clear()          # running changes the state

    def run_async(self, params, n, graph=True):
        self.run(params, n)

    def sync(self):
        pass


def test_autotune_leaves_initial_state_timed_and_cached(tmp_path, monkeypatch):
    from rhyme_amd import autotune
    monkeypatch.setenv("RHYME_AMD_TILE_CACHE", str(tmp_path / "tiles.json"))
    p = hydro.make_params(3)
    init = [np.full(3, float(i)) for i in range(4)]
    b = _FakeBatch([[8, 8, 8]] * 4)
    best, _ = autotune.tune(b, p, init, steps=1, candidates=autotune.CANDIDATES[:2])
    assert b.runs > 0 and b.tiles == tuple(best)
    assert sorted(b.uploads) == [0, 1, 2, 3]
    # same box set again: the cached shape, and still the initial state
    b2 = _FakeBatch([[8, 8, 8]] * 4)
    best2, _ = autotune.tune(b2, p, init, steps=1, candidates=autotune.CANDIDATES[:2])
    assert b2.runs == 0 and tuple(best2) == tuple(best) and b2.tiles == tuple(best)
    assert sorted(b2.uploads) == [0, 1, 2, 3]
