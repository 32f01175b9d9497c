"""Serving loop: a stream of independent hydro jobs, continuously batched.

Each job is one box (2D or 3D) with its own initial state, slope limiter,
face types and step count.  Jobs arrive over time; `HydroServer.serve`
admits them in arrival order while the resident set fits an HBM budget,
keeps every admitted job's state resident on the device, advances all
resident jobs together (one pool per dimensionality -- rhy_pool_* in
include/rhyme_hip.h -- each on its own stream, so 2D and 3D quanta overlap)
for a quantum of steps, then retires the jobs that reached their step count
and refills the freed room from the queue before the next quantum.

A box in a pool evolves exactly as it would alone (its own ghost fill and
CFL dt), so every job's result is bit-identical to running it by itself
with hydro.run; the reference has no counterpart (it advances one
configuration per process, src/main.f90:106-177).

Failure detection: with `check_every` = N, every N quanta the device
diagnostics (rhy_batch_diagnostics, the sanity-check quantities) of each
resident job are read, and a job whose state holds a NaN or a negative
density or total energy is ended at once (`failed`), its slot freed; the
other jobs are not touched.  With `diagnose`, every retired job carries its
final diagnostics.

Admission: a job is admitted when the bytes of all resident jobs plus its
own (state planes + flux scratch, batch.box_bytes) stay within the budget and
its pool's arena has a free block large enough.  With admission="fifo" a
job that does not fit holds every later one (first come, first served).
With admission="backfill" (the default) the jobs behind it that fit are
admitted past it -- up to `lookahead` of them per admission pass -- until
it has waited `max_bypass_s` seconds since it first failed to fit; from then
on nothing passes it, so a large job is delayed by at most that long plus
the time the resident set needs to drain.  Each pool's arena is sized
to the whole budget, so the budget is the only limit that binds except for
fragmentation -- and the device holds two arenas of the budget's size (one
per dimensionality), 2 x budget_bytes of HBM, of which at most the budget is
ever in use; the constructor fails with the library's out-of-memory error
when the device cannot hold both.
"""
import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import hydro, ics
from .batch import PoolFull, SamrPool, box_bytes
from .hydro import FP64, ncmp_for


@dataclass
class Job:
    """One request: box shape and spacing, initial state, limiter, faces."""
    jid: int
    ndim: int
    dims: list
    dx: list
    limiter: int
    bc_types: list
    steps: int
    arrival: float = 0.0        # seconds after the start of serve()
    kind: str = ""
    seed: int = 0
    ic: dict = None             # painted initial state (rhy_ic_t fields), or
    cells: np.ndarray = None    # an explicit one, hydro.box_shape layout
    # filled in by the server
    admitted: float = None
    finished: float = None
    done: int = 0
    slot: int = None
    result: np.ndarray = None
    dts: list = field(default_factory=list)
    diag: dict = None           # rhy_diag_t of the final (or failing) state
    failed: str = None          # why the job was ended early, if it was
    blocked_since: float = None  # when it first failed to fit as the head job

    def zones(self):
        return int(np.prod(self.dims))

    def make_cells(self, ncmp=None):
        """Host array of the job's initial state (its painted `ic`)."""
        return ics.paint(self.ic, self.ndim, self.dims, ncmp)


def serving_jobs(njobs=1000, rate=40.0, seed=0, steps=100, n2d=(64, 512), n3d=(32, 128)):
    """The serving workload: Poisson arrivals at `rate` jobs/s, half 2D boxes
    n x n with n log-uniform in n2d, half 3D cubes n^3 with n log-uniform in
    n3d, dx = 1/n; seeded two-state (outflow faces) or blast (reflective
    faces) initial state, described by its painting parameters (ics.ic_*):
    the device paints it at admission; limiter drawn uniformly from the
    four."""
    g = ics.rng(seed)
    gaps = g.exponential(1.0 / rate, size=njobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    ndims = np.array([2] * (njobs // 2) + [3] * (njobs - njobs // 2))
    g.shuffle(ndims)
    jobs = []
    for i in range(njobs):
        nd = int(ndims[i])
        lo, hi = n2d if nd == 2 else n3d
        n = int(round(lo * (hi / lo) ** g.uniform()))
        kind = "blast" if g.uniform() < 0.5 else "two_state"
        bc = hydro.REFLECTIVE if kind == "blast" else hydro.OUTFLOW
        seed_i = int(g.integers(1 << 30))
        ic = ics.ic_blast(seed_i) if kind == "blast" else ics.ic_two_state(nd, seed_i)
        jobs.append(Job(jid=i, ndim=nd, dims=[n] * nd, dx=[1.0 / n] * nd,
                        limiter=int(g.integers(1, 5)), bc_types=[bc] * 6, steps=steps,
                        arrival=float(arrivals[i]), kind=kind, seed=seed_i, ic=ic))
    return jobs


def job_params(ndim, courant=0.2, solver=hydro.MEMORY_INTENSIVE, tolerance=1e-6,
               n_iteration=100, limiter=hydro.MINMOD, bc_types=None):
    """Parameters shared by every job of one dimensionality (the pool's);
    limiter and faces are per job."""
    return hydro.make_params(
        ndim, rp=hydro.RiemannProblem(irs=hydro.Irs(n_iteration=n_iteration,
                                                    tolerance=tolerance)),
        sl=hydro.SlopeLimiter(type=limiter, w=0.0), mh=hydro.MusclHancock(solver_type=solver),
        cfl=hydro.Cfl(courant),
        bc=hydro.SamrBc(types=bc_types or [hydro.REFLECTIVE] * 6))


def run_alone(job, params, precision=FP64):
    """The job by itself on one box through the drop-in path (hydro.run)."""
    p = type(params).from_buffer_copy(params)
    p.limiter = job.limiter
    for f in range(6):
        p.bc_types[f] = int(job.bc_types[f])
    if job.cells is not None:
        box = hydro.SamrBox(job.ndim, job.dims, precision=precision,
                            cells=np.array(job.cells, dtype=np.float64))
        box.upload()
    else:
        box = hydro.SamrBox(job.ndim, job.dims, precision=precision)
        box.paint(job.ic)
    dts = hydro.run(p, box, job.dx, job.steps)
    out = box.download().copy()
    box.close()
    return out, dts


class HydroServer:
    """Continuous batching of 2D and 3D jobs under one HBM budget.

    keep_results: True downloads every finished job's state into job.result,
    False none, a set of job ids only those jobs'."""

    def __init__(self, budget_bytes, precision=FP64, quantum=10, params=None, max_boxes=4096,
                 tiles=None, keep_results=True, record_dts=False, check_every=0,
                 diagnose=False, admission="backfill", lookahead=64, max_bypass_s=1.0):
        if admission not in ("fifo", "backfill"):
            raise ValueError("admission must be 'fifo' or 'backfill'")
        self.admission = admission
        self.lookahead = int(lookahead)
        self.max_bypass_s = float(max_bypass_s)
        self.bypassed = 0  # admissions past a waiting head job
        self.budget = int(budget_bytes)
        self.precision = precision
        self.quantum = int(quantum)
        self.params = params or {2: job_params(2), 3: job_params(3)}
        self.pools = {}
        for nd in (2, 3):
            self.pools[nd] = SamrPool(nd, self.budget, max_boxes=max_boxes,
                                      ncmp=self.params[nd].ncmp, precision=precision)
            if tiles and nd in tiles:
                self.pools[nd].set_tiles(*tiles[nd])
        self.keep_results = keep_results
        self.record_dts = record_dts
        self.check_every = int(check_every)
        self.diagnose = diagnose
        self.failures = 0
        self.resident = {2: {}, 3: {}}  # slot -> job
        self.resident_bytes = 0
        self.peak_bytes = 0
        self.deferrals = 0
        self.quanta = 0
        # host seconds spent admitting (incl. uploads), running quanta
        # (enqueue to sync), retiring (incl. downloads) and idle
        self.seconds = dict(admit=0.0, quanta=0.0, retire=0.0, idle=0.0)

    def close(self):
        for p in self.pools.values():
            p.close()

    def job_bytes(self, job):
        return box_bytes(job.ndim, job.dims, self.params[job.ndim].ncmp, self.precision)

    def _try_admit(self, job, now):
        nb = self.job_bytes(job)
        if nb > self.budget:
            raise ValueError("job %d needs %d bytes, more than the whole budget" % (job.jid, nb))
        if self.resident_bytes + nb > self.budget:
            return False
        pool = self.pools[job.ndim]
        try:
            if job.cells is not None:
                slot = pool.add(job.dims, job.dx, job.cells, job.limiter, job.bc_types)
            else:
                slot = pool.add_ic(job.dims, job.dx, job.ic, job.limiter, job.bc_types)
        except PoolFull:
            return False
        job.slot = slot
        job.admitted = now
        self.resident[job.ndim][slot] = job
        self.resident_bytes += nb
        self.peak_bytes = max(self.peak_bytes, self.resident_bytes)
        return True

    @staticmethod
    def unhealthy(diag):
        """Why a state fails the sanity checks (None if it does not)."""
        if any(diag["nan_count"]):
            return "NaN in components %s" % [c for c, n in enumerate(diag["nan_count"]) if n]
        if diag["neg_rho"]:
            return "negative density (%d cells)" % diag["neg_rho"]
        if diag["neg_e_tot"]:
            return "negative total energy (%d cells)" % diag["neg_e_tot"]
        return None

    def _retire(self, job, now):
        pool = self.pools[job.ndim]
        if self.diagnose and job.diag is None:
            job.diag = pool.diagnostics(job.slot)
        if self.keep_results is True or (self.keep_results and job.jid in self.keep_results):
            job.result = pool.download(job.slot)
        pool.remove(job.slot)
        del self.resident[job.ndim][job.slot]
        self.resident_bytes -= self.job_bytes(job)
        job.finished = now
        job.slot = None

    def serve(self, jobs, clock=time.perf_counter, sleep=time.sleep):
        """Serve `jobs` to completion in arrival order; returns them with
        admitted/finished times (seconds from the start) filled in.  A job's
        state comes from its `cells` (uploaded at admission) or is painted on
        the device from its `ic`."""
        queue = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        t0 = clock()
        left = len(queue)
        while left:
            ta = clock()
            now = ta - t0
            # admission: in arrival order; with backfill, arrived jobs that
            # fit pass a head job that does not, for up to max_bypass_s
            while queue and queue[0].arrival <= now:
                if not self._try_admit(queue[0], now):
                    self.deferrals += 1
                    head = queue[0]
                    if head.blocked_since is None:
                        head.blocked_since = now
                    if (self.admission == "backfill" and
                            now - head.blocked_since < self.max_bypass_s):
                        passed = []
                        for i in range(1, min(len(queue), self.lookahead + 1)):
                            j = queue[i]
                            if j.arrival > now:
                                break
                            if self._try_admit(j, now):
                                passed.append(i)
                        for i in reversed(passed):
                            del queue[i]
                        self.bypassed += len(passed)
                    break
                queue.popleft()
            tb = clock()
            self.seconds["admit"] += tb - ta
            busy = [nd for nd in (2, 3) if self.resident[nd]]
            if not busy:
                if queue:
                    sleep(max(0.0, queue[0].arrival - (clock() - t0)))
                self.seconds["idle"] += clock() - tb
                continue
            # one quantum on every busy pool: enqueue all, then wait for all
            steps = {}
            for nd in busy:
                res = self.resident[nd]
                q = min(self.quantum, min(j.steps - j.done for j in res.values()))
                steps[nd] = q
                self.pools[nd].run_async(self.params[nd], q, graph=True)
            for nd in busy:
                self.pools[nd].sync()
            self.quanta += 1
            tc = clock()
            self.seconds["quanta"] += tc - tb
            now = tc - t0
            check = self.check_every and self.quanta % self.check_every == 0
            for nd in busy:
                q = steps[nd]
                dts = self.pools[nd].last_dts(q) if self.record_dts else None
                for slot, job in list(self.resident[nd].items()):
                    if dts is not None:
                        job.dts.extend(dts[:, slot].tolist())
                    job.done += q
                    if check:
                        job.diag = self.pools[nd].diagnostics(slot)
                        job.failed = self.unhealthy(job.diag)
                    if job.failed or job.done >= job.steps:
                        self.failures += job.failed is not None
                        self._retire(job, now)
                        left -= 1
            self.seconds["retire"] += clock() - tc
        return jobs


def latency_report(jobs):
    """Completion latency (finish - arrival) percentiles and throughput of
    the jobs that completed (failed ones are counted apart)."""
    done = [j for j in jobs if not j.failed]
    lat = np.array([j.finished - j.arrival for j in done])
    wait = np.array([j.admitted - j.arrival for j in done])
    span = max(j.finished for j in jobs) - min(j.arrival for j in jobs)
    zu = float(sum(j.zones() * j.done for j in jobs))
    return dict(
        jobs=len(jobs), failed=len(jobs) - len(done), p50_ms=1e3 * float(np.percentile(lat, 50)),
        p99_ms=1e3 * float(np.percentile(lat, 99)), max_ms=1e3 * float(lat.max()),
        wait_p50_ms=1e3 * float(np.percentile(wait, 50)),
        wait_p99_ms=1e3 * float(np.percentile(wait, 99)),
        makespan_s=float(span), zone_updates=zu, zone_updates_per_s=zu / span)
