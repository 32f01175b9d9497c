"""Device diagnostics (rhy_box_diagnostics, rhy_batch_diagnostics): NaN
counts per component, negative density / total energy counts, minima and
their first positions in Fortran element order, interior sums and the
reference's masked totals, checked against numpy on states with planted
NaNs, negatives and ties; and the serving loop ending a job whose state
turns bad without disturbing the jobs beside it."""
import math

import numpy as np
import pytest

from rhyme_amd import hydro, ics, serve
from rhyme_amd.batch import SamrPool

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def expected(cells, ndim, dims):
    """numpy restatement of rhy_diag_t for a host array."""
    inner = tuple(slice(2, 2 + n) for n in dims[::-1])
    rho = cells[0][inner]
    e = cells[ndim + 1][inner]
    out = {"nan_count": [int(np.isnan(cells[c][inner]).sum()) for c in range(cells.shape[0])],
           "neg_rho": int((rho < 0).sum()), "neg_e_tot": int((e < 0).sum())}
    for name, a in (("rho", rho), ("e_tot", e)):
        f = a.ravel()                                   # C order (z, y, x): x fastest
        ok = ~np.isnan(f)
        m = np.min(f[ok])
        first = int(np.flatnonzero(ok & (f == m))[0])   # first in Fortran element order
        pos = np.unravel_index(first, a.shape)[::-1]    # (i, j, k), 0-based
        out["min_" + name] = float(m)
        out["min_%s_at" % name] = [int(p) + 1 for p in pos]
    out["mass"] = math.fsum(rho[~np.isnan(rho)].ravel()) if not np.isnan(rho).any() else np.nan
    out["energy"] = math.fsum(e.ravel()) if not np.isnan(e).any() else np.nan
    allr, alle = cells[0].ravel(), cells[ndim + 1].ravel()
    out["mass_masked"] = math.fsum(allr[allr > 0])
    out["energy_masked"] = math.fsum(alle[alle > 0])
    return out


def close(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return abs(a - b) <= 2 * np.spacing(abs(b))


def planted(ndim, dims, seed):
    cells = ics.random_smooth(ndim, dims, seed=seed)
    g = ics.rng(seed + 100)
    shape = cells.shape
    inner_idx = lambda: tuple([int(g.integers(0, ndim + 2))] +
                              [int(g.integers(2, 2 + n)) for n in dims[::-1]])
    for _ in range(3):
        cells[inner_idx()] = np.nan
    # negatives in rho and e_tot, with a tie for the minimum of rho
    k1 = tuple(int(g.integers(2, 2 + n)) for n in dims[::-1])
    k2 = tuple(int(g.integers(2, 2 + n)) for n in dims[::-1])
    cells[(0,) + k1] = -0.25
    cells[(0,) + k2] = -0.25
    cells[(ndim + 1,) + k2] = -1.5
    assert cells.shape == shape
    return cells


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.MIXED])
@pytest.mark.parametrize("ndim,dims", [(1, [70]), (2, [19, 23]), (3, [13, 9, 11])])
def test_box_diagnostics_match_numpy(ndim, dims, precision):
    say("diag", ndim, dims, precision)
    cells = planted(ndim, dims, seed=7 * ndim)
    box = hydro.SamrBox(ndim, dims, precision=precision, cells=cells.copy())
    box.upload()
    got = box.diagnostics()
    ref_cells = cells.astype(np.float32).astype(np.float64) if precision else cells
    want = expected(ref_cells, ndim, dims)
    assert got["nan_count"] == want["nan_count"]
    assert got["neg_rho"] == want["neg_rho"] and got["neg_e_tot"] == want["neg_e_tot"]
    for k in ("min_rho", "min_e_tot"):
        assert got[k] == want[k]
        assert got[k + "_at"] == want[k + "_at"], (k, got[k + "_at"], want[k + "_at"])
    for k in ("mass_masked", "energy_masked"):
        assert close(got[k], want[k]), (k, got[k], want[k])
    box.close()


def test_clean_state_sums_are_exactly_rounded():
    dims = [33, 17, 21]
    cells = ics.random_smooth(3, dims, seed=3)
    box = hydro.SamrBox(3, dims, cells=cells.copy())
    box.upload()
    got = box.diagnostics()
    want = expected(cells, 3, dims)
    assert sum(got["nan_count"]) == 0 and got["neg_rho"] == 0
    for k in ("mass", "energy", "mass_masked", "energy_masked"):
        assert close(got[k], want[k]), (k, got[k], want[k])
    # repeatable bit for bit
    assert box.diagnostics() == got
    box.close()


def test_pool_diagnostics_and_failed_job():
    jobs = serve.serving_jobs(njobs=8, rate=1e6, seed=31, steps=8, n2d=(12, 30),
                              n3d=(8, 14))
    for j in jobs:
        j.cells = j.make_cells()
    bad = jobs[3]
    inner = (0,) + tuple(slice(2, -2) for _ in range(bad.ndim))
    view = bad.cells[inner]
    view.flat[view.size // 2] = np.nan
    srv = serve.HydroServer(1 << 30, quantum=3, record_dts=True, check_every=1, diagnose=True)
    srv.serve(jobs)
    say("failures", srv.failures, bad.failed)
    assert srv.failures == 1 and bad.failed and "NaN" in bad.failed
    assert bad.done == 3                      # ended at the first check
    for j in jobs:
        if j is bad:
            continue
        assert j.failed is None and j.diag is not None and sum(j.diag["nan_count"]) == 0
        ref, rdts = serve.run_alone(j, srv.params[j.ndim])
        np.testing.assert_array_equal(j.result, ref)
        np.testing.assert_array_equal(np.array(j.dts), rdts)
    rep = serve.latency_report(jobs)
    assert rep["failed"] == 1
    srv.close()
