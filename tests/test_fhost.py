"""The Fortran host modules (rhyme_amd/fhost/rhyme_host.f90), which keep the
reference's module interfaces (rhyme_cfl_time_step, rhyme_samr_bc_*,
rhyme_muscl_hancock_solve, ...), driven by a main loop written like the
reference's (tests/fhost/rhy_fhost_driver.f90) and checked against the CPU
oracle.  The GPU cases call the device through ISO_C_BINDING."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from oracle import chombo as ochombo
from rhyme_amd import hydro, ics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "rhyme_amd", "fhost", "_build")


def exe(nd):
    return os.path.join(BUILD, "%dd" % nd, "rhy_fhost_driver.exe")


def write_nml(path, **kw):
    with open(path, "w") as f:
        f.write("&run\n")
        for k, v in kw.items():
            if isinstance(v, (list, tuple)):
                v = ", ".join(repr(x) for x in v)
            elif isinstance(v, float):
                v = repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
            f.write("  %s = %s\n" % (k, v))
        f.write("/\n")


def run_driver(nd, dims, cells, nsteps, p, precision=0):
    with tempfile.TemporaryDirectory() as d:
        nml = os.path.join(d, "run.nml")
        cin = os.path.join(d, "in.bin")
        cout = os.path.join(d, "out.bin")
        write_nml(nml, dims=list(dims) + [1] * (3 - nd), nsteps=nsteps,
                  state_of_matter=[k for k, v in hydro.IG_GAMMA.items() if v == p.gamma][0],
                  limiter=p.limiter, solver_type=p.solver_type, n_iteration=p.n_iteration,
                  bc_types=list(p.bc_types), kb=1.0, amu=1.0 / p.kb_amu, omega=p.omega,
                  tolerance=p.tolerance, courant=p.courant, box_length=1.0,
                  precision=precision)
        cells.tofile(cin)
        r = subprocess.run([exe(nd), nml, cin, cout], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = np.fromfile(cout, dtype=np.uint8)
    nb = 8 * (cells.size + nsteps)
    f64 = raw[:nb].view(np.float64)
    out = f64[:cells.size].reshape(cells.shape)
    dts = f64[cells.size:]
    payload = raw[nb:].view(np.float32)
    return out, dts, payload


def test_fhost_built():
    for nd in (1, 2, 3):
        assert os.path.exists(exe(nd)), "run __graft_entry__.build()"


def test_fhost_reports_missing_device():
    """Without a usable device the host modules log an error and the
    driver exits with status 3 (never a silent CPU fallback)."""
    from rhyme_amd import _lib
    if _lib.load().rhy_init(0) == _lib.RHY_OK:
        pytest.skip("a gfx950 device is present")
    p = hydro.make_params(1, cfl=hydro.Cfl(0.8))
    cells, _ = ics.two_state_1d(64, seed=0)
    with tempfile.TemporaryDirectory() as d:
        nml = os.path.join(d, "run.nml")
        write_nml(nml, dims=[64, 1, 1], nsteps=1)
        cin = os.path.join(d, "in.bin")
        cells.tofile(cin)
        r = subprocess.run([exe(1), nml, cin, os.path.join(d, "o.bin")], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode != 0


def _kb_amu_params(p):
    # the driver forms kb_amu = kb / amu with kb = 1, amu = 1/kb_amu; use the
    # value it will reproduce so both sides see the same bits
    q = oracle.params_from(p)
    q.kb_amu = 1.0 / (1.0 / p.kb_amu)
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("nd,dims,limiter,bc,solver", [
    (1, [256], hydro.VAN_LEER, hydro.OUTFLOW, hydro.MEMORY_INTENSIVE),
    (2, [48, 40], hydro.MINMOD, hydro.PERIODIC, hydro.CPU_INTENSIVE),
    (3, [20, 16, 24], hydro.SUPERBEE, hydro.REFLECTIVE, hydro.MEMORY_INTENSIVE),
])
def test_fhost_main_loop_vs_oracle(nd, dims, limiter, bc, solver):
    print("fhost main loop nd", nd, flush=True)
    g = 5.0 / 3.0
    p = hydro.make_params(nd, rp=hydro.RiemannProblem(irs=hydro.Irs(100, 1e-6)),
                          sl=hydro.SlopeLimiter(type=limiter),
                          mh=hydro.MusclHancock(solver_type=solver),
                          cfl=hydro.Cfl(0.2 if nd > 1 else 0.8),
                          bc=hydro.SamrBc(types=[bc] * 6))
    if nd == 3:
        cells = ics.blast_3d(dims=dims, seed=5)
    else:
        cells = ics.random_smooth(nd, dims, seed=6, gamma=g)
    nsteps = 10
    out, dts, payload = run_driver(nd, dims, cells.copy(), nsteps, p)
    ref = cells.copy()
    q = _kb_amu_params(p)
    rdts = oracle.run(q, dims, ref, [1.0 / n for n in dims], nsteps)
    sl = (slice(0, nd + 2),) + tuple(slice(2, -2) for _ in range(nd))
    a, b = out[sl], ref[sl]
    for c in range(nd + 2):
        e = np.abs(a[c] - b[c]).sum() / np.abs(b[c]).sum()
        assert e <= 1e-12, (c, e)
    assert np.max(np.abs(dts - rdts) / rdts) <= 1e-12
    # the level's snapshot payload, packed on the device through the Fortran
    # binding: the restated Chombo writer's data of the final state
    want = ochombo.level_data(nd, [out], out.shape[0])
    np.testing.assert_array_equal(payload.view(np.uint32), want.view(np.uint32))
