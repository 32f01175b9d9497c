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


# ------------------------------------------------ unit procedures + errors ----

def units_exe(nd):
    return os.path.join(BUILD, "%dd" % nd, "rhy_fhost_units.exe")


def run_units(nd, mode, lines, som=2):
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.txt")
        fout = os.path.join(d, "out.bin")
        with open(fin, "w") as f:
            f.write("\n".join(lines) + "\n")
        env = dict(os.environ, RHY_SOM=str(som))
        r = subprocess.run([units_exe(nd), mode, fin, fout], capture_output=True, text=True,
                           timeout=120, env=env)
        return r, np.fromfile(fout, dtype=np.float64)


def fmt(*xs):
    return " ".join(repr(float(x)) for x in xs)


def within_sig(val, exp, n):
    power = int(np.log10(abs(val))) if val != 0 else 0
    return abs(val - exp) <= 10.0 ** (power - n + 1)


@pytest.mark.gpu
def test_fhost_irs_exec_p_stars():
    """rhyme_irs_exec through the Fortran host, as rhyme_irs_exec_test.f90:36-58
    calls it: gamma 5/3, c = sqrt(5./3.*p/rho) with an FP32 5./3., tol 1e-10,
    the first 1000 rows of p_stars.txt, 6 significant figures."""
    rows = np.loadtxt(os.path.join(ROOT, "tests", "golden", "p_stars_subset.csv"),
                      delimiter=",")[:1000]
    c53 = float(np.float32(5.0) / np.float32(3.0))
    lines = ["%d 100000 1e-10" % len(rows)]
    for rl, ul, pl, rr, ur, pr, _ in rows:
        lines.append(fmt(rl, ul, pl, np.sqrt(c53 * pl / rl), rr, ur, pr, np.sqrt(c53 * pr / rr)))
    r, out = run_units(1, "irs", lines, som=1)
    assert r.returncode == 0, r.stdout + r.stderr
    out = out.reshape(-1, 3)
    bad = [i for i in range(len(rows)) if not within_sig(out[i, 0], rows[i, 6], 6)]
    assert not bad, bad[:10]


TORO = None


def toro_cases():
    import json
    return json.load(open(os.path.join(ROOT, "tests", "golden", "toro5.json")))["cases"]


def fortran_sides(p, l, r, nd, axis):
    s = oracle.rp_solution(oracle.params_from(p), l, r)
    return fmt(s.l_rho, *list(s.l_v)[:nd], s.l_p, s.l_cs, s.r_rho, *list(s.r_v)[:nd], s.r_p,
               s.r_cs) + " %d" % axis


def cons1(g, rho, v, pr, nd, axis=1):
    vv = [0.0] * nd
    vv[axis - 1] = v
    s = 0.0
    for x in vv:
        s = s + x * x
    return np.array([rho] + [rho * x for x in vv] + [.5 * rho * s + pr / (g - 1.0)])


@pytest.mark.gpu
@pytest.mark.parametrize("nd,axis", [(1, 1), (3, 2)])
def test_fhost_star_toro_five(nd, axis):
    """rhyme_riemann_problem_star through the Fortran host on Toro's five
    tests (rhyme_riemann_problem_star_test.f90:31-35), along axis `axis`."""
    g = 7.0 / 5.0
    p = hydro.make_params(nd, thermo=hydro.Thermo(state_of_matter=hydro.DIATOMIC))
    cases = toro_cases()
    lines = ["%d 1000 1e-8" % len(cases)]
    for c in cases:
        l = cons1(g, c["left"][0], c["left"][1], c["left"][2], nd, axis)
        r = cons1(g, c["right"][0], c["right"][1], c["right"][2], nd, axis)
        lines.append(fortran_sides(p, l, r, nd, axis))
    res, out = run_units(nd, "star", lines)
    assert res.returncode == 0, res.stdout + res.stderr
    out = out.reshape(-1, 18)
    for c, o in zip(cases, out):
        n = c["sig_figs"]
        assert within_sig(o[0], c["p_star"], n) and within_sig(o[1], c["u_star"], n), c["name"]
        assert bool(o[2]) == c["left_is_shock"] and bool(o[9]) == c["right_is_shock"]
        lr = o[3] if o[2] else o[5]
        rr = o[10] if o[9] else o[12]
        assert within_sig(lr, c["left_rho_star"], n) and within_sig(rr, c["right_rho_star"], n)


@pytest.mark.gpu
def test_fhost_sampling_and_solve_sod():
    """rhyme_riemann_problem_sampling and rhyme_riemann_problem_solve (dx != 0)
    through the Fortran host on the Sod profile of
    rhyme_riemann_problem_sampling_test.f90:37-66 (16/14/16/15 figures); the
    two agree bit for bit."""
    g = 7.0 / 5.0
    p = hydro.make_params(1, thermo=hydro.Thermo(state_of_matter=hydro.DIATOMIC))
    l = cons1(g, 1.0, 0.0, 1.0, 1)
    r = cons1(g, 0.125, 0.0, 0.1, 1)
    rows = np.loadtxt(os.path.join(ROOT, "tests", "golden", "sod_t0.2_sampling.csv"),
                      delimiter=",")
    lines = ["%d 1000 1e-8" % len(rows), fortran_sides(p, l, r, 1, 1), fmt(*l), fmt(*r)]
    for i in range(len(rows)):
        lines.append(fmt(-.5 + i / 499.0, 0.2))
    res, out = run_units(1, "sampling", lines)
    assert res.returncode == 0, res.stdout + res.stderr
    out = out.reshape(-1, 2, 3)
    assert np.array_equal(out[:, 0], out[:, 1])
    op = oracle.params_from(p)
    for i, (x, rho, v, pr) in enumerate(rows):
        u = out[i, 0]
        assert within_sig(rho, u[0], 16) and within_sig(rho * v, u[1], 14), i
        assert within_sig(pr, oracle.pressure(op, u), 16), i


@pytest.mark.gpu
@pytest.mark.parametrize("nd", [1, 3])
def test_fhost_errors_reported(nd):
    """Every failure of a host procedure reaches the caller: logger%err where
    the reference's interface takes a logger, the host error count otherwise;
    the box's cells and the caller's output arrays are left as they were,
    and a valid step afterwards still advances the box."""
    res, out = run_units(nd, "errors", ["1 1"])
    assert res.returncode == 0, res.stdout + res.stderr
    assert out.size == 10, out
    assert out[0] == 1 and out[1] == 1            # logger%err, cells unchanged
    assert out[2] == 1 and out[3] == 1            # host error, cells unchanged
    assert out[4] == -1.0 and out[5] == 3         # cfl dt -1, cfl + bc errors
    assert out[6] == 1 and out[7] == 7            # u unchanged, 4 more errors
    assert out[8] == 1 and out[9] == 1            # no new logger error, box advanced
    assert "unknown slope limiter" in res.stderr
    assert "axis out of range" in res.stderr
