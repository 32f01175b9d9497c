"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product path
(rhyme_amd, librhyme_hip.so) never imports, links or calls it.

The C library restates the reference Fortran hot path (rhyme_muscl_hancock,
rhyme_riemann_problem + rhyme_irs, rhyme_slope_limiter, rhyme_cfl,
rhyme_samr_bc, rhyme_ideal_gas/thermo_base/hydro_base); see rhyme_oracle.c for
the per-function file:line citations.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_CMP = 16


class OrcParams(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("limiter", C.c_int32),
        ("solver_type", C.c_int32),
        ("n_iteration", C.c_int32),
        ("gamma", C.c_double),
        ("kb_amu", C.c_double),
        ("omega", C.c_double),
        ("tolerance", C.c_double),
        ("w_vacuum_rho", C.c_double),
        ("w_vacuum_p", C.c_double),
        ("courant", C.c_double),
        ("bc_types", C.c_int32 * 6),
        ("ncmp", C.c_int32),
        ("reserved", C.c_int32),
        ("cons_inflows", (C.c_double * MAX_CMP) * 6),
    ]


class OrcRpSolution(C.Structure):
    _fields_ = [
        ("l_rho", C.c_double), ("l_v", C.c_double * 3), ("l_p", C.c_double),
        ("l_cs", C.c_double),
        ("r_rho", C.c_double), ("r_v", C.c_double * 3), ("r_p", C.c_double),
        ("r_cs", C.c_double),
        ("star_p", C.c_double), ("star_u", C.c_double),
        ("l_is_shock", C.c_int32), ("r_is_shock", C.c_int32),
        ("l_shock_rho", C.c_double), ("l_shock_speed", C.c_double),
        ("l_fan_rho", C.c_double), ("l_fan_cs", C.c_double),
        ("l_fan_speedH", C.c_double), ("l_fan_speedT", C.c_double),
        ("r_shock_rho", C.c_double), ("r_shock_speed", C.c_double),
        ("r_fan_rho", C.c_double), ("r_fan_cs", C.c_double),
        ("r_fan_speedH", C.c_double), ("r_fan_speedT", C.c_double),
        ("l_f", C.c_double), ("r_f", C.c_double),
    ]


_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        pp = C.POINTER(OrcParams)
        L.orc_pressure.restype = C.c_double
        L.orc_pressure.argtypes = [pp, dp]
        L.orc_speed_of_sound.restype = C.c_double
        L.orc_speed_of_sound.argtypes = [pp, dp]
        L.orc_flux.argtypes = [pp, dp, C.c_int, dp]
        L.orc_prim_vars_to_cons.argtypes = [pp, C.c_double, dp, C.c_double, dp]
        L.orc_slope_limiter.argtypes = [pp, dp, dp, dp, dp]
        L.orc_half_step.argtypes = [pp, dp, dp, C.c_int, C.c_double, C.c_double, dp, dp]
        L.orc_irs_guess.argtypes = [pp, C.c_double, dp, C.c_double, C.c_double,
                                    C.c_double, dp, C.c_double, C.c_double, C.c_int, dp]
        L.orc_irs_exec.restype = C.c_double
        L.orc_irs_exec.argtypes = [pp, C.c_double, dp, C.c_double, C.c_double, dp,
                                   C.c_double, dp, C.c_double, C.c_double, dp,
                                   C.c_int, ip]
        L.orc_nonlinear_wave_function.argtypes = [pp, C.c_double, C.c_double,
                                                  C.c_double, C.c_double, dp, dp]
        L.orc_rp_star.argtypes = [pp, C.POINTER(OrcRpSolution), C.c_int]
        L.orc_rp_sampling.argtypes = [pp, C.POINTER(OrcRpSolution), C.c_int,
                                      C.c_double, C.c_double, dp]
        L.orc_rp_solve.argtypes = [pp, dp, dp, C.c_double, C.c_double, C.c_int, dp]
        L.orc_rp_solve_batch.argtypes = [pp, C.c_int64, dp, dp, ip, C.c_double, dp]
        L.orc_limiter_batch.argtypes = [pp, C.c_int64, dp, dp, dp, dp]
        L.orc_half_step_batch.argtypes = [pp, C.c_int64, dp, dp, ip, C.c_double,
                                          C.c_double, dp, dp]
        L.orc_flux_batch.argtypes = [pp, C.c_int64, dp, ip, dp]
        L.orc_cfl_time_step.restype = C.c_double
        L.orc_cfl_time_step.argtypes = [pp, ip, dp, dp]
        L.orc_bc_set_boundaries.argtypes = [pp, ip, dp]
        L.orc_bc_set_face.argtypes = [pp, ip, dp, C.c_int]
        L.orc_bc_init_ghosts.argtypes = [pp, ip, dp]
        L.orc_mh_solve.restype = C.c_int
        L.orc_mh_solve.argtypes = [pp, ip, dp, dp, C.c_double]
        L.orc_mh_solve_fluxes.restype = C.c_int
        L.orc_mh_solve_fluxes.argtypes = [pp, ip, dp, dp, C.c_double, dp]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [pp, ip, dp, dp, C.c_int32, dp]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def params_from(p):
    """Copy a dict/obj with the rhy_params fields into an OrcParams."""
    o = OrcParams()
    for name, _ in OrcParams._fields_:
        if name in ("bc_types", "cons_inflows", "reserved"):
            continue
        setattr(o, name, getattr(p, name) if not isinstance(p, dict) else p[name])
    bct = p.bc_types if not isinstance(p, dict) else p["bc_types"]
    for i in range(6):
        o.bc_types[i] = int(bct[i])
    inf = p.cons_inflows if not isinstance(p, dict) else p.get("cons_inflows")
    if inf is not None:
        inf = np.asarray(inf, dtype=np.float64)
        for f in range(6):
            for c in range(min(MAX_CMP, inf.shape[1])):
                o.cons_inflows[f][c] = inf[f, c]
    return o


def _nv(p):
    return p.ndim + 2


def pressure(p, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_pressure(C.byref(p), _dp(u))


def speed_of_sound(p, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_speed_of_sound(C.byref(p), _dp(u))


def flux(p, u, axis):
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros(_nv(p))
    lib().orc_flux(C.byref(p), _dp(u), axis, _dp(f))
    return f


def prim_vars_to_cons(p, rho, v, pr):
    v = np.ascontiguousarray(np.resize(np.asarray(v, dtype=np.float64), 3))
    u = np.zeros(_nv(p))
    lib().orc_prim_vars_to_cons(C.byref(p), rho, _dp(v), pr, _dp(u))
    return u


def slope_limiter(p, ul, u, ur):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ul, u, ur)]
    d = np.zeros(_nv(p))
    lib().orc_slope_limiter(C.byref(p), _dp(a[0]), _dp(a[1]), _dp(a[2]), _dp(d))
    return d


def half_step(p, u, delta, axis, dx, dt):
    u = np.ascontiguousarray(u, dtype=np.float64)
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    l = np.zeros(_nv(p))
    r = np.zeros(_nv(p))
    lib().orc_half_step(C.byref(p), _dp(u), _dp(delta), axis, dx, dt, _dp(l), _dp(r))
    return l, r


def _v3(v):
    out = np.zeros(3)
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    out[:len(v)] = v
    return out


def irs_guess(p, l_rho, l_v, l_p, l_cs, r_rho, r_v, r_p, r_cs, axis=1):
    lv, rv = _v3(l_v), _v3(r_v)
    g = np.zeros(6)
    lib().orc_irs_guess(C.byref(p), l_rho, _dp(lv), l_p, l_cs, r_rho, _dp(rv),
                        r_p, r_cs, axis, _dp(g))
    return g


def irs_exec(p, l_rho, l_v, l_p, l_cs, r_rho, r_v, r_p, r_cs, axis=1):
    lv, rv = _v3(l_v), _v3(r_v)
    lf = C.c_double(0.0)
    rf = C.c_double(0.0)
    n = C.c_int32(0)
    ps = lib().orc_irs_exec(C.byref(p), l_rho, _dp(lv), l_p, l_cs, C.byref(lf),
                            r_rho, _dp(rv), r_p, r_cs, C.byref(rf), axis, C.byref(n))
    return ps, lf.value, rf.value, n.value


def wave_function(p, rho, pk, cs, pstar):
    f = C.c_double(0.0)
    fp = C.c_double(0.0)
    lib().orc_nonlinear_wave_function(C.byref(p), rho, pk, cs, pstar, C.byref(f),
                                      C.byref(fp))
    return f.value, fp.value


def rp_solution(p, l, r):
    """riemann_problem_solution_t sides from conserved l, r (the factory's
    rhyme_riemann_problem_factory_set_sides: v = m/rho, p = calc_p, cs = calc_cs)."""
    s = OrcRpSolution()
    nd = p.ndim
    l = np.asarray(l, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    s.l_rho, s.r_rho = l[0], r[0]
    for d in range(nd):
        s.l_v[d] = l[1 + d] / l[0]
        s.r_v[d] = r[1 + d] / r[0]
    s.l_p, s.r_p = pressure(p, l), pressure(p, r)
    s.l_cs, s.r_cs = speed_of_sound(p, l), speed_of_sound(p, r)
    return s


def rp_star(p, s, axis=1):
    lib().orc_rp_star(C.byref(p), C.byref(s), axis)
    return s


def rp_sampling(p, s, axis, dx, dt):
    u = np.zeros(_nv(p))
    lib().orc_rp_sampling(C.byref(p), C.byref(s), axis, dx, dt, _dp(u))
    return u


def rp_solve(p, l, r, dx, dt, axis):
    l = np.ascontiguousarray(l, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    u = np.zeros(_nv(p))
    lib().orc_rp_solve(C.byref(p), _dp(l), _dp(r), dx, dt, axis, _dp(u))
    return u


def rp_solve_batch(p, l, r, axis, dt):
    l = np.ascontiguousarray(l, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    u = np.zeros_like(l)
    lib().orc_rp_solve_batch(C.byref(p), len(axis), _dp(l), _dp(r), _ip(axis), dt, _dp(u))
    return u


def limiter_batch(p, ul, u, ur):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ul, u, ur)]
    d = np.zeros_like(a[1])
    lib().orc_limiter_batch(C.byref(p), a[1].shape[0], _dp(a[0]), _dp(a[1]), _dp(a[2]), _dp(d))
    return d


def half_step_batch(p, u, delta, axis, dx, dt):
    u = np.ascontiguousarray(u, dtype=np.float64)
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    l = np.zeros_like(u)
    r = np.zeros_like(u)
    lib().orc_half_step_batch(C.byref(p), u.shape[0], _dp(u), _dp(delta), _ip(axis),
                              dx, dt, _dp(l), _dp(r))
    return l, r


def flux_batch(p, u, axis):
    u = np.ascontiguousarray(u, dtype=np.float64)
    axis = np.ascontiguousarray(axis, dtype=np.int32)
    f = np.zeros_like(u)
    lib().orc_flux_batch(C.byref(p), u.shape[0], _dp(u), _ip(axis), _dp(f))
    return f


def _dims(dims):
    d = np.ones(3, dtype=np.int32)
    d[:len(dims)] = dims
    return d


def _check_cells(p, dims, cells):
    assert cells.dtype == np.float64 and cells.flags["C_CONTIGUOUS"]
    ext = [n + 4 for n in dims] + [1] * (3 - len(dims))
    need = p.ncmp * ext[0] * ext[1] * ext[2]
    assert cells.size == need, (cells.size, need)


def cfl_time_step(p, dims, cells, dx):
    _check_cells(p, dims, cells)
    dx = np.ascontiguousarray(np.resize(np.asarray(dx, dtype=np.float64), 3))
    return lib().orc_cfl_time_step(C.byref(p), _ip(_dims(dims)), _dp(cells), _dp(dx))


def bc_set_boundaries(p, dims, cells):
    _check_cells(p, dims, cells)
    lib().orc_bc_set_boundaries(C.byref(p), _ip(_dims(dims)), _dp(cells))


def bc_set_face(p, dims, cells, face):
    _check_cells(p, dims, cells)
    lib().orc_bc_set_face(C.byref(p), _ip(_dims(dims)), _dp(cells), face)


def bc_init_ghosts(p, dims, cells):
    _check_cells(p, dims, cells)
    lib().orc_bc_init_ghosts(C.byref(p), _ip(_dims(dims)), _dp(cells))


def mh_solve(p, dims, cells, dx, dt):
    _check_cells(p, dims, cells)
    dx = np.ascontiguousarray(np.resize(np.asarray(dx, dtype=np.float64), 3))
    rc = lib().orc_mh_solve(C.byref(p), _ip(_dims(dims)), _dp(cells), _dp(dx), dt)
    if rc:
        raise RuntimeError("orc_mh_solve failed: %d" % rc)


def mh_solve_fluxes(p, dims, cells, dx, dt):
    """mh_solve that also returns the step's intercell fluxes, an array of
    shape (ndim, nvar) + the ghosted box (the face between a cell and its
    neighbour above along an axis stored at the cell)."""
    _check_cells(p, dims, cells)
    nd = p.ndim
    dx = np.ascontiguousarray(np.resize(np.asarray(dx, dtype=np.float64), 3))
    fr = np.zeros((nd, nd + 2) + cells.shape[1:])
    rc = lib().orc_mh_solve_fluxes(C.byref(p), _ip(_dims(dims)), _dp(cells), _dp(dx), dt,
                                   _dp(fr))
    if rc:
        raise RuntimeError("orc_mh_solve_fluxes failed: %d" % rc)
    return fr


def run(p, dims, cells, dx, nsteps):
    """Advance `cells` in place by nsteps main-loop steps; returns the dts."""
    _check_cells(p, dims, cells)
    dx = np.ascontiguousarray(np.resize(np.asarray(dx, dtype=np.float64), 3))
    dts = np.zeros(max(nsteps, 1))
    rc = lib().orc_run(C.byref(p), _ip(_dims(dims)), _dp(cells), _dp(dx), nsteps, _dp(dts))
    if rc:
        raise RuntimeError("orc_run failed: %d" % rc)
    return dts[:nsteps]


def num_threads():
    return lib().orc_num_threads()
