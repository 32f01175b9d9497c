"""ctypes binding of librhyme_hip.so (the C ABI declared in include/rhyme_hip.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback; if the library is missing or cannot find a gfx950 device, calls
raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RHYME_HIP_LIB selects another build of the same library (e.g. one compiled
# with different tile constants); the default is the in-tree build.
LIB_PATH = os.environ.get("RHYME_HIP_LIB") or os.path.join(HERE, "librhyme_hip.so")

RHY_OK = 0
RHY_ERR_ARG, RHY_ERR_HIP, RHY_ERR_OOM, RHY_ERR_SHAPE, RHY_ERR_NODEV, RHY_ERR_SOLVER = \
    -1, -2, -3, -4, -5, -6
MAX_CMP = 16

# Every symbol include/rhyme_hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "rhy_init", "rhy_last_error", "rhy_version",
    "rhy_box_create", "rhy_box_resize", "rhy_box_destroy", "rhy_box_upload",
    "rhy_box_download", "rhy_box_sync",
    "rhy_bc_set_boundaries", "rhy_bc_set_face", "rhy_cfl_time_step",
    "rhy_mh_solve", "rhy_run", "rhy_run_async", "rhy_box_last_dts",
    "rhy_box_set_profiling", "rhy_box_kernel_stats",
    "rhy_riemann_batch", "rhy_limiter_batch", "rhy_half_step_batch",
    "rhy_flux_batch", "rhy_box_paint", "rhy_box_diagnostics",
    "rhy_batch_create", "rhy_batch_destroy", "rhy_batch_bytes", "rhy_batch_upload",
    "rhy_batch_download", "rhy_batch_set_tiles", "rhy_batch_run_async", "rhy_batch_sync",
    "rhy_batch_last_dts", "rhy_batch_last_error", "rhy_batch_diagnostics",
    "rhy_pool_create", "rhy_pool_box_bytes", "rhy_pool_add", "rhy_pool_add_ic",
    "rhy_pool_remove", "rhy_pool_free_bytes",
    "rhy_box_last_deferred", "rhy_batch_last_deferred", "rhy_debug_defer",
    "rhy_batch_cfl_dts", "rhy_batch_run_dt_async", "rhy_amr_prolong",
    "rhy_amr_restrict",
    "rhy_box_snapshot_pack", "rhy_box_snapshot_unpack", "rhy_batch_snapshot_length",
    "rhy_batch_snapshot_pack", "rhy_batch_snapshot_unpack",
    "rhy_build_options", "rhy_riemann_solve_batch", "rhy_riemann_star_batch",
    "rhy_riemann_sampling_batch", "rhy_riemann_region_batch", "rhy_irs_exec_batch",
    "rhy_irs_guess_batch", "rhy_pow_batch", "rhy_mh_solve_noalloc", "rhy_box_draw",
    "rhy_amr_flux_register", "rhy_amr_reflux", "rhy_amr_coarse_step", "rhy_amr_regrid",
    "rhy_amr_prolong_slot", "rhy_amr_restrict_slot", "rhy_amr_reflux_slot",
    "rhy_amr_regrid_slot",
]
# size of the static device scratch of rhy_mh_solve_noalloc
NOALLOC_SCRATCH_BYTES = 64 << 20

# rhy_build_options bits
OPT_POW_FAST, OPT_CONTACT_EXACT, OPT_LDS_WALK, OPT_FAST_RIEMANN = 1, 2, 4, 8
# rhy_pow_batch policies
POW_PATH, POW_LIBM, POW_MIXED = 0, 1, 2
# rhy_riemann_region_batch region ids
W_L, W_R, W_STARL_SHO, W_STARR_SHO, W_STARL_FAN, W_STARR_FAN, W_LFAN, W_RFAN = range(1, 9)


class RhyParams(C.Structure):
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


class RhyRpSolution(C.Structure):
    """rhy_rp_solution_t: riemann_problem_solution_t flattened."""
    _fields_ = [
        ("l_rho", C.c_double), ("l_v", C.c_double * 3), ("l_p", C.c_double),
        ("l_cs", C.c_double),
        ("r_rho", C.c_double), ("r_v", C.c_double * 3), ("r_p", C.c_double),
        ("r_cs", C.c_double),
        ("star_p", C.c_double), ("star_u", C.c_double),
        ("l_is_shock", C.c_int32), ("r_is_shock", C.c_int32),
        ("l_shock_rho", C.c_double), ("l_shock_speed", C.c_double),
        ("l_fan_rho", C.c_double), ("l_fan_cs", C.c_double), ("l_fan_speedH", C.c_double),
        ("l_fan_speedT", C.c_double),
        ("r_shock_rho", C.c_double), ("r_shock_speed", C.c_double),
        ("r_fan_rho", C.c_double), ("r_fan_cs", C.c_double), ("r_fan_speedH", C.c_double),
        ("r_fan_speedT", C.c_double),
        ("l_f", C.c_double), ("r_f", C.c_double),
    ]


RHY_IC_TWO_STATE, RHY_IC_BLAST = 1, 2


class RhyIC(C.Structure):
    """rhy_ic_t: a seeded initial state painted on the device."""
    _fields_ = [
        ("kind", C.c_int32),
        ("axis", C.c_int32),
        ("gamma", C.c_double),
        ("rho_l", C.c_double),
        ("p_l", C.c_double),
        ("rho_r", C.c_double),
        ("p_r", C.c_double),
        ("radius", C.c_double),
        ("jitter", C.c_double),
        ("p0", C.c_double),
        ("energy", C.c_double),
        ("seed", C.c_uint64),
    ]


RHY_DRAW_SPHERE, RHY_DRAW_CUBOID = 1, 2
DRAW_MAX_SHAPES, DRAW_MAX_HARMONICS = 256, 8


class RhyShape(C.Structure):
    """rhy_shape_t: a sharp sphere or cuboid of a drawing."""
    _fields_ = [
        ("kind", C.c_int32),
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("pad", C.c_int32),
        ("centre", C.c_double * 3),
        ("radius", C.c_double),
        ("rho", C.c_double),
        ("v", C.c_double * 3),
        ("p", C.c_double),
    ]


class RhyHarmonic(C.Structure):
    """rhy_harmonic_t: a harmonic perturbation and its host profile."""
    _fields_ = [
        ("axis", C.c_int32),
        ("pad", C.c_int32),
        ("profile", C.POINTER(C.c_double)),
        ("rho", C.c_double),
        ("v", C.c_double * 3),
        ("p", C.c_double),
    ]


class RhyDrawing(C.Structure):
    """rhy_drawing_t: canvas colour, shapes and perturbations."""
    _fields_ = [
        ("rho", C.c_double),
        ("v", C.c_double * 3),
        ("p", C.c_double),
        ("nshapes", C.c_int32),
        ("nharmonics", C.c_int32),
        ("shapes", C.POINTER(RhyShape)),
        ("harmonics", C.POINTER(RhyHarmonic)),
    ]


class RhyDiag(C.Structure):
    """rhy_diag_t: the diagnostics of one box."""
    _fields_ = [
        ("nan_count", C.c_int64 * MAX_CMP),
        ("neg_rho", C.c_int64),
        ("neg_e_tot", C.c_int64),
        ("min_rho", C.c_double),
        ("min_e_tot", C.c_double),
        ("min_rho_at", C.c_int32 * 3),
        ("min_e_tot_at", C.c_int32 * 3),
        ("mass", C.c_double),
        ("energy", C.c_double),
        ("mass_masked", C.c_double),
        ("energy_masked", C.c_double),
    ]

    def as_dict(self, ncmp, ndim):
        return dict(nan_count=list(self.nan_count)[:ncmp], neg_rho=self.neg_rho,
                    neg_e_tot=self.neg_e_tot, min_rho=self.min_rho, min_e_tot=self.min_e_tot,
                    min_rho_at=list(self.min_rho_at)[:ndim],
                    min_e_tot_at=list(self.min_e_tot_at)[:ndim], mass=self.mass,
                    energy=self.energy, mass_masked=self.mass_masked,
                    energy_masked=self.energy_masked)


def make_ic(ic):
    """RhyIC from a dict of its fields (missing fields 0)."""
    out = RhyIC()
    for k, v in ic.items():
        setattr(out, k, v)
    return out


def make_drawing(d):
    """RhyDrawing from the dict layout of ics.draw; returns it with the
    arrays it points to (keep them alive while the call runs)."""
    import numpy as np
    shapes = d.get("shapes", [])
    harms = d.get("harmonics", [])
    sa = (RhyShape * max(1, len(shapes)))()
    for i, sh in enumerate(shapes):
        s = sa[i]
        s.kind = sh["kind"]
        for a in range(3):
            s.lo[a] = int(sh.get("lo", [0, 0, 0])[a])
            s.hi[a] = int(sh.get("hi", [0, 0, 0])[a])
            s.centre[a] = float(sh.get("centre", [0.0, 0.0, 0.0])[a])
            s.v[a] = float(sh["v"][a]) if a < len(sh["v"]) else 0.0
        s.radius = float(sh.get("radius", 0.0))
        s.rho = float(sh["rho"])
        s.p = float(sh["p"])
    ha = (RhyHarmonic * max(1, len(harms)))()
    profiles = []
    for i, h in enumerate(harms):
        prof = np.ascontiguousarray(h["profile"], dtype=np.float64)
        profiles.append(prof)
        x = ha[i]
        x.axis = int(h["axis"])
        x.profile = prof.ctypes.data_as(C.POINTER(C.c_double))
        x.rho = float(h["rho"])
        for a in range(3):
            x.v[a] = float(h["v"][a]) if a < len(h["v"]) else 0.0
        x.p = float(h["p"])
    out = RhyDrawing()
    out.rho = float(d["rho"])
    for a in range(3):
        out.v[a] = float(d["v"][a]) if a < len(d["v"]) else 0.0
    out.p = float(d["p"])
    out.nshapes = len(shapes)
    out.nharmonics = len(harms)
    out.shapes = C.cast(sa, C.POINTER(RhyShape))
    out.harmonics = C.cast(ha, C.POINTER(RhyHarmonic))
    return out, (sa, ha, profiles)


class RhyError(RuntimeError):
    pass


_lib = None
_by_path = {}


def load():
    """Load librhyme_hip.so and declare its signatures (no device needed)."""
    global _lib
    if _lib is not None:
        return _lib
    _lib = open_library(LIB_PATH)
    return _lib


class using:
    """Route every call made inside the block to another build of the library
    (e.g. one compiled with other build options), loaded beside the default
    one; handles made inside the block belong to that build and must be
    closed inside it."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self._saved = _lib
        _lib = open_library(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


def open_library(path):
    """A ctypes handle of the library at `path`, signatures declared."""
    if path in _by_path:
        return _by_path[path]
    if not os.path.exists(path):
        raise RhyError("%s is not built (run __graft_entry__.build())" % path)
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    pp = C.POINTER(RhyParams)
    sig = {
        "rhy_init": (C.c_int, [C.c_int32]),
        "rhy_last_error": (C.c_char_p, []),
        "rhy_version": (C.c_int32, []),
        "rhy_box_create": (C.c_int, [C.c_int32, ip, C.c_int32, C.c_int32, C.POINTER(vp)]),
        "rhy_box_resize": (C.c_int, [vp, ip]),
        "rhy_box_destroy": (C.c_int, [vp]),
        "rhy_box_upload": (C.c_int, [vp, dp]),
        "rhy_box_download": (C.c_int, [vp, dp]),
        "rhy_box_sync": (C.c_int, [vp]),
        "rhy_bc_set_boundaries": (C.c_int, [vp, pp]),
        "rhy_bc_set_face": (C.c_int, [vp, pp, C.c_int32]),
        "rhy_cfl_time_step": (C.c_int, [vp, pp, dp, dp]),
        "rhy_mh_solve": (C.c_int, [vp, pp, dp, C.c_double]),
        "rhy_run": (C.c_int, [vp, pp, dp, C.c_int32, dp]),
        "rhy_run_async": (C.c_int, [vp, pp, dp, C.c_int32]),
        "rhy_box_last_dts": (C.c_int, [vp, C.c_int32, dp]),
        "rhy_box_set_profiling": (C.c_int, [vp, C.c_int32]),
        "rhy_box_kernel_stats": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_double)]),
        "rhy_riemann_batch": (C.c_int, [pp, C.c_int64, dp, dp, ip, dp]),
        "rhy_limiter_batch": (C.c_int, [pp, C.c_int64, dp, dp, dp, dp]),
        "rhy_half_step_batch": (C.c_int, [pp, C.c_int64, dp, dp, ip, C.c_double,
                                          C.c_double, dp, dp]),
        "rhy_flux_batch": (C.c_int, [pp, C.c_int64, dp, ip, dp]),
        "rhy_box_paint": (C.c_int, [vp, C.POINTER(RhyIC)]),
        "rhy_box_draw": (C.c_int, [vp, C.POINTER(RhyParams), C.POINTER(RhyDrawing)]),
        "rhy_box_diagnostics": (C.c_int, [vp, C.POINTER(RhyDiag)]),
        "rhy_box_last_deferred": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "rhy_debug_defer": (C.c_int, [C.c_int32, C.c_int64]),
        "rhy_box_snapshot_pack": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "rhy_box_snapshot_unpack": (C.c_int, [vp, C.POINTER(C.c_float)]),
    }
    sp = C.POINTER(RhyRpSolution)
    sig.update({
        "rhy_build_options": (C.c_int32, []),
        "rhy_riemann_solve_batch": (C.c_int, [pp, C.c_int64, dp, dp, ip, dp, dp, dp]),
        "rhy_riemann_star_batch": (C.c_int, [pp, C.c_int64, sp, ip]),
        "rhy_riemann_sampling_batch": (C.c_int, [pp, C.c_int64, sp, ip, dp, dp, dp]),
        "rhy_riemann_region_batch": (C.c_int, [pp, C.c_int64, sp, ip, ip, dp, dp]),
        "rhy_irs_exec_batch": (C.c_int, [pp, C.c_int64, dp, dp, dp, dp, dp]),
        "rhy_irs_guess_batch": (C.c_int, [pp, C.c_int64, dp, dp, dp]),
        "rhy_pow_batch": (C.c_int, [C.c_int32, C.c_int64, dp, dp, dp]),
        "rhy_mh_solve_noalloc": (C.c_int, [pp, C.c_int32, ip, C.c_int32, C.c_int32, dp,
                                           C.c_double, vp, vp]),
    })
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _by_path[path] = L
    return L


class debug_defer:
    """Test hook (rhy_debug_defer) as a context manager: inside, the flux
    kernels also leave every `every`-th face (by low-cell offset) to the
    fixup kernel, and the deferred-face list holds at most `cap` entries
    (cap >= 0; past it the fixup recomputes every face)."""

    def __init__(self, every, cap=-1):
        self.every, self.cap = int(every), int(cap)

    def __enter__(self):
        check(load().rhy_debug_defer(self.every, self.cap), "rhy_debug_defer")
        return self

    def __exit__(self, *exc):
        check(load().rhy_debug_defer(0, -1), "rhy_debug_defer")
        return False


def check(rc, what):
    if rc != RHY_OK:
        msg = load().rhy_last_error()
        raise RhyError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def dx3(dx):
    out = np.ones(3, dtype=np.float64)
    dx = np.atleast_1d(np.asarray(dx, dtype=np.float64))
    out[:len(dx)] = dx
    return out


def dims3(dims):
    out = np.ones(3, dtype=np.int32)
    out[:len(dims)] = dims
    return out
