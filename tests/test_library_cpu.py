"""CPU-side checks of the product library: librhyme_hip.so loads, exports
every entry point include/rhyme_hip.h declares, the host mirror packs
rhy_params_t with the header's layout, and the product fails loudly (no CPU
fallback) when no gfx950 device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from rhyme_amd import _lib, hydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rhyme_hip.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rhy_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_what_binding_binds():
    assert sorted(_lib.EXPORTS) == declared()


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_params_layout_matches_header():
    # int32 x4, double x7, int32 x6, int32 ncmp, int32 reserved, double[6][16]
    assert C.sizeof(_lib.RhyParams) == 4 * 4 + 8 * 7 + 4 * 6 + 4 + 4 + 8 * 6 * 16
    assert _lib.RhyParams.cons_inflows.offset == 4 * 4 + 8 * 7 + 4 * 6 + 8
    import oracle
    assert C.sizeof(oracle.OrcParams) == C.sizeof(_lib.RhyParams)


def test_ic_layout_matches_header():
    # int32 kind, int32 axis, double x9, uint64 seed
    assert C.sizeof(_lib.RhyIC) == 4 + 4 + 8 * 9 + 8
    assert _lib.RhyIC.gamma.offset == 8 and _lib.RhyIC.seed.offset == 80
    txt = open(HEADER).read()
    assert "double radius, jitter, p0, energy;" in txt and "uint64_t seed;" in txt


def test_make_params_defaults_follow_reference():
    p = hydro.make_params(3)
    assert p.limiter == hydro.MINMOD and p.omega == 0.0        # slope_limiter_t
    assert p.solver_type == hydro.MEMORY_INTENSIVE             # muscl_hancock_t
    assert p.n_iteration == 10000 and p.tolerance == 1e-8      # irs_t
    assert p.courant == 0.8                                    # cfl_t
    assert list(p.bc_types) == [hydro.REFLECTIVE] * 6          # samr_bc_t
    assert p.gamma == 5.0 / 3.0 and p.ncmp == 9


def test_box_layout():
    assert hydro.box_shape(3, [4, 5, 6], 9) == (9, 10, 9, 8)
    assert hydro.box_shape(2, [4, 5], 8) == (8, 9, 8)
    assert hydro.box_shape(1, [7], 7) == (7, 11)


def test_fails_loudly_without_device():
    """No GPU in this container: creating a box must raise, not fall back."""
    L = _lib.load()
    if L.rhy_init(0) == _lib.RHY_OK:
        pytest.skip("a gfx950 device is present")
    with pytest.raises(_lib.RhyError):
        hydro.SamrBox(3, [8, 8, 8])
    p = hydro.make_params(3)
    with pytest.raises(_lib.RhyError):
        hydro.riemann_batch(p, np.ones((1, 5)), np.ones((1, 5)), np.ones(1, np.int32))
