"""rhyme_amd: MI355X (gfx950) implementation of Rhyme's hydro hot path.

The MUSCL-Hancock step on SAMR boxes (slope limiting, half-step
extrapolation, exact iterative Riemann solve, flux update, CFL dt and ghost
fill) runs in hand-written HIP kernels behind the C ABI of
include/rhyme_hip.h (librhyme_hip.so).  `rhyme_amd.hydro` mirrors the
reference's Fortran module interfaces on the host; `rhyme_amd.ics` builds the
seeded benchmark initial conditions.
"""
from . import _lib  # noqa: F401
from .hydro import *  # noqa: F401,F403
