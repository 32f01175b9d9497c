"""CPU restatement of the reference's Chombo level payload and its restart.

TEST INFRASTRUCTURE ONLY, like the rest of oracle/.

* level_data: rhyme_chombo_write_level_data
  (src/rhyme_chombo/src/rhyme_chombo_write_level_data.f90:41-71): for box b,
  for component v, real(reshape(cells(1:nx,1:ny,1:nz,v), [nx*ny*nz])) -- the
  interior in Fortran element order (i fastest), converted to default real
  (IEEE binary32, round to nearest even), appended box after box.
* load_level: rhyme_initial_condition_load_rhyme
  (src/rhyme_initial_condition/src/rhyme_initial_condition_load_rhyme.f90:57-76):
  the same slices reshaped to the box dims and widened to real(8).

Pinned by tests/test_chombo_cpu.py: element order against an explicit
column-major index loop, the binary32 rounding against known answers (ties
to even, overflow, subnormals), and the round trip.
"""
import numpy as np


def _interior(ndim, cells):
    return cells[(slice(None),) + (slice(2, -2),) * ndim]


def level_data(ndim, boxes_cells, ncmp):
    """The FP32 payload of a level from its boxes' host arrays
    (ncmp, [z,] [y,] x) with 2 ghost layers."""
    parts = []
    for cells in boxes_cells:
        inner = _interior(ndim, cells)
        for v in range(ncmp):
            # C order over (z, y, x) is Fortran order over (i, j, k)
            with np.errstate(over="ignore"):      # beyond FLT_MAX: +-inf, as real() gives
                parts.append(np.ascontiguousarray(inner[v]).ravel().astype(np.float32))
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.float32)


def load_level(ndim, ncmp, dims_list, data):
    """Interiors (ncmp, [z,] [y,] x) in float64 from a level payload."""
    out, ofs = [], 0
    for dims in dims_list:
        blen = int(np.prod(dims))
        shape = tuple(reversed(list(dims)))
        arr = np.empty((ncmp,) + shape)
        for v in range(ncmp):
            arr[v] = data[ofs + v * blen: ofs + (v + 1) * blen].astype(np.float64).reshape(shape)
        out.append(arr)
        ofs += ncmp * blen
    return out
