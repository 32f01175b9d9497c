"""Chombo snapshot pieces around the device payload.

The reference writes a snapshot with rhyme_chombo_write_samr
(src/rhyme_chombo/src/rhyme_chombo_write_samr.f90:3-33): headers
(rhyme_chombo_write_headers.f90:15-64), then per level one FP32 dataset
"data:datatype=0" and a "boxes" table (rhyme_chombo_write_level_data.f90:41-71).
A restart reads them back (rhyme_initial_condition_load_rhyme.f90:20-76).

The dataset payload is produced and consumed on the device
(SamrBox.snapshot / SamrBatch.snapshot and load_snapshot, C:
rhy_*_snapshot_pack / _unpack).  This module holds the host-side rest: the
boxes table, the header attributes and the split of a level payload into
boxes.  The HDF5 container itself (files, groups, attribute and dataset
writes) belongs to the host's HDF5 library, which this image does not have.
"""
import numpy as np

from .hydro import ncmp_for

BOXES_HEADERS = ("lo_i", "lo_j", "lo_k", "hi_i", "hi_j", "hi_k")


def component_labels(ndim, nspe=3):
    """(attribute name, label) per component: component_<i-1> with the
    reference's cid labels (rhyme_units.f90:45-53, rhyme_chombo.f90:143-150)."""
    names = ["rho", "rho_u", "rho_v", "rho_w"][:1 + ndim] + ["e_tot", "temp"]
    names += ["ntr_frac_%d" % s for s in range(nspe)]
    return [("component_%d" % i, n) for i, n in enumerate(names)]


def boxes_table(ndim, left_edges, dims):
    """The level's "boxes" table: int32 [nboxes][6], (left edge - 1) then
    (right edge - 1), zero-padded to three axes (write_level_data.f90:55-56)."""
    t = np.zeros((len(dims), 6), dtype=np.int32)
    for b, (le, d) in enumerate(zip(left_edges, dims)):
        for a in range(ndim):
            t[b, a] = le[a] - 1
            t[b, 3 + a] = le[a] + d[a] - 2
    return t


def headers(ndim, base_grid, level_dx, ref_ratios, iteration, time, units=None, nspe=3):
    """Attribute values the reference writes, by group (write_headers.f90:15-64).
    level_dx / ref_ratios: one entry per level.  units: strings of the
    '/Chombo_global' unit attributes (length_unit, density_unit, ...)."""
    out = {}
    for l, (dx, r) in enumerate(zip(level_dx, ref_ratios)):
        out["/level_%d" % l] = {"dx": np.array(list(dx)[:ndim] + [0.0] * (3 - ndim)),
                                "ref_ratio": int(r)}
    shape = [0, 0, 0] + [g - 1 for g in list(base_grid)[:ndim]] + [1] * (3 - ndim)
    out["/level_0"]["prob_domain"] = dict(zip(BOXES_HEADERS, shape))
    root = {"ProblemDomain": np.array(list(base_grid)[:ndim] + [1] * (3 - ndim), dtype=np.int32),
            "num_levels": len(level_dx), "num_components": ncmp_for(ndim, nspe)}
    root.update(dict(component_labels(ndim, nspe)))
    root.update({"iteration": int(iteration), "time": float(time)})
    out["/"] = root
    glob = {"SpaceDim": 3}
    glob.update(units or {})
    out["/Chombo_global"] = glob
    return out


def split_level(ndim, table, data, nspe=3):
    """A level payload cut into its boxes, from the boxes table as the
    restart reads it (load_rhyme.f90:37-76): per box (left edges, dims,
    payload slice of ncmp * prod(dims) values)."""
    ncmp = ncmp_for(ndim, nspe)
    out, ofs = [], 0
    for row in np.asarray(table):
        dims = [int(row[3 + a] - row[a] + 1) for a in range(ndim)]
        n = ncmp * int(np.prod(dims))
        if ofs + n > len(data):
            raise ValueError("payload shorter than its boxes table")
        out.append(([int(row[a]) + 1 for a in range(ndim)], dims, data[ofs:ofs + n]))
        ofs += n
    if ofs != len(data):
        raise ValueError("payload longer than its boxes table")
    return out
