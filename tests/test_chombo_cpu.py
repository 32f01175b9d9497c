"""Snapshot payload restatement (oracle/chombo.py) and the host-side Chombo
pieces (rhyme_amd/chombo.py), on the CPU."""
import numpy as np
import pytest

from oracle import chombo as och
from rhyme_amd import chombo, hydro, ics


@pytest.mark.parametrize("ndim,dims_list", [(1, [[7], [3]]), (2, [[5, 3], [2, 4]]),
                                            (3, [[3, 2, 4], [2, 2, 2]])])
def test_level_data_element_order(ndim, dims_list):
    """Each value sits where Fortran's reshape of cells(1:nx,1:ny,1:nz,v)
    puts it: offset of box b + (v-1)*blen + (i-1) + nx*((j-1) + ny*(k-1))
    (rhyme_chombo_write_level_data.f90:60-69), by an explicit index loop."""
    ncmp = hydro.ncmp_for(ndim)
    g = np.random.default_rng(ndim)
    boxes = [g.uniform(-3, 3, size=hydro.box_shape(ndim, d, ncmp)) for d in dims_list]
    data = och.level_data(ndim, boxes, ncmp)
    ofs = 0
    for cells, d in zip(boxes, dims_list):
        d3 = list(d) + [1] * (3 - ndim)
        blen = d3[0] * d3[1] * d3[2]
        for v in range(ncmp):
            for k in range(1, d3[2] + 1):
                for j in range(1, d3[1] + 1):
                    for i in range(1, d3[0] + 1):
                        idx = (i + 1, j + 1, k + 1)[:ndim]
                        want = np.float32(cells[(v,) + tuple(reversed(idx))])
                        got = data[ofs + v * blen + (i - 1) + d3[0] * ((j - 1) + d3[1] * (k - 1))]
                        assert got == want and got.dtype == np.float32
        ofs += ncmp * blen
    assert ofs == data.size


def test_binary32_rounding_known_answers():
    """real() of a real(8): IEEE binary32, round to nearest, ties to even."""
    x = np.array([1.0 + 2.0 ** -24,             # tie, even neighbour is 1
                  1.0 + 3 * 2.0 ** -24,         # tie, even neighbour is 1 + 2^-22
                  1.0 + 2.0 ** -24 + 2.0 ** -40,  # just above the tie: up
                  0.1, -2.5e-45, 1e-46, 3.4028234663852886e38 + 1e30, -1e39, 0.0, -0.0])
    cells = np.zeros(hydro.box_shape(1, [x.size], hydro.ncmp_for(1)))
    cells[0, 2:-2] = x
    got = och.level_data(1, [cells], hydro.ncmp_for(1))[:x.size].view(np.uint32)
    want = np.array([0x3F800000, 0x3F800002, 0x3F800001, 0x3DCCCCCD, 0x80000002, 0x00000000,
                     0x7F7FFFFF, 0xFF800000, 0x00000000, 0x80000000], dtype=np.uint32)
    # FLT_MAX + 1e30 is below FLT_MAX + half an ulp (2^103): rounds to FLT_MAX;
    # -1e39 overflows to -inf; 1e-46 is below half the smallest subnormal: +0
    np.testing.assert_array_equal(got, want)


def test_round_trip_and_split():
    ndim, dims = 2, [[6, 4], [3, 5], [8, 2]]
    ncmp = hydro.ncmp_for(ndim)
    boxes = [ics.random_smooth(ndim, d, seed=i) for i, d in enumerate(dims)]
    data = och.level_data(ndim, boxes, ncmp)
    back = och.load_level(ndim, ncmp, dims, data)
    for cells, b in zip(boxes, back):
        inner = cells[(slice(None),) + (slice(2, -2),) * ndim]
        np.testing.assert_array_equal(b, inner.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(och.level_data(ndim, [np.pad(b, ((0, 0),) + ((2, 2),) * ndim)
                                                        for b in back], ncmp), data)
    left = [[1, 1], [7, 1], [1, 5]]
    table = chombo.boxes_table(ndim, left, dims)
    np.testing.assert_array_equal(table, [[0, 0, 0, 5, 3, 0], [6, 0, 0, 8, 4, 0],
                                          [0, 4, 0, 7, 5, 0]])
    parts = chombo.split_level(ndim, table, data)
    assert [p[0] for p in parts] == left and [p[1] for p in parts] == dims
    assert sum(p[2].size for p in parts) == data.size
    with pytest.raises(ValueError):
        chombo.split_level(ndim, table, data[:-1])
    with pytest.raises(ValueError):
        chombo.split_level(ndim, table, np.concatenate([data, data[:1]]))


def test_headers_and_labels():
    """write_headers.f90:15-64 and the labels of rhyme_chombo.f90:143-150."""
    assert [l for _, l in chombo.component_labels(2)] == [
        "rho", "rho_u", "rho_v", "e_tot", "temp", "ntr_frac_0", "ntr_frac_1", "ntr_frac_2"]
    assert chombo.component_labels(3)[4] == ("component_4", "e_tot")
    assert len(chombo.component_labels(1)) == hydro.ncmp_for(1)
    h = chombo.headers(2, [64, 32], [[1 / 64, 1 / 32], [1 / 128, 1 / 64]], [2, 2], 17, 0.25,
                       {"length_unit": "m"})
    np.testing.assert_array_equal(h["/level_1"]["dx"], [1 / 128, 1 / 64, 0.0])
    assert h["/level_0"]["prob_domain"] == dict(lo_i=0, lo_j=0, lo_k=0, hi_i=63, hi_j=31, hi_k=1)
    np.testing.assert_array_equal(h["/"]["ProblemDomain"], [64, 32, 1])
    assert h["/"]["num_levels"] == 2 and h["/"]["num_components"] == 8
    assert h["/"]["component_2"] == "rho_v" and h["/"]["iteration"] == 17
    assert h["/Chombo_global"] == {"SpaceDim": 3, "length_unit": "m"}
