"""The host mirror of the device drawing (ics.draw), pinned on small grids
against a direct cell-by-cell restatement of rhyme_drawing's operations:
sphere membership of cell centres q - 1/2, inclusive 1-based cuboid ranges,
later shapes over earlier ones, and the harmonic perturbation added last in
conserved form (rhyme_drawing_apply_perturbations.f90:51-60, 115, 184)."""
import math

import numpy as np

from rhyme_amd import ics


def loop_draw(d, ndim, dims, gamma):
    """Cell-by-cell: colour from the canvas and the shapes, then the
    perturbation, each conserved value formed as conv_prim_to_cons does."""
    def cons(w):
        s = 0.0
        for a in range(ndim):
            s = s + w[1 + a] * w[1 + a]
        return [w[0]] + [w[0] * w[1 + a] for a in range(ndim)] + \
            [.5 * w[0] * s + w[-1] / (gamma - 1.0)]

    def colour(c):
        return [c["rho"]] + list(c["v"])[:ndim] + [c["p"]]
    out = ics.draw(d, ndim, dims, gamma) * 0.0
    for idx in np.ndindex(*dims[::-1]):
        q = [i + 1 for i in idx[::-1]]  # 1-based x, y, z
        w = colour(d)
        for sh in d["shapes"]:
            if sh["kind"] == ics.SPHERE:
                r2 = sum((q[a] - 0.5 - sh["centre"][a]) ** 2 for a in range(ndim))
                inside = r2 <= sh["radius"] ** 2
            else:
                inside = all(sh["lo"][a] <= q[a] <= sh["hi"][a] for a in range(ndim))
            if inside:
                w = colour(sh)
        u = cons(w)
        if d["harmonics"]:
            h = [0.0] * (ndim + 2)
            for hm in d["harmonics"]:
                a = hm["profile"][q[hm["axis"]] - 1]
                base = colour(hm)
                h = [h[c] + a * base[c] for c in range(ndim + 2)]
            up = cons(h)
            u = [u[c] + up[c] for c in range(ndim + 2)]
        for c in range(ndim + 2):
            out[(c,) + tuple(i + 2 for i in idx)] = u[c]
    return out


def test_draw_matches_cell_loop():
    g = np.random.default_rng(3)
    for ndim, dims in ((1, [13]), (2, [9, 7]), (3, [6, 5, 7])):
        shapes = [dict(kind=ics.SPHERE, centre=[3.0, 2.5, 3.5], radius=2.0, rho=2.0,
                       v=[0.5, -0.5, 0.25], p=3.0),
                  dict(kind=ics.CUBOID, lo=[2, 1, 3], hi=[4, 3, 5], rho=0.5, v=[0.0, 1.0, 0.0],
                       p=0.25),
                  dict(kind=ics.SPHERE, centre=[1.0, 1.0, 1.0], radius=1.5, rho=4.0,
                       v=[0.0, 0.0, 0.0], p=0.1)]
        harms = [dict(axis=a, profile=g.uniform(-0.1, 0.1, dims[a]), rho=0.5, v=[0.1, 0.2, 0.3],
                      p=1.0) for a in range(ndim)]
        d = dict(rho=1.0, v=[0.1, 0.0, -0.2], p=1.5, shapes=shapes, harmonics=harms)
        for gamma in (5.0 / 3.0, 7.0 / 5.0):
            np.testing.assert_array_equal(ics.draw(d, ndim, dims, gamma),
                                          loop_draw(d, ndim, dims, gamma))


def test_harmonic_on_pressure_adds_to_energy_only():
    """With base (0, 0, p = 1) the perturbation adds A cos(kx) / (gamma - 1)
    to e_tot and leaves rho and the momenta as painted."""
    n, gamma = 64, 5.0 / 3.0
    d = ics.spheres_2d_drawing(n, seed=4, nspheres=3)
    plain = ics.draw(dict(d, harmonics=[]), 2, [n, n], gamma)
    pert = ics.draw(d, 2, [n, n], gamma)
    np.testing.assert_array_equal(pert[:3], plain[:3])
    prof = d["harmonics"][0]["profile"]
    np.testing.assert_array_equal(pert[3, 2:-2, 2:-2],
                                  plain[3, 2:-2, 2:-2] + (prof / (gamma - 1.0))[None, :])


def test_harmonic_profile_is_the_references_cosine():
    """kx = 2 pi / lambda * (q - 1/2) (:115), A cos(kx), against math.cos."""
    prof = ics.harmonic_profile(100, 0.05, 32.0)
    for q in range(1, 101):
        want = 0.05 * math.cos(2.0 * math.pi / 32.0 * (q - 0.5))
        assert abs(prof[q - 1] - want) <= 1e-17
    # a period of 32 cells: q and q + 32 sit at the same phase
    np.testing.assert_allclose(prof[:68], prof[32:], rtol=0, atol=1e-15)
