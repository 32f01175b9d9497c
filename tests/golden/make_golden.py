#!/usr/bin/env python3
"""Extract the known-answer data that the reference's own tests hold for the
hydro hot path into small fixtures under tests/golden/.

Inputs (read from /root/reference, data files only; no reference code runs):
  src/rhyme_riemann_problem/src/rhyme_irs/tests/assets/p_stars.txt
      rows (rho_l, u_l, p_l, rho_r, u_r, p_r, p_star) as read by
      rhyme_irs_exec_test.f90:46 (the file's header labels are mislabeled;
      the test reads the left state first).  Kept: the first 1000 rows (what
      the reference test checks) and every 7th row after them.
  src/rhyme_riemann_problem/tests/sod_shock_tube_t_0_2_analytical.txt
      500 samples (x, rho, u, p) of the Sod problem at t = 0.2 (gamma 1.4),
      read by rhyme_riemann_problem_sampling_test.f90:40-66.
  tests/Sod-s-shock/analytical_solutions/sod_s_shock_{10,20,30,40}.csv
      256 samples (x, rho, vx, p) of the 1D Sod tube at t = 0.1..0.4 for the
      configuration tests/Sod-s-shock/sod_s_shock_1d/sod_s_shock_1d.conf.
The Toro test values transcribed from
src/rhyme_riemann_problem/tests/rhyme_irs_tests_factory.f90:11-184 are
written to toro5.json.
"""
import json
import os

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def p_stars():
    path = os.path.join(REF, "src/rhyme_riemann_problem/src/rhyme_irs/tests/assets/p_stars.txt")
    rows = np.loadtxt(path, delimiter=",", skiprows=1)
    keep = np.concatenate([rows[:1000], rows[1000::7]])
    np.savetxt(os.path.join(HERE, "p_stars_subset.csv"), keep, delimiter=",",
               header="rho_l,u_l,p_l,rho_r,u_r,p_r,p_star (first 1000 rows, then every 7th)",
               fmt="%.17g")
    return len(keep)


def sod_sampling():
    path = os.path.join(REF, "src/rhyme_riemann_problem/tests/sod_shock_tube_t_0_2_analytical.txt")
    rows = np.loadtxt(path, skiprows=12)
    np.savetxt(os.path.join(HERE, "sod_t0.2_sampling.csv"), rows, delimiter=",",
               header="x,rho,u,p (Sod, gamma 1.4, t = 0.2; sampled at x in [-0.5, 0.5])",
               fmt="%.17g")
    return len(rows)


def sod_tube():
    n = 0
    for t in (10, 20, 30, 40):
        path = os.path.join(REF, "tests/Sod-s-shock/analytical_solutions/sod_s_shock_%d.csv" % t)
        rows = np.loadtxt(path, delimiter=",", skiprows=1)
        np.savetxt(os.path.join(HERE, "sod_tube_t0.%d.csv" % (t // 10)), rows, delimiter=",",
                   header="x,rho,vx,p (1D Sod tube on [0,1], gamma 1.4, t = 0.%d)" % (t // 10),
                   fmt="%.17g")
        n += len(rows)
    return n


def toro5():
    # (name, sig figs, left (rho, u, p), right (rho, u, p), p*, u*,
    #  left is_shock, left rho*, right is_shock, right rho*)
    cases = [
        ("Sod", 6, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745,
         False, 0.42632, True, 0.26557),
        ("123", 4, (1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.0,
         False, 0.02185, False, 0.02185),
        ("left blast wave", 6, (1.0, 0.0, 1e3), (1.0, 0.0, 1e-2), 460.894, 19.5975,
         False, 0.57506, True, 5.99924),
        ("right blast wave", 6, (1.0, 0.0, 1e-2), (1.0, 0.0, 1e2), 46.0950, -6.19633,
         True, 5.99242, False, 0.57511),
        ("two shocks collision", 3, (5.99924, 19.5975, 460.894), (5.99924, -6.19633, 46.0950),
         1691.64, 8.68975, True, 14.2823, True, 31.0426),
    ]
    out = [dict(name=c[0], sig_figs=c[1], left=c[2], right=c[3], p_star=c[4], u_star=c[5],
                left_is_shock=c[6], left_rho_star=c[7], right_is_shock=c[8],
                right_rho_star=c[9]) for c in cases]
    json.dump(dict(gamma="diatomic (7/5)", source="rhyme_irs_tests_factory.f90:11-184",
                   cases=out), open(os.path.join(HERE, "toro5.json"), "w"), indent=1)
    return len(out)


if __name__ == "__main__":
    print("p_stars rows", p_stars())
    print("sod sampling rows", sod_sampling())
    print("sod tube rows", sod_tube())
    print("toro cases", toro5())
