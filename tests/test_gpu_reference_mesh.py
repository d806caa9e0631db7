"""The GPU sweep on the reference's own mesh: regression/OpenParEM3D/WR90/straight_study/straight.msh (30 tets),
committed as tests/golden/wr90_straight.json, at 9 GHz and ND1-ND3.

S is checked against the oracle's direct solve on the same mesh (the solve tolerance) and against the
reference's own S from results.ods (sheet p, row 1: the unrefined mesh) at the discretization level that the
oracle itself meets there (tests/test_oracle.py::test_reference_s_on_reference_mesh): the reference drives with
its discrete 2D eigenmode and integrates port overlaps on an L2_p projection, here the mode is the analytic TE10.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

pytestmark = pytest.mark.gpu

REF_TOL = {2: 0.03, 3: 0.005}


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-8)])
def test_reference_mesh_s(p, precision, tol):
    import wr90
    from openparem_amd.ports import te10
    from openparem_amd.solver import Solver
    from oracle.pipeline import Problem

    print(f"reference mesh p={p} {precision}", flush=True)
    m, d = wr90.load()
    f = d["freq"]
    s = Solver(m, p)
    try:
        assert s.ndof == d["dof_counts"][str(p)]
        r = s.sweep(np.array([f]), precision=precision, tol=1e-12 if precision == "fp64" else 1e-10, restart=300)
    finally:
        s.close()
    assert r["converged"].all(), (r["converged"], r["iters"])
    pr = Problem(m, p)
    o = pr.solve_point(f, {a: te10(m, a, f) for a in pr.ports})
    dS = np.abs(r["S"][0] - o["S"]).max()
    print(f"|dS| vs oracle {dS:.2e}", flush=True)
    assert dS < tol, (dS, r["S"][0], o["S"])
    if p in REF_TOL:
        s11, s21 = wr90.reference_s(d, p)
        e11, e21 = abs(r["S"][0][0, 0] - s11), abs(r["S"][0][1, 0] - s21)
        print(f"vs results.ods: |dS11| {e11:.2e} |dS21| {e21:.2e}", flush=True)
        assert e11 < REF_TOL[p] and e21 < REF_TOL[p]
