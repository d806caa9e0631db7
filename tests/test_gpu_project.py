"""The reference's WR90 straight_study project read from its own files (mesh, ports file, project, materials
database) and swept on the GPU: the same S as the hand-built mesh of the same file (tests/golden/wr90.py) and the
reference's results.ods S at the discretization level of the analytic port mode."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

pytestmark = pytest.mark.gpu

REF_TOL = {2: 0.03, 3: 0.005}


@pytest.mark.parametrize("p", [1, 2, 3])
def test_ingested_project_sweep(p):
    import wr90
    from openparem_amd.project import load_project
    from openparem_amd.solver import Solver

    proj = load_project(os.path.join(HERE, "golden", "wr90_straight_study", "straight_order_2.proj"))
    print(f"ingested straight project ND{p}", flush=True)
    s = Solver(proj.mesh, p, boundaries=proj.boundary_values)
    try:
        r = s.sweep(proj.frequencies, precision="fp64", tol=1e-12, restart=30, precond="lu")
    finally:
        s.close()
    m, d = wr90.load()
    s = Solver(m, p)
    try:
        r0 = s.sweep(proj.frequencies, precision="fp64", tol=1e-12, restart=30, precond="lu")
    finally:
        s.close()
    assert r["converged"].all() and r0["converged"].all()
    dS = np.abs(r["S"] - r0["S"]).max()
    print(f"|dS| vs the hand-built mesh {dS:.2e}", flush=True)
    assert dS < 1e-10
    if p in REF_TOL:
        s11, s21 = wr90.reference_s(d, p)
        assert abs(r["S"][0][0, 0] - s11) < REF_TOL[p] and abs(r["S"][0][1, 0] - s21) < REF_TOL[p]
