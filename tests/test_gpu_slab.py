"""Two materials in one guide: a slab across the whole WR90 cross-section between air sections (teflon, er 2.2;
lossy teflon, tan delta 0.05; a magnetic slab, mu_r 2).

The GPU sweep (four frequencies in one batch, FP64 and mixed) is checked against the oracle's direct solve at
9 GHz and against the transmission-line closed form at every frequency (tests/test_oracle.py pins the oracle
to that closed form; the remaining difference is discretization error, 2.1e-3 to 2.6e-3 at ND2 on this mesh over
8-11 GHz for teflon, where |S11| falls from 0.60 to the slab's half-wave transparency near 11 GHz).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

pytestmark = pytest.mark.gpu

NX, NY, NZ, L, A = 6, 3, 18, 30e-3, 22.86e-3


@pytest.mark.parametrize("er2,tand2,mur2", [(2.2, 0.0, 1.0), (2.2, 0.05, 1.0), (1.0, 0.0, 2.0)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-8)])
def test_slab_sweep(precision, tol, er2, tand2, mur2):
    from openparem_amd.mesh import waveguide_mesh
    from openparem_amd.ports import te10
    from openparem_amd.solver import Solver
    from oracle.pipeline import Problem
    from test_oracle import slab_closed_form

    print(f"slab sweep {precision} er {er2} tand {tand2} mur {mur2}", flush=True)
    m = waveguide_mesh(nx=NX, ny=NY, nz=NZ, seed=3, slab=(6, 12, er2, tand2, mur2))
    freqs = np.array([8e9, 9e9, 10e9, 11e9])
    s = Solver(m, 2)
    try:
        r = s.sweep(freqs, precision=precision, tol=1e-12 if precision == "fp64" else 1e-10, restart=150,
                    batch=len(freqs))
    finally:
        s.close()
    assert r["converged"].all(), (r["converged"], r["iters"])
    pr = Problem(m, 2)
    o = pr.solve_point(9e9, {q: te10(m, q, 9e9) for q in pr.ports})
    dS = np.abs(r["S"][1] - o["S"]).max()
    print(f"|dS| vs oracle at 9 GHz {dS:.2e}", flush=True)
    assert dS < tol
    zs = np.linspace(0, L, NZ + 1)
    for i, f in enumerate(freqs):
        s11, s21 = slab_closed_form(f, A, L, er2, zs[6], zs[12], tand2, mur2)
        S = r["S"][i]
        e = max(abs(S[0, 0] - s11), abs(S[1, 0] - s21), abs(S[1, 1] - s11))
        print(f"  {f / 1e9:.0f} GHz: |S11| {abs(S[0, 0]):.4f} (closed form {abs(s11):.4f}), |dS| {e:.2e}", flush=True)
        assert e < 5e-3
