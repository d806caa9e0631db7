"""The per-call entry points a host FEM code binds one at a time, against the oracle.

* ``opem_solve``  ~ solveComplexLinearSystem (solveComplexLinearSystem.c:100-209): lift, port elimination, GMRES
  preconditioned by a sparse LU of the caller's matrix, and the Jacobi / no-preconditioner retry rungs
  (:133-168);
* ``opem_hfield`` ~ solveHfield (:235-315) on the H mass matrix;
* ``opem_port_project`` ~ Mode::calculateSplits (port.cpp:3524-3568).

The matrices come from the oracle's assembly of a TE10 guide, so the comparisons are with the oracle's own
excitation (`oracle.solve.excite`), direct solve, H projection and splits; no formula is restated here.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from openparem_amd.mesh import waveguide_mesh
from openparem_amd.ports import te10
from oracle import assemble as oasm
from oracle import solve as oslv
from oracle.pipeline import Problem

pytestmark = pytest.mark.gpu

FREQ = 10e9


def _guide(p):
    """Two-port guide at 10 GHz: the PEC-eliminated A (what the reference hands solveComplexLinearSystem), port 1's
    projected TE10 values, and the oracle's x, h and splits for that excitation."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    pr = Problem(m, p)
    modes = {a: te10(m, a, FREQ) for a in pr.ports}
    A = oasm.system_matrix(pr.space, FREQ, {a: modes[a]["gamma"] for a in pr.ports}, Ke=pr.Ke, Me=pr.Me)
    P, Q = oasm.h_matrices(pr.space, FREQ, Me=pr.Me)
    drv = pr.ports[0]
    tr = pr.traces[drv]
    xm = tr.project(modes[drv]["E"](tr.pts), fixed_zero=np.intersect1d(tr.port_dofs, pr.pec))
    At, b = oslv.excite(A, pr.pec, tr.port_dofs, xm)
    x = oslv.direct_solve(At, b)
    h = oslv.h_field(P, Q, x)
    o = pr.solve_point(FREQ, modes)
    return dict(m=m, pr=pr, modes=modes, A1=oslv.eliminate(A, pr.pec), port_dofs=tr.port_dofs, xm=xm, At=At, b=b,
                x=x, h=h, P=P, Q=Q, o=o)


@pytest.fixture(scope="module")
def guide():
    return _guide(2)


@pytest.fixture(scope="module")
def guide1():
    return _guide(1)


@pytest.fixture(scope="module")
def dev():
    from openparem_amd.solver import Device
    d = Device()
    yield d
    d.close()


def _x0(g):
    x0 = np.zeros(g["A1"].shape[0], np.complex128)
    x0[g["port_dofs"]] = g["xm"]
    return x0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_solve_matches_direct(guide, dev, prec):
    """LU rung: the solution of the lifted, port-eliminated system equals the oracle's direct solve."""
    from openparem_amd.solver import solve
    g = guide
    x, (rn, it, cv, rung) = solve(dev, g["A1"], g["port_dofs"], _x0(g), precision=prec, tol=1e-12)
    print(f"opem_solve {prec}: iters {it} rnorm {rn:.3e} rung {rung}", flush=True)
    assert cv == 1 and rung == 0 and it <= (4 if prec == "fp64" else 20)
    bn = np.linalg.norm(g["b"])
    assert rn <= 1e-12 * bn
    assert abs(np.linalg.norm(g["b"] - g["At"] @ x) - rn) <= 1e-3 * rn + 1e-15 * bn
    assert np.abs(x - g["x"]).max() < 1e-10 * np.abs(g["x"]).max()
    assert np.array_equal(x[g["port_dofs"]], g["xm"])


@pytest.mark.parametrize("first_rung", [1, 2])
def test_solve_lower_rungs(guide1, dev, first_rung):
    """The Jacobi and unpreconditioned rungs reach the same solution on the ND1 guide; its restart (400) exceeds the
    count of rows that are neither PEC nor port, so GMRES is not restarted before it converges."""
    from openparem_amd.solver import solve
    g = guide1
    n = g["A1"].shape[0]
    assert n - len(g["pr"].pec) <= 400
    x, (rn, it, cv, rung) = solve(dev, g["A1"], g["port_dofs"], _x0(g), restart=400, maxit=4 * n, tol=1e-10,
                                  first_rung=first_rung)
    print(f"rung {first_rung}: iters {it} rnorm {rn:.3e}", flush=True)
    assert cv == 1 and rung == first_rung and it > 4
    assert np.abs(x - g["x"]).max() < 1e-7 * np.abs(g["x"]).max()


def test_solve_reports_nonconvergence(guide, dev):
    """Running out of iterations is a flag with the last iterate, not an error (solveComplexLinearSystem.c:204-206);
    the ladder stops there (KSP_DIVERGED_ITS is not a breakdown)."""
    from openparem_amd.solver import solve
    g = guide
    x, (rn, it, cv, rung) = solve(dev, g["A1"], g["port_dofs"], _x0(g), restart=10, maxit=5, first_rung=2)
    assert cv == 0 and it == 5 and rung == 2
    assert np.isfinite(x).all() and rn > 1e-10 * np.linalg.norm(g["b"])
    assert abs(np.linalg.norm(g["b"] - g["At"] @ x) - rn) <= 1e-6 * rn


def test_solve_without_ports_is_plain_solve(dev):
    """No Dirichlet rows and b = 0 give x = 0; with the lift taken away the entry is a plain sparse solve."""
    from openparem_amd.solver import solve
    rng = np.random.default_rng(3)
    n = 300
    A = sp.random(n, n, density=0.02, random_state=4, data_rvs=lambda k: rng.standard_normal(k) + 0j)
    A = (A + A.T + sp.diags(np.full(n, 8.0 + 1j))).tocsr()
    x, (rn, it, cv, rung) = solve(dev, A, np.zeros(0, np.int32), np.ones(n, np.complex128))
    assert cv == 1 and it == 0 and rn == 0.0 and np.abs(x).max() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hfield_matches_oracle(guide, dev, prec):
    """P h = j Q x with the oracle's x: the oracle's Jacobi-CG H field to its own tolerance."""
    from openparem_amd.solver import hfield
    g = guide
    b = 1j * (g["Q"] @ g["x"])
    h, (rn, it, cv, rung) = hfield(dev, g["P"], b, precision=prec, tol=1e-13)
    print(f"opem_hfield {prec}: iters {it} rnorm {rn:.3e}", flush=True)
    assert cv == 1 and rung == 0 and it <= 25
    assert rn <= 1e-13 * np.linalg.norm(b)
    assert np.abs(h - g["h"]).max() < 1e-10 * np.abs(g["h"]).max()


def test_port_project_matches_oracle_splits(guide):
    """C+, C-, and the four inner products of both ports for the oracle's x and h equal the oracle's splits."""
    from openparem_amd.solver import Solver
    g = guide
    pr = g["pr"]
    s = Solver(g["m"], 2)
    try:
        for k, a in enumerate(pr.ports):
            r = s.port_project(a, g["modes"][a], g["x"], g["h"])
            o = g["o"]
            assert abs(r["Cp"] - o["Cp"][0, k]) < 1e-11 * max(1.0, abs(o["Cp"][0, k]))
            assert abs(r["Cm"] - o["Cm"][0, k]) < 1e-11 * max(1.0, abs(o["Cm"][0, k]))
            tr = pr.traces[a]
            Es, Hs = tr.field(g["x"]), tr.field(g["h"])
            Cp, Cm = oslv.splits(tr, *_mode_fields(pr, g["modes"], a), Es, Hs)
            assert abs(r["Cp"] - Cp) < 1e-11 * max(1.0, abs(Cp)) and abs(r["Cm"] - Cm) < 1e-11 * max(1.0, abs(Cm))
            Em, Hm = _mode_fields(pr, g["modes"], a)
            for key, ref in (("e0", tr.inner(Em, Es)), ("e2", tr.inner(Em, Em)), ("h0", tr.inner(Hm, Hs)),
                             ("h2", tr.inner(Hm, Hm))):
                assert abs(r[key] - ref) <= 1e-11 * abs(ref), key
    finally:
        s.close()


def _mode_fields(pr, modes, a):
    """The oracle's modal fields at port a's quadrature points: E projected with the port's PEC DOFs held at zero,
    H projected freely (the oracle's solve_point)."""
    tr = pr.traces[a]
    n = pr.space.ndof
    xe = np.zeros(n, np.complex128)
    xe[tr.port_dofs] = tr.project(modes[a]["E"](tr.pts), fixed_zero=np.intersect1d(tr.port_dofs, pr.pec))
    xh = np.zeros(n, np.complex128)
    xh[tr.port_dofs] = tr.project(modes[a]["H"](tr.pts))
    return tr.field(xe), tr.field(xh)
