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
    x, (rn, it, cv, rung, flags) = solve(dev, g["A1"], g["port_dofs"], _x0(g), precision=prec, tol=1e-12)
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
    x, (rn, it, cv, rung, flags) = solve(dev, g["A1"], g["port_dofs"], _x0(g), restart=400, maxit=4 * n, tol=1e-10,
                                  first_rung=first_rung)
    print(f"rung {first_rung}: iters {it} rnorm {rn:.3e}", flush=True)
    assert cv == 1 and rung == first_rung and it > 4
    assert np.abs(x - g["x"]).max() < 1e-7 * np.abs(g["x"]).max()


def test_solve_reports_nonconvergence(guide, dev):
    """Running out of iterations is a flag with the last iterate, not an error (solveComplexLinearSystem.c:204-206);
    the ladder stops there (KSP_DIVERGED_ITS is not a breakdown)."""
    from openparem_amd.solver import solve
    g = guide
    x, (rn, it, cv, rung, flags) = solve(dev, g["A1"], g["port_dofs"], _x0(g), restart=10, maxit=5, first_rung=2)
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
    x, (rn, it, cv, rung, flags) = solve(dev, A, np.zeros(0, np.int32), np.ones(n, np.complex128))
    assert cv == 1 and it == 0 and rn == 0.0 and np.abs(x).max() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hfield_matches_oracle(guide, dev, prec):
    """P h = j Q x with the oracle's x: the oracle's Jacobi-CG H field to its own tolerance."""
    from openparem_amd.solver import hfield
    g = guide
    b = 1j * (g["Q"] @ g["x"])
    h, (rn, it, cv, rung, flags) = hfield(dev, g["P"], b, precision=prec, tol=1e-13)
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


def test_solve_flags_zero_pivots_of_a_singular_matrix(guide, dev):
    """A DOF whose row and column are zero makes the matrix singular: the factorization meets an exactly zero pivot
    and reports it (OPEM_SOLVE_SINGULAR), the system stays consistent (its b entry is zero), and the other DOFs
    equal the oracle's direct solve with that DOF removed."""
    from openparem_amd.solver import solve
    from oracle import solve as oslv
    g = guide
    A = g["A1"].tolil()
    n = A.shape[0]
    free = np.setdiff1d(np.arange(n), np.concatenate([g["pr"].pec, g["port_dofs"]]))
    k = int(free[len(free) // 2])
    A[k, :] = 0
    A[:, k] = 0
    A = A.tocsr()
    A.eliminate_zeros()
    # keep the pattern structurally symmetric with an explicit zero diagonal
    A = (A + sp.csr_matrix(([0.0j], ([k], [k])), shape=A.shape)).tocsr()
    x, (rn, it, cv, rung, flags) = solve(dev, A, g["port_dofs"], _x0(g), tol=1e-12)
    print(f"singular: iters {it} rnorm {rn:.3e} rung {rung} flags {flags}", flush=True)
    assert cv == 1 and rung == 0 and flags & 2
    At, b = oslv.excite(g["A1"], np.union1d(g["pr"].pec, [k]), g["port_dofs"], g["xm"])
    xo = oslv.direct_solve(At, b)
    mask = np.ones(n, bool)
    mask[k] = False
    assert np.abs(x[mask] - xo[mask]).max() < 1e-9 * np.abs(xo).max()


STAG_CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, "tests")
from test_gpu_seams import _guide, _x0
from openparem_amd.mesh import waveguide_mesh
from openparem_amd.solver import Device, Solver, solve
g = _guide(2)
d = Device()
x, (rn, it, cv, rung, flags) = solve(d, g["A1"], g["port_dofs"], _x0(g), precision="mixed", tol=1e-12)
d.close()
m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
with Solver(m, 2) as s:
    r = s.sweep(np.array([9e9, 11e9]), precision="mixed", tol=1e-11, batch=2, restart=30, precond="lu",
                orth_passes=1)
print(json.dumps({"x_err": float(np.abs(x - g["x"]).max() / np.abs(g["x"]).max()), "flags": flags, "cv": cv,
                  "it": it, "S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()],
                  "fallback": r["fallback"].tolist(), "converged": r["converged"].tolist(),
                  "iters": r["iters"].tolist()}), flush=True)
"""


def test_fp32_stagnation_hands_over_to_fp64():
    """Host-forced stagnation of the FP32 refinement (OPEM_FORCE_STAGNATION marks every system after its first
    cycle), in a child process: opem_solve and the mixed sweep finish every system with FP64 GMRES, report it
    (OPEM_SOLVE_FP64_FALLBACK, sweep `fallback`), converge, and give the oracle's x and the FP64 path's S."""
    import json
    import os
    import subprocess
    import sys
    from openparem_amd.solver import Solver
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    with Solver(m, 2) as s:
        r64 = s.sweep(np.array([9e9, 11e9]), precision="fp64", tol=1e-12, batch=1, restart=30, precond="lu")
    env = dict(os.environ, OPEM_FORCE_STAGNATION="1")
    env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    p = subprocess.run([sys.executable, "-c", STAG_CHILD], cwd=root, env=env, capture_output=True, text=True,
                       timeout=200)
    assert p.returncode == 0, p.stderr[-2000:]
    d = json.loads(p.stdout.strip().splitlines()[-1])
    print(f"forced stagnation: solve flags {d['flags']} iters {d['it']} x err {d['x_err']:.2e}; sweep fallback "
          f"{d['fallback']} iters {d['iters']}", flush=True)
    assert d["cv"] == 1 and d["flags"] & 1 and d["x_err"] < 1e-9
    assert np.all(np.array(d["fallback"]) == 1) and np.all(np.array(d["converged"]) == 1)
    S = np.array([complex(a, b) for a, b in d["S"]]).reshape(r64["S"].shape)
    assert np.abs(S - r64["S"]).max() < 1e-9
