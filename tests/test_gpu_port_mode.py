"""opem_port_mode (the 2D port eigenproblem on the GPU path) against the oracle (oracle/port_mode.py) and the
reference's own logged OpenParEM2D results, and the 3D S it drives.

* The reference's straight_study mesh (30 tets, ports of 4 triangles) at 9 GHz, ND1-ND3: gamma equals the logged
  beta/1000 to its 8 digits and the oracle's eigenvalue to 1e-12; E_t and H_t at the port points equal the
  oracle's fields up to one complex scale shared by E and H; Z_PV along the centre line matches the log at ND1.
* The 3D sweep driven by these modes equals the oracle's direct solve driven by the oracle's modes (1e-9), and is
  closer to the reference's own S (results.ods) than the analytic TE10 drive: with the discrete port mode the
  remaining difference is the reference's L2-projected overlap integrals (port.cpp:3482-3486).
* A lossy shielded microstrip (substrate tan d = 0.02, a surface-impedance trace crossing the port): gamma and
  the fields against the oracle.
"""
import os
import sys

import numpy as np
import pytest

from openparem_amd.mesh import microstrip_mesh
from oracle import assemble as oasm
from oracle.pipeline import Problem
from oracle.port_mode import PortMode

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

pytestmark = pytest.mark.gpu

LOGGED_BETA_K = {1: 0.10038888, 2: 0.12939113, 3: 0.12915073}


def _scale_fit(a, b):
    """Complex c minimising |a - c b| and the relative residual."""
    c = np.vdot(b, a) / np.vdot(b, b)
    return c, np.abs(a - c * b).max() / np.abs(a).max()


def _oracle_at_points(pm, r, pts):
    return pm.eval_at(r["et"], pts), pm.eval_at(r["ht"], pts)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_port_mode_reference_mesh(p):
    import wr90
    from openparem_amd.solver import Solver
    m, d = wr90.load()
    f = d["freq"]
    sp_ = oasm.Space(m, p)
    with Solver(m, p) as s:
        for attr in (2, 3):
            zc = m.xyz[:, 2].max() if attr == 2 else m.xyz[:, 2].min()
            path = (np.array([d["a"] / 2, 0, zc]), np.array([d["a"] / 2, d["b"], zc]))
            md = s.port_mode(attr, f, path)
            pm = PortMode(sp_, attr)
            r = pm.solve(f)
            print(f"p={p} port {attr}: gamma {md['gamma']} oracle {r['gamma']} Z {md['Z']}", flush=True)
            assert round(md["gamma"].imag / 1000, 8) == LOGGED_BETA_K[p]
            assert abs(md["gamma"] - r["gamma"]) < 1e-12 * abs(r["gamma"])
            Eo, Ho = _oracle_at_points(pm, r, s.port_points(attr))
            c, rel = _scale_fit(md["E"].ravel(), Eo.ravel())
            assert rel < 1e-9
            assert np.abs(md["H"].ravel() - c * Ho.ravel()).max() < 1e-9 * np.abs(md["H"]).max()
            assert md["V"].real > 0 and abs(md["V"].imag) < 1e-12 * md["V"].real
            assert abs(np.linalg.norm(md["E"], axis=1).max() - 1.0) < 1e-12   # peak |E_t| of 1
            if p == 1:
                assert abs(md["Z"].real - 674.4495) < 1e-4


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sweep_driven_by_2d_modes(p):
    """The 3D S driven by opem_port_mode equals the oracle's direct solve driven by the oracle's 2D modes, and moves
    toward the reference's results.ods S (row 1 of sheet p: this mesh)."""
    import wr90
    from openparem_amd.ports import te10
    from openparem_amd.solver import Solver
    m, d = wr90.load()
    f = d["freq"]
    paths = {}
    for attr in (2, 3):
        zc = m.xyz[:, 2].max() if attr == 2 else m.xyz[:, 2].min()
        paths[attr] = (np.array([d["a"] / 2, 0, zc]), np.array([d["a"] / 2, d["b"], zc]))
    with Solver(m, p) as s:
        r = s.sweep(np.array([f]), precision="fp64", tol=1e-12, restart=60, precond="lu",
                    mode_fn=s.port_mode_provider(paths))
        rt = s.sweep(np.array([f]), precision="fp64", tol=1e-12, restart=60, precond="lu")
    assert r["converged"].all()
    pr = Problem(m, p)
    modes = {}
    for attr in pr.ports:
        pm = PortMode(pr.space, attr)
        o = pm.solve(f)
        V = pm.voltage(o["et"], *paths[attr])
        P = pm.power(o["et"], o["ht"])
        modes[attr] = dict(E=lambda pts, pm=pm, o=o: pm.eval_at(o["et"], pts.reshape(-1, 3)).reshape(pts.shape),
                           H=lambda pts, pm=pm, o=o: pm.eval_at(o["ht"], pts.reshape(-1, 3)).reshape(pts.shape),
                           gamma=o["gamma"], V=V, Z=0.5 * V * np.conj(V) / np.conj(P))
    so = pr.solve_point(f, modes)["S"]
    dS = np.abs(r["S"][0] - so).max()
    s11, s21 = wr90.reference_s(d, p)
    e2d = max(abs(r["S"][0][0, 0] - s11), abs(r["S"][0][1, 0] - s21))
    ete = max(abs(rt["S"][0][0, 0] - s11), abs(rt["S"][0][1, 0] - s21))
    print(f"p={p}: |dS| vs oracle {dS:.2e}; vs results.ods: 2D modes {e2d:.2e}, analytic TE10 {ete:.2e}", flush=True)
    assert dS < 1e-9
    assert e2d < 1.5e-3 and e2d < ete


def test_port_mode_lossy_microstrip():
    from openparem_amd.solver import Solver
    m = microstrip_mesh(nx=8, ny=5, nz=4, n_sub=2, seed=6)
    g = m.geometry
    f = 6e9
    sp_ = oasm.Space(m, 2)
    with Solver(m, 2, boundaries={4: 5.8e7}) as s:
        for attr in (2, 3):
            z = 0.0 if attr == 2 else g["length"]
            path = (np.array([g["W"] / 2, g["h_sub"], z]), np.array([g["W"] / 2, 0.0, z]))
            md = s.port_mode(attr, f, path)
            pm = PortMode(sp_, attr)
            r = pm.solve(f)
            eps_eff = -(md["gamma"] / (2 * np.pi * f * np.sqrt(4e-7 * np.pi * 8.8541878176e-12))) ** 2
            print(f"microstrip port {attr}: gamma {md['gamma']} oracle {r['gamma']} eps_eff {eps_eff} Z {md['Z']}",
                  flush=True)
            assert abs(md["gamma"] - r["gamma"]) < 1e-11 * abs(r["gamma"])
            assert md["gamma"].real > 0 and 1.0 < eps_eff.real < 4.4
            Eo, Ho = _oracle_at_points(pm, r, s.port_points(attr))
            c, rel = _scale_fit(md["E"].ravel(), Eo.ravel())
            assert rel < 1e-8
            assert np.abs(md["H"].ravel() - c * Ho.ravel()).max() < 1e-8 * np.abs(md["H"]).max()
            assert 10 < md["Z"].real < 200
