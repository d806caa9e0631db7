"""The oracle's 2D port eigenproblem (oracle/port_mode.py) pinned by the reference's own OpenParEM2D results and by
the closed form.

* On the reference's straight_study mesh (straight.msh, 30 tets; each port a rectangle of 4 triangles) at 9 GHz,
  OpenParEM2D logged beta/1000 = 0.10038888, 0.12939113, 0.12915073 at ND1-ND3 and Zo (Z_PV along the centre line)
  = 674.4495 at ND1 (regression/OpenParEM3D/WR90/straight_study/straight_order_{1,2,3}.log, the port "Results"
  tables); the discrete spaces are basis-independent, so the same eigenvalue must come out to the 8 logged digits.
* On refined guide cross-sections the TE10 beta converges to sqrt(k0^2 eps_r - (pi/a)^2) at the rate of the order.
"""
import os
import sys

import numpy as np
import pytest

from openparem_amd.mesh import waveguide_mesh
from oracle import assemble as oasm
from oracle.port_mode import PortMode, k0_of

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

LOGGED_BETA_K = {1: 0.10038888, 2: 0.12939113, 3: 0.12915073}
LOGGED_ZO = {1: 674.4495}


@pytest.mark.parametrize("p", [1, 2, 3])
def test_port_beta_matches_openparem2d_log(p):
    import wr90
    m, d = wr90.load()
    sp_ = oasm.Space(m, p)
    for attr in (2, 3):
        pm = PortMode(sp_, attr)
        r = pm.solve(d["freq"])
        beta_k = r["gamma"].imag / 1000
        print(f"p={p} port {attr}: beta/1000 {beta_k:.10f} (log {LOGGED_BETA_K[p]}), unknowns {pm.nnd + pm.nh}")
        assert round(beta_k, 8) == LOGGED_BETA_K[p]
        assert abs(r["gamma"].real) < 1e-9 * abs(r["gamma"].imag)
        if p == 2:
            assert pm.nnd + pm.nh == 37     # "matrix size: 37" in straight_order_2.log
        if p in LOGGED_ZO:
            zc = m.xyz[:, 2].max() if attr == 2 else m.xyz[:, 2].min()
            V = pm.voltage(r["et"], [d["a"] / 2, 0, zc], [d["a"] / 2, d["b"], zc])
            P = pm.power(r["et"], r["ht"])
            Z = 0.5 * V * np.conj(V) / np.conj(P)
            print(f"  Zpv {Z.real:.7f} (log {LOGGED_ZO[p]})")
            assert abs(Z.real - LOGGED_ZO[p]) < 1e-4 and abs(Z.imag) < 1e-6


@pytest.mark.parametrize("p", [1, 2])
def test_te10_beta_converges_to_closed_form(p):
    f = 10e9
    a = 22.86e-3
    k0 = k0_of(f)
    exact = np.sqrt(k0 ** 2 - (np.pi / a) ** 2)
    errs = []
    for n in (1, 2, 3):
        m = waveguide_mesh(nx=4 * n, ny=2 * n, nz=1, jitter=0.0, seed=1)
        pm = PortMode(oasm.Space(m, p), 2)
        r = pm.solve(f)
        errs.append(abs(r["gamma"].imag - exact) / exact)
    rates = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(np.array([2.0, 1.5]))
    print(f"p={p}: relative beta errors {errs}, rates {rates}")
    assert errs[-1] < (5e-3 if p == 1 else 1e-5)
    assert np.all(rates > 2 * p - 0.3)      # eigenvalue error O(h^{2p})


def test_te10_field_shape():
    """The ND2 mode on a 8 x 4 guide cross-section: E_t along the short side, sin(pi x / a) across, E_z = 0."""
    m = waveguide_mesh(nx=8, ny=4, nz=1, jitter=0.0, seed=1)
    pm = PortMode(oasm.Space(m, 2), 2)
    r = pm.solve(10e9)
    E = pm.field_nd(r["et"]).reshape(-1, 3)
    pts = pm.pts.reshape(-1, 3)
    a = m.geometry["a"]
    ref = np.sin(np.pi * pts[:, 0] / a)
    c = np.vdot(ref, E[:, 1]) / np.vdot(ref, ref)
    assert np.abs(E[:, 1] - c * ref).max() < 1e-2 * abs(c)
    assert np.abs(E[:, 0]).max() < 1e-2 * abs(c) and np.abs(E[:, 2]).max() < 1e-12 * abs(c)
    assert np.abs(r["ez"]).max() < 1e-9 * np.abs(r["et"]).max()
