"""Quasi-TEM port mode of the shielded microstrip (port-mode provider, input side of the path).

The reference solves the full 2D port eigenproblem in OpenParEM2D (``src/OpenParEM2D/fem2D.cpp:2221-2343``,
``eigensolve.c:629-1017``).  This provider uses the quasi-static limit instead, from a P1 finite-element
electrostatic solve on the port's own triangles:

    div(eps_r grad phi) = 0,  phi = 1 on the trace, 0 on PEC ground/walls, natural on the open lid;
    E_t = -grad phi,  eps_eff = W(eps_r) / W(1)  (stored energies),  gamma = j k0 sqrt(eps_eff),
    H_t = gamma/(j omega mu0) n_out x E_t,  V = 1,  Z = V V* / (2 conj(P)),  P = 1/2 int (E x H*).n.

eps_r is complex (1 - j tan(delta)) in the substrate, so gamma picks up the dielectric attenuation.
Sample points arrive grouped per port face (ascending boundary-face index, the same count per face),
which is how the field is evaluated: piecewise constant per face.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MU0 = 4e-7 * np.pi
EPS0 = 8.8541878176e-12


def _solve_potential(mesh, attr, er_map):
    g = mesh.geometry
    faces = np.flatnonzero(mesh.bface_attr == attr)
    tri = mesh.bfaces[faces].astype(np.int64)
    xy = mesh.xyz[:, :2]
    verts, loc = np.unique(tri.ravel(), return_inverse=True)
    loc = loc.reshape(-1, 3)
    P = xy[verts]
    nv = verts.size
    # owner material per face: substrate below the interface
    cy = P[loc].mean(axis=1)[:, 1]
    eps = np.where(cy < g["h_sub"], er_map[2], er_map[1])
    e1 = P[loc[:, 1]] - P[loc[:, 0]]
    e2 = P[loc[:, 2]] - P[loc[:, 0]]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    # P1 gradients
    x, y = P[loc, 0], P[loc, 1]
    b = np.stack([y[:, 1] - y[:, 2], y[:, 2] - y[:, 0], y[:, 0] - y[:, 1]], 1) / (2 * area)[:, None]
    c = np.stack([x[:, 2] - x[:, 1], x[:, 0] - x[:, 2], x[:, 1] - x[:, 0]], 1) / (2 * area)[:, None]
    sgn = np.sign((e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]))[:, None]
    b, c = b * sgn, c * sgn

    def stiff(coef):
        Ke = coef[:, None, None] * area[:, None, None] * (b[:, :, None] * b[:, None, :] + c[:, :, None] * c[:, None, :])
        r = np.repeat(loc, 3, 1).ravel()
        cc = np.tile(loc, (1, 3)).ravel()
        return sp.csr_matrix((Ke.ravel(), (r, cc)), shape=(nv, nv))

    tol = 1e-9
    on_trace = (np.abs(P[:, 1] - g["h_sub"]) < tol) & (P[:, 0] >= g["x0"] - tol) & (P[:, 0] <= g["x1"] + tol)
    on_pec = (np.abs(P[:, 1]) < tol) | (np.abs(P[:, 0]) < tol) | (np.abs(P[:, 0] - g["W"]) < tol)
    fixed = on_trace | on_pec
    phi_fixed = np.where(on_trace, 1.0, 0.0)
    free = ~fixed

    def solve(coef):
        K = stiff(coef).astype(np.complex128)
        rhs = -K[free][:, fixed] @ phi_fixed[fixed]
        phi = phi_fixed.astype(np.complex128)
        phi[free] = spla.spsolve(K[free][:, free].tocsc(), rhs)
        return phi, K

    phi, K = solve(eps)
    phi0, K0 = solve(np.ones_like(eps))
    W = phi @ (K @ phi)        # sum eps_r |grad phi|^2 area (complex eps_r, bilinear)
    W0 = (phi0 @ (K0 @ phi0)).real
    eps_eff = W / W0
    Ex = -(phi[loc] * b).sum(1)
    Ey = -(phi[loc] * c).sum(1)
    return dict(faces=faces, Ex=Ex, Ey=Ey, area=area, eps_eff=eps_eff)


_CACHE = {}


def quasi_tem(mesh, attr, freq):
    key = (id(mesh), attr)
    if key not in _CACHE:
        er_map = {a: m["er"] * (1 - 1j * m["tand"]) for a, m in mesh.materials.items()}
        _CACHE[key] = _solve_potential(mesh, attr, er_map)
    d = _CACHE[key]
    k0 = 2 * np.pi * freq * np.sqrt(MU0 * EPS0)
    omega = 2 * np.pi * freq
    gamma = 1j * k0 * np.sqrt(d["eps_eff"])
    if gamma.real < 0:
        gamma = -gamma
    n = np.asarray(mesh.port_normal[attr], dtype=np.float64)
    yw = gamma / (1j * omega * MU0)
    nf = d["faces"].size

    def E(pts):
        flat = pts.reshape(-1, 3)
        nq = flat.shape[0] // nf
        assert nq * nf == flat.shape[0], "points must be grouped per port face"
        e = np.zeros(flat.shape, np.complex128)
        e[:, 0] = np.repeat(d["Ex"], nq)
        e[:, 1] = np.repeat(d["Ey"], nq)
        return e.reshape(pts.shape)

    def H(pts):
        return yw * np.cross(n, E(pts))

    # with H_t = yw n x E_t and E_t . n = 0:  (E x H*) . n = conj(yw) |E_t|^2
    P = 0.5 * np.conj(yw) * np.sum(d["area"] * (np.abs(d["Ex"]) ** 2 + np.abs(d["Ey"]) ** 2))
    V = 1.0 + 0j
    Z = 0.5 * V * np.conj(V) / np.conj(P)
    return dict(E=E, H=H, gamma=complex(gamma), V=V, Z=complex(Z), eps_eff=complex(d["eps_eff"]))
