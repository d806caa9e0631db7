"""ORACLE (test infrastructure only): radiation currents, far field and pattern metrics in FP64 numpy.

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this module; the
product path never does.

Restated from the reference (src/OpenParEM3D):

* ``Boundary::calculateRadiationCurrents`` (port.cpp:1043-1265): on each radiation boundary, J = n x H_t and
  M = -n x E_t at the centres of the boundary triangles, refined until no edge is longer than ``limit``
  wavelengths (lambda = 2 pi / beta, beta = 2 pi f mu0 / eta, fem3D.cpp:1926-1929), each with its area.  The
  reference refines with MFEM's bisection (GeneralRefinement); here a face whose longest edge is L is split into
  s*s similar triangles, s = ceil(L / (limit lambda)).  Both are midpoint rules of the same surface integrals; the
  product and this oracle use the same split, so their currents agree entry by entry.
* ``BoundaryDatabase::calculateFarField`` (port.cpp:8077-8160; Balanis, Advanced Engineering Electromagnetics,
  1989, pp. 288-289): N = sum J e^{j beta rhat.r'} dA, L likewise with M, then
  E_theta = -k (L_phi + eta N_theta), E_phi = k (L_theta - eta N_phi), H_theta = k (N_phi - L_theta/eta),
  H_phi = -k (N_theta + L_phi/eta), k = j beta e^{-j beta r}/(4 pi r).
* ``Sphere::create/refine/setMetrics/allocateAreasToPoints`` (pattern.cpp:618-800) with the spherical-triangle
  area of ``get_areaSphereTriangle`` (pattern.cpp:45-57, here by the Van Oosterom-Strackee formula for the
  spherical excess instead of the sum of lune angles), ``Circle::create`` (pattern.cpp:1468-1504).
* Metrics: ``OPEMpoint::get_power`` (pattern.hpp:98), ``Sphere::getRadiatedPower`` (pattern.cpp:858-868),
  ``calculateIsotropicGain`` (:1144-1159), ``calculateDirectivity`` (:1161-1176),
  ``Pattern::calculateRadiationEfficiency`` (:2529-2535), ``BoundaryDatabase::calculateAcceptedPower``
  (port.cpp:7475-7497).
"""
from __future__ import annotations

import numpy as np

from . import nd

MU0 = 4e-7 * np.pi


# ----------------------------------------------------------------------------- currents
def split_centres(s):
    """Barycentrics (n,3) of the centres of the s*s similar triangles of a uniform split, in the order
    i = 0..s-1, j = 0..s-1-i: the triangle with corners (i,j),(i+1,j),(i,j+1), then, when i+j <= s-2, the one
    with corners (i+1,j),(i,j+1),(i+1,j+1); (i,j) index the lattice along vertex 1 and vertex 2."""
    out = []
    for i in range(s):
        for j in range(s - i):
            x, y = (3 * i + 1) / (3.0 * s), (3 * j + 1) / (3.0 * s)
            out.append((1.0 - x - y, x, y))
            if i + j <= s - 2:
                x, y = (3 * i + 2) / (3.0 * s), (3 * j + 2) / (3.0 * s)
                out.append((1.0 - x - y, x, y))
    return np.asarray(out)


def split_count(v, limit_len):
    """s for one face (sorted vertices v (3,3)): ceil(longest edge / limit_len), at least 1."""
    L = 0.0
    for a, b in ((0, 1), (0, 2), (1, 2)):
        d = v[b] - v[a]
        L = max(L, float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    return max(1, int(np.ceil(L / limit_len)))


def radiation_currents(space, attrs, x, h, freq, eta, limit):
    """Currents on the boundary faces whose attribute is in ``attrs``, in increasing face index.

    x, h: E and H DOF vectors (complex, global numbering).  Returns dict xyz (n,3), J (n,3), M (n,3), area (n,).
    """
    mesh = space.mesh
    lam = eta / (MU0 * freq)
    sel = np.flatnonzero(np.isin(mesh.bface_attr, list(attrs)))
    vsorted = space.bface_coords()
    basis = nd.tri_local_basis(space.p)
    tets = mesh.tets
    xyz, J, M, area = [], [], [], []
    for f in sel:
        v = vsorted[f]
        s = split_count(v, limit * lam)
        Lc = split_centres(s)
        g, a, n = nd.tri_gradients(v[None])
        # outward: away from the owning tet's vertex that is not on the face
        owner = tets[space.bface_owner[f]]
        opp = [q for q in owner if q not in set(mesh.bfaces[f].tolist())][0]
        n = n[0]
        if np.dot(n, v.mean(axis=0) - mesh.xyz[opp]) < 0:
            n = -n
        phi, _ = nd.eval_basis(basis, Lc, g)                        # (1,nq,nd,3)
        dofs = space.bface_dofs[f]
        Et = np.einsum("qak,a->qk", phi[0], x[dofs])
        Ht = np.einsum("qak,a->qk", phi[0], h[dofs])
        xyz.append(Lc @ v)
        J.append(np.cross(n[None, :], Ht))
        M.append(-np.cross(n[None, :], Et))
        area.append(np.full(Lc.shape[0], a[0] / (s * s)))
    return dict(xyz=np.concatenate(xyz), J=np.concatenate(J), M=np.concatenate(M), area=np.concatenate(area))


# ----------------------------------------------------------------------------- far field
def far_field(cur, center, beta, eta, pts):
    """(npts, 4) complex: E_phi, E_theta, H_phi, H_theta at the points (the order of OPEMpoint::set_fields)."""
    center = np.asarray(center, float)
    P = np.asarray(pts, float) - center
    r = np.linalg.norm(P, axis=1)
    rh = P / r[:, None]
    ct = P[:, 2] / r
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = np.arctan2(P[:, 1], P[:, 0])
    cp, sp = np.cos(ph), np.sin(ph)
    rp = cur["xyz"] - center
    w = np.exp(1j * beta * (rh @ rp.T)) * cur["area"][None, :]
    N = w @ cur["J"]
    Lm = w @ cur["M"]

    def th(V):
        return V[:, 0] * ct * cp + V[:, 1] * ct * sp - V[:, 2] * st

    def phc(V):
        return -V[:, 0] * sp + V[:, 1] * cp

    k = 1j * beta * np.exp(-1j * beta * r) / (4 * np.pi * r)
    Nt, Np, Lt, Lp = th(N), phc(N), th(Lm), phc(Lm)
    Et = -k * (Lp + eta * Nt)
    Ep = k * (Lt - eta * Np)
    Ht = k * (Np - Lt / eta)
    Hp = -k * (Nt + Lp / eta)
    return np.stack([Ep, Et, Hp, Ht], axis=1)


def point_power(F):
    """0.5 (E_theta H_phi* - E_phi H_theta*) per point (complex)."""
    return 0.5 * (F[:, 1] * np.conj(F[:, 2]) - F[:, 0] * np.conj(F[:, 3]))


# ----------------------------------------------------------------------------- sphere and circle
_ICO_X = 0.525731112119133606
_ICO_Z = 0.850650808352039932
_ICO_V = [(-_ICO_X, 0, _ICO_Z), (_ICO_X, 0, _ICO_Z), (-_ICO_X, 0, -_ICO_Z), (_ICO_X, 0, -_ICO_Z),
          (0, _ICO_Z, _ICO_X), (0, _ICO_Z, -_ICO_X), (0, -_ICO_Z, _ICO_X), (0, -_ICO_Z, -_ICO_X),
          (_ICO_Z, _ICO_X, 0), (-_ICO_Z, _ICO_X, 0), (_ICO_Z, -_ICO_X, 0), (-_ICO_Z, -_ICO_X, 0)]
_ICO_T = [(0, 4, 1), (0, 9, 4), (9, 5, 4), (4, 5, 8), (4, 8, 1), (8, 10, 1), (8, 3, 10), (5, 3, 8), (5, 2, 3),
          (2, 7, 3), (7, 10, 3), (7, 6, 10), (7, 11, 6), (11, 0, 6), (0, 1, 6), (6, 1, 10), (9, 0, 11), (9, 11, 2),
          (9, 2, 5), (7, 2, 11)]


def _excess(a, b, c):
    """Spherical excess of the triangle of directions a, b, c (unit vectors)."""
    num = abs(np.dot(a, np.cross(b, c)))
    den = 1.0 + np.dot(a, b) + np.dot(b, c) + np.dot(c, a)
    return 2.0 * np.arctan2(num, den)


def _unit(p, c):
    d = p - c
    return d / np.linalg.norm(d)


def sphere(radius, center, refine):
    """Points (vertices in creation order, then one centre per triangle), point areas (vertices only; centres
    carry 0), triangles (vertex indices) and triangle areas, as Sphere::create, refine(refine), setMetrics and
    allocateAreasToPoints build them: vertices about the origin, centres and areas about ``center``."""
    center = np.asarray(center, float)
    pts = [np.asarray(p, float) * radius for p in _ICO_V]
    tris = [list(t) for t in _ICO_T]
    for _ in range(refine):
        mids = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in mids:
                m = 0.5 * (pts[i] + pts[j])
                m = m / (np.linalg.norm(m) / radius)
                pts.append(m)
                mids[key] = len(pts) - 1
            return mids[key]

        n0 = len(tris)
        for t in range(n0):
            a, b, c = tris[t]
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            tris.append([ab, b, bc])
            tris.append([ab, bc, ca])
            tris.append([ca, bc, c])
            tris[t] = [a, ab, ca]
    nv = len(pts)
    tri_area = np.empty(len(tris))
    centres = []
    for t, (a, b, c) in enumerate(tris):
        m = (pts[a] - center + pts[b] - center + pts[c] - center) / 3.0
        m = m / (np.linalg.norm(m) / radius) + center
        centres.append(m)
        tri_area[t] = radius ** 2 * _excess(_unit(pts[a], center), _unit(pts[b], center), _unit(pts[c], center))
    P = np.array(pts + centres)
    parea = np.zeros(P.shape[0])
    for t, (a, b, c) in enumerate(tris):
        ctr = _unit(P[nv + t], center)
        A, B, Cc = P[a], P[b], P[c]
        mab, mbc, mca = _unit(0.5 * (A + B), center), _unit(0.5 * (B + Cc), center), _unit(0.5 * (Cc + A), center)
        ua, ub, uc = _unit(A, center), _unit(B, center), _unit(Cc, center)
        parea[a] += radius ** 2 * (_excess(ua, mab, ctr) + _excess(ua, ctr, mca))
        parea[b] += radius ** 2 * (_excess(ub, mbc, ctr) + _excess(ub, ctr, mab))
        parea[c] += radius ** 2 * (_excess(uc, mca, ctr) + _excess(uc, ctr, mbc))
    return dict(pts=P, point_area=parea, tris=np.asarray(tris), tri_area=tri_area, nvert=nv)


def circle(radius, center, theta, phi, latitude, npoints):
    """Circle::create (pattern.cpp:1468-1504): npoints on the latitude circle of the x-y plane about ``center``,
    then rotated by theta about y and phi about z (rotate_point, pattern.cpp:1302-1310), about the origin."""
    c = np.asarray(center, float)
    ang = np.arange(npoints) * (2 * np.pi / npoints)
    x = radius * np.cos(ang) * np.cos(latitude) + c[0]
    y = radius * np.sin(ang) * np.cos(latitude) + c[1]
    z = np.full(npoints, radius * np.sin(latitude) + c[2])
    xp = np.cos(theta) * x + np.sin(theta) * z
    zp = -np.sin(theta) * x + np.cos(theta) * z
    return np.stack([np.cos(phi) * xp - np.sin(phi) * y, np.sin(phi) * xp + np.cos(phi) * y, zp], axis=1)


# ----------------------------------------------------------------------------- metrics
def metrics(F, point_area, total_area, accepted_power):
    """Radiated power (complex), per-point gain and directivity (real ratios), their maxima in dB and the
    radiation efficiency, from the far fields F (npts,4) of a sphere."""
    p = point_power(F)
    prad = np.sum(p * point_area)
    gain = np.real(p / (accepted_power / total_area))
    direc = np.real(p / (prad / total_area))
    with np.errstate(divide="ignore", invalid="ignore"):
        gmax = np.max(10 * np.log10(gain))
        dmax = np.max(10 * np.log10(direc))
    return dict(radiated_power=prad, gain=gain, directivity=direc, gain_db=gmax, directivity_db=dmax,
                efficiency=float(np.real(prad / accepted_power)))


def accepted_power(Cp, Cm, Pz, sport):
    """BoundaryDatabase::calculateAcceptedPower (port.cpp:7475-7497) for driving set ``sport`` (0-based):
    |Cm|^2 Pz of the driving set itself minus, for every other driving set i, |Cp|^2 Pz; the splits are those of
    the driving port's mode, Cp[i][k] / Cm[i][k] = split at S-port k when driving set i (the layout of
    opem_sweep's Cp, Cm), so the terms read Cm[sport][sport] and Cp[i][sport]."""
    pin = Cm[sport, sport] * np.conj(Cm[sport, sport]) * Pz
    pout = sum(Cp[i, sport] * np.conj(Cp[i, sport]) * Pz for i in range(Cp.shape[0]) if i != sport)
    return pin - pout
