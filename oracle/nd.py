"""ORACLE (test infrastructure only): Nedelec ND_p (p=1..3) basis on tetrahedra and triangles.

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import this package; the product path never does.

Space: first-kind Nedelec of order p (MFEM ``ND_FECollection(p, 3)``, used by
``src/OpenParEM3D/fem3D.cpp:65-77``), dimension p(p+2)(p+3)/2 = 6/20/45.

Basis (Arnold-Falk-Winther geometric decomposition, "Geometric decompositions and
local bases for spaces of finite element differential forms", CMAME 2009, Thm 4.4):

    phi = lambda^alpha * (lambda_i grad(lambda_j) - lambda_j grad(lambda_i)),  i < j,
    |alpha| = p-1,  alpha_m = 0 for m < i,

with every element's vertices taken in increasing *global* vertex id.  Each function
is owned by the sub-simplex {i,j} U supp(alpha); its tangential trace on a face only
involves that face's barycentrics, so the shared-face functions coincide from both
sides and the global space is H(curl)-conforming.

Here the functions are evaluated pointwise in physical coordinates and integrated
with collapsed-coordinate Gauss-Jacobi quadrature (Stroud conical product) of a
degree well above 2p, which is a different route from the product's exact
reference-tensor contraction.
"""
from __future__ import annotations

import itertools
from functools import lru_cache

import numpy as np
from scipy.special import roots_jacobi

TET_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
TET_FACES = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
TRI_EDGES = [(0, 1), (0, 2), (1, 2)]


def _multi_indices(n_vars, total):
    """All alpha in N^n_vars with |alpha| = total, lexicographically decreasing in alpha[0]."""
    out = []
    for c in itertools.combinations_with_replacement(range(n_vars), total):
        a = [0] * n_vars
        for k in c:
            a[k] += 1
        out.append(tuple(a))
    return out


@lru_cache(maxsize=None)
def entity_patterns(p, nverts):
    """Basis patterns owned by a sub-simplex with `nverts` vertices (2 edge, 3 face, 4 cell).

    A pattern is (i, j, alpha) in the entity's own vertex positions (sorted global order).
    """
    pats = []
    for i, j in itertools.combinations(range(nverts), 2):
        for alpha in _multi_indices(nverts, p - 1):
            if any(alpha[m] for m in range(i)):
                continue
            supp = {m for m in range(nverts) if alpha[m]} | {i, j}
            if len(supp) != nverts:
                continue
            pats.append((i, j, alpha))
    return tuple(pats)


def ndofs_per_entity(p):
    return len(entity_patterns(p, 2)), len(entity_patterns(p, 3)), len(entity_patterns(p, 4))


@lru_cache(maxsize=None)
def tet_local_basis(p):
    """Element-local basis list [(i, j, alpha4)], ordered edges (TET_EDGES) -> faces -> cell."""
    out = []
    for ent in TET_EDGES:
        for i, j, a in entity_patterns(p, 2):
            out.append(_lift(ent, i, j, a, 4))
    for ent in TET_FACES:
        for i, j, a in entity_patterns(p, 3):
            out.append(_lift(ent, i, j, a, 4))
    for i, j, a in entity_patterns(p, 4):
        out.append(_lift((0, 1, 2, 3), i, j, a, 4))
    return tuple(out)


@lru_cache(maxsize=None)
def tri_local_basis(p):
    out = []
    for ent in TRI_EDGES:
        for i, j, a in entity_patterns(p, 2):
            out.append(_lift(ent, i, j, a, 3))
    for i, j, a in entity_patterns(p, 3):
        out.append(_lift((0, 1, 2), i, j, a, 3))
    return tuple(out)


def _lift(ent, i, j, a, n):
    alpha = [0] * n
    for k, v in enumerate(ent):
        alpha[v] += a[k]
    return (ent[i], ent[j], tuple(alpha))


# ----------------------------------------------------------------------------- quadrature
def _gj01(n, alpha):
    """Gauss-Jacobi on [0,1] with weight (1-t)^alpha."""
    x, w = roots_jacobi(n, alpha, 0.0)
    return (x + 1.0) / 2.0, w / 2.0 ** (alpha + 1)


@lru_cache(maxsize=None)
def tet_rule(n):
    """Collapsed (Duffy) product rule on the unit tet; exact for degree <= 2n-1. Returns (L (nq,4), w)."""
    u, wu = _gj01(n, 2.0)
    v, wv = _gj01(n, 1.0)
    s, ws = _gj01(n, 0.0)
    U, V, S = np.meshgrid(u, v, s, indexing="ij")
    W = (wu[:, None, None] * wv[None, :, None] * ws[None, None, :]).ravel()
    x = U.ravel()
    y = (V * (1 - U)).ravel()
    z = (S * (1 - U) * (1 - V)).ravel()
    L = np.stack([1 - x - y - z, x, y, z], axis=1)
    return L, W


@lru_cache(maxsize=None)
def tri_rule(n):
    u, wu = _gj01(n, 1.0)
    s, ws = _gj01(n, 0.0)
    U, S = np.meshgrid(u, s, indexing="ij")
    W = (wu[:, None] * ws[None, :]).ravel()
    x = U.ravel()
    y = (S * (1 - U)).ravel()
    return np.stack([1 - x - y, x, y], axis=1), W


# ----------------------------------------------------------------------------- geometry
def tet_gradients(v):
    """Barycentric gradients (ne,4,3) and signed volume*6 (det J) for vertices v (ne,4,3)."""
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)  # columns
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                      # rows are grad(lambda_1..3)
    g = np.empty((v.shape[0], 4, 3))
    g[:, 1:] = Jinv
    g[:, 0] = -Jinv.sum(axis=1)
    return g, det


def tri_gradients(v):
    """Surface barycentric gradients (nf,3,3), area (nf,) and unit normal (nf,3) of triangles v (nf,3,3)."""
    e1 = v[:, 1] - v[:, 0]
    e2 = v[:, 2] - v[:, 0]
    nrm = np.cross(e1, e2)
    a2 = np.linalg.norm(nrm, axis=1)
    n = nrm / a2[:, None]
    # grad lambda_1 = (e2 x n)/|e1 x e2| ... in-plane gradients
    g1 = np.cross(e2, n) / a2[:, None]
    g2 = np.cross(n, e1) / a2[:, None]
    g = np.stack([-(g1 + g2), g1, g2], axis=1)
    return g, 0.5 * a2, n


# ----------------------------------------------------------------------------- evaluation
def eval_basis(basis, L, g):
    """Values and curls of the basis at barycentric points.

    basis: list of (i,j,alpha); L: (nq,nv) barycentrics; g: (ne,nv,3) gradients.
    Returns phi (ne,nq,nd,3) and curl (ne,nq,nd,3).
    """
    ne, nq, nd = g.shape[0], L.shape[0], len(basis)
    phi = np.empty((ne, nq, nd, 3))
    curl = np.empty((ne, nq, nd, 3))
    for a, (i, j, alpha) in enumerate(basis):
        la = np.prod(L ** np.asarray(alpha)[None, :], axis=1)            # (nq,)
        w = L[None, :, i, None] * g[:, None, j, :] - L[None, :, j, None] * g[:, None, i, :]
        phi[:, :, a, :] = la[None, :, None] * w
        grad_la = np.zeros((ne, nq, 3))
        for m, am in enumerate(alpha):
            if am == 0:
                continue
            e = list(alpha)
            e[m] -= 1
            lm = am * np.prod(L ** np.asarray(e)[None, :], axis=1)
            grad_la += lm[None, :, None] * g[:, None, m, :]
        gij = np.cross(g[:, i], g[:, j])                                 # (ne,3)
        curl[:, :, a, :] = np.cross(grad_la, w) + 2.0 * la[None, :, None] * gij[:, None, :]
    return phi, curl


def element_matrices(v, p, nquad=None):
    """Curl-curl K_e and mass M_e (unit coefficients) for tets with *sorted* vertex coords v (ne,4,3)."""
    basis = tet_local_basis(p)
    L, w = tet_rule(nquad or p + 2)
    g, det = tet_gradients(v)
    phi, curl = eval_basis(basis, L, g)
    wd = np.sqrt(w[None, :] * np.abs(det)[:, None])[:, :, None, None]     # (ne,nq,1,1)
    ne, nq, nd_ = phi.shape[:3]
    P = (phi * wd).transpose(0, 1, 3, 2).reshape(ne, nq * 3, nd_)
    C = (curl * wd).transpose(0, 1, 3, 2).reshape(ne, nq * 3, nd_)
    M = np.matmul(P.transpose(0, 2, 1), P)
    K = np.matmul(C.transpose(0, 2, 1), C)
    return K, M


def face_matrices(v, p, nquad=None):
    """Boundary mass  int E_t.T dS  and surface curl-curl  int curl_s E curl_s T dS on triangles (nf,3,3)."""
    basis = tri_local_basis(p)
    L, w = tri_rule(nquad or p + 2)
    g, area, n = tri_gradients(v)
    phi, curl = eval_basis(basis, L, g)
    wd = w[None, :] * (2.0 * area)[:, None]
    M = np.einsum("eq,eqak,eqbk->eab", wd, phi, phi, optimize=True)
    cs = np.einsum("eqak,ek->eqa", curl, n)
    K = np.einsum("eq,eqa,eqb->eab", wd, cs, cs, optimize=True)
    return K, M


def mixed_curl_matrices(v, p, nquad=None):
    """Q_e[a,b] = int curl(phi_b) . phi_a  (MFEM MixedVectorCurlIntegrator: trial b, test a)."""
    basis = tet_local_basis(p)
    L, w = tet_rule(nquad or p + 2)
    g, det = tet_gradients(v)
    phi, curl = eval_basis(basis, L, g)
    wd = np.sqrt(w[None, :] * np.abs(det)[:, None])[:, :, None, None]
    ne, nq, nd_ = phi.shape[:3]
    P = (phi * wd).transpose(0, 1, 3, 2).reshape(ne, nq * 3, nd_)
    C = (curl * wd).transpose(0, 1, 3, 2).reshape(ne, nq * 3, nd_)
    return np.matmul(P.transpose(0, 2, 1), C)
