"""ORACLE (test infrastructure only): DOF numbering and FP64 global assembly.

Restates, from the mathematics, what the reference obtains from MFEM:
* ND_p space + true-DOF numbering      (``src/OpenParEM3D/fem3D.cpp:65-77``)
* Re/Im(A) = curl-curl(1/mu_r) + mass(-k0^2 eps) + port ABC + impedance/radiation
  boundary mass                        (``fem3D.cpp:103-190``, ``port.cpp:1009-1041,5783-5890``)
* P = ND mass(1), Q = mixed curl(1/(omega mu))   (``fem3D.cpp:192-225``)
* merge of Re/Im into one complex CSR (``csr.c:171-223``)

DOF numbering convention (shared with the product because vectors are compared
entry by entry): every tetrahedron's vertices are taken in increasing global id;
entities (6 edges, 4 faces, 1 cell per element, in that local order) are numbered
in order of first appearance while traversing elements in input order; an entity
with k DOFs gets k consecutive global DOFs in the order of its basis patterns.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import nd

MU0 = 4e-7 * np.pi
EPS0 = 8.8541878176e-12   # reference value, src/OpenParEMCommon (eps0 used to build k0 in OpenParEM3D.cpp:373)
C0 = 1.0 / np.sqrt(MU0 * EPS0)


class Space:
    """ND_p space on a TetMesh: element->DOF map, boundary-face->DOF map."""

    def __init__(self, mesh, p):
        self.p = p
        self.mesh = mesh
        tets = np.sort(mesh.tets.astype(np.int64), axis=1)
        self.tets_sorted = tets
        ne = tets.shape[0]
        ke, kf, kc = nd.ndofs_per_entity(p)
        self.k_ent = (ke, kf, kc)
        nv = mesh.nv
        # entity keys (int64): edges a*nv+b, faces packed in 3*21 bits
        ekeys = np.stack([tets[:, i] * nv + tets[:, j] for i, j in nd.TET_EDGES], axis=1)      # (ne,6)
        fkeys = np.stack([(tets[:, i] << 42) | (tets[:, j] << 21) | tets[:, k] for i, j, k in nd.TET_FACES], 1)
        self.edge_keys, e_first, e_inv = np.unique(ekeys.ravel(), return_index=True, return_inverse=True)
        self.face_keys, f_first, f_inv = np.unique(fkeys.ravel(), return_index=True, return_inverse=True)
        self.elem_edges = e_inv.reshape(ne, 6)
        self.elem_faces = f_inv.reshape(ne, 4)
        # traversal position of first appearance: element*11 + slot
        e_pos = (e_first // 6) * 11 + (e_first % 6)
        f_pos = (f_first // 4) * 11 + 6 + (f_first % 4)
        c_pos = np.arange(ne) * 11 + 10
        kind = np.concatenate([np.zeros(e_pos.size, np.int64), np.ones(f_pos.size, np.int64),
                               np.full(ne, 2, np.int64)])
        ids = np.concatenate([np.arange(e_pos.size), np.arange(f_pos.size), np.arange(ne)])
        pos = np.concatenate([e_pos, f_pos, c_pos])
        cnt = np.array([ke, kf, kc])[kind]
        order = np.argsort(pos, kind="stable")
        offs = np.zeros(order.size, np.int64)
        offs[order] = np.concatenate([[0], np.cumsum(cnt[order])[:-1]])
        self.ndof = int(cnt.sum())
        self.edge_off = offs[kind == 0]
        self.face_off = offs[kind == 1]
        self.cell_off = offs[kind == 2]
        self.nedges, self.nfaces = e_pos.size, f_pos.size
        # element -> global DOFs in element-local basis order
        cols = []
        for s in range(6):
            for t in range(ke):
                cols.append(self.edge_off[self.elem_edges[:, s]] + t)
        for s in range(4):
            for t in range(kf):
                cols.append(self.face_off[self.elem_faces[:, s]] + t)
        for t in range(kc):
            cols.append(self.cell_off + t)
        self.elem_dofs = np.stack(cols, axis=1) if cols else np.zeros((ne, 0), np.int64)
        # boundary faces -> DOFs (triangle basis order)
        bf = np.sort(mesh.bfaces.astype(np.int64), axis=1)
        self.bfaces_sorted = bf
        bek = np.stack([bf[:, i] * nv + bf[:, j] for i, j in nd.TRI_EDGES], axis=1)
        bfk = (bf[:, 0] << 42) | (bf[:, 1] << 21) | bf[:, 2]
        be = np.searchsorted(self.edge_keys, bek)
        bff = np.searchsorted(self.face_keys, bfk)
        assert np.all(self.edge_keys[be] == bek) and np.all(self.face_keys[bff] == bfk)
        cols = []
        for s in range(3):
            for t in range(ke):
                cols.append(self.edge_off[be[:, s]] + t)
        for t in range(kf):
            cols.append(self.face_off[bff] + t)
        self.bface_dofs = np.stack(cols, axis=1)
        # owning tet of each boundary face (for material-dependent port coefficients)
        self.bface_owner = self._face_owner(f_inv.reshape(ne, 4), bff)

    def _face_owner(self, elem_faces, bff):
        flat = elem_faces.ravel()
        order = np.argsort(flat, kind="stable")
        first = np.searchsorted(flat[order], bff)
        return order[first] // 4

    def sorted_coords(self):
        return self.mesh.xyz[self.tets_sorted]

    def bface_coords(self):
        return self.mesh.xyz[self.bfaces_sorted]

    def dofs_on_bdr(self, attrs):
        sel = np.isin(self.mesh.bface_attr, list(attrs))
        return np.unique(self.bface_dofs[sel].ravel())


def _scatter(rows_map, vals, n):
    nd_ = rows_map.shape[1]
    r = np.repeat(rows_map, nd_, axis=1).ravel()
    c = np.tile(rows_map, (1, nd_)).ravel()
    return sp.csr_matrix((vals.ravel(), (r, c)), shape=(n, n))


def pattern(space):
    """Shared CSR pattern of every matrix on the space (all element couplings)."""
    ne, nd_ = space.elem_dofs.shape
    A = _scatter(space.elem_dofs, np.ones((ne, nd_, nd_)), space.ndof)
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64)


def volume_matrices(space):
    """Unit-coefficient element K_e, M_e for the whole mesh (FP64)."""
    return nd.element_matrices(space.sorted_coords(), space.p)


def assemble_volume(space, coef_k, coef_m, Ke=None, Me=None):
    """sum_e coef_k[e] K_e + coef_m[e] M_e, complex CSR (coefs may be complex arrays per element)."""
    if Ke is None:
        Ke, Me = volume_matrices(space)
    vals = coef_k[:, None, None] * Ke + coef_m[:, None, None] * Me
    A = _scatter(space.elem_dofs, vals.astype(np.complex128), space.ndof)
    A.sum_duplicates()
    return A


def assemble_faces(space, sel, coef_k, coef_m):
    """sum over selected boundary faces of coef_k curl_s-curl_s + coef_m tangential mass."""
    if not np.any(sel):
        return sp.csr_matrix((space.ndof, space.ndof), dtype=np.complex128)
    Kf, Mf = nd.face_matrices(space.bface_coords()[sel], space.p)
    vals = np.asarray(coef_k)[..., None, None] * Kf + np.asarray(coef_m)[..., None, None] * Mf
    A = _scatter(space.bface_dofs[sel], vals.astype(np.complex128), space.ndof)
    A.sum_duplicates()
    return A


def materials_per_element(mesh, key):
    return np.array([mesh.materials[a][key] for a in mesh.tet_attr], dtype=np.float64)


def system_matrix(space, freq, port_gamma, rs_of_attr=None, eta_rad=None, Ke=None, Me=None):
    """Complex A(omega) with port ABCs on *all* ports (see `solve.excite`).

    fem3D.cpp:103-190: Re = K/mu_r - k0^2 er' M + ABC(Re); Im = -k0^2 er'' M + ABC(Im) + omega mu0/Rs M_s.
    er'' = -er' tan(delta)  (OpenParEMmaterials.cpp:672: eps = er' eps0 (1 - j tand)).
    ABC on non-driving port (port.cpp:5783-5890): curl-curl coefficient 1/(mu_r gamma), mass -k0^2 eps_r/gamma.
    Because a driven port's DOFs are all eliminated, adding its ABC block changes no retained entry, so one
    A(omega) serves every excitation.
    """
    mesh = space.mesh
    k0 = 2 * np.pi * freq * np.sqrt(MU0 * EPS0)
    omega = 2 * np.pi * freq
    er = materials_per_element(mesh, "er")
    td = materials_per_element(mesh, "tand")
    mur = materials_per_element(mesh, "mur")
    ec = er * (1 - 1j * td)
    A = assemble_volume(space, 1.0 / mur, -k0 * k0 * ec, Ke, Me)
    owner_attr = mesh.tet_attr[space.bface_owner]
    for attr, kind in mesh.bdr_kind.items():
        sel = mesh.bface_attr == attr
        if kind == "port":
            g = port_gamma[attr]
            own = owner_attr[sel]
            er_f = np.array([mesh.materials[a]["er"] * (1 - 1j * mesh.materials[a]["tand"]) for a in own])
            mu_f = np.array([mesh.materials[a]["mur"] for a in own])
            A = A + assemble_faces(space, sel, 1.0 / (mu_f * g), -k0 * k0 * er_f / g)
        elif kind == "zs":
            rs = rs_of_attr[attr]
            A = A + assemble_faces(space, sel, 0.0, 1j * omega * MU0 / rs)
        elif kind == "radiation":
            eta = eta_rad if eta_rad is not None else np.sqrt(MU0 / EPS0)
            A = A + assemble_faces(space, sel, 0.0, 1j * omega * MU0 / eta)
    A.sort_indices()
    return A


def h_matrices(space, freq, Me=None):
    """P (mass, coefficient 1) and Q with 1/(omega mu) (fem3D.cpp:192-225); Q is real, used as jQ."""
    mesh = space.mesh
    omega = 2 * np.pi * freq
    mu = materials_per_element(mesh, "mur") * MU0
    v = space.sorted_coords()
    if Me is None:
        _, Me = nd.element_matrices(v, space.p)
    Qe = nd.mixed_curl_matrices(v, space.p)
    P = _scatter(space.elem_dofs, Me, space.ndof)
    Q = _scatter(space.elem_dofs, Qe / (omega * mu)[:, None, None], space.ndof)
    P.sum_duplicates()
    Q.sum_duplicates()
    return P.tocsr(), Q.tocsr()
