"""ORACLE (test infrastructure only): excitation, elimination, Krylov solve, H projection, S.

Follows:
* ``solveComplexLinearSystem.c:61-65``   eliminatePEC = MatZeroRowsColumns(diag 1)
* ``solveComplexLinearSystem.c:100-209`` b = -A x, b[port] = x[port], MatZeroRowsColumns(port), KSP solve
* ``fem3D.cpp:610-648`` + ``solveComplexLinearSystem.c:211-315``  P h = j Q x
* ``port.cpp:3463-3489,3524-3568``      tangential inner products and C+/C- splits
* ``port.cpp:7499-7606`` + ``Zsolve.c:299-436``  N^2 x N^2 S solve
* ``results.cpp:401-419``               forced reciprocity S <- (S + S^T)/2
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import nd


def eliminate(A, dofs, diag=1.0):
    """MatZeroRowsColumns(A, dofs, diag): zero rows and columns, put `diag` on their diagonal."""
    n = A.shape[0]
    m = np.ones(n)
    m[dofs] = 0.0
    D = sp.diags(m)
    return (D @ A @ D + sp.diags((1.0 - m) * diag)).tocsr()


def excite(A, pec_dofs, port_dofs, x_port):
    """Eliminated system (A_t, b) for one excitation, in the reference's order of operations.

    A is first PEC-eliminated, then b = -A x with x holding only the driven-port values,
    b[port] = x[port], then the port rows/columns are eliminated.
    """
    n = A.shape[0]
    A1 = eliminate(A, pec_dofs)
    x = np.zeros(n, np.complex128)
    x[port_dofs] = x_port
    b = -(A1 @ x)
    b[port_dofs] = x_port
    A2 = eliminate(A1, port_dofs)
    return A2, b


def direct_solve(A, b):
    return spla.splu(A.tocsc()).solve(b)


class BlockJacobi:
    """Element-patch restricted additive Schwarz: for each element, invert the assembled
    matrix restricted to the element's free DOFs; each DOF takes its correction from
    its owning element (the first element, in traversal order, that contains it)."""

    def __init__(self, A, elem_dofs, dirichlet_mask):
        A = A.tocsr()
        ne, nd_ = elem_dofs.shape
        n = A.shape[0]
        owner = np.full(n, -1, np.int64)
        flat = elem_dofs.ravel()
        first = np.unique(flat, return_index=True)
        owner[first[0]] = first[1] // nd_
        self.elem_dofs = elem_dofs
        self.owned = owner[elem_dofs] == np.arange(ne)[:, None]          # (ne,nd)
        self.free = ~dirichlet_mask[elem_dofs]
        blocks = np.empty((ne, nd_, nd_), np.complex128)
        for a in range(nd_):
            for b_ in range(nd_):
                blocks[:, a, b_] = np.asarray(A[elem_dofs[:, a], elem_dofs[:, b_]]).ravel()
        fm = self.free[:, :, None] & self.free[:, None, :]
        eye = np.broadcast_to(np.eye(nd_), blocks.shape)
        blocks = np.where(fm, blocks, eye)
        self.inv = np.linalg.inv(blocks)
        self.dmask = dirichlet_mask

    def __call__(self, r):
        rl = r[self.elem_dofs] * self.free
        zl = np.einsum("eab,eb->ea", self.inv, rl)
        z = np.zeros_like(r)
        sel = self.owned
        z[self.elem_dofs[sel]] = zl[sel]
        z[self.dmask] = r[self.dmask]
        return z


class VertexPatchASM:
    """Additive Schwarz over vertex patches (Arnold, Falk, Winther, Numer. Math. 85 (2000)):
    patch v = DOFs of the edges, faces and cells incident to vertex v, minus Dirichlet DOFs;
    M^{-1} r = sum_v R_v^T (R_v A R_v^T)^{-1} R_v r.  Patches of equal size are inverted and
    applied as one batched dense product."""

    def __init__(self, A, space, dirichlet_mask):
        A = A.tocsr()
        p = space.p
        ke, kf, kc = nd.ndofs_per_entity(p)
        ents = []
        for e in nd.TET_EDGES:
            ents += [e] * ke
        for f in nd.TET_FACES:
            ents += [f] * kf
        ents += [(0, 1, 2, 3)] * kc
        ts = space.tets_sorted
        rows, cols = [], []
        for a, ent in enumerate(ents):
            d = space.elem_dofs[:, a]
            for k in ent:
                rows.append(ts[:, k])
                cols.append(d)
        v = np.concatenate(rows)
        d = np.concatenate(cols)
        keep = ~dirichlet_mask[d]
        pairs = np.unique(np.stack([v[keep], d[keep]], 1), axis=0)       # (vertex, dof), sorted
        verts, start, counts = np.unique(pairs[:, 0], return_index=True, return_counts=True)
        self.groups = []
        self.n = n = A.shape[0]
        self.dmask = dirichlet_mask
        A.sort_indices()
        keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.indptr)) * n + A.indices
        for size in np.unique(counts):
            sel = np.flatnonzero(counts == size)
            idx = pairs[start[sel][:, None] + np.arange(size)[None, :], 1]  # (np_, size)
            q = (idx[:, :, None].astype(np.int64) * n + idx[:, None, :]).ravel()
            pos = np.searchsorted(keys, q)
            pos = np.minimum(pos, keys.size - 1)
            hit = keys[pos] == q
            blocks = np.where(hit, A.data[pos], 0.0).reshape(len(sel), size, size)
            self.groups.append((idx, np.linalg.inv(blocks)))

    def __call__(self, r):
        z = np.zeros(self.n, np.complex128)
        for idx, inv in self.groups:
            zi = np.matmul(inv, r[idx][..., None])[..., 0].ravel()
            f = idx.ravel()
            z += np.bincount(f, zi.real, self.n) + 1j * np.bincount(f, zi.imag, self.n)
        z[self.dmask] = r[self.dmask]
        return z


def gmres(A, b, M=None, x0=None, restart=50, rtol=1e-10, maxit=20000):
    """Right-preconditioned restarted GMRES(m) with CGS2 Arnoldi; converged on the true residual."""
    n = b.size
    M = M or (lambda v: v)
    x = np.zeros(n, np.complex128) if x0 is None else x0.astype(np.complex128).copy()
    bn = np.linalg.norm(b)
    it = 0
    hist = []
    while it < maxit:
        r = b - A @ x
        beta = np.linalg.norm(r)
        hist.append(beta / bn)
        if beta <= rtol * bn:
            return x, it, True, hist
        V = np.zeros((restart + 1, n), np.complex128)
        Z = np.zeros((restart, n), np.complex128)
        H = np.zeros((restart + 1, restart), np.complex128)
        V[0] = r / beta
        g = np.zeros(restart + 1, np.complex128)
        g[0] = beta
        k = 0
        for k in range(restart):
            Z[k] = M(V[k])
            w = A @ Z[k]
            for _ in range(2):
                h = np.conj(V[: k + 1] @ np.conj(w))
                w = w - h @ V[: k + 1]
                H[: k + 1, k] += h
            H[k + 1, k] = np.linalg.norm(w)
            V[k + 1] = w / H[k + 1, k] if H[k + 1, k] != 0 else 0
            it += 1
            y, *_ = np.linalg.lstsq(H[: k + 2, : k + 1], g[: k + 2], rcond=None)
            res = np.linalg.norm(g[: k + 2] - H[: k + 2, : k + 1] @ y)
            if res <= rtol * bn * 0.5 or it >= maxit:
                break
        x = x + y @ Z[: k + 1]
    r = b - A @ x
    return x, it, np.linalg.norm(r) <= rtol * bn, hist


def h_field(P, Q, x, rtol=1e-14):
    """h = P^{-1} (j Q x)   (fem3D.cpp:610-648), P SPD: Jacobi-preconditioned CG, FP64."""
    b = 1j * (Q @ x)
    d = 1.0 / P.diagonal()
    h = np.zeros_like(b)
    r = b.copy()
    z = d * r
    p = z.copy()
    rz = np.vdot(r, z).real
    bn = np.linalg.norm(b)
    for _ in range(20 * b.size):
        q = P @ p
        a = rz / np.vdot(p, q).real
        h += a * p
        r -= a * q
        if np.linalg.norm(r) <= rtol * bn:
            break
        z = d * r
        rz2 = np.vdot(r, z).real
        p = z + (rz2 / rz) * p
        rz = rz2
    return h


# ----------------------------------------------------------------------------- ports
class PortTrace:
    """Tangential trace of ND_p on one port's triangles, with quadrature."""

    def __init__(self, space, attr, nquad=None):
        self.space = space
        self.sel = np.flatnonzero(space.mesh.bface_attr == attr)
        v = space.bface_coords()[self.sel]
        L, w = nd.tri_rule(nquad or space.p + 2)
        g, area, n = nd.tri_gradients(v)
        self.phi, _ = nd.eval_basis(nd.tri_local_basis(space.p), L, g)     # (nf,nq,nd,3)
        self.wq = w[None, :] * (2 * area)[:, None]                            # (nf,nq)
        self.pts = np.einsum("qk,fkd->fqd", L, v)                             # (nf,nq,3)
        self.dofs = space.bface_dofs[self.sel]                                 # (nf,nd)
        self.port_dofs = np.unique(self.dofs.ravel())

    def field(self, x):
        """Tangential field of global DOF vector x at quadrature points (nf,nq,3)."""
        return np.einsum("fqak,fa->fqk", self.phi, x[self.dofs])

    def inner(self, fa, fb):
        """int conj(A_t).B_t dS  (tangentialInnerProduct, port.cpp:3463-3489; exact quadrature)."""
        return np.sum(self.wq[..., None] * np.conj(fa) * fb)

    def project(self, fvals, fixed_zero=()):
        """L2 projection of sampled tangential field values (nf,nq,3) onto the trace space.

        Returns values on self.port_dofs; DOFs in `fixed_zero` are held at 0.
        """
        loc = {d: i for i, d in enumerate(self.port_dofs)}
        li = np.vectorize(loc.get)(self.dofs)
        nd_ = self.dofs.shape[1]
        Me = np.einsum("fq,fqak,fqbk->fab", self.wq, self.phi, self.phi)
        be = np.einsum("fq,fqak,fqk->fa", self.wq, self.phi, fvals)
        npd = self.port_dofs.size
        M = sp.csr_matrix((Me.ravel(), (np.repeat(li, nd_, 1).ravel(), np.tile(li, (1, nd_)).ravel())),
                          shape=(npd, npd))
        b = np.zeros(npd, np.complex128)
        np.add.at(b, li.ravel(), be.ravel())
        free = ~np.isin(self.port_dofs, np.asarray(fixed_zero, dtype=np.int64))
        out = np.zeros(npd, np.complex128)
        Mf = M[free][:, free].tocsc().astype(np.complex128)
        out[free] = spla.splu(Mf).solve(b[free])
        return out


def splits(trace, Em, Hm, Es, Hs):
    """C+ and C- for one port/mode/excitation (port.cpp:3524-3568)."""
    e0 = trace.inner(Em, Es)
    e2 = trace.inner(Em, Em)
    h0 = trace.inner(Hm, Hs)
    h2 = trace.inner(Hm, Hm)
    return 0.5 * (e0 / e2 + h0 / h2), 0.5 * (e0 / e2 - h0 / h2)


def calculate_s(Cp, Cm, V, Z):
    """S from splits (port.cpp:7499-7606).

    Cp[i][k], Cm[i][k]: split of S-port k when driving set i.  Unknown s = vec(S) row-major;
    row (i,j): sum_k Cm[i][k] V_k/sqrt(Z_k) S[j][k] = Cp[i][j] V_j/sqrt(Z_j).
    """
    N = len(V)
    w = np.asarray(V, np.complex128) / np.sqrt(np.asarray(Z, np.complex128))
    a = np.zeros((N * N, N * N), np.complex128)
    b = np.zeros(N * N, np.complex128)
    for i in range(N):
        for j in range(N):
            b[i * N + j] = Cp[i][j] * w[j]
            for k in range(N):
                a[i * N + j, j * N + k] = Cm[i][k] * w[k]
    s = np.linalg.solve(a, b)
    return s.reshape(N, N)


def force_reciprocal(S):
    return 0.5 * (S + S.T)
