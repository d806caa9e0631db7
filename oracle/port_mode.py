"""ORACLE (test infrastructure only): the 2D wave-port eigenproblem on a port's triangles.

Restates OpenParEM2D's port-mode solve (``src/OpenParEM2D/fem2D.cpp:2221-2343`` matrix builds,
``eigensolve.c:629-1017`` eigenproblem and post-processing, ``Hsolve.c`` H from E):

* spaces: ND_p (the tangential trace of the 3D ND_p space on the port's boundary triangles, numbered by the 3D
  space's boundary-face DOFs) for e_t and H1_p (continuous P_p) for e_z;
* with Tt = int N.N, St = int curl N curl N, G = int N . grad phi, Sz = int grad phi . grad phi, Tz = int phi phi,
  the generalized eigenproblem (eigensolve.c:776-812)
      A = [St/mu - k0^2 eps Tt, 0; 0, 0],  B = [Tt/mu, G/mu; G^T/mu, Sz/mu - k0^2 eps Tz],  A x = gamma^2 B x,
  with x = (gamma e_t, e_z) for fields (e_t + n e_z) e^{-gamma s}, s along the port's outward normal n
  (postProcess, eigensolve.c:451-505: E_t = x_t / gamma);
* the mode kept is the eigenvalue of smallest real part (EPS_SMALLEST_REAL, eigensolve.c:852) among the finite,
  non-null ones; gamma = sqrt(gamma^2) with alpha, beta >= 0 as in eigensolve.c:446-466;
* PEC and surface-impedance conductors are Dirichlet for e_t and e_z (the ports' perimeter on PEC walls, a
  microstrip trace crossing the port); other boundaries are natural;
* H_t = n x (gamma e_t + grad e_z) / (j omega mu), projected onto ND_p (Hsolve.c: Mt h = Zt/Cz terms);
* V = line integral of E_t along a path, Z = V V* / (2 conj(P)), P = 1/2 int (E_t x H_t*) . n (Z_PV,
  fem2D.cpp:1819).

Dense generalized eigensolve (scipy.linalg.eig) up to 600 free unknowns; above, inverse iteration with SuperLU of
A - sigma B at sigma = -1.05 k0^2 max(eps_r mu_r) and Rayleigh-quotient steps (the product's algorithm; B is
indefinite, so ARPACK's shift-invert mode does not apply).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import nd
from .assemble import EPS0, MU0


def k0_of(freq):
    return 2 * np.pi * freq * np.sqrt(MU0 * EPS0)


def h1_local_basis(p):
    """Hierarchical P_p basis on a triangle with vertices in increasing global id: vertex functions lambda_i, edge
    functions lambda_a lambda_b (lambda_a - lambda_b)^k (a < b, k < p - 1), bubbles lambda_0 lambda_1 lambda_2
    lambda_1^i lambda_2^j (i + j <= p - 3).  Entries: ('v', i) / ('e', a, b, k) / ('f', i, j)."""
    out = [("v", 0), ("v", 1), ("v", 2)]
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for k in range(p - 1):
            out.append(("e", a, b, k))
    for s in range(p - 2):
        for i in range(s + 1):
            out.append(("f", s - i, i))
    return out


def eval_h1(basis, L, g):
    """Values (nq, nh) and in-plane gradients (nf, nq, nh, 3) of the H1 basis at barycentric points L (nq, 3) for
    barycentric gradients g (nf, 3, 3)."""
    nq, nh = L.shape[0], len(basis)
    val = np.empty((nq, nh))
    dval = np.empty((nq, nh, 3))   # derivative with respect to lambda_0..2
    for a, b_ in enumerate(basis):
        d = np.zeros((nq, 3))
        if b_[0] == "v":
            i = b_[1]
            val[:, a] = L[:, i]
            d[:, i] = 1.0
        elif b_[0] == "e":
            _, i, j, k = b_
            t = L[:, i] - L[:, j]
            tk = t ** k
            dtk = k * t ** (k - 1) if k > 0 else np.zeros(nq)
            val[:, a] = L[:, i] * L[:, j] * tk
            d[:, i] = L[:, j] * tk + L[:, i] * L[:, j] * dtk
            d[:, j] = L[:, i] * tk - L[:, i] * L[:, j] * dtk
        else:
            _, pi_, pj = b_
            c = L[:, 0] * L[:, 1] * L[:, 2]
            m = L[:, 1] ** pi_ * L[:, 2] ** pj
            val[:, a] = c * m
            dc = np.stack([L[:, 1] * L[:, 2], L[:, 0] * L[:, 2], L[:, 0] * L[:, 1]], axis=1)
            dm = np.zeros((nq, 3))
            if pi_ > 0:
                dm[:, 1] = pi_ * L[:, 1] ** (pi_ - 1) * L[:, 2] ** pj
            if pj > 0:
                dm[:, 2] = pj * L[:, 1] ** pi_ * L[:, 2] ** (pj - 1)
            d = dc * m[:, None] + c[:, None] * dm
        dval[:, a, :] = d
    grad = np.einsum("qal,fld->fqad", dval, g)
    return val, grad


class PortMode:
    """The 2D eigenproblem of one port of a 3D space (oracle Space), materials from the faces' owner tets."""

    def __init__(self, space, attr, conductor_attrs=None, nquad=None):
        m = space.mesh
        p = space.p
        self.space, self.attr, self.p = space, attr, p
        if conductor_attrs is None:
            conductor_attrs = [a for a, k in m.bdr_kind.items() if k in ("pec", "zs")]
        self.sel = np.flatnonzero(m.bface_attr == attr)
        vs = space.bfaces_sorted[self.sel]                                  # (nf,3) global vertex ids, ascending
        v = m.xyz[vs]
        L, w = nd.tri_rule(nquad or p + 2)
        g, area, nrm = nd.tri_gradients(v)
        self.n = np.asarray(m.port_normal[attr], dtype=np.float64)
        phi, curl = nd.eval_basis(nd.tri_local_basis(p), L, g)              # (nf,nq,ndf,3)
        self.phi = phi
        self.cs = np.einsum("fqak,fk->fqa", curl, nrm)                     # scalar curls (any consistent sign)
        self.wq = w[None, :] * (2 * area)[:, None]
        self.pts = np.einsum("qk,fkd->fqd", L, v)
        self.L, self.g, self.vs, self.v = L, g, vs, v
        # ND: the 3D space's boundary-face DOFs of the port, numbered locally
        self.nd_dofs = space.bface_dofs[self.sel]
        self.nd_glob = np.unique(self.nd_dofs.ravel())
        self.nd_loc = np.searchsorted(self.nd_glob, self.nd_dofs)
        # H1 numbering: vertices, edges (sorted global pair), bubbles
        hb = h1_local_basis(p)
        self.h1_basis = hb
        self.hval, self.hgrad = eval_h1(hb, L, g)
        keys = []
        for f in range(len(self.sel)):
            row = []
            for b_ in hb:
                if b_[0] == "v":
                    row.append(("v", int(vs[f, b_[1]])))
                elif b_[0] == "e":
                    row.append(("e", int(vs[f, b_[1]]), int(vs[f, b_[2]]), b_[3]))
                else:
                    row.append(("f", int(self.sel[f]), b_[1], b_[2]))
            keys.append(row)
        uniq = {}
        for row in keys:
            for k in row:
                uniq.setdefault(k, len(uniq))
        self.h1_keys = list(uniq)
        self.h1_loc = np.array([[uniq[k] for k in row] for row in keys], dtype=np.int64)
        self.nh = len(uniq)
        self.nnd = self.nd_glob.size
        # Dirichlet sets: ND DOFs on conductor faces; H1 vertices and edges of conducting edges
        cond_faces = np.flatnonzero(np.isin(m.bface_attr, list(conductor_attrs)))
        cond_dofs = np.unique(space.bface_dofs[cond_faces].ravel()) if cond_faces.size else np.zeros(0, np.int64)
        self.nd_dir = np.isin(self.nd_glob, cond_dofs)
        cedges = set()
        for f in cond_faces:
            a, b, c = (int(x) for x in space.bfaces_sorted[f])
            cedges.update({(a, b), (a, c), (b, c)})
        self.h1_dir = np.zeros(self.nh, bool)
        for k, i in uniq.items():
            if k[0] == "v":
                self.h1_dir[i] = any(k[1] in e for e in cedges)
            elif k[0] == "e":
                self.h1_dir[i] = (k[1], k[2]) in cedges
        # vertices on a conductor but on no conducting port edge are kept free (point contacts)
        pe = set()
        for f in range(len(self.sel)):
            a, b, c = (int(x) for x in vs[f])
            pe.update({(a, b), (a, c), (b, c)})
        ce = cedges & pe
        for k, i in uniq.items():
            if k[0] == "v":
                self.h1_dir[i] = any(k[1] in e for e in ce)
        # materials of the owner tets
        owner = space.bface_owner[self.sel]
        ta = m.tet_attr[owner]
        self.er = np.array([m.materials[int(a)]["er"] * (1 - 1j * m.materials[int(a)]["tand"]) for a in ta])
        self.mur = np.array([m.materials[int(a)]["mur"] for a in ta], dtype=np.float64)
        self._elem()

    def _elem(self):
        wq, phi, cs, hv, hg = self.wq, self.phi, self.cs, self.hval, self.hgrad
        self.Tt_e = np.einsum("fq,fqak,fqbk->fab", wq, phi, phi)
        self.St_e = np.einsum("fq,fqa,fqb->fab", wq, cs, cs)
        self.G_e = np.einsum("fq,fqak,fqbk->fab", wq, phi, hg)
        self.Sz_e = np.einsum("fq,fqak,fqbk->fab", wq, hg, hg)
        self.Tz_e = np.einsum("fq,qa,qb->fab", wq, hv, hv)
        rphi = np.cross(self.n[None, None, None, :], phi)                 # n x N
        rgrad = np.cross(self.n[None, None, None, :], hg)                 # n x grad phi
        self.Zt_e = np.einsum("fq,fqak,fqbk->fab", wq, phi, rphi)
        self.Cz_e = np.einsum("fq,fqak,fqbk->fab", wq, phi, rgrad)

    def _asm(self, Me, ri, ci_, coef, shape):
        nf, a, b = Me.shape
        vals = (Me * coef[:, None, None]).ravel()
        r = np.repeat(ri, b, axis=1).reshape(nf, a, b) if ri.ndim == 2 else None
        rows = np.broadcast_to(ri[:, :, None], (nf, a, b)).ravel()
        cols = np.broadcast_to(ci_[:, None, :], (nf, a, b)).ravel()
        return sp.csr_matrix((vals, (rows, cols)), shape=shape)

    def matrices(self, freq):
        k0 = k0_of(freq)
        nn, nh = self.nnd, self.nh
        inv_mu = 1.0 / self.mur
        keps = k0 * k0 * self.er
        Tt_mu = self._asm(self.Tt_e, self.nd_loc, self.nd_loc, inv_mu.astype(complex), (nn, nn))
        Tt_eps = self._asm(self.Tt_e, self.nd_loc, self.nd_loc, keps, (nn, nn))
        St = self._asm(self.St_e, self.nd_loc, self.nd_loc, inv_mu.astype(complex), (nn, nn))
        G = self._asm(self.G_e, self.nd_loc, self.h1_loc, inv_mu.astype(complex), (nn, nh))
        Sz = self._asm(self.Sz_e, self.h1_loc, self.h1_loc, inv_mu.astype(complex), (nh, nh))
        Tz = self._asm(self.Tz_e, self.h1_loc, self.h1_loc, keps, (nh, nh))
        A = sp.bmat([[St - Tt_eps, None], [None, sp.csr_matrix((nh, nh))]]).tocsr()
        B = sp.bmat([[Tt_mu, G], [G.T, Sz - Tz]]).tocsr()
        return A, B

    def solve(self, freq, dense_limit=600):
        """gamma, e_t (values on nd_glob), e_z (H1 values), h_t (values on nd_glob) of the dominant mode."""
        k0 = k0_of(freq)
        A, B = self.matrices(freq)
        free = np.concatenate([~self.nd_dir, ~self.h1_dir])
        Af = A[free][:, free]
        Bf = B[free][:, free]
        nfree = int(free.sum())
        epsmax = float(np.max(self.er.real * self.mur))
        if nfree <= dense_limit:
            lam, X = sla.eig(Af.toarray(), Bf.toarray())
            ok = np.isfinite(lam) & (np.abs(lam) > 1e-6 * k0 * k0)
            idx = np.flatnonzero(ok)
            i = idx[np.argmin(lam[idx].real)]
            lam2, x = lam[i], X[:, i]
        else:
            lam2, x = self._inverse_iteration(Af.tocsc(), Bf.tocsc(), -1.05 * k0 * k0 * epsmax)
        gamma = np.sqrt(lam2 + 0j)
        if abs(gamma.imag) > abs(gamma.real):
            if gamma.imag < 0:
                gamma = -gamma
        elif gamma.real < 0:
            gamma = -gamma
        full = np.zeros(free.size, np.complex128)
        full[free] = x
        et = full[:self.nnd] / gamma
        ez = full[self.nnd:]
        ht = self.h_from_e(freq, gamma, et, ez)
        return dict(gamma=complex(gamma), gamma2=complex(lam2), et=et, ez=ez, ht=ht)

    @staticmethod
    def _inverse_iteration(A, B, sigma, tol=1e-12, maxit=300):
        """Eigenpair of A x = lam B x nearest sigma: inverse iteration with (A - sigma B)^-1 B (B is indefinite,
        so ARPACK's shift-invert mode does not apply), then Rayleigh-quotient steps (x^T A x / x^T B x)."""
        n = A.shape[0]
        x = (1.0 + 0.1 * np.sin(1.0 + np.arange(n))) + 0.05j * np.cos(2.0 * np.arange(n))
        lu = spla.splu((A - sigma * B).tocsc())
        lam = sigma
        for it in range(maxit):
            z = lu.solve(B @ x)
            prev, lam = lam, (z @ (A @ z)) / (z @ (B @ z))
            x = z / np.linalg.norm(z)
            if it > 0 and abs(lam - prev) <= 1e-9 * abs(lam):
                break
        for _ in range(3):
            z = spla.splu((A - lam * B).tocsc()).solve(B @ x)
            prev, lam = lam, (z @ (A @ z)) / (z @ (B @ z))
            x = z / np.linalg.norm(z)
            if abs(lam - prev) <= tol * abs(lam):
                break
        return lam, x

    def h_from_e(self, freq, gamma, et, ez):
        """H_t = n x (gamma e_t + grad e_z) / (j omega mu0 mur), L2-projected onto the port's ND_p."""
        omega = 2 * np.pi * freq
        nn, nh = self.nnd, self.nh
        Tt = self._asm(self.Tt_e, self.nd_loc, self.nd_loc, np.ones(len(self.sel), complex), (nn, nn))
        coef = 1.0 / (1j * omega * MU0 * self.mur)
        Zt = self._asm(self.Zt_e, self.nd_loc, self.nd_loc, coef, (nn, nn))
        Cz = self._asm(self.Cz_e, self.nd_loc, self.h1_loc, coef, (nn, nh))
        rhs = gamma * (Zt @ et) + Cz @ ez
        return spla.spsolve(Tt.tocsc(), rhs)

    def field_nd(self, vals):
        """Tangential field of ND values (on nd_glob) at the quadrature points (nf, nq, 3)."""
        return np.einsum("fqak,fa->fqk", self.phi, vals[self.nd_loc])

    def power(self, et, ht):
        """P = 1/2 int (E_t x H_t*) . n dA (exact quadrature)."""
        E, H = self.field_nd(et), self.field_nd(ht)
        return 0.5 * np.sum(self.wq * np.einsum("fqk,k->fq", np.cross(E, np.conj(H)), self.n))

    def eval_at(self, vals, pts):
        """Tangential field of ND values at arbitrary points on the port (nearest-containing triangle)."""
        out = np.zeros((len(pts), 3), np.complex128)
        for k, x in enumerate(pts):
            f, lam = self._locate(x)
            L = lam[None, :]
            phi, _ = nd.eval_basis(nd.tri_local_basis(self.p), L, self.g[f:f + 1])
            out[k] = np.einsum("ak,a->k", phi[0, 0], vals[self.nd_loc[f]])
        return out

    def _locate(self, x):
        best, bf, bl = -1e300, 0, None
        for f in range(len(self.sel)):
            v = self.v[f]
            T = np.stack([v[1] - v[0], v[2] - v[0]], axis=1)
            st, *_ = np.linalg.lstsq(T, x - v[0], rcond=None)
            lam = np.array([1 - st.sum(), st[0], st[1]])
            if lam.min() > best:
                best, bf, bl = lam.min(), f, lam
        return bf, bl

    def voltage(self, et, a, b, nseg=64):
        """V = int_a^b E_t . dl (Gauss-Legendre on nseg pieces of the segment)."""
        a, b = np.asarray(a, float), np.asarray(b, float)
        xg, wg = np.polynomial.legendre.leggauss(6)
        ts, ws = [], []
        for s in range(nseg):
            t0, t1 = s / nseg, (s + 1) / nseg
            ts.extend(t0 + (xg + 1) / 2 * (t1 - t0))
            ws.extend(wg / 2 * (t1 - t0))
        ts, ws = np.array(ts), np.array(ws)
        pts = a[None, :] + ts[:, None] * (b - a)[None, :]
        E = self.eval_at(et, pts)
        return complex(np.sum(ws * (E @ (b - a))))
