"""Host-side mirror of the reference's per-frequency solve path over the C ABI (libopem.so).

Call surface, in the reference's order (``src/OpenParEM3D/OpenParEM3D.cpp:367-566``):

    s = Solver(mesh, order)                       # fem3D::build_fe_spaces + MFEM assembly on GPU
    r = s.sweep(freqs, precision="fp64"|"mixed")  # per frequency: build_A, eliminatePEC, fillX,
                                                  # solveComplexLinearSystem, buildHgrids,
                                                  # calculateSplits, calculateS, forceReciprocal
    r["S"][f]                                     # S matrix per frequency (S-port order)

Nonconvergence is reported in r["converged"], never raised (solveComplexLinearSystem.c:204-206).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BDR, check, cplx_view, f64p, i32p, i64p, lib, ptr
from .ports import provider_for


class Solver:
    def __init__(self, mesh, order, device=0, materials=None, boundaries=None):
        L = lib()
        self.mesh = mesh
        self.order = order
        self.ctx = C.c_void_p()
        rc = L.opem_create(C.byref(self.ctx), device)
        if rc != 0:
            raise _lib.OpemError(f"opem_create failed (rc={rc}): no usable GPU device {device}")
        xyz = np.ascontiguousarray(mesh.xyz, np.float64)
        tets = np.ascontiguousarray(mesh.tets, np.int32)
        ta = np.ascontiguousarray(mesh.tet_attr, np.int32)
        bf = np.ascontiguousarray(mesh.bfaces, np.int32)
        ba = np.ascontiguousarray(mesh.bface_attr, np.int32)
        check(L.opem_mesh_set(self.ctx, xyz.shape[0], ptr(xyz, C.c_double), tets.shape[0], ptr(tets, C.c_int32),
                              ptr(ta, C.c_int32), bf.shape[0], ptr(bf, C.c_int32), ptr(ba, C.c_int32)),
              self.ctx, "opem_mesh_set")
        nd, nz = C.c_int64(), C.c_int64()
        check(L.opem_space_build(self.ctx, order, C.byref(nd), C.byref(nz)), self.ctx, "opem_space_build")
        self.ndof, self.nnz = nd.value, nz.value
        for attr, m in (materials or mesh.materials).items():
            check(L.opem_material_set(self.ctx, int(attr), m["er"], m["tand"], m["mur"]), self.ctx,
                  "opem_material_set")
        bvals = boundaries or {}
        for attr, kind in mesh.bdr_kind.items():
            val = bvals.get(attr, 0.0)
            if kind == "zs" and not val:
                val = mesh.conductivity.get(attr, 5.8e7) if hasattr(mesh, "conductivity") else 5.8e7
            check(L.opem_boundary_set(self.ctx, int(attr), BDR[kind], float(val)), self.ctx, "opem_boundary_set")
        self.ports = sorted(a for a, k in mesh.bdr_kind.items() if k == "port")

    def close(self):
        if self.ctx:
            lib().opem_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ inspection
    def space(self):
        nt, nbf = self.mesh.nt, self.mesh.bfaces.shape[0]
        nd = self.order * (self.order + 2) * (self.order + 3) // 2
        ndf = self.order * (self.order + 2)
        ed = np.empty((nt, nd), np.int32)
        bd = np.empty((nbf, ndf), np.int32)
        rp = np.empty(self.ndof + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        check(lib().opem_space_get(self.ctx, ptr(ed, C.c_int32), ptr(bd, C.c_int32), ptr(rp, C.c_int32),
                                   ptr(ci, C.c_int32)), self.ctx, "opem_space_get")
        return dict(elem_dofs=ed, bface_dofs=bd, rowptr=rp, colind=ci)

    def base_matrix(self, which, attr):
        out = np.empty(self.nnz, np.float64)
        w = {"K": 0, "M": 1, "Q": 2}[which]
        check(lib().opem_base_matrix(self.ctx, w, attr, ptr(out, C.c_double)), self.ctx, "opem_base_matrix")
        return out

    def system_matrix(self, freq, gammas):
        g = cplx_view(np.asarray([gammas[a] for a in self.ports], np.complex128))
        out = np.empty(self.nnz, np.complex128)
        check(lib().opem_system_matrix(self.ctx, freq, ptr(g, C.c_double), ptr(out.view(np.float64), C.c_double)),
              self.ctx, "opem_system_matrix")
        return out

    def sweep_apply(self, freqs, gammas, x, precision="fp64"):
        """y[f] = A(freqs[f]) x[f] through the batched sweep operator (gammas[f] = {attr: gamma})."""
        freqs = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        g = np.asarray([[gm[a] for a in self.ports] for gm in gammas], np.complex128)
        g = cplx_view(np.ascontiguousarray(g))
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        check(lib().opem_sweep_apply(self.ctx, 0 if precision == "fp64" else 1, freqs.size, ptr(freqs, C.c_double),
                                     ptr(g, C.c_double), ptr(x.view(np.float64), C.c_double),
                                     ptr(y.view(np.float64), C.c_double)), self.ctx, "opem_sweep_apply")
        return y

    def port_points(self, attr):
        n = C.c_int64()
        check(lib().opem_port_points(self.ctx, attr, C.byref(n), ptr(None, C.c_double)), self.ctx, "port_points")
        pts = np.empty((n.value, 3), np.float64)
        check(lib().opem_port_points(self.ctx, attr, C.byref(n), ptr(pts, C.c_double)), self.ctx, "port_points")
        return pts

    def port_project(self, attr, mode, x, h):
        """opem_port_project: {C+, C-, e0, e2, h0, h2} of port `attr` for the E and H DOF vectors x, h, with the
        modal fields of `mode` (a provider dict with callables E, H) sampled at the port's points."""
        pts = self.port_points(attr)
        Ev = cplx_view(mode["E"](pts).ravel())
        Hv = cplx_view(mode["H"](pts).ravel())
        out = np.zeros(6, np.complex128)
        check(lib().opem_port_project(self.ctx, int(attr), ptr(Ev, C.c_double), ptr(Hv, C.c_double),
                                      ptr(cplx_view(x), C.c_double), ptr(cplx_view(h), C.c_double),
                                      ptr(out.view(np.float64), C.c_double)), self.ctx, "opem_port_project")
        return dict(zip(("Cp", "Cm", "e0", "e2", "h0", "h2"), out.tolist()))

    def port_mode(self, attr, freq, vpath=None):
        """opem_port_mode: the dominant mode of port ``attr`` at ``freq`` from the 2D port eigenproblem.  Returns
        gamma, E and H (complex (npts, 3) at the port's points), V and Z; ``vpath`` = (start, end) of a voltage path
        in the port (V = 1 without one)."""
        n = C.c_int64()
        check(lib().opem_port_points(self.ctx, attr, C.byref(n), ptr(None, C.c_double)), self.ctx, "port_points")
        E = np.zeros((n.value, 3), np.complex128)
        H = np.zeros_like(E)
        g, V, Z = (np.zeros(2) for _ in range(3))
        vp = None if vpath is None else np.ascontiguousarray(np.concatenate([np.ravel(vpath[0]), np.ravel(vpath[1])]),
                                                              np.float64)
        check(lib().opem_port_mode(self.ctx, int(attr), float(freq), ptr(vp, C.c_double), ptr(g, C.c_double),
                                   ptr(E.view(np.float64), C.c_double), ptr(H.view(np.float64), C.c_double),
                                   ptr(V, C.c_double), ptr(Z, C.c_double)), self.ctx, "opem_port_mode")
        return dict(gamma=complex(g[0], g[1]), E=E, H=H, V=complex(V[0], V[1]), Z=complex(Z[0], Z[1]))

    def port_mode_provider(self, vpaths=None):
        """A ``mode_fn`` for ``sweep`` that takes every port's mode from the 2D port eigenproblem (opem_port_mode);
        ``vpaths``: {attr: (start, end)} voltage paths."""
        pts = {a: self.port_points(a) for a in self.ports}

        def mode_fn(mesh, attr, freq):
            md = self.port_mode(attr, freq, None if vpaths is None else vpaths.get(attr))

            def sampled(values):
                def at(p):
                    if p.shape != pts[attr].shape or not np.array_equal(p, pts[attr]):
                        raise ValueError("2D port modes are sampled at the port's own points only")
                    return values
                return at
            return dict(E=sampled(md["E"]), H=sampled(md["H"]), gamma=md["gamma"], V=md["V"], Z=md["Z"])
        return mode_fn

    def spmv_bench(self, precision="fp32", ngroups=1, R=2, nrep=20):
        """(mean kernel seconds, algorithmic bytes) of the batched SpMV on this pattern."""
        t, b = C.c_double(), C.c_double()
        check(lib().opem_spmv_bench(self.ctx, 0 if precision == "fp64" else 1, ngroups, R, nrep, C.byref(t),
                                    C.byref(b)), self.ctx, "opem_spmv_bench")
        return t.value, b.value

    def csr_spmv_bench(self, precision="fp64", row_group=0, lanes_per_row=0, nrep=50):
        """(mean kernel seconds, bytes the kernel must move) of the single-matrix complex CSR SpMV with explicit
        values on this pattern (each launch on its own copy of the matrix).  row_group 0 = rows grouped as the
        pattern allows (ND2 pairs, ND3 triples), 1 = one row at a time."""
        t, b = C.c_double(), C.c_double()
        check(lib().opem_csr_spmv_bench(self.ctx, 0 if precision == "fp64" else 1, row_group, lanes_per_row, nrep,
                                        C.byref(t), C.byref(b)), self.ctx, "opem_csr_spmv_bench")
        return t.value, b.value

    def krylov_bench(self, precision="fp32", S=128, k=76, nrep=5, ns=1):
        """((projection s, update s), (projection bytes, update bytes)) of the batched Gram-Schmidt kernels with
        ns new vectors per pass."""
        t = (C.c_double * 2)()
        b = (C.c_double * 2)()
        check(lib().opem_krylov_bench(self.ctx, 0 if precision == "fp64" else 1, S, k, ns, nrep, t, b), self.ctx,
              "opem_krylov_bench")
        return (t[0], t[1]), (b[0], b[1])

    # ------------------------------------------------------------------ hot path
    def sweep(self, freqs, precision="fp64", batch=0, restart=150, maxit=20000, tol=1e-10, reciprocal=True,
              h_tol=1e-13, want_fields=False, mode_fn=None, precond="patch", coarse=0, orth_passes=2,
              precond_share=1, poly_degree=0, arnoldi_step=1, warm_start=0, mem_fraction=0.0):
        """One ``opem_sweep``: S (forced reciprocal unless ``reciprocal=False``) for every frequency, all port
        excitations.  ``precision`` "fp64" or "mixed" (FP32 inner GMRES, FP64 refinement); ``batch`` frequencies per
        batched solve (0: all); GMRES(``restart``) to the true residual ``tol``; ``precond`` "patch" (vertex
        patches), "element" (element blocks) or "lu" (sparse LU of each frequency's matrix, FP32 factors in mixed
        precision); ``precond_share`` adjacent frequencies per set of patch inverses or per LU factor (their mean
        matrix);
        ``arnoldi_step`` Krylov vectors per Gram-Schmidt pass; ``warm_start`` bisection levels over the frequency
        index (interpolated starting iterates); ``mem_fraction`` the share of device memory the batch planner may
        fill (0: three quarters; contexts sweeping on one GPU at once divide it).  ``mode_fn(mesh, attr, f)`` supplies each port's modal E, H, gamma,
        V, Z (default: the mesh family's provider).  Returns S, S_raw, Cp, Cm, iters, true_res, converged,
        fallback (FP64 GMRES finished a system whose FP32 refinement stagnated), singular (the system's LU batch met
        exactly zero pivots) and, with ``want_fields``, the E and H DOFs x, h [freq][excitation][ndof]."""
        freqs = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        nf, np_ = freqs.size, len(self.ports)
        mode_fn = mode_fn or provider_for(self.mesh)
        pts = [self.port_points(a) for a in self.ports]
        E, H = [], []
        gam = np.empty((nf, np_), np.complex128)
        V = np.empty((nf, np_), np.complex128)
        Z = np.empty((nf, np_), np.complex128)
        for i, f in enumerate(freqs):
            for k, a in enumerate(self.ports):
                md = mode_fn(self.mesh, a, f)
                E.append(md["E"](pts[k]).ravel())
                H.append(md["H"](pts[k]).ravel())
                gam[i, k], V[i, k], Z[i, k] = md["gamma"], md["V"], md["Z"]
        E = cplx_view(np.concatenate(E))
        H = cplx_view(np.concatenate(H))
        pa = np.asarray(self.ports, np.int32)
        gv, Vv, Zv = cplx_view(gam), cplx_view(V), cplx_view(Z)
        sin = _lib.SweepIn(nfreq=nf, freqs=ptr(freqs, C.c_double), nports=np_, port_attr=ptr(pa, C.c_int32),
                           mode_E=ptr(E, C.c_double), mode_H=ptr(H, C.c_double), gamma=ptr(gv, C.c_double),
                           V=ptr(Vv, C.c_double), Z=ptr(Zv, C.c_double),
                           precision=_lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64,
                           batch=int(batch), restart=int(restart), maxit=int(maxit), tol=float(tol),
                           reciprocal=int(bool(reciprocal)), h_tol=float(h_tol),
                           precond={"patch": 0, "element": 1, "lu": 2}[precond], coarse_size=int(coarse), orth_passes=int(orth_passes),
                           precond_share=int(precond_share), poly_degree=int(poly_degree),
                           arnoldi_step=int(arnoldi_step), warm_start=int(warm_start),
                           mem_fraction=float(mem_fraction))
        S = np.zeros((nf, np_, np_), np.complex128)
        Sr = np.zeros_like(S)
        Cp = np.zeros_like(S)
        Cm = np.zeros_like(S)
        it = np.zeros((nf, np_), np.int32)
        res = np.zeros((nf, np_), np.float64)
        cv = np.zeros((nf, np_), np.int32)
        fb = np.zeros((nf, np_), np.int32)
        sg = np.zeros((nf, np_), np.int32)
        x = np.zeros((nf, np_, self.ndof), np.complex128) if want_fields else None
        h = np.zeros((nf, np_, self.ndof), np.complex128) if want_fields else None
        fv = lambda a: ptr(a.view(np.float64), C.c_double) if a is not None else ptr(None, C.c_double)
        sout = _lib.SweepOut(S=fv(S), S_raw=fv(Sr), Cp=fv(Cp), Cm=fv(Cm), iters=ptr(it, C.c_int32),
                             true_res=ptr(res, C.c_double), converged=ptr(cv, C.c_int32), x_out=fv(x), h_out=fv(h),
                             fallback=ptr(fb, C.c_int32), singular=ptr(sg, C.c_int32))
        check(lib().opem_sweep(self.ctx, C.byref(sin), C.byref(sout)), self.ctx, "opem_sweep")
        return dict(S=S, S_raw=Sr, Cp=Cp, Cm=Cm, iters=it, true_res=res, converged=cv, x=x, h=h, fallback=fb,
                    singular=sg,
                    seconds=dict(total=sout.seconds_total, setup=sout.seconds_setup, solve=sout.seconds_solve,
                                 post=sout.seconds_post))


def solve(ctx_holder, A, port_dof, x, precision="fp64", restart=30, maxit=10000, tol=1e-10, first_rung=0):
    """opem_solve on a scipy CSR matrix A (complex): x carries the Dirichlet values on port_dof.  Returns the
    solution and (rnorm, iters, converged, rung, flags) with flags the OPEM_SOLVE_* bits (1: FP64 continuation after
    FP32 stagnation, 2: zero pivots met)."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, np.int32)
    ci = np.ascontiguousarray(A.indices, np.int32)
    v = cplx_view(A.data)
    pdof = np.ascontiguousarray(port_dof, np.int32)
    xx = np.ascontiguousarray(x, np.complex128).copy()
    o = _lib.SolveOpts(precision=_lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64, restart=int(restart),
                       maxit=int(maxit), tol=float(tol), first_rung=int(first_rung))
    rn, it, cv, rg, fl = C.c_double(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().opem_solve(ctx_holder.ctx, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double),
                           pdof.size, ptr(pdof, C.c_int32), C.byref(o), ptr(xx.view(np.float64), C.c_double),
                           C.byref(rn), C.byref(it), C.byref(cv), C.byref(rg), C.byref(fl)), ctx_holder.ctx,
          "opem_solve")
    return xx, (rn.value, it.value, cv.value, rg.value, fl.value)


def hfield(ctx_holder, P, b, precision="fp64", restart=30, maxit=10000, tol=1e-10, first_rung=0):
    """opem_hfield: P h = b for a scipy CSR P.  Returns h and (rnorm, iters, converged, rung, flags)."""
    P = P.tocsr().astype(np.complex128)
    P.sort_indices()
    n = P.shape[0]
    rp = np.ascontiguousarray(P.indptr, np.int32)
    ci = np.ascontiguousarray(P.indices, np.int32)
    v = cplx_view(P.data)
    bb = cplx_view(b)
    h = np.zeros(n, np.complex128)
    o = _lib.SolveOpts(precision=_lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64, restart=int(restart),
                       maxit=int(maxit), tol=float(tol), first_rung=int(first_rung))
    rn, it, cv, rg, fl = C.c_double(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().opem_hfield(ctx_holder.ctx, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double),
                            ptr(bb, C.c_double), C.byref(o), ptr(h.view(np.float64), C.c_double), C.byref(rn),
                            C.byref(it), C.byref(cv), C.byref(rg), C.byref(fl)), ctx_holder.ctx, "opem_hfield")
    return h, (rn.value, it.value, cv.value, rg.value, fl.value)


LU_STAT_KEYS = ("gemm_s", "gemm_flops", "gemm_launches", "panel_s", "panel_launches", "apply_s", "apply_bytes",
                "apply_launches", "factor_wall_s", "matrices_factored", "matrix_applies", "event_timing",
                "gemm128_s", "gemm128_flops", "gemm128_launches")


def lu_stats(solver, reset=False):
    """opem_lu_stats: sparse-LU kernel timings accumulated on the solver's context (dict)."""
    out = np.zeros(15, np.float64)
    check(lib().opem_lu_stats(solver.ctx, int(bool(reset)), ptr(out, C.c_double)), solver.ctx, "opem_lu_stats")
    return dict(zip(LU_STAT_KEYS, out.tolist()))


def spmv(ctx_solver, rowptr, colind, vals, x, precision="fp64"):
    """y = A x on the device (complex CSR parity helper)."""
    n = rowptr.size - 1
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    v = cplx_view(vals)
    xv = cplx_view(x)
    y = np.empty(n, np.complex128)
    check(lib().opem_spmv(ctx_solver.ctx, 0 if precision == "fp64" else 1, n, ptr(rp, C.c_int32),
                          ptr(ci, C.c_int32), ptr(v, C.c_double), ptr(xv, C.c_double),
                          ptr(y.view(np.float64), C.c_double)), ctx_solver.ctx, "opem_spmv")
    return y


class Device:
    """Bare context (no mesh) for kernel-level calls."""

    def __init__(self, device=0):
        self.ctx = C.c_void_p()
        rc = lib().opem_create(C.byref(self.ctx), device)
        if rc != 0:
            raise _lib.OpemError(f"opem_create failed (rc={rc})")

    def close(self):
        if self.ctx:
            lib().opem_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def element_matrices(self, order, prec, verts, er, tand, k0, nrep=1, want=True):
        ne = verts.shape[0]
        nd = order * (order + 2) * (order + 3) // 2
        v = np.ascontiguousarray(verts, np.float64)
        e = np.ascontiguousarray(er, np.float64)
        t = np.ascontiguousarray(tand, np.float64)
        out = np.empty((ne, 2, nd, nd), np.float64) if want else None
        sec = C.c_double()
        check(lib().opem_element_matrices(self.ctx, order, prec, ne, ptr(v, C.c_double), ptr(e, C.c_double),
                                          ptr(t, C.c_double), k0, nrep, ptr(out, C.c_double) if want else
                                          ptr(None, C.c_double), C.byref(sec)), self.ctx, "opem_element_matrices")
        return out, sec.value
