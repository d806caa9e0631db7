"""Error behaviour of the C ABI on the GPU: misuse returns < 0 with a message, leaves the context usable, and
nothing is written to the caller's outputs.  The reference's PETSc seam reports misuse the same way (a nonzero
PetscErrorCode from `solveComplexLinearSystem.c:100-209`), and Krylov nonconvergence is a flag, not an error
(`solveComplexLinearSystem.c:204-206`)."""
import ctypes as C

import numpy as np
import pytest

from openparem_amd import _lib
from openparem_amd._lib import BDR, OpemError, lib, ptr
from openparem_amd.mesh import waveguide_mesh

pytestmark = pytest.mark.gpu


def _err(L, ctx):
    return L.opem_last_error(ctx).decode()


def _dummy_sweep_in(nports, port_attr, freqs, **kw):
    """A SweepIn whose pointers are valid but whose mode arrays are placeholders: every case below fails its
    argument checks before the library reads them."""
    keep = [np.zeros(2, np.float64) for _ in range(5)]
    pa = np.asarray(port_attr, np.int32)
    fr = np.asarray(freqs, np.float64)
    keep += [pa, fr]
    d = dict(nfreq=fr.size, freqs=ptr(fr, C.c_double), nports=nports, port_attr=ptr(pa, C.c_int32),
             mode_E=ptr(keep[0], C.c_double), mode_H=ptr(keep[1], C.c_double), gamma=ptr(keep[2], C.c_double),
             V=ptr(keep[3], C.c_double), Z=ptr(keep[4], C.c_double), precision=_lib.PREC_FP64, batch=0,
             restart=30, maxit=100, tol=1e-10, reciprocal=1, h_tol=1e-13, precond=0, coarse_size=0,
             orth_passes=2, precond_share=1, poly_degree=0, arnoldi_step=1, warm_start=0)
    d.update(kw)
    return _lib.SweepIn(**d), keep


def test_misuse_is_reported_and_leaves_the_context_usable():
    L = lib()
    m = waveguide_mesh(nx=3, ny=2, nz=4, seed=5)
    ctx = C.c_void_p()
    assert L.opem_create(C.byref(ctx), -1) == -2
    assert L.opem_create(None, 0) == -1
    assert L.opem_create(C.byref(ctx), 0) == 0
    S = np.full((1, 2, 2), 7.0 + 7.0j, np.complex128)
    out = _lib.SweepOut(S=ptr(S.view(np.float64), C.c_double))
    try:
        nd, nz = C.c_int64(), C.c_int64()
        print("misuse: ordering of the set-up calls", flush=True)
        assert L.opem_space_build(ctx, 1, C.byref(nd), C.byref(nz)) == -2 and "mesh not set" in _err(L, ctx)
        sin, keep = _dummy_sweep_in(2, [2, 3], [9e9])
        assert L.opem_sweep(ctx, C.byref(sin), C.byref(out)) == -2 and "space not built" in _err(L, ctx)
        assert L.opem_sweep(ctx, None, C.byref(out)) == -2 and "bad sweep arguments" in _err(L, ctx)
        assert L.opem_sweep(None, C.byref(sin), C.byref(out)) == -1
        xyz = np.ascontiguousarray(m.xyz, np.float64)
        tets = np.ascontiguousarray(m.tets, np.int32)
        ta = np.ascontiguousarray(m.tet_attr, np.int32)
        bf = np.ascontiguousarray(m.bfaces, np.int32)
        ba = np.ascontiguousarray(m.bface_attr, np.int32)
        assert L.opem_mesh_set(ctx, xyz.shape[0], ptr(xyz, C.c_double), tets.shape[0], ptr(tets, C.c_int32),
                               ptr(ta, C.c_int32), bf.shape[0], ptr(bf, C.c_int32), ptr(ba, C.c_int32)) == 0
        assert L.opem_space_build(ctx, 4, C.byref(nd), C.byref(nz)) == -2 and "order" in _err(L, ctx)
        assert L.opem_space_build(ctx, 1, C.byref(nd), C.byref(nz)) == 0
        print("misuse: materials, boundaries and ports", flush=True)
        assert L.opem_sweep(ctx, C.byref(sin), C.byref(out)) == -2 and "material not set" in _err(L, ctx)
        vals = np.empty(nz.value, np.float64)
        assert L.opem_base_matrix(ctx, 0, 9, ptr(vals, C.c_double)) == -2
        assert "unknown volume attribute" in _err(L, ctx)
        assert L.opem_material_set(ctx, 9, 1.0, 0.0, 1.0) == 0  # a material the mesh does not use is kept unused
        assert L.opem_material_set(ctx, 1, -1.0, 0.0, 1.0) == -2 and "bad material" in _err(L, ctx)
        assert L.opem_material_set(ctx, 1, 1.0, -0.1, 1.0) == -2 and "bad material" in _err(L, ctx)
        assert L.opem_material_set(ctx, 1, 1.0, 0.0, 1.0) == 0
        assert L.opem_boundary_set(ctx, 1, 99, 0.0) == -2 and "bad boundary kind" in _err(L, ctx)
        assert L.opem_boundary_set(ctx, 1, BDR["zs"], 0.0) == -2 and "conductivity" in _err(L, ctx)
        for a, k in m.bdr_kind.items():
            assert L.opem_boundary_set(ctx, a, BDR[k], 0.0) == 0
        npts = C.c_int64()
        assert L.opem_port_points(ctx, 1, C.byref(npts), ptr(None, C.c_double)) == -2
        assert "not a port" in _err(L, ctx)
        cases = [
            (dict(nports=1, port_attr=[1]), "not an OPEM_BDR_PORT"),
            (dict(nports=1, port_attr=[2]), "every port boundary must be an S-port"),
            (dict(nports=0, port_attr=[2]), "nports must be 1..4"),
            (dict(nports=2, port_attr=[2, 2]), "duplicate"),
            (dict(freqs=[9e9, -1e9]), "frequencies must be positive"),
            (dict(freqs=[float("nan")]), "frequencies must be positive"),
            (dict(precision=5), "bad precision"),
            (dict(precond=3), "bad preconditioner"),
            (dict(coarse_size=129), "coarse_size"),
            (dict(poly_degree=33), "poly_degree"),
            (dict(arnoldi_step=5), "arnoldi_step"),
            (dict(warm_start=-1), "warm_start"),
            (dict(mem_fraction=1.5), "mem_fraction"),
            (dict(mem_fraction=float("nan")), "mem_fraction"),
            (dict(precond_share=-1), "precond_share"),
        ]
        for kw, msg in cases:
            a = dict(nports=2, port_attr=[2, 3], freqs=[9e9])
            a.update(kw)
            sin, keep = _dummy_sweep_in(a.pop("nports"), a.pop("port_attr"), a.pop("freqs"), **a)
            rc = L.opem_sweep(ctx, C.byref(sin), C.byref(out))
            print(f"misuse: {kw} -> rc={rc} {_err(L, ctx)!r}", flush=True)
            assert rc == -2 and msg in _err(L, ctx), (kw, rc, _err(L, ctx))
        assert (S == 7.0 + 7.0j).all()  # no failed call wrote the caller's output
    finally:
        assert L.opem_destroy(ctx) == 0
    assert L.opem_destroy(None) == -1


def test_solver_raises_and_recovers():
    """After failed sweeps the same solver returns the same S as before them; nonconvergence is a flag."""
    from openparem_amd.solver import Solver
    m = waveguide_mesh(nx=3, ny=2, nz=4, seed=5)
    f = np.array([9e9, 11e9])
    with Solver(m, 1) as s:
        r0 = s.sweep(f, precision="mixed", batch=2, restart=40, precond_share=2, arnoldi_step=2)
        assert r0["converged"].all()
        with pytest.raises(OpemError, match=r"rc=-2.*arnoldi_step"):
            s.sweep(f, precision="mixed", batch=2, precond_share=3, arnoldi_step=7)
        with pytest.raises(OpemError, match="poly_degree"):
            s.sweep(f, precision="mixed", batch=2, precond_share=3, poly_degree=40)
        r1 = s.sweep(f, precision="mixed", batch=2, restart=40, precond_share=2, arnoldi_step=2)
        np.testing.assert_array_equal(r1["S"], r0["S"])
        np.testing.assert_array_equal(r1["iters"], r0["iters"])
        r2 = s.sweep(f, precision="fp64", restart=4, maxit=2)  # too few iterations: a flag, not an error
        assert not r2["converged"].any() and (r2["iters"] <= 4).all()


def test_spmv_rejects_bad_csr_before_upload():
    """opem_spmv on a caller's CSR: rowptr[0] != 0, a negative rowptr[n], a column out of range and an unknown
    precision are rejected on the host (-2) before anything reaches the device; y is left untouched."""
    from openparem_amd.solver import Device
    L = lib()
    dev = Device()
    try:
        n = 6
        ci = np.arange(n, dtype=np.int32)
        v = np.zeros(2 * n, np.float64)
        v[0::2] = 1.0  # identity (1 + 0j on the diagonal)
        x = np.ones(2 * n, np.float64)  # 1 + 1j
        for rp, prec, msg in ((np.arange(1, n + 2, dtype=np.int32), 0, "rowptr[0]"),
                              (np.array([0, 1, 2, 3, 4, 5, -6], np.int32), 0, "monotone"),
                              (np.arange(n + 1, dtype=np.int32), 2, "bad precision")):
            y = np.full(2 * n, 5.0)
            rc = L.opem_spmv(dev.ctx, prec, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double),
                             ptr(x, C.c_double), ptr(y, C.c_double))
            print(f"spmv misuse {msg}: rc={rc} {_err(L, dev.ctx)!r}", flush=True)
            assert rc == -2 and msg in _err(L, dev.ctx)
            assert (y == 5.0).all()
        bad = ci.copy()
        bad[3] = n
        y = np.full(2 * n, 5.0)
        rp = np.arange(n + 1, dtype=np.int32)
        assert L.opem_spmv(dev.ctx, 0, n, ptr(rp, C.c_int32), ptr(bad, C.c_int32), ptr(v, C.c_double),
                           ptr(x, C.c_double), ptr(y, C.c_double)) == -2
        assert "colind" in _err(L, dev.ctx) and (y == 5.0).all()
        assert L.opem_spmv(dev.ctx, 0, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double),
                           ptr(x, C.c_double), ptr(y, C.c_double)) == 0
        assert np.array_equal(y, x)
    finally:
        dev.close()


def test_solve_entries_reject_misuse():
    """opem_solve / opem_hfield / opem_port_project: out-of-range port DOFs, bad options and a context without a
    mesh are rejected (-2) before any device work; x and h keep their contents; the context stays usable."""
    from openparem_amd.solver import Device
    L = lib()
    dev = Device()
    try:
        n = 6
        rp = np.arange(n + 1, dtype=np.int32)
        ci = np.arange(n, dtype=np.int32)
        v = np.zeros(2 * n, np.float64)
        v[0::2] = 2.0  # 2 I
        x = np.full(2 * n, 5.0)
        outs = (C.c_double(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32())
        refs = [C.byref(o) for o in outs]

        def solve(port, opts):
            return L.opem_solve(dev.ctx, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double), port.size,
                                ptr(port, C.c_int32), C.byref(opts) if opts is not None else None,
                                ptr(x, C.c_double), *refs)

        ok = _lib.SolveOpts(precision=_lib.PREC_FP64)
        for port, opts, msg in ((np.array([n], np.int32), ok, "port DOF out of range"),
                                (np.array([1], np.int32), _lib.SolveOpts(precision=2), "bad precision"),
                                (np.array([1], np.int32), _lib.SolveOpts(restart=401), "restart"),
                                (np.array([1], np.int32), _lib.SolveOpts(first_rung=3), "bad solve options"),
                                (np.array([1], np.int32), _lib.SolveOpts(tol=-1.0), "bad solve options")):
            rc = solve(port, opts)
            print(f"solve misuse {msg}: rc={rc} {_err(L, dev.ctx)!r}", flush=True)
            assert rc == -2 and msg in _err(L, dev.ctx) and (x == 5.0).all()
        h = np.full(2 * n, 5.0)
        assert L.opem_hfield(dev.ctx, n, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(v, C.c_double), None, None,
                             ptr(h, C.c_double), *refs) == -2 and (h == 5.0).all()
        sp6 = np.zeros(12)
        z = np.zeros(2 * n)
        assert L.opem_port_project(dev.ctx, 2, ptr(z, C.c_double), ptr(z, C.c_double), ptr(z, C.c_double),
                                   ptr(z, C.c_double), ptr(sp6, C.c_double)) == -2
        assert "space not built" in _err(L, dev.ctx)
        # the context still solves: 2 I x = b with the Dirichlet row 1 = 5 + 5j, every other row b = 0
        assert solve(np.array([1], np.int32), None) == 0
        assert outs[2].value == 1 and x[2] == 5.0 and x[3] == 5.0 and (np.delete(x, [2, 3]) == 0.0).all()
    finally:
        dev.close()
