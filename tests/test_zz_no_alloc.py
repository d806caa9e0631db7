"""The allocation-free SpMV entry point (opem_spmv_noalloc) against opem_spmv, through the C ABI.

opem_spmv_noalloc stages the caller's matrix and vectors in a fixed device array the library holds per GPU instead
of allocating; it runs the same complex CSR kernel, so its output must equal opem_spmv's bit for bit.  Two calls in
a row, the second on a smaller matrix, check that the staging area carries nothing from one call to the next.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import assemble as oasm
from openparem_amd._lib import lib, ptr
from openparem_amd.mesh import waveguide_mesh
from openparem_amd.solver import Device, spmv

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _noalloc(dev, A, x, precision):
    L = lib()
    rp = np.ascontiguousarray(A.indptr, np.int32)
    ci = np.ascontiguousarray(A.indices, np.int32)
    v = np.ascontiguousarray(A.data, np.complex128)
    xx = np.ascontiguousarray(x, np.complex128)
    y = np.full(A.shape[0], np.nan + 0j, np.complex128)
    rc = L.opem_spmv_noalloc(dev.ctx, 0 if precision == "fp64" else 1, A.shape[0], ptr(rp, C.c_int32),
                             ptr(ci, C.c_int32), ptr(v.view(np.float64), C.c_double),
                             ptr(xx.view(np.float64), C.c_double), ptr(y.view(np.float64), C.c_double))
    return rc, y


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_spmv_on_two_calls(precision):
    """System matrix of a 4x3x5 guide at ND2 (paired rows), then a smaller ND1 guide (single rows)."""
    rng = np.random.default_rng(7)
    cases = []
    for nx, ny, nz, order in ((4, 3, 5, 2), (3, 2, 3, 1)):
        m = waveguide_mesh(nx=nx, ny=ny, nz=nz, seed=9)
        o = oasm.Space(m, order)
        A = oasm.system_matrix(o, 9e9, {2: 130j, 3: 130j}).tocsr()
        A.sort_indices()
        x = rng.standard_normal(o.ndof) + 1j * rng.standard_normal(o.ndof)
        cases.append((A, x))
    assert cases[1][0].shape[0] < cases[0][0].shape[0]
    dev = Device()
    try:
        for A, x in cases:
            print(f"noalloc {precision} n={A.shape[0]} nnz={A.nnz}", flush=True)
            ref = spmv(dev, A.indptr, A.indices, A.data, x, precision)
            rc, y = _noalloc(dev, A, x, precision)
            assert rc == 0, lib().opem_last_error(dev.ctx)
            assert np.array_equal(y, ref), np.abs(y - ref).max()
            exact = A @ x
            tol = 1e-13 if precision == "fp64" else 2e-6
            assert np.abs(y - exact).max() / np.abs(exact).max() < tol
    finally:
        dev.close()


def test_noalloc_rejects_oversize_and_bad_rowptr():
    """Host-side rejections: staged bytes beyond the 64 MiB area, rowptr[0] != 0; the caller's y is untouched."""
    dev = Device()
    try:
        n = 2_000_000  # diagonal: 4(n+1) + 4n + 16n + 32n bytes > 64 MiB
        A = sp.identity(n, dtype=np.complex128, format="csr")
        x = np.ones(n, np.complex128)
        rc, y = _noalloc(dev, A, x, "fp64")
        assert rc == -2 and "staging bytes" in lib().opem_last_error(dev.ctx).decode()
        assert np.isnan(y).all()
        B = sp.identity(8, dtype=np.complex128, format="csr")
        B.indptr = B.indptr + 1
        rc, y = _noalloc(dev, B, np.ones(8, np.complex128), "fp64")
        assert rc == -2 and "rowptr[0]" in lib().opem_last_error(dev.ctx).decode()
        assert np.isnan(y).all()
        # still usable afterwards
        A = sp.random(500, 500, density=0.02, random_state=3, format="csr") * (1 + 2j)
        A.sort_indices()
        x = np.arange(500) * (1 - 0.5j)
        rc, y = _noalloc(dev, A, x, "fp64")
        assert rc == 0 and np.allclose(y, A @ x, rtol=1e-13, atol=1e-12)
    finally:
        dev.close()
