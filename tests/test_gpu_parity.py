"""GPU path (libopem.so through the C ABI) against the FP64 oracle on the same seeded inputs."""
import numpy as np
import pytest
import scipy.sparse as sp

from openparem_amd.mesh import microstrip_mesh, random_tets, waveguide_mesh
from openparem_amd.ports import te10
from oracle import assemble as oasm
from oracle import nd as ond
from oracle.pipeline import Problem

pytestmark = pytest.mark.gpu

COPPER = 5.813e7


def _solver(mesh, p):
    from openparem_amd.solver import Solver
    bnd = {a: COPPER for a, k in mesh.bdr_kind.items() if k == "zs"}
    return Solver(mesh, p, boundaries=bnd)


def _csr(s, vals):
    sp_ = s.space()
    return sp.csr_matrix((vals, sp_["colind"], sp_["rowptr"]), shape=(s.ndof, s.ndof))


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6), (2, 3e-2)])
def test_element_matrices(p, prec, tol):
    from openparem_amd.solver import Device
    print(f"element matrices p={p} prec={prec}", flush=True)
    v, er, td = random_tets(300, seed=11)
    k0 = 2 * np.pi * 5e9 / 299792458.0 * 1e-2  # unit-cube tets: scale k0 so both terms matter
    dev = Device()
    try:
        out, _ = dev.element_matrices(p, prec, v, er, td, k0)
    finally:
        dev.close()
    K, M = ond.element_matrices(v, p)
    ref_re = K - (k0 * k0 * er)[:, None, None] * M
    ref_im = (k0 * k0 * er * td)[:, None, None] * M
    ref = np.stack([ref_re, ref_im], 1)
    err = np.linalg.norm((out - ref).reshape(300, -1), axis=1) / np.linalg.norm(ref.reshape(300, -1), axis=1)
    assert err.max() < tol, err.max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_numbering_and_pattern(p):
    m = waveguide_mesh(nx=3, ny=2, nz=4, seed=5)
    s = _solver(m, p)
    try:
        g = s.space()
    finally:
        s.close()
    o = oasm.Space(m, p)
    assert s.ndof == o.ndof
    np.testing.assert_array_equal(g["elem_dofs"], o.elem_dofs)
    np.testing.assert_array_equal(g["bface_dofs"], o.bface_dofs)
    rp, ci = oasm.pattern(o)
    np.testing.assert_array_equal(g["rowptr"], rp)
    np.testing.assert_array_equal(g["colind"], ci)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_base_matrices(p):
    m = microstrip_mesh(nx=6, ny=4, nz=4, n_sub=1, seed=2)
    s = _solver(m, p)
    try:
        mats = {(w, a): _csr(s, s.base_matrix(w, a)) for w in "KMQ" for a in (1, 2)}
    finally:
        s.close()
    o = oasm.Space(m, p)
    v = o.sorted_coords()
    K, M = ond.element_matrices(v, p)
    Q = ond.mixed_curl_matrices(v, p)
    for a in (1, 2):
        sel = (m.tet_attr == a).astype(float)
        for w, E in (("K", K), ("M", M), ("Q", Q)):
            ref = oasm._scatter(o.elem_dofs, E * sel[:, None, None], o.ndof)
            d = abs(mats[(w, a)] - ref).max()
            assert d <= 1e-12 * abs(ref).max(), (w, a, d)


@pytest.mark.parametrize("p", [1, 2])
def test_system_matrix_microstrip(p):
    m = microstrip_mesh(nx=6, ny=4, nz=4, n_sub=1, seed=2)
    f = 5e9
    gam = {2: 0.3 + 180j, 3: 0.2 + 150j}
    s = _solver(m, p)
    try:
        A = _csr(s, s.system_matrix(f, gam))
    finally:
        s.close()
    o = oasm.Space(m, p)
    rs = np.sqrt(np.pi * f * oasm.MU0 / COPPER)
    ref = oasm.system_matrix(o, f, gam, rs_of_attr={4: rs})
    d = abs(A - ref).max()
    assert d <= 1e-11 * abs(ref).max(), d


@pytest.mark.parametrize("nf", [3, 12, 13, 70])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_sweep_operator(nf, prec, tol):
    """The batched operator of the solve (one pass of the frequency-independent matrices for every frequency:
    lossy substrate, impedance trace, radiation lid, two port ABCs) against A(f) x from the assembled values
    (themselves pinned to the oracle above).  nf = 3 takes the narrow form, 13 the one-system-per-lane tiles,
    12 and 70 the paired tiles (70: a partial last tile); the error is measured against sum_j |A_ij| |x_j|
    per row."""
    m = microstrip_mesh(nx=6, ny=4, nz=4, n_sub=1, seed=2)
    fr = np.linspace(2e9, 9e9, nf)
    gams = [{2: 0.3 + 180j * f / 5e9, 3: 0.2 + 150j * f / 5e9} for f in fr]
    rng = np.random.default_rng(nf)
    s = _solver(m, 2)
    try:
        x = rng.standard_normal((nf, s.ndof)) + 1j * rng.standard_normal((nf, s.ndof))
        y = s.sweep_apply(fr, gams, x, prec)
        A = [_csr(s, s.system_matrix(f, g)) for f, g in zip(fr, gams)]
    finally:
        s.close()
    for i in range(nf):
        ref = A[i] @ x[i]
        bound = abs(A[i]) @ np.abs(x[i])
        err = (np.abs(y[i] - ref) / bound.max()).max()
        assert err < tol, (i, err)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-14), ("fp32", 2e-6)])
def test_spmv(prec, tol):
    from openparem_amd.solver import Device, spmv
    m = waveguide_mesh(nx=4, ny=3, nz=5, seed=9)
    o = oasm.Space(m, 2)
    A = oasm.system_matrix(o, 9e9, {2: 130j, 3: 130j})
    rng = np.random.default_rng(0)
    x = rng.standard_normal(o.ndof) + 1j * rng.standard_normal(o.ndof)
    dev = Device()
    try:
        y = spmv(dev, A.indptr, A.indices, A.data, x, prec)
    finally:
        dev.close()
    ref = A @ x
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("mean_len", [3, 30, 90])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_spmv_ragged(mean_len, prec, tol):
    """Empty rows, one-entry rows and rows several passes long (up to 300 entries), at mean row lengths that
    select 8, 16 and 32 lanes per row; checked against scipy's CSR product, entry by entry."""
    from openparem_amd.solver import Device, spmv
    print(f"spmv ragged mean_len={mean_len} {prec}", flush=True)
    rng = np.random.default_rng(mean_len)
    n = 3000
    lens = rng.poisson(mean_len, n)
    lens[rng.choice(n, 200, replace=False)] = 0
    lens[rng.choice(n, 50, replace=False)] = 1
    lens[rng.choice(n, 20, replace=False)] = 300
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, rows.size)
    vals = rng.standard_normal(rows.size) + 1j * rng.standard_normal(rows.size)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dev = Device()
    try:
        y = spmv(dev, A.indptr, A.indices, A.data, x, prec)
    finally:
        dev.close()
    ref = A @ x
    bound = abs(A) @ np.abs(x)
    assert np.all(y[np.diff(A.indptr) == 0] == 0)
    err = (np.abs(y - ref) / np.maximum(bound, 1e-300)).max()
    assert err < tol, err


def _grouped_csr(rng, ngroups, G, mean_len, n_cols=None, break_one=False):
    """Random complex CSR whose rows come in groups of G with one column list (as ND2/ND3 patterns do), with
    empty groups and groups several passes long."""
    n = ngroups * G
    n_cols = n_cols or n
    lens = rng.poisson(mean_len, ngroups)
    lens[rng.choice(ngroups, ngroups // 15, replace=False)] = 0
    lens[rng.choice(ngroups, 10, replace=False)] = 200
    rowptr = [0]
    cols = []
    for L in lens:
        c = np.sort(rng.choice(n_cols, L, replace=False)).astype(np.int32)
        for _ in range(G):
            cols.append(c)
            rowptr.append(rowptr[-1] + L)
    colind = np.concatenate(cols)
    if break_one:  # one row of one group gets a different first column: the grouping no longer applies
        k = int(np.argmax(lens > 2)) * G + G - 1
        row = colind[rowptr[k]:rowptr[k + 1]]
        row[0] = np.setdiff1d(np.arange(n_cols), row)[0]
        row.sort()
    vals = rng.standard_normal(colind.size) + 1j * rng.standard_normal(colind.size)
    return sp.csr_matrix((vals, colind, np.array(rowptr)), shape=(n, n_cols))


@pytest.mark.parametrize("G,mean_len,break_one", [(2, 40, False), (3, 60, False), (2, 40, True), (3, 5, False)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_spmv_row_groups(G, mean_len, break_one, prec, tol):
    """Rows in groups of 2 or 3 sharing one column list (the grouped kernel), and the same with one row changed
    (the one-row kernel), against scipy's CSR product entry by entry; empty groups give zero rows."""
    from openparem_amd.solver import Device, spmv
    print(f"spmv row groups G={G} mean_len={mean_len} break={break_one} {prec}", flush=True)
    rng = np.random.default_rng(100 * G + mean_len + break_one)
    A = _grouped_csr(rng, 1500, G, mean_len, break_one=break_one)
    x = rng.standard_normal(A.shape[1]) + 1j * rng.standard_normal(A.shape[1])
    dev = Device()
    try:
        y = spmv(dev, A.indptr, A.indices, A.data, x, prec)
    finally:
        dev.close()
    ref = A @ x
    bound = abs(A) @ np.abs(x)
    assert np.all(y[np.diff(A.indptr) == 0] == 0)
    err = (np.abs(y - ref) / np.maximum(bound, 1e-300)).max()
    assert err < tol, err


@pytest.mark.parametrize("p,G", [(1, 1), (2, 2), (3, 3)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_csr_spmv_bench_bytes(p, G, prec):
    """The timing entry runs on the context's pattern, groups its rows as ND_p allows (1, 2, 3) and reports the
    bytes of one launch; misuse (a group size the pattern does not admit, bad lanes) is an error."""
    from openparem_amd._lib import OpemError
    from openparem_amd.solver import Solver
    m = waveguide_mesh(nx=4, ny=3, nz=5, seed=9)
    s = Solver(m, p)
    try:
        sv = 16 if prec == "fp64" else 8
        t, b = s.csr_spmv_bench(prec, nrep=4)
        assert t > 0
        assert b == s.nnz * sv + 4 * (s.nnz // G) + 4 * (s.ndof + 1) + 2 * s.ndof * sv
        t1, b1 = s.csr_spmv_bench(prec, row_group=1, nrep=4)
        assert t1 > 0 and b1 == s.nnz * (sv + 4) + 4 * (s.ndof + 1) + 2 * s.ndof * sv
        with pytest.raises(OpemError, match="rc=-2"):
            s.csr_spmv_bench(prec, row_group=1, lanes_per_row=12)
        with pytest.raises(OpemError, match="rc=-2"):
            s.csr_spmv_bench(prec, row_group=3 if G != 3 else 2)
    finally:
        s.close()


@pytest.mark.parametrize("precond", ["patch", "element"])
@pytest.mark.parametrize("p", [1, 2])
def test_sweep_waveguide_fp64(p, precond):
    """E, H fields and S of the GPU sweep against the oracle's direct solve on a small guide."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    f = np.array([9e9, 11e9])
    s = _solver(m, p)
    try:
        r = s.sweep(f, precision="fp64", tol=1e-12, restart=60, want_fields=True, precond=precond)
    finally:
        s.close()
    assert r["converged"].all(), r
    pr = Problem(m, p)
    from openparem_amd.ports import te10
    for i, fi in enumerate(f):
        modes = {a: te10(m, a, fi) for a in pr.ports}
        o = pr.solve_point(fi, modes)
        for k in range(2):
            xo, ho = o["info"][k]["x"], o["info"][k]["h"]
            ex = np.abs(r["x"][i, k] - xo).max() / np.abs(xo).max()
            eh = np.abs(r["h"][i, k] - ho).max() / np.abs(ho).max()
            print(f"f={fi:.3g} exc={k} |dx|={ex:.3e} |dh|={eh:.3e}", flush=True)
            assert ex < 1e-9 and eh < 1e-9
        print("Cp", np.abs(r["Cp"][i] - o["Cp"]).max(), "Cm", np.abs(r["Cm"][i] - o["Cm"]).max(), flush=True)
        dS = np.abs(r["S"][i] - o["S"]).max()
        assert dS < 1e-9, (fi, dS, r["S"][i], o["S"])


def test_sweep_mixed_vs_fp64():
    """Mixed (FP32 inner, FP64 refinement, batched) against the plain FP64 path and both against the oracle."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    f = np.array([8e9, 10e9, 12e9])
    s = _solver(m, 2)
    try:
        r64 = s.sweep(f, precision="fp64", tol=1e-10, batch=1)
        r32 = s.sweep(f, precision="mixed", tol=1e-10, batch=3)
    finally:
        s.close()
    assert r64["converged"].all() and r32["converged"].all()
    assert np.abs(r64["S"] - r32["S"]).max() < 1e-6
    assert (r32["true_res"] <= 1e-10).all()
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert np.abs(r64["S"][i] - o["S"]).max() < 1e-9
        assert np.abs(r32["S"][i] - o["S"]).max() < 1e-8


def test_reference_mesh_gpu():
    """GPU S on the reference's own 30-tet WR90 mesh: equal to the oracle, and to the reference's
    recorded S (results.ods) at the discretization level of the port mode (see test_oracle)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import wr90
    from openparem_amd.ports import te10
    m, d = wr90.load()
    for p, tol in ((2, 0.03), (3, 0.005)):
        s = _solver(m, p)
        try:
            # tiny, strongly indefinite system: full GMRES (restart > N) avoids restart stagnation
            r = s.sweep(np.array([9e9]), precision="fp64", tol=1e-10, restart=400)
        finally:
            s.close()
        assert r["converged"].all(), (r["true_res"], r["iters"])
        pr = Problem(m, p)
        o = pr.solve_point(9e9, {a: te10(m, a, 9e9) for a in pr.ports})
        assert np.abs(r["S"][0] - o["S"]).max() < 1e-7
        s11, s21 = wr90.reference_s(d, p)
        assert abs(r["S"][0, 0, 0] - s11) < tol and abs(r["S"][0, 1, 0] - s21) < tol


@pytest.mark.parametrize("coarse,passes", [(0, 2), (32, 1)])
def test_sweep_microstrip(coarse, passes):
    """Lossy substrate, surface-impedance trace, radiation lid, quasi-TEM ports: GPU vs oracle."""
    from openparem_amd.microstrip_mode import quasi_tem
    m = microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)
    f = np.array([3e9, 6e9])
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision="mixed", tol=1e-11, batch=2, coarse=coarse, orth_passes=passes)
    finally:
        s.close()
    assert r["converged"].all()
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        rs = np.sqrt(np.pi * fi * oasm.MU0 / COPPER)
        o = pr.solve_point(fi, {a: quasi_tem(m, a, fi) for a in pr.ports}, rs={4: rs})
        assert np.abs(r["S"][i] - o["S"]).max() < 1e-8


def test_c1_guide_against_oracle_fixture():
    """The 12x6x18 WR90 guide at 10 GHz, ND2, 53 280 DOF: S against the oracle's direct solve (the fixture
    tests/golden/c1_guide_10ghz.json, written by tests/golden/make_c1.py), plain FP64 to 1e-10 and mixed to 1e-8,
    and against the closed form: angle S21 = -beta L = 88.0085 degrees, |S11| = 0 up to discretization."""
    import json
    import os
    from openparem_amd import workloads
    d = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_guide_10ghz.json")))
    w = workloads.make("waveguide-c1")
    s = _solver(w["mesh"], w["order"])
    try:
        assert s.ndof == d["ndof"] == 53280
        r64 = s.sweep(w["freqs"], precision="fp64", tol=1e-12, restart=150)
        rmx = s.sweep(w["freqs"], precision="mixed", tol=1e-10, restart=150, orth_passes=1, arnoldi_step=3)
    finally:
        s.close()
    so = np.array([complex(*v) for v in d["S"]]).reshape(2, 2)
    assert r64["converged"].all() and rmx["converged"].all()
    e64, emx = np.abs(r64["S"][0] - so).max(), np.abs(rmx["S"][0] - so).max()
    print(f"C1: |dS| fp64 {e64:.2e} mixed {emx:.2e}, iterations {r64['iters'].tolist()} {rmx['iters'].tolist()}",
          flush=True)
    assert e64 < 1e-10 and emx < 1e-8
    assert abs(np.angle(r64["S"][0, 1, 0], deg=True) - 88.0085) < 0.1
    assert abs(r64["S"][0, 0, 0]) < 5e-3


def test_sweep_microstrip_low_k0h():
    """Low k0 h (0.3 and 0.6 GHz on a few-mm grid): the patch inverses are dominated by their near-kernel
    gradient part, which FP16 hi planes alone round away; with the lo planes the mixed solve must converge to
    the FP64 S without a blow-up.  The FP32 inner vectors themselves are dominated by gradient components here
    (about 1e4 x the rest at 0.3 GHz), so the inner solve needs several times the FP64 iterations (measured
    1641 vs 320 at 0.3 GHz, 433 vs 299 at 0.6 GHz); the bound only excludes a stall."""
    m = microstrip_mesh(nx=10, ny=6, nz=12, n_sub=2, seed=8)
    f = np.array([0.3e9, 0.6e9])
    s = _solver(m, 2)
    try:
        rm = s.sweep(f, precision="mixed", tol=1e-10, batch=2, orth_passes=1, arnoldi_step=3, restart=100,
                     maxit=4000)
        r64 = s.sweep(f, precision="fp64", tol=1e-10, batch=1, restart=100, maxit=4000)
    finally:
        s.close()
    assert r64["converged"].all() and rm["converged"].all()
    assert np.isfinite(rm["S"]).all()
    assert np.abs(rm["S"] - r64["S"]).max() < 1e-7
    assert rm["iters"].max() <= 8 * r64["iters"].max()


@pytest.mark.parametrize("cross,prec", [(False, "mixed"), (False, "fp64"), (True, "mixed"), (True, "fp64")])
def test_sweep_tee_junctions(cross, prec):
    """Three ports (H-plane tee) and four (cross): the batched sweep's N x N S against the oracle's direct
    solve."""
    from openparem_amd.mesh import tee_mesh
    from openparem_amd.ports import te10
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=cross)
    f = np.array([9e9, 10e9])
    n = 4 if cross else 3
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision=prec, tol=1e-11, batch=2, orth_passes=1 if prec == "mixed" else 2,
                    precond_share=2, arnoldi_step=3 if prec == "mixed" else 1)
    finally:
        s.close()
    assert r["converged"].all() and r["S"].shape == (2, n, n)
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert np.abs(r["S"][i] - o["S"]).max() < 1e-8


def test_warm_start_three_ports():
    """Warm start with three excitations per frequency (H-plane tee, five frequencies): S against the oracle."""
    from openparem_amd.mesh import tee_mesh
    from openparem_amd.ports import te10
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3)
    f = np.linspace(7.5e9, 11.5e9, 5)
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision="mixed", tol=1e-11, batch=5, orth_passes=1, precond_share=2, arnoldi_step=3,
                    warm_start=1)
    finally:
        s.close()
    assert r["converged"].all()
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert np.abs(r["S"][i] - o["S"]).max() < 1e-8


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_sweep_one_port(prec):
    """One excitation per frequency (guide shorted at z = L): S11 against the oracle and |S11| = 1 (the lossless
    one-port); the mixed sweep batches nine frequencies on one inverse set with a warm start."""
    from openparem_amd.ports import te10
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4, short=True)
    f = np.linspace(8e9, 12e9, 9)
    s = _solver(m, 2)
    try:
        if prec == "fp64":
            r = s.sweep(f, precision="fp64", tol=1e-11, batch=1, restart=60)
        else:
            r = s.sweep(f, precision="mixed", tol=1e-11, batch=9, orth_passes=1, precond_share=9, arnoldi_step=3,
                        warm_start=1)
    finally:
        s.close()
    assert r["converged"].all() and r["S"].shape == (9, 1, 1)
    assert np.abs(np.abs(r["S"][:, 0, 0]) - 1).max() < 1e-8
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert abs(r["S"][i, 0, 0] - o["S"][0, 0]) < 1e-8, (fi, r["S"][i], o["S"])


def test_sweep_nd3():
    """ND3 (vertex patches exceed 100 DOFs: element-block Jacobi is used)."""
    from openparem_amd.ports import te10
    m = waveguide_mesh(nx=3, ny=2, nz=4, seed=7)
    s = _solver(m, 3)
    try:
        r = s.sweep(np.array([10e9]), precision="fp64", tol=1e-11, restart=200)
    finally:
        s.close()
    assert r["converged"].all()
    pr = Problem(m, 3)
    o = pr.solve_point(10e9, {a: te10(m, a, 10e9) for a in pr.ports})
    assert np.abs(r["S"][0] - o["S"]).max() < 1e-8


@pytest.fixture(scope="module")
def nine_point_fp64():
    """Nine frequencies solved one at a time in FP64 on the GPU and by the oracle's direct solve: the GPU's S must
    equal the oracle's to 1e-9, and the batched runs below are checked against the oracle's S (r1["S"]; the
    GPU's one-at-a-time S is r1["S_gpu"])."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    f = np.linspace(8e9, 12e9, 9)
    s = _solver(m, 2)
    try:
        r1 = s.sweep(f, precision="fp64", tol=1e-10, batch=1)
    finally:
        s.close()
    assert r1["converged"].all()
    pr = Problem(m, 2)
    So = np.array([pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})["S"] for fi in f])
    assert np.abs(r1["S"] - So).max() < 1e-9
    r1 = dict(r1, S_gpu=r1["S"], S=So)
    return m, f, r1


@pytest.mark.parametrize("share,prec", [(1, "mixed"), (2, "mixed"), (4, "mixed"), (5, "mixed"), (8, "mixed"),
                                        (16, "mixed"), (3, "fp64")])
def test_batched_sweep_shared_patch_inverses(nine_point_fp64, share, prec):
    """Nine frequencies in one batch with one set of patch inverses per `share` adjacent frequencies (the
    last set covering fewer when 9 % share != 0): same S as one-at-a-time FP64.  Shares 5, 8 and 16 put 10,
    16 and 18 systems on one inverse set (one MFMA wavefront per 8 systems; the last wavefront of a set holds
    8 systems, 2 systems, or none)."""
    m, f, r1 = nine_point_fp64
    s = _solver(m, 2)
    try:
        rb = s.sweep(f, precision=prec, tol=1e-10, batch=9, orth_passes=1, precond_share=share)
    finally:
        s.close()
    assert rb["converged"].all()
    assert np.abs(rb["S"] - r1["S"]).max() < 1e-6


@pytest.mark.parametrize("prec,levels", [("mixed", 1), ("mixed", 2), ("fp64", 1)])
def test_warm_start(nine_point_fp64, prec, levels):
    """Bisection over the frequency index: every 2^L-th frequency from the Dirichlet lift, then the midpoints,
    each started from the interpolation of its solved neighbours (the last by extrapolation): same S as
    one-at-a-time FP64.  In FP64 (one cycle to the outer tolerance) the warm-started frequencies need fewer
    iterations; mixed precision trades iterations for shorter inner cycles on warm starts (1e-3 per cycle), which
    pays on long cycles only, so no iteration bound is asserted there."""
    m, f, r1 = nine_point_fp64
    s = _solver(m, 2)
    try:
        kw = dict(precision=prec, tol=1e-10, batch=9, orth_passes=1 if prec == "mixed" else 2, precond_share=1)
        rc = s.sweep(f, warm_start=0, want_fields=True, **kw)
        rw = s.sweep(f, warm_start=levels, want_fields=True, **kw)
        r8 = s.sweep(f[:8], warm_start=levels, **kw)  # even count: the last frequency is extrapolated
    finally:
        s.close()
    assert rw["converged"].all() and r8["converged"].all()
    assert np.abs(rw["S"] - r1["S"]).max() < 1e-6
    # fields come back in the caller's frequency order although the batches solve them out of order
    for key in ("x", "h"):
        d = np.abs(rw[key] - rc[key]).max(axis=2) / np.abs(rc[key]).max(axis=2)
        assert d.max() < 1e-6, key
    assert np.abs(r8["S"] - r1["S"][:8]).max() < 1e-6
    if prec == "fp64":
        assert rw["iters"][1::2].sum() < rc["iters"][1::2].sum()


def test_shared_patch_inverse_limits(nine_point_fp64):
    """An inverse set serves at most 32 systems in mixed precision (MFMA apply) and 8 in FP64; warm_start is 0..4."""
    from openparem_amd._lib import OpemError
    m, f, _ = nine_point_fp64
    s = _solver(m, 2)
    try:
        with pytest.raises(OpemError, match="<= 8"):
            s.sweep(f, precision="fp64", tol=1e-10, batch=9, precond_share=5)
        with pytest.raises(OpemError, match="<= 32"):
            s.sweep(f, precision="mixed", tol=1e-10, batch=9, precond_share=17)
        with pytest.raises(OpemError, match="warm_start"):
            s.sweep(f, precision="mixed", tol=1e-10, batch=9, warm_start=5)
    finally:
        s.close()


def test_sweep_splits_large_batches():
    """130 frequencies x 2 ports with batch = 0 (all) exceed the 256 systems of one batch: the library splits the
    sweep (128 + 2 frequencies) instead of refusing it, with the S of batches of 13."""
    m = waveguide_mesh(nx=3, ny=2, nz=4, seed=5)
    f = np.linspace(8e9, 12e9, 130)
    s = _solver(m, 1)
    try:
        ra = s.sweep(f, precision="fp64", batch=0, restart=60, tol=1e-11)
        rb = s.sweep(f, precision="fp64", batch=13, restart=60, tol=1e-11)
    finally:
        s.close()
    assert ra["converged"].all() and rb["converged"].all()
    assert np.abs(ra["S"] - rb["S"]).max() < 1e-9
    pr = Problem(m, 1)
    for i in (0, 64, 127, 128, 129):  # the first batch, its last frequency, and the split-off batch
        o = pr.solve_point(f[i], {a: te10(m, a, f[i]) for a in pr.ports})
        assert np.abs(ra["S"][i] - o["S"]).max() < 1e-9


@pytest.mark.parametrize("deg,prec", [(3, "mixed"), (4, "mixed"), (5, "fp64")])
def test_polynomial_preconditioner(nine_point_fp64, deg, prec):
    """GMRES-polynomial preconditioning (Krylov space of B p(B), roots from a deg-step Arnoldi run per system):
    the batched nine-frequency sweep reaches the same S as one-at-a-time FP64 without it."""
    m, f, r1 = nine_point_fp64
    s = _solver(m, 2)
    try:
        rb = s.sweep(f, precision=prec, tol=1e-10, batch=9, orth_passes=1, precond_share=2, poly_degree=deg)
    finally:
        s.close()
    assert rb["converged"].all()
    assert np.abs(rb["S"] - r1["S"]).max() < 1e-6


@pytest.mark.parametrize("step,prec,restart", [(2, "mixed", 150), (2, "mixed", 31), (2, "fp64", 150),
                                               (3, "mixed", 150), (3, "mixed", 31), (4, "mixed", 150),
                                               (4, "mixed", 30), (4, "fp64", 150)])
def test_block_arnoldi(nine_point_fp64, step, prec, restart):
    """s Krylov vectors per Gram-Schmidt pass (block projection, Cholesky QR of the block, Hessenberg columns from
    the raw s-step relation; restarts 30 and 31 end cycles on shorter blocks or single steps): same S as
    one-at-a-time FP64 with the one-step Arnoldi, and a comparable iteration count."""
    m, f, r1 = nine_point_fp64
    s = _solver(m, 2)
    try:
        rb = s.sweep(f, precision=prec, tol=1e-10, batch=9, orth_passes=1, precond_share=2, arnoldi_step=step,
                     restart=restart)
    finally:
        s.close()
    assert rb["converged"].all()
    assert np.abs(rb["S"] - r1["S"]).max() < 1e-6
    assert rb["iters"].max() <= 2 * r1["iters"].max() + 2 * restart
