"""Sparse-LU-preconditioned sweeps (precond "lu": nested-dissection multifrontal factor on the GPU, the counterpart of
the PCCHOLESKY/PCLU preconditioner of `solveComplexLinearSystem.c:133-168`) against the FP64 oracle's direct solve.

A factor of the system's own matrix makes GMRES converge in a few iterations, so these tests also bound the
iteration counts: a wrong factor (a dropped update, a misplaced pivot, a wrong root block) shows up as many
iterations long before it shows up as a wrong S.
"""
import json
import os

import numpy as np
import pytest

from openparem_amd.mesh import microstrip_mesh, tee_mesh, waveguide_mesh
from openparem_amd.ports import te10
from oracle import assemble as oasm
from oracle.pipeline import Problem

pytestmark = pytest.mark.gpu

COPPER = 5.813e7
HERE = os.path.dirname(os.path.abspath(__file__))


def _solver(mesh, p):
    from openparem_amd.solver import Solver
    bnd = {a: COPPER for a, k in mesh.bdr_kind.items() if k == "zs"}
    return Solver(mesh, p, boundaries=bnd)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_lu_sweep_guide_against_oracle(p):
    """Two-port guide at ND1-ND3, two frequencies: S and the E and H DOFs against the oracle; FP64 factors need one
    GMRES cycle of at most a few iterations, FP32 factors a few refinement cycles."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    f = np.array([9e9, 11e9])
    print(f"lu guide p={p}", flush=True)
    s = _solver(m, p)
    try:
        r64 = s.sweep(f, precision="fp64", tol=1e-12, batch=2, restart=30, precond="lu", want_fields=True)
        rmx = s.sweep(f, precision="mixed", tol=1e-11, batch=2, restart=30, precond="lu", orth_passes=1)
    finally:
        s.close()
    print(f"iterations fp64 {r64['iters'].tolist()} mixed {rmx['iters'].tolist()}", flush=True)
    assert r64["converged"].all() and rmx["converged"].all()
    assert r64["iters"].max() <= 6 and rmx["iters"].max() <= 20
    pr = Problem(m, p)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert np.abs(r64["S"][i] - o["S"]).max() < 1e-9
        assert np.abs(rmx["S"][i] - o["S"]).max() < 1e-8
        for k in range(len(pr.ports)):
            xo, ho = o["info"][k]["x"], o["info"][k]["h"]
            assert np.abs(r64["x"][i, k] - xo).max() < 1e-10 * np.abs(xo).max()
            assert np.abs(r64["h"][i, k] - ho).max() < 1e-9 * np.abs(ho).max()


def test_lu_one_port_short():
    """One port (guide shorted at z = L): the port root has no free rows in the only excitation."""
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4, short=True)
    f = np.linspace(8e9, 12e9, 4)
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision="mixed", tol=1e-11, batch=4, restart=30, precond="lu", precond_share=2)
    finally:
        s.close()
    print(f"one-port iterations {r['iters'].ravel().tolist()}", flush=True)
    assert r["converged"].all() and np.abs(np.abs(r["S"][:, 0, 0]) - 1).max() < 1e-8
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert abs(r["S"][i, 0, 0] - o["S"][0, 0]) < 1e-8


@pytest.mark.parametrize("cross", [False, True])
def test_lu_tee_and_cross(cross):
    """Three and four ports: every excitation factors the port rows of the other ports."""
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=cross)
    f = np.array([9e9, 10e9])
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision="fp64", tol=1e-12, batch=2, restart=30, precond="lu")
    finally:
        s.close()
    print(f"{'cross' if cross else 'tee'} iterations {r['iters'].tolist()}", flush=True)
    assert r["converged"].all() and r["iters"].max() <= 6
    pr = Problem(m, 2)
    for i, fi in enumerate(f):
        o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
        assert np.abs(r["S"][i] - o["S"]).max() < 1e-9


@pytest.mark.parametrize("share", [4, 8])
def test_lu_cross_many_right_hand_sides(share):
    """Four ports and FP32 factors shared by 4 or 8 close frequencies: 16 or 32 right-hand sides per factor, so the
    apply runs with 2 or 4 row groups of 8 right-hand sides per thread.  S against the oracle's direct solve."""
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=True)
    f = np.linspace(9.0e9, 9.35e9, 8)
    s = _solver(m, 2)
    try:
        r = s.sweep(f, precision="mixed", tol=1e-11, batch=8, restart=30, precond="lu", orth_passes=1,
                    precond_share=share)
    finally:
        s.close()
    print(f"cross share {share}: iterations {r['iters'].tolist()}", flush=True)
    assert r["converged"].all() and r["iters"].max() <= 40
    pr = Problem(m, 2)
    for i in (0, 3, 7):
        o = pr.solve_point(f[i], {a: te10(m, a, f[i]) for a in pr.ports})
        assert np.abs(r["S"][i] - o["S"]).max() < 1e-8


@pytest.fixture(scope="module")
def band():
    d = json.load(open(os.path.join(HERE, "golden", "microstrip_band.json")))
    return d, np.array(d["freqs"]), np.array([[complex(*v) for v in S] for S in d["S"]]).reshape(-1, 2, 2)


@pytest.mark.parametrize("prec,share", [("fp64", 1), ("mixed", 1), ("mixed", 2), ("mixed", 4)])
def test_lu_microstrip_band(band, prec, share):
    """The reduced microstrip box over 1-20 GHz (multimode upper band included) against the oracle's S fixture."""
    from openparem_amd.solver import Solver
    d, f, So = band
    m = microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)
    s = Solver(m, 2, boundaries={4: d["conductivity"]})
    try:
        assert s.ndof == d["ndof"]
        r = s.sweep(f, precision=prec, tol=1e-11, batch=len(f), restart=40, precond="lu", precond_share=share,
                    orth_passes=1 if prec == "mixed" else 2)
    finally:
        s.close()
    dS = np.abs(r["S"] - So).max(axis=(1, 2))
    print(f"band {prec} share {share}: iterations {r['iters'].ravel().tolist()} |dS| {dS.max():.2e}", flush=True)
    assert r["converged"].all()
    assert dS.max() < (1e-9 if prec == "fp64" else 1e-8)
    if share == 1:
        assert r["iters"].max() <= (6 if prec == "fp64" else 30)


DEEP_CHILD = r"""
import json
import numpy as np
from openparem_amd.mesh import waveguide_mesh
from openparem_amd.solver import Solver, lu_stats
m = waveguide_mesh(nx=6, ny=3, nz=12, seed=7)
f = np.array([9e9, 10e9, 11e9, 12e9])
with Solver(m, 2) as s:
    r = s.sweep(f, precision="mixed", tol=1e-11, batch=4, restart=30, precond="lu", orth_passes=1, precond_share=1)
    st = lu_stats(s)
print(json.dumps({"S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()], "ndof": s.ndof,
                  "converged": int(r["converged"].all()), "iters": r["iters"].tolist(),
                  "deep_launches": st["gemm128_launches"], "small_launches": st["gemm_launches"]}), flush=True)
"""


def test_lu_deep_trailing_updates_128_tile_kernel():
    """The 128 x 128 three-multiplication kernel of the deep FP32 trailing updates (k_dense_gemm_t128), forced onto
    every deep update (OPEM_LU_DEEP_GEMM=2, edge tiles included) and switched off (=0), in child processes: each FP32
    factor of a frequency's own matrix must bring GMRES to 1e-11 in a few iterations (a wrong Schur complement
    would not), the iteration counts of the two kernels must agree, and S must equal the FP64 path's."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    m = waveguide_mesh(nx=6, ny=3, nz=12, seed=7)
    f = np.array([9e9, 10e9, 11e9, 12e9])
    s = _solver(m, 2)
    try:
        r64 = s.sweep(f, precision="fp64", tol=1e-12, batch=1, restart=30, precond="lu")
    finally:
        s.close()
    assert r64["converged"].all()
    out = {}
    for mode in ("0", "2"):
        env = dict(os.environ, OPEM_LU_DEEP_GEMM=mode, OPEM_LU_TIMING="1")
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        print(f"deep gemm mode {mode}", flush=True)
        p = subprocess.run([sys.executable, "-c", DEEP_CHILD], cwd=root, env=env, capture_output=True, text=True,
                           timeout=150)
        assert p.returncode == 0, p.stderr[-2000:]
        d = json.loads(p.stdout.strip().splitlines()[-1])
        print(f"mode {mode}: ndof {d['ndof']} iters {d['iters']} deep launches {d['deep_launches']} "
              f"64-tile launches {d['small_launches']}", flush=True)
        S = np.array([complex(a, b) for a, b in d["S"]]).reshape(r64["S"].shape)
        assert d["converged"] == 1
        assert np.abs(S - r64["S"]).max() < 1e-8
        assert max(max(it) for it in d["iters"]) <= 12
        out[mode] = d
    assert out["0"]["deep_launches"] == 0 and out["2"]["deep_launches"] > 0
    it0 = np.array(out["0"]["iters"])
    it2 = np.array(out["2"]["iters"])
    assert np.abs(it0 - it2).max() <= 2


APPLY_CHILD = r"""
import json
import numpy as np
from openparem_amd.mesh import tee_mesh
from openparem_amd.solver import Solver
m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)
f = np.linspace(9.0e9, 9.3e9, 4)
with Solver(m, 2) as s:
    r = s.sweep(f, precision="mixed", tol=1e-11, batch=4, restart=30, precond="lu", orth_passes=1, precond_share=2)
print(json.dumps({"S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()],
                  "converged": int(r["converged"].all()), "iters": r["iters"].tolist()}), flush=True)
"""


def test_lu_apply_mfma_matches_vector_kernel():
    """The MFMA apply of FP32 factors (k_lu_apply_mfma) against the vector kernel (OPEM_LU_APPLY_MFMA=0), in child
    processes, on a three-port tee with factors shared by 2 frequencies (6 right-hand sides per factor: the root
    kinds take every third): the same iteration counts within one, and S within 1e-9 of each other and of the
    oracle's direct solve."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    out = {}
    for mode in ("0", "1"):
        env = dict(os.environ, OPEM_LU_APPLY_MFMA=mode)
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        print(f"apply mfma {mode}", flush=True)
        p = subprocess.run([sys.executable, "-c", APPLY_CHILD], cwd=root, env=env, capture_output=True, text=True,
                           timeout=150)
        assert p.returncode == 0, p.stderr[-2000:]
        d = json.loads(p.stdout.strip().splitlines()[-1])
        print(f"mode {mode}: iters {d['iters']}", flush=True)
        assert d["converged"] == 1
        out[mode] = d
    S0 = np.array([complex(a, b) for a, b in out["0"]["S"]]).reshape(4, 3, 3)
    S1 = np.array([complex(a, b) for a, b in out["1"]["S"]]).reshape(4, 3, 3)
    assert np.abs(S0 - S1).max() < 1e-9
    assert np.abs(np.array(out["0"]["iters"]) - np.array(out["1"]["iters"])).max() <= 1
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)
    pr = Problem(m, 2)
    f = np.linspace(9.0e9, 9.3e9, 4)
    o = pr.solve_point(f[2], {a: te10(m, a, f[2]) for a in pr.ports})
    assert np.abs(S1[2] - o["S"]).max() < 1e-8


SYM_CHILD = r"""
import json
import numpy as np
from openparem_amd.mesh import tee_mesh, waveguide_mesh
from openparem_amd.solver import Solver, lu_stats
out = {}
for name, m in (("tee", tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)),
                ("guide", waveguide_mesh(nx=6, ny=3, nz=12, seed=7))):
    f = np.array([9e9, 11e9])
    with Solver(m, 2) as s:
        r = s.sweep(f, precision="fp64", tol=1e-12, batch=2, restart=30, precond="lu")
        rm = s.sweep(f, precision="mixed", tol=1e-11, batch=2, restart=30, precond="lu", orth_passes=1)
        st = lu_stats(s)
    out[name] = {"S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()],
                 "Sm": [[float(z.real), float(z.imag)] for z in rm["S"].ravel()], "shape": list(r["S"].shape),
                 "converged": int(r["converged"].all() and rm["converged"].all()), "iters": r["iters"].tolist(),
                 "iters_mixed": rm["iters"].tolist(), "flops": st["gemm_flops"] + st["gemm128_flops"]}
print(json.dumps(out), flush=True)
"""


def test_lu_symmetric_factorization_against_oracle():
    """The symmetric factorization (lower tiles of every trailing update, rows of each outer block mirrored from its
    columns, pivots inside the outer block, lower-triangle extend-add and root gather) forced onto the FP64 factors
    too (OPEM_LU_SYM=2) and switched off (=0), in child processes, on the three-port tee (port root gathered per
    excitation) and a guide whose fronts span several 128-column outer blocks: an FP64 factor of the frequency's own
    matrix is exact, so GMRES must stop within a few iterations (a stale mirrored row would not), S must match the
    oracle's direct solve, and the symmetric form must count fewer trailing-update flops."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    out = {}
    for mode in ("0", "2"):
        env = dict(os.environ, OPEM_LU_SYM=mode, OPEM_LU_TIMING="1")
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        print(f"symmetric factorization mode {mode}", flush=True)
        p = subprocess.run([sys.executable, "-c", SYM_CHILD], cwd=root, env=env, capture_output=True, text=True,
                           timeout=150)
        assert p.returncode == 0, p.stderr[-2000:]
        out[mode] = json.loads(p.stdout.strip().splitlines()[-1])
        for name, d in out[mode].items():
            print(f"mode {mode} {name}: iters fp64 {d['iters']} mixed {d['iters_mixed']} flops {d['flops']:.3e}",
                  flush=True)
            assert d["converged"] == 1
            assert max(max(it) for it in d["iters"]) <= 6
            assert max(max(it) for it in d["iters_mixed"]) <= 20
    for name, m in (("tee", tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)),
                    ("guide", waveguide_mesh(nx=6, ny=3, nz=12, seed=7))):
        assert out["2"][name]["flops"] < 0.8 * out["0"][name]["flops"]
        pr = Problem(m, 2)
        f = np.array([9e9, 11e9])
        for mode in ("0", "2"):
            d = out[mode][name]
            S = np.array([complex(a, b) for a, b in d["S"]]).reshape(d["shape"])
            Sm = np.array([complex(a, b) for a, b in d["Sm"]]).reshape(d["shape"])
            for i, fi in enumerate(f):
                o = pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})
                assert np.abs(S[i] - o["S"]).max() < 1e-9
                assert np.abs(Sm[i] - o["S"]).max() < 1e-8


GRAPH_CHILD = r"""
import json
import numpy as np
from openparem_amd.mesh import tee_mesh
from openparem_amd.solver import Solver
m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)
f = np.linspace(9.0e9, 9.3e9, 4)
out = []
with Solver(m, 2) as s:
    for rep in range(2):  # the second sweep replays the graphs the first one captured
        r = s.sweep(f, precision="mixed", tol=1e-11, batch=4, restart=30, precond="lu", orth_passes=1,
                    precond_share=2)
        out.append({"S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()], "iters": r["iters"].tolist(),
                    "converged": int(r["converged"].all())})
print(json.dumps(out), flush=True)
"""


def test_lu_apply_graph_replay_matches_launches():
    """The apply's level sweeps replayed from captured HIP graphs (the default while kernel timing is off) against
    the same launches issued one by one (OPEM_LU_GRAPH=0), in child processes: two sweeps each on the three-port
    tee, so every apply of the first sweep after the capture and every apply of the second replays a graph between
    the Krylov solver's own launches and host synchronisations; the iteration counts and S must be identical, and S
    must match the oracle's direct solve."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    out = {}
    for mode in ("0", "1"):
        env = {k: v for k, v in os.environ.items() if k != "OPEM_LU_TIMING"}
        env["OPEM_LU_GRAPH"] = mode
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        print(f"apply graph {mode}", flush=True)
        p = subprocess.run([sys.executable, "-c", GRAPH_CHILD], cwd=root, env=env, capture_output=True, text=True,
                           timeout=150)
        assert p.returncode == 0, p.stderr[-2000:]
        out[mode] = json.loads(p.stdout.strip().splitlines()[-1])
        for d in out[mode]:
            print(f"mode {mode}: iters {d['iters']}", flush=True)
            assert d["converged"] == 1
    for rep in range(2):
        assert out["0"][rep]["iters"] == out["1"][rep]["iters"]
        assert out["0"][rep]["S"] == out["1"][rep]["S"]
    S = np.array([complex(a, b) for a, b in out["1"][1]["S"]]).reshape(4, 3, 3)
    m = tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=False)
    pr = Problem(m, 2)
    f = np.linspace(9.0e9, 9.3e9, 4)
    for i in (0, 3):
        o = pr.solve_point(f[i], {a: te10(m, a, f[i]) for a in pr.ports})
        assert np.abs(S[i] - o["S"]).max() < 1e-8
