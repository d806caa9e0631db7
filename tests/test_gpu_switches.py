"""The library's diagnostic switches (environment variables read once per process) select measured kernel variants:
each one runs a batched mixed sweep in a child process and must give the oracle's S (which the plain FP64
one-frequency-at-a-time GPU solve matches to 1e-9), to the |dS| < 1e-6 bar of the mixed path.  One child at a
time; each prints one JSON line."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from openparem_amd.mesh import waveguide_mesh

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json
import numpy as np
from openparem_amd.mesh import waveguide_mesh
from openparem_amd.solver import Solver
m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
f = np.linspace(8e9, 12e9, 9)
with Solver(m, 2) as s:
    r = s.sweep(f, precision="mixed", tol=1e-11, batch=9, orth_passes=1, precond_share=4, arnoldi_step=3)
print(json.dumps({"S": [[float(z.real), float(z.imag)] for z in r["S"].ravel()],
                  "converged": int(r["converged"].all()), "iters": r["iters"].tolist()}), flush=True)
"""


@pytest.fixture(scope="module")
def reference_s():
    """The oracle's S at the nine frequencies (direct solve), after checking the GPU's plain FP64 one-frequency-at-
    a-time S against it to 1e-9."""
    from openparem_amd.ports import te10
    from openparem_amd.solver import Solver
    from oracle.pipeline import Problem
    m = waveguide_mesh(nx=4, ny=2, nz=6, seed=4)
    f = np.linspace(8e9, 12e9, 9)
    with Solver(m, 2) as s:
        r = s.sweep(f, precision="fp64", tol=1e-10, batch=1)
    assert r["converged"].all()
    pr = Problem(m, 2)
    So = np.array([pr.solve_point(fi, {a: te10(m, a, fi) for a in pr.ports})["S"] for fi in f])
    assert np.abs(r["S"] - So).max() < 1e-9
    return So


@pytest.mark.parametrize("env", [
    {},                                              # the defaults, through the same child mechanism
    {"OPEM_SPMV_PAIR": "0"},                         # one-system-per-lane tiles of the batched operator
    {"OPEM_SPMV_PAIR": "0", "OPEM_SPMV_WIDE": "1"},  # row-per-wavefront form of the batched operator
    {"OPEM_PATCH_MFMA": "0"},                        # FP16 patch inverses on the vector kernel
    {"OPEM_PATCH_SPLIT_KAPPA": "0"},                 # every FP16 block carries lo planes
    {"OPEM_PATCH_SPLIT_KAPPA": "1e30"},              # no block carries lo planes
    {"OPEM_XCD_REMAP": "1"},                         # XCD-contiguous workgroup order of the single-matrix SpMV
    {"OPEM_SPMV_TUNE": "0"},                         # the fixed rule for the batched operator's form
], ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()) or "defaults")
def test_switch_gives_the_same_s(reference_s, env):
    penv = dict(os.environ)
    penv["PYTHONPATH"] = ROOT + (os.pathsep + penv["PYTHONPATH"] if penv.get("PYTHONPATH") else "")
    penv.update(env)
    print(f"switches {env}", flush=True)
    p = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=penv, capture_output=True, text=True,
                       timeout=150)
    assert p.returncode == 0, p.stderr[-2000:]
    d = json.loads([ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1])
    S = np.array([complex(*z) for z in d["S"]]).reshape(reference_s.shape)
    err = np.abs(S - reference_s).max()
    print(f"  |dS| {err:.2e}, iterations {d['iters']}", flush=True)
    assert d["converged"] == 1 and err < 1e-6


def _child(env):
    penv = dict(os.environ)
    penv["PYTHONPATH"] = ROOT + (os.pathsep + penv["PYTHONPATH"] if penv.get("PYTHONPATH") else "")
    penv.update(env)
    p = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=penv, capture_output=True, text=True,
                       timeout=150)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads([ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1]), p.stderr


def test_tuned_operator_is_bit_identical():
    """The timed choice of the batched operator's form changes no bit: the sweep with tuning (which prints its
    timings and choice) and the sweep with the fixed rule return identical S and iteration counts."""
    print("tuned vs fixed-rule operator", flush=True)
    tuned, err = _child({"OPEM_SPMV_TUNE": "2"})
    fixed, _ = _child({"OPEM_SPMV_TUNE": "0"})
    lines = [ln for ln in err.splitlines() if "batched operator" in ln]
    for ln in lines:
        print("  " + ln, flush=True)
    assert lines, "tuning did not run"
    assert tuned["iters"] == fixed["iters"]
    assert tuned["S"] == fixed["S"]
