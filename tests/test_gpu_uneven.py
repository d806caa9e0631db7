"""A reduced uneven set (the mixed-mesh configuration at test size): guides, a microstrip, a tee and a cross at ND1,
ND2 and ND3 with 2, 3 and 4 ports, each swept with the sparse-LU preconditioner in mixed precision and compared
with the oracle's direct solve at every frequency.  One context per instance, as the full set runs."""
import json
import os

import numpy as np
import pytest

from openparem_amd.mesh import microstrip_mesh, tee_mesh, waveguide_mesh
from openparem_amd.ports import te10
from oracle.pipeline import Problem

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

# (name, mesh, order, frequencies): the guide and junction bands are the TE10 band of WR90
CASES = [
    ("guide ND1", lambda: waveguide_mesh(nx=5, ny=3, nz=8, seed=21), 1, np.linspace(7e9, 12.5e9, 4)),
    ("guide ND3", lambda: waveguide_mesh(nx=3, ny=2, nz=4, seed=22), 3, np.linspace(8e9, 12e9, 3)),
    ("tee ND3", lambda: tee_mesh(h=5e-3, ny=2, length=40e-3, arm=15e-3), 3, np.array([9e9, 11e9])),
    ("cross ND2", lambda: tee_mesh(h=4e-3, ny=2, length=50e-3, arm=20e-3, cross=True), 2,
     np.linspace(8e9, 12e9, 4)),
]


@pytest.mark.parametrize("name,make,order,freqs", CASES, ids=[c[0] for c in CASES])
def test_uneven_instance_against_oracle(name, make, order, freqs):
    from openparem_amd.solver import Solver
    m = make()
    s = Solver(m, order)
    try:
        r = s.sweep(freqs, precision="mixed", tol=1e-11, batch=len(freqs), restart=30, precond="lu",
                    precond_share=2, orth_passes=1)
        ndof, nports = s.ndof, len(s.ports)
    finally:
        s.close()
    pr = Problem(m, order)
    assert pr.space.ndof == ndof
    dS = []
    for i, f in enumerate(freqs):
        o = pr.solve_point(f, {a: te10(m, a, f) for a in pr.ports})
        dS.append(np.abs(r["S"][i] - o["S"]).max())
    print(f"{name}: {ndof} DOF, {nports} ports, iterations {r['iters'].ravel().tolist()}, max |dS| {max(dS):.2e}",
          flush=True)
    assert r["converged"].all()
    assert max(dS) < 1e-8


def test_uneven_microstrip_band_share_group_of_one():
    """The microstrip of the set's second family across 1-20 GHz with a factor per frequency (share 1, batch 3: the
    last batch holds a single frequency) against the oracle's S fixture."""
    from openparem_amd.solver import Solver
    d = json.load(open(os.path.join(HERE, "golden", "microstrip_band.json")))
    f = np.array(d["freqs"])
    So = np.array([[complex(*v) for v in S] for S in d["S"]]).reshape(-1, 2, 2)
    m = microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)
    s = Solver(m, 2, boundaries={4: d["conductivity"]})
    try:
        r = s.sweep(f[:7], precision="mixed", tol=1e-11, batch=3, restart=30, precond="lu", precond_share=1,
                    orth_passes=1)
    finally:
        s.close()
    assert r["converged"].all()
    assert np.abs(r["S"] - So[:7]).max() < 1e-8


def test_uneven_instances_sharing_the_gpu():
    """The four reduced instances swept at the same time from four host threads, each on its own context (own HIP
    stream and workspace, a quarter of the planner's memory share), as the uneven set runs them: every instance's S
    against the oracle's direct solve at every frequency, and against the same instance swept alone."""
    import threading

    from openparem_amd.solver import Solver
    meshes = [make() for _, make, _, _ in CASES]

    def sweep(k, out, frac):
        _, _, order, freqs = CASES[k]
        s = Solver(meshes[k], order)
        try:
            out[k] = s.sweep(freqs, precision="mixed", tol=1e-11, batch=len(freqs), restart=30, precond="lu",
                             precond_share=2, orth_passes=1, mem_fraction=frac)
        except Exception as e:  # noqa: BLE001 - checked below
            out[k] = e
        finally:
            s.close()

    alone, together = {}, {}
    for k in range(len(CASES)):
        sweep(k, alone, 0.0)
    print("instances swept alone", flush=True)
    th = [threading.Thread(target=sweep, args=(k, together, 0.75 / len(CASES))) for k in range(len(CASES))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    print("instances swept together", flush=True)
    for k, (name, _, order, freqs) in enumerate(CASES):
        assert not isinstance(together[k], Exception), (name, together[k])
        r, r1 = together[k], alone[k]
        assert r["converged"].all()
        assert np.abs(r["S"] - r1["S"]).max() < 1e-9
        pr = Problem(meshes[k], order)
        for i, f in enumerate(freqs):
            o = pr.solve_point(f, {a: te10(meshes[k], a, f) for a in pr.ports})
            dS = np.abs(r["S"][i] - o["S"]).max()
            print(f"{name} {f / 1e9:.2f} GHz: |dS| vs oracle {dS:.2e}", flush=True)
            assert dS < 1e-8
