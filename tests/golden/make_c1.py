"""Write the oracle's S for the 12x6x18 WR90 guide at 10 GHz (ND2, 53 280 DOF) as a fixture.

    python tests/golden/make_c1.py

The oracle's sparse direct solve takes about 1.5 minutes on one CPU core, too long for a test, so its result is
committed (tests/golden/c1_guide_10ghz.json) and the GPU test compares against it.  The mesh is rebuilt from its
seed by ``openparem_amd.workloads.make("waveguide-c1")``; the fixture records ndof so a changed mesh
generator shows up as a mismatch rather than as a parity failure.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from openparem_amd import workloads  # noqa: E402
from openparem_amd.ports import te10  # noqa: E402
from oracle.pipeline import Problem  # noqa: E402


def main():
    w = workloads.make("waveguide-c1")
    m, f = w["mesh"], float(w["freqs"][0])
    pr = Problem(m, w["order"])
    o = pr.solve_point(f, {a: te10(m, a, f) for a in pr.ports})
    pair = lambda z: [[float(v.real), float(v.imag)] for v in z.ravel()]
    out = dict(workload="waveguide-c1", freq=f, order=w["order"], ndof=int(pr.space.ndof),
               S=pair(o["S"]), S_raw=pair(o["S_raw"]))
    with open(os.path.join(HERE, "c1_guide_10ghz.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
