"""Write the oracle's S for the reduced microstrip box across 1-20 GHz as a fixture.

    python tests/golden/make_microstrip_band.py

The 40 x 20 x 60 mm shielded-microstrip box (lossy FR4-like substrate, copper trace as a surface impedance,
radiation lid) on an 8 x 5 x 8 grid at ND2 (13 642 DOF), eight frequencies spaced evenly over 1-20 GHz: the
band of the batched-sweep configuration, including the multimode upper band where the box's cross-section carries
several propagating modes.  The oracle's sparse direct solve takes about 15 s per frequency on one CPU core, so its
S is committed (tests/golden/microstrip_band.json) and the GPU tests compare against it.  ndof is recorded so a
changed mesh generator shows up as a mismatch rather than as a parity failure.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from openparem_amd.mesh import microstrip_mesh  # noqa: E402
from openparem_amd.microstrip_mode import quasi_tem  # noqa: E402
from oracle import assemble as oasm  # noqa: E402
from oracle.pipeline import Problem  # noqa: E402

COPPER = 5.813e7


def main():
    m = microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)
    freqs = np.linspace(1e9, 20e9, 8)
    pr = Problem(m, 2)
    pair = lambda z: [[float(v.real), float(v.imag)] for v in np.ravel(z)]
    S = []
    for f in freqs:
        rs = np.sqrt(np.pi * f * oasm.MU0 / COPPER)
        o = pr.solve_point(float(f), {a: quasi_tem(m, a, float(f)) for a in pr.ports}, rs={4: rs})
        S.append(pair(o["S"]))
        print(f"{f / 1e9:.3f} GHz |S| {np.abs(o['S']).round(4).tolist()}", flush=True)
    out = dict(mesh="microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)", order=2, ndof=int(pr.space.ndof),
               freqs=[float(f) for f in freqs], conductivity=COPPER, S=S)
    with open(os.path.join(HERE, "microstrip_band.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
