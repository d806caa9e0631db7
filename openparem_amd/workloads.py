"""Named synthetic workloads (meshes, orders, frequency plans) used by tests, smoke and the benchmark.

* ``waveguide-c1``   air-filled 22.86 x 10.16 x 30 mm guide, 12x6x18 Kuhn grid, ND2, 10 GHz.
* ``microstrip-c3``  shielded microstrip box 40x20x60 mm, substrate 1.5 mm (er 4.4, tand 0.02), copper trace
                     2.8 mm (surface impedance), radiation lid, 36x18x75 grid, ND2 (~1.9 M DOF), 5 GHz.
* ``microstrip-c4``  same geometry on a 22x11x50 grid (~0.48 M DOF), 64 frequencies 1-20 GHz x 2 ports.
* ``waveguide-sweep`` the c1 guide on a 24x12x36 grid (~0.43 M ND2 DOF), 64 frequencies over the TE10-only
                     band 7-12.5 GHz x 2 ports = 128 systems on one pattern.
"""
from __future__ import annotations

import numpy as np

from .mesh import microstrip_mesh, waveguide_mesh

COPPER = 5.813e7


def make(name):
    if name == "waveguide-c1":
        return dict(mesh=waveguide_mesh(), order=2, freqs=np.array([10e9]), boundaries={})
    if name == "microstrip-c3":
        m = microstrip_mesh(nx=36, ny=18, nz=75, seed=3)
        return dict(mesh=m, order=2, freqs=np.array([5e9]), boundaries={4: COPPER})
    if name == "microstrip-c4":
        m = microstrip_mesh(nx=22, ny=11, nz=50, seed=4)
        return dict(mesh=m, order=2, freqs=np.linspace(1e9, 20e9, 64), boundaries={4: COPPER})
    if name == "waveguide-sweep":
        m = waveguide_mesh(nx=24, ny=12, nz=36, seed=11)
        return dict(mesh=m, order=2, freqs=np.linspace(7e9, 12.5e9, 64), boundaries={})
    if name.startswith("microstrip-small"):
        m = microstrip_mesh(nx=8, ny=5, nz=8, n_sub=2, seed=6)
        return dict(mesh=m, order=2, freqs=np.linspace(2e9, 10e9, 4), boundaries={4: COPPER})
    raise ValueError(name)
