"""Named synthetic workloads (meshes, orders, frequency plans) used by tests, smoke and the benchmark.

* ``waveguide-c1``   air-filled 22.86 x 10.16 x 30 mm guide, 12x6x18 Kuhn grid, ND2, 10 GHz.
* ``microstrip-c3``  shielded microstrip box 40x20x60 mm, substrate 1.5 mm (er 4.4, tand 0.02), copper trace
                     2.8 mm (surface impedance), radiation lid, 36x18x75 grid, ND2 (~1.9 M DOF), 5 GHz.
* ``microstrip-c4``  same geometry on a 22x11x50 grid (~0.48 M DOF), 64 frequencies 1-20 GHz x 2 ports.
* ``waveguide-sweep`` the c1 guide on a 24x12x36 grid (~0.43 M ND2 DOF), 64 frequencies over the TE10-only
                     band 7-12.5 GHz x 2 ports = 128 systems on one pattern.

``c5_instances()`` is a seeded mixed set (16 instances): twelve guide or microstrip instances, ND1/ND2 (one small
ND3), target DOF count log-uniform in 50 k - 1.5 M, 8-32 frequencies, 2 ports; then four junctions of guides: an
H-plane tee (3 ports) and a cross (4 ports) at ND2, a finer tee at ND1 (148 k DOF, 24 frequencies) and a coarser
cross at ND2 (77 k DOF, 32 frequencies).  Guides, tees and crosses sweep the TE10 band 7-12.5 GHz; microstrips
sweep 1-20 GHz, the band of the single-pattern sweep, multimode upper band included.
"""
from __future__ import annotations

import numpy as np

from .mesh import microstrip_mesh, tee_mesh, waveguide_mesh

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


# ND DOFs per hexahedral grid cell (6 tetrahedra) at orders 1, 2, 3 on these box meshes
_DOF_PER_CELL = {1: 8.0, 2: 41.0, 3: 117.0}


def c5_instances(seed=5):
    """The mixed set: list of dicts with family, order, mesh, freqs, boundaries (see the module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(12):
        family = "waveguide" if i % 2 == 0 else "microstrip"
        order = 3 if i == 10 else (1 if i % 3 == 0 else 2)
        target = 6e4 if order == 3 else float(np.exp(rng.uniform(np.log(5e4), np.log(1.5e6))))
        nf = int(rng.choice([8, 16, 24, 32]))
        cells = target / _DOF_PER_CELL[order]
        if family == "waveguide":
            f = (cells / (12 * 6 * 18)) ** (1.0 / 3.0)
            m = waveguide_mesh(nx=max(4, round(12 * f)), ny=max(2, round(6 * f)), nz=max(4, round(18 * f)),
                               seed=100 + i)
            out.append(dict(family=family, order=order, mesh=m, freqs=np.linspace(7e9, 12.5e9, nf),
                            boundaries={}))
        else:
            f = (cells / (22 * 11 * 50)) ** (1.0 / 3.0)
            m = microstrip_mesh(nx=max(8, round(22 * f)), ny=max(5, round(11 * f)), nz=max(6, round(50 * f)),
                                seed=100 + i)
            out.append(dict(family=family, order=order, mesh=m, freqs=np.linspace(1e9, 20e9, nf),
                            boundaries={4: COPPER}))
    for cross, h, nf, order in ((False, 1.6e-3, 16, 2), (True, 1.8e-3, 8, 2), (False, 1.0e-3, 24, 1),
                                (True, 2.4e-3, 32, 2)):
        m = tee_mesh(h=h, ny=max(2, round(10.16e-3 / h)), cross=cross)
        out.append(dict(family="cross" if cross else "tee", order=order, mesh=m,
                        freqs=np.linspace(7e9, 12.5e9, nf), boundaries={}))
    return out
