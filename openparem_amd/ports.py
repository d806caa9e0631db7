"""Port-mode provider: modal fields sampled at port points, gamma, V and Z.

The reference obtains these from the OpenParEM2D eigen-solve (``src/OpenParEM3D/port.cpp:5081-5160``,
``load_modeMetrics`` :5302-5453).  Here the rectangular-guide TE10 mode is given in closed form.
Conventions:

* E_t is normalised to a peak of 1 (``FieldSet::scaleSolution``, port.cpp:2291-2335).
* H_t is the field of the wave leaving the domain through the port:
  H_t = gamma/(j omega mu) n_out x E_t  (Faraday's law for an e^{-gamma n} dependence).
* V = line integral of E across the guide centre (bottom to top wall);
  Z = V V* / (2 conj(P)) with P = 1/2 int (E x H*) . n  (Z_PV, OpenParEM2D/fem2D.cpp:1819).
"""
from __future__ import annotations

import numpy as np

MU0 = 4e-7 * np.pi
EPS0 = 8.8541878176e-12


def k0_of(freq):
    return 2 * np.pi * freq * np.sqrt(MU0 * EPS0)


def te10(mesh, attr, freq):
    a, b = mesh.geometry["a"], mesh.geometry["b"]
    mat = mesh.materials[1]
    er = mat["er"] * (1 - 1j * mat["tand"])
    mur = mat["mur"]
    k0 = k0_of(freq)
    kc = np.pi / a
    gamma = np.sqrt(kc * kc - k0 * k0 * er * mur + 0j)
    if gamma.real < 0 or (gamma.real == 0 and gamma.imag < 0):
        gamma = -gamma
    omega = 2 * np.pi * freq
    n = np.asarray(mesh.port_normal[attr], dtype=np.float64)
    y_over = gamma / (1j * omega * MU0 * mur)

    # cross-section corner and width axis of this port (x from 0 on the straight guide); E_t along y
    pg = mesh.geometry.get("ports", {}).get(attr)
    origin = np.zeros(3) if pg is None else np.asarray(pg["origin"], dtype=np.float64)
    u = np.array([1.0, 0.0, 0.0]) if pg is None else np.asarray(pg["u"], dtype=np.float64)

    def E(pts):
        e = np.zeros(pts.shape, np.complex128)
        e[..., 1] = np.sin(kc * ((pts - origin) @ u))
        return e

    def H(pts):
        return y_over * np.cross(n, E(pts))

    V = complex(b)
    P = 0.5 * np.conj(y_over) * (a * b / 2)
    Z = 0.5 * V * np.conj(V) / np.conj(P)
    return dict(E=E, H=H, gamma=complex(gamma), V=V, Z=complex(Z))


def tem_pmc_sides(mesh, attr, freq):
    """TEM mode of a guide with PEC walls at y = 0, b and magnetic walls at x = 0, a: uniform E_y (peak 1),
    gamma = j k0 sqrt(er mur), same conventions as te10 (V across the guide, Z = V V* / (2 conj(P)))."""
    a, b = mesh.geometry["a"], mesh.geometry["b"]
    mat = mesh.materials[1]
    er = mat["er"] * (1 - 1j * mat["tand"])
    mur = mat["mur"]
    gamma = np.sqrt(-k0_of(freq) ** 2 * er * mur + 0j)
    if gamma.real < 0 or (gamma.real == 0 and gamma.imag < 0):
        gamma = -gamma
    omega = 2 * np.pi * freq
    n = np.asarray(mesh.port_normal[attr], dtype=np.float64)
    y_over = gamma / (1j * omega * MU0 * mur)

    def E(pts):
        e = np.zeros(pts.shape, np.complex128)
        e[..., 1] = 1.0
        return e

    def H(pts):
        return y_over * np.cross(n, E(pts))

    V = complex(b)
    P = 0.5 * np.conj(y_over) * (a * b)
    Z = 0.5 * V * np.conj(V) / np.conj(P)
    return dict(E=E, H=H, gamma=complex(gamma), V=V, Z=complex(Z))


def provider_for(mesh):
    if mesh.name.startswith("waveguide") and mesh.name.endswith("_pmc"):
        return tem_pmc_sides
    if mesh.name.startswith("waveguide") or mesh.name.startswith("tee"):
        return te10
    if mesh.name.startswith("microstrip"):
        from .microstrip_mode import quasi_tem
        return quasi_tem
    raise ValueError(f"no port-mode provider for mesh {mesh.name}")
