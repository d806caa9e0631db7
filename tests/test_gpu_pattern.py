"""Radiation patterns on the GPU path against the oracle (oracle/pattern.py) and closed forms.

* k_far_field_part / k_far_field_sum (opem_far_field) vs the oracle's far-field sum on random currents and points
  (1e-12), and on a Hertzian dipole's exact near field over a closed cube vs the dipole's closed-form far field
  (equivalence principle, midpoint-rule error 2e-3).
* opem_radiation_currents on solved fields (an open-ended WR90 guide radiating through a first-order radiation
  boundary, ND1 and ND2; the microstrip's radiation lid) vs the oracle's currents from the same DOF vectors.
* The whole pattern (currents, sphere, far field, metrics) of the open-ended guide vs the oracle's.
"""
import numpy as np
import pytest

from openparem_amd.mesh import microstrip_mesh, waveguide_mesh
from oracle import assemble as oasm
from oracle import pattern as opat

pytestmark = pytest.mark.gpu

ETA0 = np.sqrt(4e-7 * np.pi / 8.8541878176e-12)
MU0 = 4e-7 * np.pi


def _solver(m, p, **kw):
    from openparem_amd.solver import Solver
    return Solver(m, p, **kw)


def test_far_field_kernel_matches_oracle():
    m = waveguide_mesh(nx=2, ny=1, nz=2, seed=1, end="radiation")
    rng = np.random.default_rng(11)
    f = 5e9
    beta = 2 * np.pi * f * MU0 / ETA0
    lam = 2 * np.pi / beta
    for ncur, npts in ((3000, 700), (17, 5), (1, 1)):
        cur = dict(xyz=rng.uniform(-0.05, 0.05, (ncur, 3)),
                   J=rng.normal(size=(ncur, 3)) + 1j * rng.normal(size=(ncur, 3)),
                   M=rng.normal(size=(ncur, 3)) + 1j * rng.normal(size=(ncur, 3)),
                   area=rng.uniform(1e-6, 2e-6, ncur))
        d = rng.normal(size=(npts, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        center = np.array([0.01, 0.0, -0.02])
        P = center + 1e5 * lam * d
        with _solver(m, 1) as s:
            from openparem_amd.pattern import far_field
            F = far_field(s, cur, center, beta, ETA0, P)
        Fo = opat.far_field(cur, center, beta, ETA0, P)
        print(f"ncur {ncur} npts {npts}: max |dF| / max |F| = {np.abs(F - Fo).max() / np.abs(Fo).max():.2e}",
              flush=True)
        assert np.abs(F - Fo).max() < 1e-12 * np.abs(Fo).max()


def test_far_field_dipole_closed_form():
    from test_oracle_pattern import cube_currents, dipole_far
    from openparem_amd.pattern import far_field
    f = 3e9
    k = 2 * np.pi * f * MU0 / ETA0
    lam = 2 * np.pi / k
    cur = cube_currents(0.2 * lam, 24, k, ETA0)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P = 1e5 * lam * d
    m = waveguide_mesh(nx=2, ny=1, nz=2, seed=1, end="radiation")
    with _solver(m, 1) as s:
        F = far_field(s, cur, np.zeros(3), k, ETA0, P)
    Et, Hp = dipole_far(P, k, ETA0)
    sc = np.abs(Et).max()
    assert np.abs(F[:, 1] - Et).max() < 2e-3 * sc and np.abs(F[:, 0]).max() < 2e-3 * sc
    assert np.abs(F[:, 2] - Hp).max() < 2e-3 * sc / ETA0 and np.abs(F[:, 3]).max() < 2e-3 * sc / ETA0
    assert np.abs(F - opat.far_field(cur, np.zeros(3), k, ETA0, P)).max() < 1e-12 * sc


def _open_guide_fields(p, f=10e9):
    m = waveguide_mesh(nx=4, ny=2, nz=5, seed=3, end="radiation")
    s = _solver(m, p)
    r = s.sweep(np.array([f]), precision="fp64", tol=1e-12, restart=60, precond="lu", want_fields=True)
    assert r["converged"].all()
    return m, s, r


@pytest.mark.parametrize("p", [1, 2])
def test_radiation_currents_match_oracle(p):
    from openparem_amd.pattern import radiation_currents
    f = 10e9
    m, s, r = _open_guide_fields(p, f)
    try:
        x, h = r["x"][0, 0], r["h"][0, 0]
        for limit in (0.5, 0.05):
            cur = radiation_currents(s, x, h, f, limit)
            rad = [a for a, k in m.bdr_kind.items() if k == "radiation"]
            o = opat.radiation_currents(oasm.Space(m, p), rad, x, h, f, ETA0, limit)
            assert cur["eta"] == pytest.approx(ETA0, rel=1e-15)
            assert cur["xyz"].shape == o["xyz"].shape
            sc = max(np.abs(o["J"]).max(), 1e-300)
            assert np.abs(cur["xyz"] - o["xyz"]).max() < 1e-15
            assert np.abs(cur["area"] - o["area"]).max() < 1e-12 * o["area"].max()
            assert np.abs(cur["J"] - o["J"]).max() < 1e-12 * sc
            assert np.abs(cur["M"] - o["M"]).max() < 1e-12 * np.abs(o["M"]).max()
            print(f"p={p} limit {limit}: {cur['area'].size} currents", flush=True)
    finally:
        s.close()


def test_radiation_currents_microstrip_lid():
    from openparem_amd.pattern import radiation_currents
    m = microstrip_mesh(nx=8, ny=5, nz=6, n_sub=2, seed=6)
    f = 12e9
    with _solver(m, 2, boundaries={4: 5.8e7}) as s:
        r = s.sweep(np.array([f]), precision="fp64", tol=1e-12, restart=60, precond="lu", want_fields=True)
        assert r["converged"].all()
        x, h = r["x"][0, 1], r["h"][0, 1]
        cur = radiation_currents(s, x, h, f, 0.05)
    o = opat.radiation_currents(oasm.Space(m, 2), [5], x, h, f, ETA0, 0.05)
    assert np.abs(cur["J"] - o["J"]).max() < 1e-12 * np.abs(o["J"]).max()
    assert np.abs(cur["M"] - o["M"]).max() < 1e-12 * np.abs(o["M"]).max()


def test_open_guide_pattern_matches_oracle():
    from openparem_amd.pattern import pattern
    from openparem_amd.ports import te10
    f = 10e9
    m, s, r = _open_guide_fields(2, f)
    try:
        md = te10(m, 2, f)
        Pz = 0.5 * md["V"] * np.conj(md["V"]) / np.conj(md["Z"])   # Z = V V* / (2 conj P)
        from openparem_amd.pattern import accepted_power
        acc = accepted_power(r["Cp"][0], r["Cm"][0], Pz, 0)
        res = pattern(s, r["x"][0, 0], r["h"][0, 0], f, acc, limit=0.05, refine=3)
    finally:
        s.close()
    # oracle: the same steps
    o_cur = opat.radiation_currents(oasm.Space(m, 2), [3], r["x"][0, 0], r["h"][0, 0], f, ETA0, 0.05)
    center = 0.5 * (m.xyz.min(axis=0) + m.xyz.max(axis=0))
    beta = 2 * np.pi * f * MU0 / ETA0
    sph = opat.sphere(1e5 * 2 * np.pi / beta, center, 3)
    # the phase beta r (6e5 rad at 1e5 wavelengths) turns a last-bit difference of the two spheres' points into a
    # 1e-10 relative phase, so the fields are compared at the product's points, the metrics (power) at the oracle's
    Fo = opat.far_field(o_cur, center, beta, ETA0, res["sphere"].pts)
    assert np.abs(res["fields"] - Fo).max() < 1e-12 * np.abs(Fo).max()
    Fo = opat.far_field(o_cur, center, beta, ETA0, sph["pts"])
    mo = opat.metrics(Fo, sph["point_area"], sph["tri_area"].sum(), acc)
    assert abs(res["radiated_power"] - mo["radiated_power"]) < 1e-11 * abs(mo["radiated_power"])
    assert abs(res["directivity_db"] - mo["directivity_db"]) < 1e-9
    assert abs(res["gain_db"] - mo["gain_db"]) < 1e-9
    # physics: the beam of an open guide end points along +z, out of the aperture
    d = res["directivity"]
    top = np.argmax(d)
    u = (res["sphere"].pts[top] - center) / np.linalg.norm(res["sphere"].pts[top] - center)
    print(f"open guide: P_rad {res['radiated_power']:.4e} accepted {acc:.4e} eff {res['efficiency']:.3f} "
          f"D {res['directivity_db']:.2f} dBi toward {u}", flush=True)
    assert u[2] > 0.9
    assert 3.0 < res["directivity_db"] < 12.0
    assert res["radiated_power"].real > 0
