"""Project ingest (host side, no GPU): the path geometry against the reference's own known-answer cases
(tests/golden/path_cases.json, extracted from src/OpenParEMCommon/path.cpp by make_path_cases.py), the gmsh reader
and the boundary marking on the reference's WR90 straight_study project (tests/golden/wr90_straight_study: mesh,
ports file, ND2 project and the regression materials database), the materials lookup rules, the MFEM reader."""
import json
import os

import numpy as np
import pytest

from openparem_amd import project as pj

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
STUDY = os.path.join(GOLD, "wr90_straight_study")


@pytest.fixture(scope="module")
def cases():
    return json.load(open(os.path.join(GOLD, "path_cases.json")))


def test_point_on_segment_reference_cases(cases):
    bad = [c for c in cases["on_line"] if pj.point_on_segment(c["p"], c["a"], c["b"], c["tol"]) != c["on"]]
    assert len(cases["on_line"]) > 100 and not bad, bad[:5]


@pytest.mark.parametrize("name", ["m", "mr", "sqr2"])
def test_point_inside_reference_cases(cases, name):
    """The M-shaped path in a z plane, the same path rotated into a y plane, and a tilted square: the reference's
    expected inside/outside answer for every test point (points on an edge or vertex count as inside)."""
    c = cases["inside"][name]
    o = pj.Outline(name, np.array(c["path"]), c["closed"])
    bad = [t for t in c["cases"] if o.inside(t["p"]) != t["inside"]]
    assert len(c["cases"]) >= 14 and not bad, bad


def test_gmsh_reader_matches_fixture():
    m = pj.read_gmsh(os.path.join(STUDY, "straight.msh"))
    d = json.load(open(os.path.join(GOLD, "wr90_straight.json")))
    assert np.array_equal(m["xyz"], np.array(d["nodes"])) and np.array_equal(m["tets"], np.array(d["tets"]))
    assert m["region_names"] == {1: "air"} and (m["tet_attr"] == 1).all()


def test_straight_project_marks_ports_and_walls():
    """Port 'input' (S-port 1) is the z = 0.1 face, 'output' the z = 0 face; every other boundary triangle is PEC;
    air from the materials database; rectangular outlines give the TE10 geometry (broad wall along x)."""
    p = pj.load_project(os.path.join(STUDY, "straight_order_2.proj"))
    m = p.mesh
    assert p.order == 2 and np.allclose(p.frequencies, [9e9]) and p.reference_impedance == 0
    assert p.port_names == ["input", "output"] and m.bdr_kind == {1: "pec", 2: "port", 3: "port"}
    c = m.xyz[m.bfaces].mean(axis=1)
    assert (np.abs(c[m.bface_attr == 2][:, 2] - 0.1) < 1e-12).all() and (np.abs(c[m.bface_attr == 3][:, 2]) < 1e-12).all()
    on_ends = (np.abs(c[:, 2]) < 1e-12) | (np.abs(c[:, 2] - 0.1) < 1e-12)
    assert ((m.bface_attr == 1) == ~on_ends).all()
    assert np.allclose(m.port_normal[2], [0, 0, 1]) and np.allclose(m.port_normal[3], [0, 0, -1])
    assert m.materials == {1: dict(er=1.0006, tand=0.0, mur=1.0)}
    g = m.geometry
    assert np.isclose(g["a"], 22.86e-3) and np.isclose(g["b"], 10.16e-3)
    assert np.allclose(g["ports"][2]["u"], [1, 0, 0]) and np.allclose(g["ports"][2]["origin"], [-0.01143, -0.00508, 0.1])


def test_ports_file_errors(tmp_path):
    src = open(os.path.join(STUDY, "straight_ports.txt")).read()
    (tmp_path / "p.txt").write_text(src.replace("path=+port2", "path=+nowhere"))
    with pytest.raises(ValueError, match="undefined path nowhere"):
        pj.read_ports(tmp_path / "p.txt")
    outlines, ports, bounds = pj.read_ports(os.path.join(STUDY, "straight_ports.txt"))
    assert [p.sports for p in ports] == [[1], [2]] and [p.paths for p in ports] == [["port1"], ["port2"]]
    outlines["port2"].points[:, 2] += 0.05  # the outline no longer lies on the mesh boundary
    m = pj.read_gmsh(os.path.join(STUDY, "straight.msh"))
    from openparem_amd.mesh import _boundary_faces
    bf, _ = _boundary_faces(m["tets"])
    with pytest.raises(ValueError, match="not assigned to the mesh boundary: output"):
        pj.mark_boundaries(m["xyz"], bf, outlines, ports, bounds)


MATS = """#OpenParEMmaterials 1.0
Material
   name=stepped
   Temperature
      temperature=any
      Frequency
         frequency=1e9
         er=4.0
         mur=1
         tand=0.01
      EndFrequency
      Frequency
         frequency=3e9
         er=3.0
         mur=1
         tand=0.03
      EndFrequency
   EndTemperature
EndMaterial
Material
   name=lossy
   Temperature
      temperature=20
      Frequency
         frequency=any
         er=2
         mur=1
         conductivity=0.5
      EndFrequency
   EndTemperature
EndMaterial
"""


def test_material_lookup_rules(tmp_path):
    (tmp_path / "m.txt").write_text(MATS)
    mats = pj.read_materials(tmp_path / "m.txt")
    eps = pj.material_eps(mats["stepped"], 2e9)         # halfway: er 3.5, tand 0.02
    assert np.isclose(eps.real / pj.EPS0, 3.5) and np.isclose(-eps.imag / eps.real, 0.02)
    assert np.isclose(pj.material_eps(mats["stepped"], 3e9).real / pj.EPS0, 3.0)  # exact block
    with pytest.raises(ValueError, match="no extrapolation"):
        pj.material_eps(mats["stepped"], 5e9)
    f = 1e9                                              # conductivity: Im eps = -sigma / omega
    assert np.isclose(pj.material_eps(mats["lossy"], f).imag, -0.5 / (2 * np.pi * f))
    with pytest.raises(ValueError, match="no data at 30"):
        pj.material_eps(mats["lossy"], f, temperature=30)


def test_wideband_model_limits():
    """The regression database's FR4 (Djordjevic-Sarkar: er_inf 4.27, delta 1.12, poles 10^4..10^12 rad/s): the
    static and optical limits and the constant slope of er per decade of angular frequency between the poles."""
    mats = pj.read_materials(os.path.join(STUDY, "global_materials.txt"))
    fr4 = mats["FR4"]
    er = lambda f: pj.material_eps(fr4, f).real / pj.EPS0
    assert abs(er(1e-3) - (4.27 + 1.12)) < 1e-3 and abs(er(1e15) - 4.27) < 1e-3
    slope = (er(1e8) - er(1e7)) / np.log10(10)
    assert abs(slope + 1.12 / 8) < 1e-6


def test_mfem_reader(tmp_path):
    txt = ["MFEM mesh v1.0", "", "dimension", "3", "", "elements", "1", "7 4 0 1 2 3", "", "boundary", "1",
           "5 2 0 1 2", "", "vertices", "4", "3", "0 0 0", "1 0 0", "0 1 0", "0 0 1"]
    (tmp_path / "t.mesh").write_text("\n".join(txt) + "\n")
    m = pj.read_mesh(tmp_path / "t.mesh")
    assert m["tets"].tolist() == [[0, 1, 2, 3]] and m["tet_attr"].tolist() == [7]
    assert m["tris"].tolist() == [[0, 1, 2]] and m["xyz"][3].tolist() == [0, 0, 1]
