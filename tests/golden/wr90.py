"""Loader for the wr90_straight.json fixture as a TetMesh (x, y shifted so the guide spans [0,a]x[0,b])."""
import json
import os

import numpy as np

from openparem_amd.mesh import TetMesh, _boundary_faces

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    d = json.load(open(os.path.join(HERE, "wr90_straight.json")))
    xyz = np.array(d["nodes"], dtype=np.float64)
    xyz[:, 0] += d["a"] / 2
    xyz[:, 1] += d["b"] / 2
    tets = np.array(d["tets"], dtype=np.int64)
    bf, _ = _boundary_faces(tets)
    c = xyz[bf].mean(axis=1)
    attr = np.ones(bf.shape[0], np.int32)
    attr[np.abs(c[:, 2] - d["length"]) < 1e-9] = 2      # S-port 1 ("input", z = 0.1)
    attr[np.abs(c[:, 2]) < 1e-9] = 3                     # S-port 2 ("output", z = 0)
    m = TetMesh(xyz=xyz, tets=tets.astype(np.int32), tet_attr=np.ones(len(tets), np.int32),
                bfaces=bf.astype(np.int32), bface_attr=attr, bdr_kind={1: "pec", 2: "port", 3: "port"},
                port_normal={2: np.array([0.0, 0.0, 1.0]), 3: np.array([0.0, 0.0, -1.0])},
                materials={1: dict(er=d["air_er"], tand=0.0, mur=1.0)}, name="waveguide_wr90_straight")
    m.geometry = dict(a=d["a"], b=d["b"], length=d["length"])
    return m, d


def reference_s(d, p):
    """Reference S11 and S21 (complex) at 9 GHz on the unrefined mesh (results.ods sheet p, row 1)."""
    row = [float(x.lstrip("#")) if x else 0.0 for x in d["s_results"][str(p)]["row"][:9]]
    s11 = 10 ** (row[1] / 20) * np.exp(1j * np.radians(row[2]))
    s21 = 10 ** (row[3] / 20) * np.exp(1j * np.radians(row[4]))
    return s11, s21
