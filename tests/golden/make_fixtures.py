"""Extract small fixtures from the reference's own regression data (run in a container that has the
reference checkout; the outputs are committed, the reference never travels).

    python tests/golden/make_fixtures.py /root/reference

Writes tests/golden/wr90_straight.json:
  * the 18-node / 30-tet mesh regression/OpenParEM3D/WR90/straight_study/straight.msh (gmsh 2.2),
  * the reference's DOF counts on it for ND1..ND4 (straight_order_{p}_iterations.txt, first data row),
  * the reference's S at 9 GHz on that unrefined mesh for ND1..ND3 (results.ods sheets 1-3, row 1),
  * the analytic exact.ods phase of S21 for the 100 mm guide (regression/OpenParEM3D/WR90/straight).
"""
import json
import os
import re
import sys
import zipfile
import xml.etree.ElementTree as ET


def read_msh(path):
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    nodes = []
    for k in range(nn):
        t = lines[i + 2 + k].split()
        nodes.append([float(t[1]), float(t[2]), float(t[3])])
    j = lines.index("$Elements")
    ne = int(lines[j + 1])
    tets, attrs = [], []
    for k in range(ne):
        t = [int(x) for x in lines[j + 2 + k].split()]
        etype, ntags = t[1], t[2]
        if etype == 4:  # 4-node tetrahedron
            tags = t[3:3 + ntags]
            tets.append([v - 1 for v in t[3 + ntags:3 + ntags + 4]])
            attrs.append(tags[0])
    return nodes, tets, attrs


def dof_count(path):
    for line in open(path):
        t = line.split()
        if len(t) >= 4 and re.match(r"^[0-9.e+]+$", t[0]) and t[1] == "1":
            return int(t[3])
    raise ValueError(path)


def ods_first_rows(path, nsheets):
    ns = {"table": "urn:oasis:names:tc:opendocument:xmlns:table:1.0",
          "text": "urn:oasis:names:tc:opendocument:xmlns:text:1.0",
          "office": "urn:oasis:names:tc:opendocument:xmlns:office:1.0"}
    root = ET.fromstring(zipfile.ZipFile(path).read("content.xml"))
    out = []
    for sheet in root.iter("{%s}table" % ns["table"]):
        rows = []
        for row in sheet.iter("{%s}table-row" % ns["table"]):
            cells = []
            for c in row.iter("{%s}table-cell" % ns["table"]):
                rep = int(c.get("{%s}number-columns-repeated" % ns["table"], "1"))
                v = c.get("{%s}value" % ns["office"])
                txt = "".join(c.itertext())
                cells.extend([v if v is not None else txt] * min(rep, 64))
            rows.append(cells)
        out.append(rows)
        if len(out) == nsheets:
            break
    return out


def main(ref):
    base = os.path.join(ref, "regression/OpenParEM3D/WR90/straight_study")
    nodes, tets, attrs = read_msh(os.path.join(base, "straight.msh"))
    dofs = {p: dof_count(os.path.join(base, f"straight_order_{p}_iterations.txt")) for p in (1, 2, 3, 4)}
    sheets = ods_first_rows(os.path.join(base, "results.ods"), 3)
    sref = {}
    for p, rows in zip((1, 2, 3), sheets):
        hdr = next(r for r in rows if any("S11" in str(x) for x in r))
        data = rows[rows.index(hdr) + 1]
        sref[p] = {"header": [str(x) for x in hdr[:12]], "row": [str(x) for x in data[:12]]}
    out = dict(source="regression/OpenParEM3D/WR90/straight_study (straight.msh, *_iterations.txt, results.ods)",
               nodes=nodes, tets=tets, tet_attr=attrs, dof_counts=dofs, s_results=sref,
               freq=9e9, air_er=1.0006, length=0.1, a=0.02286, b=0.01016)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wr90_straight.json")
    json.dump(out, open(dst, "w"), indent=1)
    print("wrote", dst, dofs)
    for p, v in sref.items():
        print(p, v)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
