"""Project ingest: the reference's project, mesh, port-definition and materials files -> a mesh the solver takes.

    p = load_project("straight.proj")          # reads the mesh, ports file and materials the project names
    with Solver(p.mesh, p.order, boundaries=p.boundary_values) as s:
        r = s.sweep(p.frequencies)

Covers, each restated from the reference's behaviour (file:line):

* gmsh 2.2 ASCII meshes with $PhysicalNames (``MeshMaterialList::loadGMSH``, src/OpenParEMCommon/mesh.cpp:386-449;
  the volume attribute of a tetrahedron is its physical tag, as MFEM reads it) and MFEM v1.0 meshes;
* the port-definition file's Path, Port and Boundary blocks (``BoundaryDatabase`` loading, port.cpp) and the marking
  of the mesh boundary (``BoundaryDatabase::assignAttributes`` / ``markMeshBoundaries``, port.cpp:6885-7010):
  every boundary triangle starts as the default (PEC, attribute 1, or the project's default surface-impedance
  material); a triangle whose three vertices lie inside a port's path takes that port's attribute, then one inside
  a non-default boundary's path takes the boundary's;
* the point-in-path test of ``Path::is_point_inside`` (src/OpenParEMCommon/path.cpp:1233-1284): on the path within
  the path tolerance, or coplanar with it and wound once around;
* the materials database (``Temperature::get_eps`` / ``get_mu`` / ``get_Rs``, OpenParEMmaterials.cpp:600-830):
  "any" or an exact frequency block, else linear interpolation between the nearest blocks; loss as a loss tangent
  or a conductivity; the Djordjevic-Sarkar wideband model for blocks without frequency entries;
* the project file keys the solve path needs (mesh file and order, port and materials files, frequency plan,
  reference impedance, tolerances, default boundary material).

Differences from the reference, by design: one boundary attribute per port (the reference gives a port one
attribute per adjacent volume region); paths of a port or boundary are taken as separate closed outlines (a
triangle inside any of them belongs to it) instead of being merged.
"""
from __future__ import annotations

import cmath
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .mesh import TetMesh, _boundary_faces

EPS0 = 8.8541878176e-12
MU0 = 4e-7 * math.pi
PATH_TOL = 1e-11  # absolute tolerance of the path geometry (src/OpenParEMCommon/path.hpp:57)


# ----------------------------------------------------------------------------------------------- meshes
def read_gmsh(path):
    """gmsh 2.2 ASCII: nodes, tetrahedra (type 4) with their physical tags, triangles (type 2), physical names."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    if not lines or lines[0] != "$MeshFormat":
        raise ValueError(f"{path}: not a gmsh mesh")
    ver = lines[1].split()
    if not ver[0].startswith("2.") or ver[1] != "0":
        raise ValueError(f"{path}: only gmsh 2.x ASCII meshes are read (found {lines[1]})")
    names = {}
    if "$PhysicalNames" in lines:
        i = lines.index("$PhysicalNames")
        for k in range(int(lines[i + 1])):
            m = re.match(r'(\d+)\s+(\d+)\s+"(.*)"', lines[i + 2 + k])
            names[(int(m.group(1)), int(m.group(2)))] = m.group(3)
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    ids = np.empty(nn, np.int64)
    xyz = np.empty((nn, 3))
    for k in range(nn):
        t = lines[i + 2 + k].split()
        ids[k] = int(t[0])
        xyz[k] = [float(t[1]), float(t[2]), float(t[3])]
    index = {int(v): k for k, v in enumerate(ids)}
    j = lines.index("$Elements")
    tets, tet_attr, tris, tri_attr = [], [], [], []
    for k in range(int(lines[j + 1])):
        t = [int(x) for x in lines[j + 2 + k].split()]
        etype, ntags = t[1], t[2]
        tags = t[3:3 + ntags]
        verts = [index[v] for v in t[3 + ntags:]]
        phys = tags[0] if tags else 1
        if etype == 4:
            tets.append(verts[:4])
            tet_attr.append(phys)
        elif etype == 2:
            tris.append(verts[:3])
            tri_attr.append(phys)
    return dict(xyz=xyz, tets=np.array(tets, np.int64).reshape(-1, 4), tet_attr=np.array(tet_attr, np.int32),
                tris=np.array(tris, np.int64).reshape(-1, 3), tri_attr=np.array(tri_attr, np.int32),
                region_names={tag: nm for (dim, tag), nm in names.items() if dim == 3})


def read_mfem(path):
    """MFEM mesh v1.0 (ASCII, 3D, tetrahedra): elements "attr 4 v0 v1 v2 v3", boundary "attr 2 v0 v1 v2",
    vertices."""
    with open(path) as fh:
        toks = [ln.split("#")[0].strip() for ln in fh]
    toks = [t for t in toks if t]
    if not toks[0].startswith("MFEM mesh v1."):
        raise ValueError(f"{path}: not an MFEM mesh")
    k = toks.index("dimension")
    if int(toks[k + 1]) != 3:
        raise ValueError(f"{path}: only 3D meshes are read")

    def block(name, width):
        s = toks.index(name)
        n = int(toks[s + 1])
        return [[int(v) for v in toks[s + 2 + i].split()[:width]] for i in range(n)]

    els = block("elements", 6)
    if any(e[1] != 4 for e in els):
        raise ValueError(f"{path}: only tetrahedral elements are read")
    bdr = block("boundary", 5)
    s = toks.index("vertices")
    nv, dim = int(toks[s + 1]), int(toks[s + 2])
    xyz = np.array([[float(v) for v in toks[s + 3 + i].split()[:dim]] for i in range(nv)])
    return dict(xyz=xyz, tets=np.array([e[2:6] for e in els], np.int64).reshape(-1, 4),
                tet_attr=np.array([e[0] for e in els], np.int32),
                tris=np.array([b[2:5] for b in bdr if b[1] == 2], np.int64).reshape(-1, 3),
                tri_attr=np.array([b[0] for b in bdr if b[1] == 2], np.int32), region_names={})


def read_mesh(path):
    with open(path) as fh:
        first = fh.readline().strip()
    if first == "$MeshFormat":
        return read_gmsh(path)
    if first.startswith("MFEM mesh v1."):
        return read_mfem(path)
    raise ValueError(f"{path}: unrecognized mesh format")


# ----------------------------------------------------------------------------------------------- path geometry
def _same(a, b, tol):
    """Relative comparison of two numbers (double_compare, src/OpenParEMCommon/misc.cpp:76-84)."""
    if a == b:
        return True
    if a == 0:
        return abs(b) < tol
    if b == 0:
        return abs(a) < tol
    if abs(a) < tol and abs(b) < tol:
        return True
    return abs((a - b) / a) < tol


def _pad3(p):
    p = [float(v) for v in p]
    return p + [0.0] * (3 - len(p))


def point_on_segment(p, a, b, tol):
    """p on the segment a-b: at an end point (coordinates equal to relative tolerance tol), or inside the box of
    the two end points widened by tol and spanning a triangle (p, a, b) of area below tol
    (is_point_on_line, src/OpenParEMCommon/path.cpp:155-163)."""
    p, a, b = _pad3(p), _pad3(a), _pad3(b)
    if all(_same(x, y, tol) for x, y in zip(p, a)) or all(_same(x, y, tol) for x, y in zip(p, b)):
        return True
    for k in range(3):
        lo, hi = min(a[k], b[k]), max(a[k], b[k])
        if p[k] < lo - tol or p[k] > hi + tol:
            return False
    u = np.subtract(a, p)
    v = np.subtract(b, p)
    return 0.5 * float(np.linalg.norm(np.cross(u, v))) < tol


@dataclass
class Outline:
    """A named path: its points (k, 3) and whether it closes on itself."""
    name: str
    points: np.ndarray
    closed: bool

    def __post_init__(self):
        # plane normal of the outline (Newell's method), used to sign the winding angles
        P = self.points
        n = np.zeros(3)
        for i in range(len(P)):
            q, r = P[i], P[(i + 1) % len(P)]
            n += np.array([(q[1] - r[1]) * (q[2] + r[2]), (q[2] - r[2]) * (q[0] + r[0]),
                           (q[0] - r[0]) * (q[1] + r[1])])
        nn = np.linalg.norm(n)
        self.normal = n / nn if nn > 0 else n

    def segments(self):
        P = self.points
        k = len(P)
        for i in range(k - 1):
            yield P[i], P[i + 1]
        if self.closed and k > 2:
            yield P[-1], P[0]

    def on_path(self, p, tol=PATH_TOL):
        return any(point_on_segment(p, a, b, tol) for a, b in self.segments())

    def interior(self, p, tol=PATH_TOL):
        """Strictly inside a closed planar outline: coplanar with it (every tetrahedron of the point and three
        consecutive outline points has volume below tol) and wound around once (the signed angles the outline's
        edges subtend at the point sum to 2 pi)."""
        P = self.points
        if not self.closed or len(P) < 3 or self.on_path(p, tol):
            return False
        p = np.asarray(_pad3(p))
        for i in range(len(P) - 2):
            vol = np.dot(np.cross(P[i] - p, P[i + 1] - p), P[i + 2] - p) / 6.0
            if abs(vol) > tol:
                return False
        total = 0.0
        for a, b in self.segments():
            u = (a - p) / np.linalg.norm(a - p)
            v = (b - p) / np.linalg.norm(b - p)
            total += math.atan2(float(np.dot(np.cross(u, v), self.normal)), float(np.dot(u, v)))
        return _same(abs(total), 2 * math.pi, tol)

    def inside(self, p, tol=PATH_TOL):
        """On the outline or strictly inside it (Path::is_point_inside)."""
        return self.on_path(p, tol) or self.interior(p, tol)


# ----------------------------------------------------------------------------------------------- port file
def _blocks(lines, start, end):
    """Top-level blocks `start` ... `end` of a keyword file, as lists of stripped lines (nested blocks kept)."""
    out, cur, depth = [], None, 0
    for ln in lines:
        if cur is None:
            if ln == start:
                cur, depth = [], 0
            continue
        if ln == end and depth == 0:
            out.append(cur)
            cur = None
            continue
        if re.match(r"^[A-Z][A-Za-z]*$", ln):
            depth += 1 if not ln.startswith("End") else -1
        cur.append(ln)
    return out


def _kv(block, depth0_only=True):
    """key=value pairs of a block at its top nesting level; repeated keys are collected in lists."""
    out, depth = {}, 0
    for ln in block:
        if re.match(r"^[A-Z][A-Za-z]*$", ln):
            depth += 1 if not ln.startswith("End") else -1
            continue
        if depth and depth0_only:
            continue
        if "=" in ln:
            k, v = ln.split("=", 1)
            out.setdefault(k.strip(), []).append(v.strip())
    return out


def _strip_comments(text):
    out = []
    for ln in text.split("\n"):
        ln = ln.split("//")[0].strip()
        if ln:
            out.append(ln)
    return out


@dataclass
class PortDef:
    name: str
    paths: list            # outline names
    sports: list           # S-port numbers of the port's modes or lines


@dataclass
class BoundaryDef:
    name: str
    kind: str              # perfect_electric_conductor | perfect_magnetic_conductor | surface_impedance | radiation
    paths: list
    material: str | None = None


def read_ports(path):
    """The port-definition file: outlines (Path), ports (Port) and boundaries (Boundary)."""
    with open(path) as fh:
        lines = _strip_comments(fh.read())
    if not lines or not lines[0].startswith("#OpenParEMports"):
        raise ValueError(f"{path}: not an OpenParEM ports file")
    outlines = {}
    for b in _blocks(lines, "Path", "EndPath"):
        kv = _kv(b)
        pts = [[float(v) for v in re.match(r"\((.*)\)", s).group(1).split(",")] for s in kv.get("point", [])]
        name = kv["name"][0]
        outlines[name] = Outline(name, np.array([_pad3(p) for p in pts]), kv.get("closed", ["false"])[0] == "true")
    ports = []
    for b in _blocks(lines, "Port", "EndPort"):
        kv = _kv(b)
        allkv = _kv(b, depth0_only=False)
        sports = sorted({int(s) for s in allkv.get("Sport", [])})
        ports.append(PortDef(kv["name"][0], [p.lstrip("+-") for p in kv.get("path", [])], sports))
    bounds = []
    for b in _blocks(lines, "Boundary", "EndBoundary"):
        kv = _kv(b)
        bounds.append(BoundaryDef(kv["name"][0], kv["type"][0], [p.lstrip("+-") for p in kv.get("path", [])],
                                  kv.get("material", [None])[0]))
    for d in ports + bounds:
        for p in d.paths:
            if p not in outlines:
                raise ValueError(f"{path}: {d.name} names an undefined path {p}")
    return outlines, ports, bounds


# ----------------------------------------------------------------------------------------------- materials
def read_materials(path):
    """Material database: {name: [temperature block, ...]}; a temperature block is {"temperature": str,
    "frequencies": [{key: str}, ...]} or, for the wideband model, {"temperature": str, "model": {key: str}}."""
    with open(path) as fh:
        lines = _strip_comments(fh.read())
    out = {}
    for b in _blocks(lines, "Material", "EndMaterial"):
        name = _kv(b)["name"][0]
        temps = []
        for t in _blocks(b, "Temperature", "EndTemperature"):
            kv = {k: v[0] for k, v in _kv(t).items()}
            freqs = [{k: v[0] for k, v in _kv(f).items()} for f in _blocks(t, "Frequency", "EndFrequency")]
            temps.append(dict(temperature=kv.pop("temperature", "any"), frequencies=freqs, model=kv))
        out[name] = temps
    return out


def _loss_key(block):
    for k in ("loss_tangent", "tand", "tandel"):
        if k in block:
            return k, True
    return "conductivity", False


def _temperature(mat, temperature, tol=1e-12):
    for t in mat:
        if t["temperature"] == "any" or _same(float(t["temperature"]), temperature, tol):
            return t
    raise ValueError(f"material has no data at {temperature} C")


def _lookup(freqs, f, get, tol):
    """Frequency-list lookup: an "any" block, an exact match, else linear interpolation between the nearest
    blocks below and above f (no extrapolation)."""
    lo = hi = None
    for blk in freqs:
        if blk["frequency"] == "any":
            return get(blk)
        fb = float(blk["frequency"])
        if _same(f, fb, tol):
            return get(blk)
        if fb < f and (lo is None or fb > lo[0]):
            lo = (fb, get(blk))
        if fb > f and (hi is None or fb < hi[0]):
            hi = (fb, get(blk))
    if lo is None or hi is None:
        raise ValueError(f"material has no data at {f} Hz (no extrapolation)")
    w = (f - lo[0]) / (hi[0] - lo[0])
    return tuple(a + w * (b - a) for a, b in zip(lo[1], hi[1]))


def material_eps(mat, f, temperature=20.0, tol=1e-12):
    """Complex permittivity eps (F/m) of a material at frequency f (Temperature::get_eps)."""
    t = _temperature(mat, temperature, tol)
    w = 2 * math.pi * f
    if t["frequencies"]:
        key, is_tand = _loss_key(t["frequencies"][0])
        er, loss = _lookup(t["frequencies"], f, lambda b: (float(b["er"]), float(b.get(key, 0.0))), tol)
        return complex(er * EPS0, -loss * er * EPS0) if is_tand else complex(er * EPS0, -loss / w)
    m = t["model"]  # Djordjevic-Sarkar: er_inf + delta_er/(m2-m1) log10((10^m2 + j w)/(10^m1 + j w)) - j sigma/w
    key, is_tand = _loss_key(m)
    er_inf, d_er, m1, m2 = (float(m[k]) for k in ("er_infinity", "delta_er", "m1", "m2"))
    loss = float(m.get(key, 0.0))
    sigma = w * er_inf * EPS0 * loss if is_tand else loss
    t1, t2 = complex(10 ** m1, w), complex(10 ** m2, w)
    return er_inf * EPS0 + d_er * EPS0 / (m2 - m1) * cmath.log(t2 / t1) / math.log(10) - 1j * sigma / w


def material_mur(mat, f, temperature=20.0, tol=1e-12):
    t = _temperature(mat, temperature, tol)
    if t["frequencies"]:
        return _lookup(t["frequencies"], f, lambda b: (float(b.get("mur", 1.0)),), tol)[0]
    return float(t["model"].get("mur", 1.0))


def material_conductivity(mat, f, temperature=20.0, tol=1e-12):
    """Conductivity (S/m) of a conductor material (the loss entry of its frequency block)."""
    t = _temperature(mat, temperature, tol)
    return _lookup(t["frequencies"], f, lambda b: (float(b["conductivity"]),), tol)[0]


# ----------------------------------------------------------------------------------------------- project
def read_project(path):
    """Keyword file of a project: {key: [values]} (comments after // dropped)."""
    out = {}
    with open(path) as fh:
        for ln in _strip_comments(fh.read()):
            if ln.startswith("#"):
                continue
            parts = ln.split(None, 1)
            out.setdefault(parts[0], []).append(parts[1].strip() if len(parts) > 1 else "")
    return out


def _frequency_plan(kv):
    f = [float(v) for v in kv.get("frequency.plan.point", [])]
    for v in kv.get("frequency.plan.linear", []):  # start, stop, step
        a, b, s = (float(x) for x in v.split(","))
        n = int(round((b - a) / s))
        f += [a + k * s for k in range(n + 1)]
    return np.array(sorted(set(f)))


def _find(base, names):
    for d, n in names:
        if not n or n.startswith("//"):
            continue
        for cand in (os.path.join(base, n), os.path.join(base, d, n)):
            if os.path.exists(cand):
                return cand
    return None


@dataclass
class Project:
    mesh: TetMesh
    order: int
    frequencies: np.ndarray
    reference_impedance: float
    tolerance: float
    iteration_limit: int
    boundary_values: dict = field(default_factory=dict)   # bdr attr -> conductivity (zs) or wave impedance
    port_names: list = field(default_factory=list)         # in S-port order
    region_names: dict = field(default_factory=dict)       # volume attr -> region (material) name


def _outward_normals(xyz, tets, bf):
    """Outward unit normal of each boundary face (away from the tetrahedron that owns it)."""
    faces = {}
    for t, tet in enumerate(tets):
        for k in range(4):
            f = tuple(sorted(np.delete(tet, k)))
            faces[f] = tet[k]
    out = np.empty((len(bf), 3))
    for i, f in enumerate(bf):
        a, b, c = xyz[f]
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        if np.dot(n, xyz[faces[tuple(sorted(f))]] - a) > 0:
            n = -n
        out[i] = n
    return out


def mark_boundaries(xyz, bf, outlines, ports, bounds, default_kind="pec"):
    """Boundary attribute of each boundary triangle, port attributes in S-port order, and {attr: (kind, def)}.

    Attribute 1 is PEC; a default surface-impedance material gets the first free attribute; ports follow in S-port
    order, then the non-default boundaries.  A triangle belongs to a port or boundary when its three vertices lie
    inside one of its outlines; a boundary overrides a port."""
    ports = sorted(ports, key=lambda p: p.sports[0] if p.sports else 0)
    nxt = 2
    kinds = {1: ("pec", None)}
    default_attr = 1
    if default_kind != "pec":
        kinds[nxt] = ("zs", default_kind)
        default_attr = nxt
        nxt += 1
    attr = np.full(len(bf), default_attr, np.int32)
    port_attr = []
    for p in ports:
        kinds[nxt] = ("port", p)
        port_attr.append(nxt)
        nxt += 1
    kind_map = {"perfect_electric_conductor": "pec", "perfect_magnetic_conductor": "pmc",
                "surface_impedance": "zs", "radiation": "radiation"}
    bnd_attr = []
    for b in bounds:
        if b.kind not in kind_map:
            raise ValueError(f"boundary {b.name}: unknown type {b.kind}")
        kinds[nxt] = (kind_map[b.kind], b)
        bnd_attr.append(nxt)
        nxt += 1
    assigned = set()
    for i, f in enumerate(bf):
        pts = xyz[f]
        for pa, p in zip(port_attr, ports):
            if any(all(outlines[o].inside(q) for q in pts) for o in p.paths):
                attr[i] = pa
                assigned.add(pa)
                break
        for ba, b in zip(bnd_attr, bounds):
            if any(all(outlines[o].inside(q) for q in pts) for o in b.paths):
                attr[i] = ba
                assigned.add(ba)
                break
    missing = [kinds[a][1].name for a in port_attr + bnd_attr if a not in assigned]
    if missing:
        raise ValueError(f"not assigned to the mesh boundary: {', '.join(missing)}")
    return attr, port_attr, kinds


def _rect_port_geometry(outline, normal):
    """Corner and broad-wall direction of a rectangular port outline (for the analytic TE10 mode)."""
    P = outline.points
    if len(P) != 4:
        return None
    edges = [(P[i], P[(i + 1) % 4]) for i in range(4)]
    lens = [np.linalg.norm(b - a) for a, b in edges]
    k = int(np.argmax(lens))
    a, b = edges[k]
    u = (b - a) / lens[k]
    return dict(origin=a, u=u, a=lens[k], b=lens[(k + 1) % 4])


def load_project(path, temperature=None):
    """Read a project file and everything it names; returns a Project whose mesh carries volume materials at the
    first frequency of the plan, boundary kinds and port normals."""
    base = os.path.dirname(os.path.abspath(path))
    kv = read_project(path)

    def one(key, default=None):
        return kv.get(key, [default])[0]

    mesh_path = os.path.join(base, one("mesh.file"))
    ports_path = os.path.join(base, one("port.definition.file"))
    order = int(one("mesh.order", "1"))
    freqs = _frequency_plan(kv)
    temp = float(temperature if temperature is not None else one("solution.temperature", "20"))
    mats = {}
    for d, n in ((one("materials.global.path", "./"), one("materials.global.name")),
                 (one("materials.local.path", "./"), one("materials.local.name"))):
        p = _find(base, [(d, n)])
        if p:
            mats.update(read_materials(p))
    m = read_mesh(mesh_path)
    outlines, ports, bounds = read_ports(ports_path)
    tets = m["tets"]
    bf, _ = _boundary_faces(tets)
    bf = bf.astype(np.int64)
    default = one("materials.default.boundary", "PEC")
    battr, port_attr, kinds = mark_boundaries(m["xyz"], bf, outlines, ports, bounds,
                                              "pec" if default == "PEC" else default)
    normals = _outward_normals(m["xyz"], tets, bf)
    f0 = float(freqs[0]) if len(freqs) else 1e9
    materials, region_names = {}, {}
    for attr in np.unique(m["tet_attr"]):
        name = m["region_names"].get(int(attr), str(int(attr)))
        region_names[int(attr)] = name
        if name not in mats:
            raise ValueError(f"region {name} (attribute {attr}) is not in the materials database")
        eps = material_eps(mats[name], f0, temp)
        materials[int(attr)] = dict(er=eps.real / EPS0, tand=-eps.imag / eps.real, mur=material_mur(mats[name], f0, temp))
    bdr_kind, bvals, port_normal, geometry = {}, {}, {}, {"ports": {}}
    for a, (kind, d) in kinds.items():
        if a not in np.unique(battr) and kind != "pec":
            continue
        bdr_kind[a] = kind
        if kind == "zs":
            mname = d if isinstance(d, str) else d.material
            if mname not in mats:
                raise ValueError(f"boundary material {mname} is not in the materials database")
            bvals[a] = material_conductivity(mats[mname], f0, temp)
        if kind == "port":
            sel = battr == a
            n = normals[sel].mean(axis=0)
            port_normal[a] = n / np.linalg.norm(n)
            g = _rect_port_geometry(outlines[d.paths[0]], port_normal[a])
            if g:
                geometry["ports"][a] = dict(origin=g["origin"], u=g["u"])
                geometry.setdefault("a", g["a"])
                geometry.setdefault("b", g["b"])
    if not (battr == 1).any():
        bdr_kind.pop(1, None)
    # the analytic TE10 port modes apply when every port outline is a rectangle (mesh names select the provider)
    stem = os.path.splitext(os.path.basename(mesh_path))[0]
    rect = len(geometry["ports"]) == len(port_attr)
    mesh = TetMesh(xyz=m["xyz"], tets=tets.astype(np.int32), tet_attr=m["tet_attr"].astype(np.int32),
                   bfaces=bf.astype(np.int32), bface_attr=battr, bdr_kind=bdr_kind, port_normal=port_normal,
                   materials=materials, name=("waveguide_" if rect else "ingested_") + stem)
    mesh.geometry = geometry
    return Project(mesh=mesh, order=order, frequencies=freqs, reference_impedance=float(one("reference.impedance", "0")),
                   tolerance=float(one("solution.3D.tolerance", "1e-10")),
                   iteration_limit=int(float(one("solution.iteration.limit", "10000"))), boundary_values=bvals,
                   port_names=[kinds[a][1].name for a in port_attr], region_names=region_names)
