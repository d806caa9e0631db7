"""Synthetic tetrahedral meshes for the per-frequency solve path.

These are the *inputs* of the solve path (the reference reads gmsh/MFEM meshes,
``src/OpenParEMCommon/mesh.cpp:386-449``); here they are generated from a seed so
every workload is reproducible without files:

* ``waveguide_mesh``  - air-filled rectangular guide with PEC walls and two wave
  ports (the WR90-style straight section used by
  ``regression/OpenParEM3D/WR90/straight*``).
* ``microstrip_mesh`` - shielded microstrip box: substrate, zero-thickness
  surface-impedance trace, PEC ground/walls, 1st-order radiation lid, two ports.
* ``tee_mesh``        - H-plane tee (three wave ports) or cross (four) junction of rectangular guides.
* ``random_tets``     - independent random affine tetrahedra (assembly-only workload).

Structured grids are split into 6 Kuhn tetrahedra per hexahedral cell (all cells
use the same main diagonal, so the split is conforming).  Grid vertices are
jittered by U(-j, +j) of the local cell size per axis, except along axes where the
vertex lies on a feature plane (outer walls, material interfaces, trace edges), so
every feature surface stays planar.

Attributes are 1-based, as in MFEM meshes.  Boundary attributes name boundary
kinds through ``TetMesh.bdr_kind``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# The 6 Kuhn tetrahedra of the unit cube: vertex paths (0,0,0) -> (1,1,1) along
# the 3! axis permutations.  Corner index c = dx + 2*dy + 4*dz.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_corners():
    tets = []
    for perm in _PERMS:
        c = [0, 0, 0]
        path = [0]
        for ax in perm:
            c[ax] = 1
            path.append(c[0] + 2 * c[1] + 4 * c[2])
        tets.append(path)
    return np.array(tets, dtype=np.int64)  # (6,4) corner ids


KUHN = _kuhn_corners()


@dataclass
class TetMesh:
    xyz: np.ndarray                 # (nv,3) float64, metres
    tets: np.ndarray                # (nt,4) int32 vertex ids
    tet_attr: np.ndarray            # (nt,) int32, 1-based volume attribute
    bfaces: np.ndarray              # (nb,3) int32 vertex ids of boundary / interface triangles
    bface_attr: np.ndarray          # (nb,) int32, 1-based boundary attribute
    bdr_kind: dict = field(default_factory=dict)   # bdr attr -> 'pec'|'port'|'zs'|'radiation'|'pmc'
    port_normal: dict = field(default_factory=dict)  # bdr attr of a port -> outward unit normal
    materials: dict = field(default_factory=dict)  # vol attr -> dict(er, tand, mur)
    name: str = ""

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.tets.shape[0]


def _grid_axis(lo, hi, n):
    return np.linspace(lo, hi, n + 1)


def _box_tets(nx, ny, nz):
    """Kuhn split of an nx*ny*nz cell grid; returns tets (6*nc,4) and cell ids."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i = i.transpose(2, 1, 0).ravel()  # cell order: x fastest, then y, then z
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    base = i + sx * j + sy * k
    corner_off = np.array([dx + sx * dy + sy * dz for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)])
    corners = base[:, None] + corner_off[None, :]          # (nc,8)
    tets = corners[:, KUHN]                                  # (nc,6,4)
    cell = np.repeat(np.arange(base.size), 6)
    return tets.reshape(-1, 4), cell, (i, j, k)


def _jitter(xs, ys, zs, feat, jitter, rng):
    """Vertex coordinates of the tensor grid with per-axis jitter off feature planes."""
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    coords = [X.transpose(2, 1, 0).ravel().copy(), Y.transpose(2, 1, 0).ravel().copy(),
              Z.transpose(2, 1, 0).ravel().copy()]
    n = [len(xs) - 1, len(ys) - 1, len(zs) - 1]
    idx = np.meshgrid(np.arange(n[0] + 1), np.arange(n[1] + 1), np.arange(n[2] + 1), indexing="ij")
    idx = [a.transpose(2, 1, 0).ravel() for a in idx]
    grids = [xs, ys, zs]
    if jitter > 0:
        r = rng.uniform(-1.0, 1.0, size=(3, coords[0].size))
        for ax in range(3):
            g = grids[ax]
            h = np.minimum(np.diff(g, prepend=g[0] - (g[1] - g[0]))[idx[ax]],
                           np.diff(g, append=g[-1] + (g[-1] - g[-2]))[idx[ax]])
            free = ~np.isin(idx[ax], np.asarray(sorted(feat[ax]), dtype=np.int64))
            coords[ax][free] += jitter * h[free] * r[ax][free]
    return np.stack(coords, axis=1)


def _boundary_faces(tets):
    """Triangles that belong to exactly one tet, plus the owning tet index."""
    loc = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    f = tets[:, loc].reshape(-1, 3)
    owner = np.repeat(np.arange(tets.shape[0]), 4)
    key = np.sort(f, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    once = cnt[inv.ravel()] == 1
    return f[once], owner[once]


def waveguide_mesh(a=22.86e-3, b=10.16e-3, length=30e-3, nx=12, ny=6, nz=18, jitter=0.10, seed=1,
                   er=1.0, tand=0.0, short=False, slab=None, pmc_sides=False, end=None):
    """Rectangular guide x in [0,a], y in [0,b], z in [0,L]; ports at z=0 (attr 2) and z=L (attr 3), or with
    ``short`` one port at z=0 and a PEC wall at z=L (attr 3).  ``slab = (k1, k2, er, tand[, mur])`` fills the
    whole cross-section between the grid planes k1 < k2 (cell indices along z, kept planar) with a second material
    (volume attribute 2).  ``end`` = "radiation" or "zs" closes z = L (attr 3) with that boundary instead of a port
    or PEC (one port at z = 0).  ``pmc_sides`` makes the side walls x = 0, a magnetic walls (attribute 4): with PEC at
    y = 0, b the guide then carries a TEM mode (uniform E_y, ``ports.tem_pmc_sides``)."""
    if end is not None and (short or end not in ("radiation", "zs")):
        raise ValueError('end must be "radiation" or "zs", without short')
    one_port = short or end is not None
    end_kind = "pec" if short else end if end else "port"
    rng = np.random.default_rng(seed)
    xs, ys, zs = _grid_axis(0, a, nx), _grid_axis(0, b, ny), _grid_axis(0, length, nz)
    feat = [{0, nx}, {0, ny}, {0, nz} | ({slab[0], slab[1]} if slab else set())]
    xyz = _jitter(xs, ys, zs, feat, jitter, rng)
    tets, _, _ = _box_tets(nx, ny, nz)
    bf, _ = _boundary_faces(tets)
    c = xyz[bf].mean(axis=1)
    tol = 1e-9
    attr = np.ones(bf.shape[0], dtype=np.int32)          # PEC walls
    if pmc_sides:
        attr[(np.abs(c[:, 0]) < tol) | (np.abs(c[:, 0] - a) < tol)] = 4
    attr[np.abs(c[:, 2]) < tol] = 2
    attr[np.abs(c[:, 2] - length) < tol] = 3
    tet_attr = np.ones(tets.shape[0], np.int32)
    materials = {1: dict(er=er, tand=tand, mur=1.0)}
    if slab:
        k1, k2, er2, tand2 = slab[:4]
        mur2 = slab[4] if len(slab) > 4 else 1.0
        if not 0 < k1 < k2 < nz:
            raise ValueError("slab planes must satisfy 0 < k1 < k2 < nz")
        zc = xyz[tets][:, :, 2].mean(axis=1)
        tet_attr[(zc > zs[k1]) & (zc < zs[k2])] = 2
        materials[2] = dict(er=er2, tand=tand2, mur=mur2)
    m = TetMesh(xyz=xyz, tets=tets.astype(np.int32), tet_attr=tet_attr,
                bfaces=bf.astype(np.int32), bface_attr=attr,
                bdr_kind={1: "pec", 2: "port", 3: end_kind, **({4: "pmc"} if pmc_sides else {})},
                port_normal={2: np.array([0.0, 0.0, -1.0])} if one_port else
                {2: np.array([0.0, 0.0, -1.0]), 3: np.array([0.0, 0.0, 1.0])},
                materials=materials,
                name=f"waveguide_{nx}x{ny}x{nz}" + ("_short" if short else "") + (f"_{end}" if end else "") + ("_slab" if slab else "") + ("_pmc" if pmc_sides else ""))
    m.geometry = dict(a=a, b=b, length=length)
    return m


def tee_mesh(a=22.86e-3, b=10.16e-3, length=60e-3, arm=25e-3, z0=None, h=2.5e-3, ny=3, cross=False):
    """H-plane tee: main guide x in [0,a], y in [0,b], z in [0,L] with ports at z=0 (attr 2) and z=L (attr 3);
    side arm x in [a, a+arm], z in [z0, z0+a] with its port at x=a+arm (attr 4); with ``cross`` a second arm
    x in [-arm, 0] with its port at x=-arm (attr 5).  A structured Kuhn grid of cell size about h (no jitter:
    the junction's reentrant edges stay on grid lines).  Each port's TE10 mode is described in
    ``geometry["ports"]`` by the corner of its cross-section, the unit width axis and a, b."""
    z0 = (length - a) / 2 if z0 is None else z0

    def seg(lo, hi):
        return max(1, int(round((hi - lo) / h)))

    xseg = ([(-arm, 0.0, seg(-arm, 0.0))] if cross else []) + [(0.0, a, seg(0, a)), (a, a + arm, seg(a, a + arm))]
    xs, _ = _graded(xseg)
    zs, _ = _graded([(0.0, z0, seg(0, z0)), (z0, z0 + a, seg(z0, z0 + a)), (z0 + a, length, seg(z0 + a, length))])
    ys = _grid_axis(0.0, b, ny)
    nx, nz = len(xs) - 1, len(zs) - 1
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    xyz_all = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()], 1)
    tets, cell, (ci, cj, ck) = _box_tets(nx, ny, nz)
    xc = 0.5 * (xs[ci] + xs[ci + 1])
    zc = 0.5 * (zs[ck] + zs[ck + 1])
    keep_cell = ((xc > 0) & (xc < a)) | ((zc > z0) & (zc < z0 + a))
    tets = tets[keep_cell[cell]]
    used = np.unique(tets)
    remap = -np.ones(xyz_all.shape[0], np.int64)
    remap[used] = np.arange(used.size)
    tets = remap[tets]
    xyz = xyz_all[used]
    bf, _ = _boundary_faces(tets)
    c = xyz[bf].mean(axis=1)
    tol = 1e-9
    attr = np.ones(bf.shape[0], dtype=np.int32)          # PEC walls
    attr[np.abs(c[:, 2]) < tol] = 2
    attr[np.abs(c[:, 2] - length) < tol] = 3
    attr[np.abs(c[:, 0] - (a + arm)) < tol] = 4
    kinds = {1: "pec", 2: "port", 3: "port", 4: "port"}
    normals = {2: np.array([0.0, 0.0, -1.0]), 3: np.array([0.0, 0.0, 1.0]), 4: np.array([1.0, 0.0, 0.0])}
    ports = {2: dict(origin=np.zeros(3), u=np.array([1.0, 0.0, 0.0])),
             3: dict(origin=np.zeros(3), u=np.array([1.0, 0.0, 0.0])),
             4: dict(origin=np.array([0.0, 0.0, z0]), u=np.array([0.0, 0.0, 1.0]))}
    if cross:
        attr[np.abs(c[:, 0] + arm) < tol] = 5
        kinds[5] = "port"
        normals[5] = np.array([-1.0, 0.0, 0.0])
        ports[5] = dict(origin=np.array([0.0, 0.0, z0]), u=np.array([0.0, 0.0, 1.0]))
    m = TetMesh(xyz=xyz, tets=tets.astype(np.int32), tet_attr=np.ones(tets.shape[0], np.int32),
                bfaces=bf.astype(np.int32), bface_attr=attr, bdr_kind=kinds, port_normal=normals,
                materials={1: dict(er=1.0, tand=0.0, mur=1.0)},
                name=f"tee_{nx}x{ny}x{nz}")
    m.geometry = dict(a=a, b=b, length=length, arm=arm, z0=z0, ports=ports)
    return m


def _graded(segments):
    """Concatenate uniform segments [(lo,hi,n),...] into one grid-line array and feature indices."""
    pts = [segments[0][0]]
    feats = {0}
    for lo, hi, n in segments:
        pts.extend(np.linspace(lo, hi, n + 1)[1:])
        feats.add(len(pts) - 1)
    return np.array(pts), feats


def microstrip_mesh(W=40e-3, H=20e-3, length=60e-3, h_sub=1.5e-3, w_trace=2.8e-3,
                    nx=36, ny=18, nz=75, n_sub=2, jitter=0.10, seed=3, er_sub=4.4, tand_sub=0.02):
    """Shielded microstrip box (x: width, y: height, z: length).

    Volume attrs: 1 air, 2 substrate.  Boundary attrs: 1 PEC (ground y=0, side walls x=0,W),
    2 port z=0, 3 port z=L, 4 surface-impedance trace (interior faces on y=h_sub), 5 radiation (y=H).
    """
    rng = np.random.default_rng(seed)
    x0, x1 = (W - w_trace) / 2, (W + w_trace) / 2
    n_tr = max(2, int(round(nx * w_trace / W)))
    n_side = max(1, (nx - n_tr) // 2)
    xs, fx = _graded([(0.0, x0, n_side), (x0, x1, n_tr), (x1, W, n_side)])
    ys, fy = _graded([(0.0, h_sub, n_sub), (h_sub, H, max(1, ny - n_sub))])
    zs = _grid_axis(0, length, nz)
    nxr, nyr = len(xs) - 1, len(ys) - 1
    feat = [fx, fy, {0, nz}]
    xyz = _jitter(xs, ys, zs, feat, jitter, rng)
    tets, cell, (ci, cj, ck) = _box_tets(nxr, nyr, nz)
    tet_attr = np.where(cj[cell] < n_sub, 2, 1).astype(np.int32)
    bf, _ = _boundary_faces(tets)
    c = xyz[bf].mean(axis=1)
    tol = 1e-9
    attr = np.ones(bf.shape[0], dtype=np.int32)
    attr[np.abs(c[:, 2]) < tol] = 2
    attr[np.abs(c[:, 2] - length) < tol] = 3
    attr[np.abs(c[:, 1] - H) < tol] = 5
    # interior trace faces: triangles of the Kuhn split lying on y = h_sub within the trace
    loc = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    f = tets[:, loc].reshape(-1, 3)
    iy = np.searchsorted(ys, h_sub - 1e-12)
    vy = np.rint((np.arange(xyz.shape[0]) // (nxr + 1)) % (nyr + 1)).astype(np.int64)
    vx = np.arange(xyz.shape[0]) % (nxr + 1)
    on = np.all(vy[f] == iy, axis=1)
    ix0, ix1 = sorted(fx)[1], sorted(fx)[2]
    inside = on & np.all((vx[f] >= ix0) & (vx[f] <= ix1), axis=1)
    tf = f[inside]
    key = np.sort(tf, axis=1)
    _, first = np.unique(key, axis=0, return_index=True)
    tf = tf[np.sort(first)]
    bf = np.concatenate([bf, tf])
    attr = np.concatenate([attr, np.full(tf.shape[0], 4, np.int32)])
    m = TetMesh(xyz=xyz, tets=tets.astype(np.int32), tet_attr=tet_attr, bfaces=bf.astype(np.int32),
                bface_attr=attr.astype(np.int32),
                bdr_kind={1: "pec", 2: "port", 3: "port", 4: "zs", 5: "radiation"},
                port_normal={2: np.array([0.0, 0.0, -1.0]), 3: np.array([0.0, 0.0, 1.0])},
                materials={1: dict(er=1.0, tand=0.0, mur=1.0),
                           2: dict(er=er_sub, tand=tand_sub, mur=1.0)},
                name=f"microstrip_{nxr}x{nyr}x{nz}")
    m.geometry = dict(W=W, H=H, length=length, h_sub=h_sub, w_trace=w_trace, x0=x0, x1=x1)
    return m


def aspect_ratio(v):
    """rho = l_max / (2*sqrt(6)*r_in) for tets v (n,4,3); 1 for a regular tetrahedron."""
    e = np.stack([v[:, j] - v[:, i] for i, j in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]], 1)
    lmax = np.linalg.norm(e, axis=2).max(axis=1)
    vol = np.abs(np.einsum("ij,ij->i", e[:, 0], np.cross(e[:, 1], e[:, 2]))) / 6.0
    area = 0.0
    for f in ([1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]):
        area = area + 0.5 * np.linalg.norm(np.cross(v[:, f[1]] - v[:, f[0]], v[:, f[2]] - v[:, f[0]]), axis=1)
    r_in = 3.0 * vol / area
    return lmax / (2.0 * np.sqrt(6.0) * r_in)


def random_tets(n, seed=7, jitter=0.2, max_aspect=8.0):
    """n independent tets: a Kuhn tet of the unit cube plus U(+-jitter) per coordinate, aspect <= max.

    Returns vertex coordinates (n,4,3) float64 and per-element er ~ U[1,10], tand ~ U[0,0.02].
    """
    rng = np.random.default_rng(seed)
    cube = np.array([[dx, dy, dz] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], dtype=np.float64)
    out = np.empty((0, 4, 3))
    while out.shape[0] < n:
        m = int((n - out.shape[0]) * 1.3) + 16
        k = rng.integers(0, 6, size=m)
        v = cube[KUHN[k]] + rng.uniform(-jitter, jitter, size=(m, 4, 3))
        v = v[aspect_ratio(v) <= max_aspect]
        out = np.concatenate([out, v])
    out = out[:n]
    er = rng.uniform(1.0, 10.0, size=n)
    tand = rng.uniform(0.0, 0.02, size=n)
    return out, er, tand
