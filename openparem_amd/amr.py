"""Mesh errors for adaptive refinement over the C ABI: the ZZ flux-recovery estimate on the GPU
(``opem_mesh_errors``) and the element marking (``opem_mark_elements``), the counterparts of the reference's
``fem3D::calculateMeshErrors`` (``src/OpenParEM3D/fem3D.cpp:1162-1380``) with ``OPEM_L2ZZErrorEstimator``."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from ._lib import check, lib, ptr


def mesh_errors(solver, hs, tol=1e-12, per_field=False):
    """Element errors (ne,) of the H fields ``hs`` (complex, (nsys, ndof) or one vector): the largest over the fields;
    with ``per_field`` also the (ne, nsys) errors of each."""
    H = np.ascontiguousarray(np.atleast_2d(hs), np.complex128)
    if H.shape[1] != solver.ndof:
        raise ValueError("H fields must have the space's DOF count")
    ne = solver.mesh.nt
    err = np.zeros(ne)
    es = np.zeros((ne, H.shape[0])) if per_field else None
    check(lib().opem_mesh_errors(solver.ctx, H.shape[0], ptr(H.view(np.float64), C.c_double), float(tol),
                                 ptr(err, C.c_double), ptr(es, C.c_double)), solver.ctx, "opem_mesh_errors")
    return (err, es) if per_field else err


def mark(errors, fraction):
    """opem_mark_elements: (element indices, largest first; the largest error)."""
    e = np.ascontiguousarray(errors, np.float64)
    n = C.c_int64(0)
    emax = C.c_double(0.0)
    out = np.zeros(max(1, e.size), np.int32)
    rc = lib().opem_mark_elements(e.size, ptr(e, C.c_double), float(fraction), C.byref(n), ptr(out, C.c_int32),
                                  C.byref(emax))
    if rc != 0:
        raise ValueError(f"opem_mark_elements failed (rc={rc})")
    return out[:n.value].copy(), emax.value


def refine(mesh, marked):
    """opem_refine_mesh: a new TetMesh with the marked tets bisected (conforming), and each new tet's parent."""
    xyz = np.ascontiguousarray(mesh.xyz, np.float64)
    tets = np.ascontiguousarray(mesh.tets, np.int32)
    ta = np.ascontiguousarray(mesh.tet_attr, np.int32)
    bf = np.ascontiguousarray(mesh.bfaces, np.int32)
    ba = np.ascontiguousarray(mesh.bface_attr, np.int32)
    mk = np.ascontiguousarray(marked, np.int32)
    nv, nt, nb = C.c_int64(0), C.c_int64(0), C.c_int64(0)
    L = lib()
    args = (xyz.shape[0], ptr(xyz, C.c_double), tets.shape[0], ptr(tets, C.c_int32), ptr(ta, C.c_int32), bf.shape[0],
            ptr(bf, C.c_int32), ptr(ba, C.c_int32), mk.size, ptr(mk, C.c_int32))
    rc = L.opem_refine_mesh(*args, C.byref(nv), ptr(None, C.c_double), C.byref(nt), ptr(None, C.c_int32),
                            ptr(None, C.c_int32), C.byref(nb), ptr(None, C.c_int32), ptr(None, C.c_int32),
                            ptr(None, C.c_int32))
    if rc != 0:
        raise ValueError(f"opem_refine_mesh failed (rc={rc})")
    X = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), np.int32)
    TA = np.zeros(nt.value, np.int32)
    B = np.zeros((nb.value, 3), np.int32)
    BA = np.zeros(nb.value, np.int32)
    par = np.zeros(nt.value, np.int32)
    rc = L.opem_refine_mesh(*args, C.byref(nv), ptr(X, C.c_double), C.byref(nt), ptr(T, C.c_int32),
                            ptr(TA, C.c_int32), C.byref(nb), ptr(B, C.c_int32), ptr(BA, C.c_int32),
                            ptr(par, C.c_int32))
    if rc != 0:
        raise ValueError(f"opem_refine_mesh failed (rc={rc})")
    out = dataclasses.replace(mesh, xyz=X, tets=T, tet_attr=TA, bfaces=B, bface_attr=BA)
    if hasattr(mesh, "geometry"):
        out.geometry = mesh.geometry
    return out, par
