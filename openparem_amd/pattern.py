"""Radiation patterns over the C ABI: radiation-boundary currents, their far field on the GPU, the pattern sphere
and circles, and the antenna metrics -- the counterparts of the reference's ``fem3D::calculateRadiationCurrents`` /
``calculateRadiationPatterns`` (``src/OpenParEM3D/fem3D.cpp:1902-2044``), ``Boundary::calculateRadiationCurrents``
(``port.cpp:1043-1265``), ``BoundaryDatabase::calculateFarField`` (``port.cpp:8077-8160``) and ``Sphere`` /
``Circle`` / ``Pattern`` (``pattern.cpp``).

Fields are complex ``[npts][4]`` in the order E_phi, E_theta, H_phi, H_theta (``OPEMpoint::set_fields``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, cplx_view, lib, ptr

MU0 = 4e-7 * np.pi


def radiation_currents(solver, x, h, freq, limit, attrs=None):
    """opem_radiation_currents: J = n x H_t, M = -n x E_t at the centres of the radiation faces split so that no
    edge exceeds ``limit`` wavelengths.  Returns dict xyz, J, M (complex), area, eta."""
    x = np.ascontiguousarray(x, np.complex128)
    h = np.ascontiguousarray(h, np.complex128)
    if x.shape != (solver.ndof,) or h.shape != (solver.ndof,):
        raise ValueError("x and h must be complex vectors of the space's DOF count")
    at = np.ascontiguousarray([] if attrs is None else attrs, np.int32)
    n = C.c_int64(0)
    eta = C.c_double(0.0)
    L = lib()
    nul = ptr(None, C.c_double)
    check(L.opem_radiation_currents(solver.ctx, at.size, ptr(at, C.c_int32), float(freq), nul, nul, float(limit),
                                    C.byref(n), nul, nul, nul, nul, C.byref(eta)), solver.ctx,
          "opem_radiation_currents")
    xyz = np.zeros((n.value, 3))
    J = np.zeros((n.value, 3), np.complex128)
    M = np.zeros((n.value, 3), np.complex128)
    area = np.zeros(n.value)
    check(L.opem_radiation_currents(solver.ctx, at.size, ptr(at, C.c_int32), float(freq),
                                    ptr(cplx_view(x), C.c_double), ptr(cplx_view(h), C.c_double), float(limit),
                                    C.byref(n), ptr(xyz, C.c_double), ptr(J.view(np.float64), C.c_double),
                                    ptr(M.view(np.float64), C.c_double), ptr(area, C.c_double), C.byref(eta)),
          solver.ctx, "opem_radiation_currents")
    return dict(xyz=xyz, J=J, M=M, area=area, eta=eta.value)


def far_field(solver, cur, center, beta, eta, pts, timing=False):
    """opem_far_field: complex (npts, 4) far fields of the currents at pts about center (GPU).  With
    ``timing`` also returns the kernels' milliseconds."""
    xyz = np.ascontiguousarray(cur["xyz"], np.float64)
    J = np.ascontiguousarray(cur["J"], np.complex128)
    M = np.ascontiguousarray(cur["M"], np.complex128)
    area = np.ascontiguousarray(cur["area"], np.float64)
    c = np.ascontiguousarray(center, np.float64)
    P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros((P.shape[0], 4), np.complex128)
    ms = C.c_double(0.0)
    check(lib().opem_far_field(solver.ctx, xyz.shape[0], ptr(xyz, C.c_double), ptr(J.view(np.float64), C.c_double),
                               ptr(M.view(np.float64), C.c_double), ptr(area, C.c_double), ptr(c, C.c_double),
                               float(beta), float(eta), P.shape[0], ptr(P, C.c_double),
                               ptr(out.view(np.float64), C.c_double), C.byref(ms)), solver.ctx, "opem_far_field")
    return (out, ms.value) if timing else out


class Sphere:
    """opem_pattern_sphere: the refined icosahedron of Sphere::create / refine / setMetrics /
    allocateAreasToPoints: ``pts`` (vertices, then triangle centres), ``point_area``, ``tris``, ``tri_area``,
    ``total_area``, ``nvert``."""

    def __init__(self, radius, center, refine):
        c = np.ascontiguousarray(center, np.float64)
        n, nt = C.c_int64(0), C.c_int64(0)
        L = lib()
        nul = ptr(None, C.c_double)
        rc = L.opem_pattern_sphere(float(radius), ptr(c, C.c_double), int(refine), C.byref(n), nul, nul,
                                   C.byref(nt), ptr(None, C.c_int32), nul)
        if rc != 0:
            raise ValueError(f"opem_pattern_sphere failed (rc={rc})")
        self.pts = np.zeros((n.value, 3))
        self.point_area = np.zeros(n.value)
        self.tris = np.zeros((nt.value, 3), np.int32)
        self.tri_area = np.zeros(nt.value)
        rc = L.opem_pattern_sphere(float(radius), ptr(c, C.c_double), int(refine), C.byref(n),
                                   ptr(self.pts, C.c_double), ptr(self.point_area, C.c_double), C.byref(nt),
                                   ptr(self.tris, C.c_int32), ptr(self.tri_area, C.c_double))
        if rc != 0:
            raise ValueError(f"opem_pattern_sphere failed (rc={rc})")
        self.radius, self.center, self.refine = float(radius), c, int(refine)
        self.nvert = int(self.tris.max()) + 1
        self.total_area = float(self.tri_area.sum())   # Sphere::getTotalArea (pattern.cpp:803-812)


def circle_points(radius, center, theta, phi, latitude, npoints):
    """opem_pattern_circle: Circle::create's points (npoints, 3)."""
    c = np.ascontiguousarray(center, np.float64)
    out = np.zeros((int(npoints), 3))
    rc = lib().opem_pattern_circle(float(radius), ptr(c, C.c_double), float(theta), float(phi), float(latitude),
                                   int(npoints), ptr(out, C.c_double))
    if rc != 0:
        raise ValueError(f"opem_pattern_circle failed (rc={rc})")
    return out


def metrics(F, point_area, total_area, accepted_power):
    """opem_pattern_metrics: radiated power, per-point gain and directivity, max gain / directivity in dB and the
    radiation efficiency."""
    F = np.ascontiguousarray(F, np.complex128)
    pa = np.ascontiguousarray(point_area, np.float64)
    acc = np.ascontiguousarray([complex(accepted_power)], np.complex128)
    prad = np.zeros(1, np.complex128)
    g = np.zeros(F.shape[0])
    d = np.zeros(F.shape[0])
    summ = np.zeros(3)
    rc = lib().opem_pattern_metrics(F.shape[0], ptr(F.view(np.float64), C.c_double), ptr(pa, C.c_double),
                                    float(total_area), ptr(acc.view(np.float64), C.c_double),
                                    ptr(prad.view(np.float64), C.c_double), ptr(g, C.c_double), ptr(d, C.c_double),
                                    ptr(summ, C.c_double))
    if rc != 0:
        raise ValueError(f"opem_pattern_metrics failed (rc={rc})")
    return dict(radiated_power=complex(prad[0]), gain=g, directivity=d, gain_db=summ[0], directivity_db=summ[1],
                efficiency=summ[2])


def accepted_power(Cp, Cm, Pz, sport):
    """opem_accepted_power for one frequency's splits (complex [nports][nports], [driving set][port])."""
    Cp = np.ascontiguousarray(Cp, np.complex128)
    Cm = np.ascontiguousarray(Cm, np.complex128)
    pz = np.ascontiguousarray([complex(Pz)], np.complex128)
    out = np.zeros(1, np.complex128)
    rc = lib().opem_accepted_power(Cp.shape[0], ptr(Cp.view(np.float64), C.c_double),
                                   ptr(Cm.view(np.float64), C.c_double), ptr(pz.view(np.float64), C.c_double),
                                   int(sport), ptr(out.view(np.float64), C.c_double))
    if rc != 0:
        raise ValueError(f"opem_accepted_power failed (rc={rc})")
    return complex(out[0])


def pattern(solver, x, h, freq, accepted, limit=0.05, refine=3, attrs=None):
    """fem3D::calculateRadiationCurrents + calculateRadiationPatterns for one solution (3D pattern): centre = the
    mesh bounding box's centre, radius = 1e5 wavelengths, the sphere refined ``refine`` times; returns the sphere,
    the fields at its points, the metrics and the currents."""
    cur = radiation_currents(solver, x, h, freq, limit, attrs)
    eta = cur["eta"]
    beta = 2 * np.pi * freq * MU0 / eta
    lam = eta / (MU0 * freq)
    xyz = solver.mesh.xyz
    center = 0.5 * (xyz.min(axis=0) + xyz.max(axis=0))
    sph = Sphere(1e5 * lam, center, refine)
    F = far_field(solver, cur, center, beta, eta, sph.pts)
    return dict(sphere=sph, fields=F, currents=cur, center=center, beta=beta, eta=eta,
                **metrics(F, sph.point_area, sph.total_area, accepted))
