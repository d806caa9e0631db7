"""S-parameter post-processing over the C ABI (host-only calls of libopem.so, no GPU context): the counterparts of
the reference's ``Result`` / ``ResultDatabase`` conversions and files (``src/OpenParEM3D/results.cpp``) and of
``process3D``'s regression comparator (``src/OpenParEM3D/process3D.cpp``).

S arrays are complex ``[nfreq][nports][nports]`` (or one ``[nports][nports]`` matrix), the layout ``Solver.sweep``
returns, which is the reference's ``Result::S`` layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import cplx_view, lib, ptr


class ResultsError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise ResultsError(f"{what} failed (rc={rc}): {lib().opem_results_error().decode()}")


def _square(S):
    S = np.ascontiguousarray(S, np.complex128)
    if S.ndim != 2 or S.shape[0] != S.shape[1]:
        raise ValueError("expected one square matrix")
    return S


def s2z(S, Zo):
    """Result::S2Z: Z = k (I + S)(I - S)^-1 k with k = diag(sqrt(Zo))."""
    S = _square(S)
    n = S.shape[0]
    zo = np.ascontiguousarray(np.broadcast_to(np.asarray(Zo, np.complex128), (n,)))
    Z = np.empty_like(S)
    _check(lib().opem_s2z(n, ptr(cplx_view(S), C.c_double), ptr(cplx_view(zo), C.c_double),
                          ptr(Z.view(np.float64), C.c_double)), "opem_s2z")
    return Z


def z2s(Z, Zo):
    """Result::Z2S: S = (Z k^-1 + k)^-1 (Z k^-1 - k)."""
    Z = _square(Z)
    n = Z.shape[0]
    zo = np.ascontiguousarray(np.broadcast_to(np.asarray(Zo, np.complex128), (n,)))
    S = np.empty_like(Z)
    _check(lib().opem_z2s(n, ptr(cplx_view(Z), C.c_double), ptr(cplx_view(zo), C.c_double),
                          ptr(S.view(np.float64), C.c_double)), "opem_z2s")
    return S


def renormalize(S, Zo, Zref):
    """Result::renormalize: S referred to the port impedances Zo, re-referred to Zref at every port."""
    S = _square(S)
    n = S.shape[0]
    zo = np.ascontiguousarray(np.broadcast_to(np.asarray(Zo, np.complex128), (n,)))
    out = np.empty_like(S)
    zr = complex(Zref)
    _check(lib().opem_renormalize(n, ptr(cplx_view(S), C.c_double), ptr(cplx_view(zo), C.c_double), zr.real,
                                  zr.imag, ptr(out.view(np.float64), C.c_double)), "opem_renormalize")
    return out


def _meta(fmt, unit, zref, nets, version):
    keep = [n.encode() for n in nets] if nets is not None else None
    arr = (C.c_char_p * len(keep))(*keep) if keep is not None else None
    m = _lib.SparamFile(version=version.encode() if version else None, format=fmt.encode(), unit=unit.encode(),
                        reference_impedance=float(zref),
                        nets=C.cast(arr, C.POINTER(C.c_char_p)) if arr is not None else None)
    return m, (keep, arr)


def _stack(freqs, S):
    f = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
    S = np.ascontiguousarray(S, np.complex128).reshape(f.size, -1)
    n = int(round(np.sqrt(S.shape[1])))
    if n * n != S.shape[1]:
        raise ValueError("S must be [nfreq][nports][nports]")
    return f, S, n


def write_csv(path, freqs, S, fmt="DB", unit="GHz", zref=50.0, nets=None, version=None):
    """ResultDatabase::saveCSV.  zref = 0 writes "not renormalized"."""
    f, S, n = _stack(freqs, S)
    m, keep = _meta(fmt, unit, zref, nets, version)
    _check(lib().opem_write_csv(str(path).encode(), C.byref(m), n, f.size, ptr(f, C.c_double),
                                ptr(S.view(np.float64), C.c_double)), "opem_write_csv")


def write_touchstone(path, freqs, S, fmt="DB", unit="GHz", zref=50.0, nets=None, version=None):
    """ResultDatabase::saveTouchstone (version 1.1)."""
    f, S, n = _stack(freqs, S)
    m, keep = _meta(fmt, unit, zref, nets, version)
    _check(lib().opem_write_touchstone(str(path).encode(), C.byref(m), n, f.size, ptr(f, C.c_double),
                                       ptr(S.view(np.float64), C.c_double)), "opem_write_touchstone")


def save_as_test(path, casename, freqs, S, casenumber=0):
    """ResultDatabase::save_as_test: appends prototype cases; returns the next case number."""
    f, S, n = _stack(freqs, S)
    c = C.c_int(casenumber)
    _check(lib().opem_save_as_test(str(path).encode(), casename.encode(), n, f.size, ptr(f, C.c_double),
                                   ptr(S.view(np.float64), C.c_double), C.byref(c)), "opem_save_as_test")
    return c.value


def process_tests(cases_path, results_csv, report_path):
    """process3D: evaluate test cases against a DB-format results CSV; returns (ncases, nevaluated, npassed)."""
    nc, ne, npas = C.c_int(), C.c_int(), C.c_int()
    _check(lib().opem_process_tests(str(cases_path).encode(), str(results_csv).encode(), str(report_path).encode(),
                                    C.byref(nc), C.byref(ne), C.byref(npas)), "opem_process_tests")
    return nc.value, ne.value, npas.value


def read_touchstone(path):
    """Touchstone 1.1 reader (option line "# <unit> S <RI|MA|DB> R <z>"): returns (freqs in Hz, S [nfreq][n][n],
    reference impedance).  One- and two-port data are stored column by column, larger ones row by row."""
    scale = {"HZ": 1.0, "KHZ": 1e3, "MHZ": 1e6, "GHZ": 1e9}
    unit, fmt, zref = "GHZ", "MA", 50.0
    nums = []
    n = None
    for line in open(path):
        if "!" in line:
            head, _, comment = line.partition("!")
            c = comment.strip().split()
            if len(c) >= 2 and c[1] == "S-parameter" and c[0].endswith("-port"):
                n = int(c[0].split("-")[0])
            line = head
        line = line.strip()
        if not line:
            continue
        if line.startswith("#"):
            t = line[1:].upper().split()
            for k, w in enumerate(t):
                if w in scale:
                    unit = w
                elif w in ("RI", "MA", "DB"):
                    fmt = w
                elif w == "R" and k + 1 < len(t):
                    zref = float(t[k + 1])
            continue
        nums.extend(float(v) for v in line.split())
    if n is None:
        raise ValueError("port count unknown: no '! N-port S-parameter data' line")
    per = 1 + 2 * n * n
    a = np.asarray(nums).reshape(-1, per)
    f = a[:, 0] * scale[unit]
    p, q = a[:, 1::2], a[:, 2::2]
    if fmt == "RI":
        v = p + 1j * q
    elif fmt == "MA":
        v = p * np.exp(1j * np.deg2rad(q))
    else:
        v = 10 ** (p / 20) * np.exp(1j * np.deg2rad(q))
    v = v.reshape(-1, n, n)
    S = np.transpose(v, (0, 2, 1)) if n < 3 else v
    return f, S.copy(), zref
