"""CPU BASELINE (test infrastructure only): build and ctypes binding of ``oracle/cpu_lu.cpp``.

The library times the per-frequency E solve on the host cores with the GPU path's algorithm (nested-dissection
multifrontal sparse LU, FP64 GMRES on the lifted, port-eliminated system; the reference's
``solveComplexLinearSystem.c:100-209`` with its factorization-strength preconditioner, :133-168).  bench.py's
``cpu_baseline`` leg and the CPU tests are its only callers; it checks nothing and ships nothing.

    python -m oracle.cpu_baseline        # builds oracle/libopem_cpu.so
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libopem_cpu.so")
# the product's host analysis (ordering, assembly tree) is compiled in, so both sides factor the same fronts
SOURCES = [os.path.join(HERE, "cpu_lu.cpp")] + [os.path.join(ROOT, "openparem_amd", "csrc", f)
                                               for f in ("space.cpp", "nd_tables.cpp", "lu_symbolic.cpp")]


def build():
    """g++ -O3 with AVX2/FMA (present on the x86-64 hosts this runs on) and pthreads; writes oracle/libopem_cpu.so."""
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-mavx2", "-mfma", "-fcx-limited-range", "-std=c++17", "-fPIC", "-shared", "-pthread", *SOURCES, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (python -m oracle.cpu_baseline)")
        L = C.CDLL(LIB)
        L.opem_cpu_error.restype = C.c_char_p
        L.opem_cpu_lu_port_solve.restype = C.c_int
        _lib = L
    return _lib


def host_threads():
    """Threads the baseline uses: the CPUs this process may run on, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def port_solve(xyz, tets, order, rowptr, colind, vals, exclude, port_group, xdir, nthreads=None, restart=30,
               maxit=200, tol=1e-10):
    """Solve every excitation of one frequency point (see opem_cpu_lu_port_solve in cpu_lu.cpp).

    xyz (nv, 3), tets (nt, 4): the mesh the pattern (rowptr, colind) was built on at ND ``order``; vals: complex
    values on that pattern; exclude (n,) bool: Dirichlet in every excitation; port_group (n,) int: excitation whose
    driven port holds the DOF (-1: none); xdir (nexc, n) complex: Dirichlet values.  Returns x (nexc, n) and a dict
    of timings, flops, iterations and relative residuals."""
    L = lib()
    xyz = np.ascontiguousarray(xyz, np.float64)
    tets = np.ascontiguousarray(tets, np.int32)
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    v = np.ascontiguousarray(vals, np.complex128)
    ex = np.ascontiguousarray(exclude, np.uint8)
    pg = np.ascontiguousarray(port_group, np.int32)
    xd = np.ascontiguousarray(xdir, np.complex128)
    n = rp.size - 1
    nexc = xd.shape[0]
    nth = host_threads() if nthreads is None else int(nthreads)
    x = np.zeros((nexc, n), np.complex128)
    info = np.zeros(8 + 2 * nexc, np.float64)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    rc = L.opem_cpu_lu_port_solve(C.c_int64(xyz.shape[0]), P(xyz, C.c_double), C.c_int64(tets.shape[0]),
                                  P(tets, C.c_int32), C.c_int(order), C.c_int64(n), P(rp, C.c_int32),
                                  P(ci, C.c_int32), P(v.view(np.float64), C.c_double), P(ex, C.c_uint8),
                                  P(pg, C.c_int32), C.c_int(nexc), P(xd.view(np.float64), C.c_double), C.c_int(nth),
                                  C.c_int(restart), C.c_int(maxit), C.c_double(tol), P(x.view(np.float64), C.c_double),
                                  P(info, C.c_double))
    if rc != 0:
        raise RuntimeError(f"opem_cpu_lu_port_solve: {L.opem_cpu_error().decode()}")
    return x, dict(analysis_s=info[0], factor_s=info[1], root_factor_s=info[2], solve_s=info[3], flops=info[4],
                   zero_pivots=int(info[5]), threads=int(info[6]), factored_dofs=int(info[7]),
                   iters=[int(info[8 + 2 * e]) for e in range(nexc)],
                   rel_res=[float(info[9 + 2 * e]) for e in range(nexc)])


if __name__ == "__main__":
    print(build())
