"""Run the roofline kernels alone at the benchmark shape (for rocprofv3 --pmc), plus a device copy of known size.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m openparem_amd.profile_kernels [--restart 300] [--ns 3]
    rocprofv3 --pmc WRITE_SIZE -- python3 -m openparem_amd.profile_kernels [--restart 300] [--ns 3]

Kernels, in dispatch order after the setup: the s-step projection / update k_mdot_s / k_mupdate_s (FP32 basis,
128 systems x (restart/2 + 1) vectors, the mean basis of a GMRES(restart) cycle, ns = 3 new vectors per pass:
one warm-up and two timed launches each), the batched
operator (FP32, 64 frequencies x 2 excitations: one warm-up and two timed launches), the single-matrix complex
CSR SpMV k_csr_spmv in FP64 and then FP32 (one launch per copy of the matrix, then two timed launches), then two 1 GiB
device-to-device copies whose FETCH_SIZE / WRITE_SIZE calibrate the counters (a copy reads and writes
exactly 1 GiB).
"""
import argparse
import ctypes as C
import json

from . import workloads
from ._lib import lib
from .solver import Solver


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--restart", type=int, default=300)
    ap.add_argument("--ns", type=int, default=3, help="Krylov vectors per Gram-Schmidt pass")
    args = ap.parse_args()
    w = workloads.make("waveguide-sweep")
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    nf, R = len(w["freqs"]), len(s.ports)
    (tm, tu), (bm, bu) = s.krylov_bench("fp32", S=nf * R, k=args.restart // 2 + 1, nrep=2, ns=args.ns)
    ts, bs = s.spmv_bench("fp32", ngroups=nf, R=R, nrep=2)
    tc64, bc64 = s.csr_spmv_bench("fp64", nrep=2)
    tc32, bc32 = s.csr_spmv_bench("fp32", nrep=2)
    print(json.dumps({"k_mdot_s": {"ms": tm * 1e3, "alg_bytes": bm}, "k_mupdate_s": {"ms": tu * 1e3, "alg_bytes": bu},
                      "spmv_sweep": {"ms": ts * 1e3, "alg_bytes": bs},
                      "csr_spmv_fp64": {"ms": tc64 * 1e3, "alg_bytes": bc64},
                      "csr_spmv_fp32": {"ms": tc32 * 1e3, "alg_bytes": bc32}}), flush=True)
    hip = C.CDLL("libamdhip64.so")
    n = 1 << 30
    a, d = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(d), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(2):
        hip.hipMemcpy(d, a, C.c_size_t(n), 3)  # device to device: 1 GiB read + 1 GiB write
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(d)
    s.close()
    lib()


if __name__ == "__main__":
    main()
