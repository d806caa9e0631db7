"""Run the batched SpMV benchmark kernel alone (for rocprofv3 --pmc) plus a 1 GiB device copy of known size.

    rocprofv3 --pmc FETCH_SIZE -- python -m openparem_amd.profile_spmv
    rocprofv3 --pmc WRITE_SIZE -- python -m openparem_amd.profile_spmv
"""
import ctypes as C

from . import workloads
from ._lib import lib
from .solver import Solver


def main():
    w = workloads.make("waveguide-sweep")
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    t, b = s.spmv_bench("fp32", ngroups=len(w["freqs"]), R=2, nrep=3)
    print(f"spmv fp32 64x2: {t * 1e3:.3f} ms/launch, algorithmic {b / 1e9:.3f} GB, {b / t / 1e9:.1f} GB/s", flush=True)
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
