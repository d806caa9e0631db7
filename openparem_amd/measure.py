"""Measurements on the microstrip configurations and the mixed set (one JSON line per step, printed as it ends).

    python -m openparem_amd.measure c3   # 1.84 M ND2 DOF at 5 GHz: mixed (FP32 inner + FP64 refinement) vs plain FP64
    python -m openparem_amd.measure c4   # 0.48 M ND2 DOF, 64 frequencies 1-20 GHz x 2 ports, in four batches of 16,
                                         # then plain FP64 at four of the frequencies

    python -m openparem_amd.measure c2   # assembly kernel alone: element matrices of 1M random tets, ND1-ND3,
                                         # FP64 / FP32 / BF16-scaled MFMA

    python -m openparem_amd.measure kernels  # the roofline kernels alone at the guide benchmark's shape (profiling)

    python -m openparem_amd.measure far_field  # the far-field kernels at the config-4 radiation lid's size

    python -m openparem_amd.measure c5   # the seeded mixed set of workloads.c5_instances(): one batched mixed
                                         # sweep per instance, plain FP64 at its middle frequency

All use vertex-patch additive Schwarz (element-block Jacobi at ND3), GMRES(300) to a true residual of 1e-10 and at
most 8000 iterations per system.  The mixed solves use three Krylov vectors per Gram-Schmidt pass; the plain FP64
solves take one frequency at a time with two classical Gram-Schmidt passes (the comparison path of bench.py).
"""
import json
import sys
import time

import numpy as np

from . import workloads
from .solver import Solver, lu_stats

OPTS = dict(restart=300, tol=1e-10, maxit=8000)


def _emit(d):
    print(json.dumps(d), flush=True)


def assembly_c2(order, n_elem, hbm_peak_gbs=8000.0, nrep=5, n_err=20000):
    """Assembly kernel alone: element matrices of n_elem random tets (jittered Kuhn tets of the unit cube, aspect
    ratio <= 8, er in [1, 10], tan delta in [0, 0.02]) on MFMA in FP64, FP32 and BF16-scaled.  Errors of FP32 and BF16 are against the FP64 kernel's matrices on the first n_err elements (the
    FP64 kernel matches the FP64 oracle to 1e-12, tests/test_gpu_parity.py::test_element_matrices).  The kernel
    writes 2 nd^2 values per element and is bound by those HBM writes at every precision."""
    from .mesh import random_tets
    from .solver import Device

    v, er, td = random_tets(n_elem, seed=7)
    k0 = 2 * np.pi * 5e9 / 299792458.0 * 1e-2
    nd = order * (order + 2) * (order + 3) // 2
    out = {}
    dev = Device()
    try:
        ref = None
        # dense MFMA peaks of MI355X, TFLOP/s
        for prec, name, osz, peak in ((0, "fp64", 8, 78.6), (1, "fp32", 4, 157.3), (2, "bf16", 4, 2500.0)):
            _, t = dev.element_matrices(order, prec, v, er, td, k0, nrep=nrep, want=False)
            got, _ = dev.element_matrices(order, prec, v[:n_err], er[:n_err], td[:n_err], k0, nrep=1)
            if ref is None:
                ref = got
                rn = np.linalg.norm(ref.reshape(n_err, -1), axis=1)
            err = np.linalg.norm((got - ref).reshape(n_err, -1), axis=1) / rn
            b = n_elem * (12 * 8 + 2 * 8 + 2 * nd * nd * osz)
            f = n_elem * (2 * 16 + 2 * 8) * nd * nd
            out[name] = {"ms": t * 1e3, "GB/s": b / t / 1e9, "TFLOP/s": f / t / 1e12, "frac_hbm": b / t / 1e9 / hbm_peak_gbs,
                         "frac_mfma": f / t / 1e12 / peak,
                         "rel_err_max_vs_fp64": float(err.max()), "rel_err_p99_vs_fp64": float(np.quantile(err, 0.99))}
    finally:
        dev.close()
    return {"elements": n_elem, "order": order, "nd": nd, "bytes_per_element": "12*8 + 2*8 + 2*nd^2*s_out",
            "flops_per_element": "2*(16 + 8)*nd^2 (MFMA contraction of 16 + 8 padded geometric factors)", **out}


def c2():
    for order in (1, 2, 3):
        t = time.perf_counter()
        r = assembly_c2(order, 1_000_000)
        r["wall_s"] = round(time.perf_counter() - t, 1)
        _emit(r)


def c3():
    w = workloads.make("microstrip-c3")
    t = time.perf_counter()
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    _emit({"workload": "microstrip-c3", "setup_s": round(time.perf_counter() - t, 2), "ndof": s.ndof, "nnz": s.nnz})
    f = w["freqs"]
    try:
        for rep in range(2):  # the second solve reuses the solver workspace (the first allocates it)
            t = time.perf_counter()
            r = s.sweep(f, precision="mixed", batch=1, orth_passes=1, arnoldi_step=3, **OPTS)
            _emit({"precision": "mixed", "rep": rep, "s": round(time.perf_counter() - t, 3),
                   "iters": r["iters"].tolist(), "converged": r["converged"].tolist()})
        t = time.perf_counter()
        r64 = s.sweep(f, precision="fp64", batch=1, orth_passes=2, **OPTS)
        _emit({"precision": "fp64", "s": round(time.perf_counter() - t, 3), "iters": r64["iters"].tolist(),
               "converged": r64["converged"].tolist(), "max_abs_dS": float(np.abs(r["S"] - r64["S"]).max()),
               "S21_dB": float(20 * np.log10(abs(r64["S"][0, 1, 0])))})
    finally:
        s.close()


def c4():
    w = workloads.make("microstrip-c4")
    t = time.perf_counter()
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    _emit({"workload": "microstrip-c4", "setup_s": round(time.perf_counter() - t, 2), "ndof": s.ndof, "nnz": s.nnz})
    f = w["freqs"]
    try:
        S, tot, done = [], 0.0, 0
        for c in range(4):
            fc = f[c * 16:(c + 1) * 16]
            t = time.perf_counter()
            # 300 MHz spacing: one inverse set per 4 frequencies spans 0.9 GHz
            r = s.sweep(fc, precision="mixed", batch=16, orth_passes=1, precond_share=4, arnoldi_step=3, **OPTS)
            dt = time.perf_counter() - t
            tot += dt
            done += int(np.all(r["converged"].astype(bool), axis=1).sum())
            S.append(r["S"])
            _emit({"batch": c, "f_GHz": [round(fc[0] / 1e9, 2), round(fc[-1] / 1e9, 2)], "s": round(dt, 2),
                   "iters_mean": float(r["iters"].mean()), "iters_max": int(r["iters"].max()),
                   "points_converged": int(np.all(r["converged"].astype(bool), axis=1).sum())})
        _emit({"mixed_total_s": round(tot, 2), "points_converged": done, "freq_points_per_s": round(done / tot, 3)})
        S = np.concatenate(S)
        for i in (0, 21, 42, 63):
            t = time.perf_counter()
            r64 = s.sweep(f[i:i + 1], precision="fp64", batch=1, orth_passes=2, **OPTS)
            _emit({"fp64_f_GHz": round(f[i] / 1e9, 2), "s": round(time.perf_counter() - t, 2),
                   "iters": r64["iters"].tolist(), "converged": r64["converged"].tolist(),
                   "max_abs_dS": float(np.abs(S[i] - r64["S"][0]).max())})
    finally:
        s.close()


def c4lu(share=1, fp64_points=(0, 21, 42, 63)):
    """Config 4 with the sparse-LU preconditioner: the 64-frequency x 2-port microstrip sweep (0.48 M ND2 DOF,
    1-20 GHz) in mixed precision (FP32 factors, FP32 GMRES inside FP64 refinement), batched as memory allows,
    `share` adjacent frequencies per factor; then plain FP64 (FP64 factor of each frequency, one at a time) at
    `fp64_points`."""
    import os
    os.environ.setdefault("OPEM_LU_STATS", "1")
    w = workloads.make("microstrip-c4")
    t = time.perf_counter()
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    _emit({"workload": "microstrip-c4", "setup_s": round(time.perf_counter() - t, 2), "ndof": s.ndof, "nnz": s.nnz})
    f = w["freqs"]
    try:
        for rep in range(2):
            t = time.perf_counter()
            r = s.sweep(f, precision="mixed", restart=30, tol=1e-10, maxit=600, precond="lu", precond_share=share,
                        orth_passes=1)
            dt = time.perf_counter() - t
            done = int(np.all(r["converged"].astype(bool), axis=1).sum())
            _emit({"rep": rep, "share": share, "s": round(dt, 2), "points_converged": done,
                   "freq_points_per_s": round(done / dt, 3), "iters_mean": float(r["iters"].mean()),
                   "iters_max": int(r["iters"].max()), "seconds": r["seconds"]})
        for i in fp64_points:
            t = time.perf_counter()
            r64 = s.sweep(f[i:i + 1], precision="fp64", batch=1, restart=30, tol=1e-10, maxit=600, precond="lu")
            _emit({"fp64_f_GHz": round(f[i] / 1e9, 2), "s": round(time.perf_counter() - t, 2),
                   "iters": r64["iters"].tolist(), "converged": r64["converged"].tolist(),
                   "max_abs_dS": float(np.abs(r["S"][i] - r64["S"][0]).max())})
    finally:
        s.close()


def lu_pmc(nfreq=8, share=4):
    """The sparse-LU kernels of the config-4 sweep at a counter-friendly size, for rocprofv3 --pmc passes: one mixed
    sweep of the first `nfreq` frequencies of the 64-point plan (FP32 factors of the mean matrix of each `share`
    adjacent frequencies) on a fresh context, then its opem_lu_stats totals (trailing-update flops, apply matrix
    bytes, launch counts) as one JSON line, then two 1 GiB device-to-device copies whose FETCH_SIZE / WRITE_SIZE
    calibrate the counters."""
    import os

    import torch

    os.environ["OPEM_LU_TIMING"] = "1"  # launch counts and flops are kept only with event timing on
    w = workloads.make("microstrip-c4")
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    try:
        r = s.sweep(w["freqs"][:nfreq], precision="mixed", restart=30, tol=1e-10, maxit=600, precond="lu",
                    precond_share=share, orth_passes=1)
        st = lu_stats(s)
        _emit({"lu_pmc": {"nfreq": nfreq, "share": share, "converged": int(r["converged"].sum()),
                          "gemm_flops": st["gemm_flops"], "gemm_launches": st["gemm_launches"],
                          "apply_bytes": st["apply_bytes"], "apply_launches": st["apply_launches"]}})
    finally:
        s.close()
    a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda:0")
    d = torch.empty_like(a)
    for _ in range(2):
        d.copy_(a)
    torch.cuda.synchronize()


def c5():
    tot_t, tot_pts, tot_f = 0.0, 0, 0
    for i, w in enumerate(workloads.c5_instances()):
        s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
        f = w["freqs"]
        try:
            # one inverse set per adjacent frequencies spanning at most about 1 GHz
            share = int(max(1, min(16, np.floor(1e9 / (f[1] - f[0]) + 1e-9)))) if len(f) > 1 else 1
            t = time.perf_counter()
            r = s.sweep(f, precision="mixed", batch=len(f), orth_passes=1, precond_share=share, arnoldi_step=3,
                        **OPTS)
            dt = time.perf_counter() - t
            done = int(np.all(r["converged"].astype(bool), axis=1).sum())
            mid = len(f) // 2
            t = time.perf_counter()
            r64 = s.sweep(f[mid:mid + 1], precision="fp64", batch=1, orth_passes=2, **OPTS)
            t64 = time.perf_counter() - t
            tot_t += dt
            tot_pts += done
            tot_f += len(f)
            _emit({"instance": i, "family": w["family"], "ports": len(s.ports), "order": w["order"], "ndof": s.ndof,
                   "freqs": len(f),
                   "share": share, "s": round(dt, 2), "points_converged": done,
                   "freq_points_per_s": round(done / dt, 3), "iters_mean": float(r["iters"].mean()),
                   "iters_max": int(r["iters"].max()), "fp64_mid_s": round(t64, 2),
                   "fp64_mid_iters": r64["iters"].tolist(), "fp64_mid_converged": r64["converged"].tolist(),
                   "max_abs_dS_mid": float(np.abs(r["S"][mid] - r64["S"][0]).max())})
        finally:
            s.close()
    _emit({"instances": i + 1, "points": tot_f, "points_converged": tot_pts, "mixed_s": round(tot_t, 2),
           "aggregate_freq_points_per_s": round(tot_pts / tot_t, 3)})


def kernels(restart=300, ns=3):
    """The roofline kernels alone at the guide benchmark's shape, for rocprofv3 --pmc and --kernel-trace runs.

    In dispatch order after the set-up: the s-step projection / update k_mdot_s / k_mupdate_s (FP32 basis, 128
    systems x (restart/2 + 1) vectors, the mean basis of a GMRES(restart) cycle, ns new vectors per pass: one
    warm-up and two timed launches each), the batched operator (FP32, 64 frequencies x 2 excitations: one warm-up
    and two timed launches), the single-matrix complex CSR SpMV k_csr_spmv in FP64 and then FP32 (one launch per
    copy of the matrix, then two timed launches), then two 1 GiB device-to-device copies whose FETCH_SIZE /
    WRITE_SIZE calibrate the counters (a copy reads and writes exactly 1 GiB)."""
    import torch

    w = workloads.make("waveguide-sweep")
    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    try:
        nf, R = len(w["freqs"]), len(s.ports)
        (tm, tu), (bm, bu) = s.krylov_bench("fp32", S=nf * R, k=restart // 2 + 1, nrep=2, ns=ns)
        ts, bs = s.spmv_bench("fp32", ngroups=nf, R=R, nrep=2)
        tc64, bc64 = s.csr_spmv_bench("fp64", nrep=2)
        tc32, bc32 = s.csr_spmv_bench("fp32", nrep=2)
        _emit({"k_mdot_s": {"ms": tm * 1e3, "alg_bytes": bm}, "k_mupdate_s": {"ms": tu * 1e3, "alg_bytes": bu},
               "spmv_sweep": {"ms": ts * 1e3, "alg_bytes": bs},
               "csr_spmv_fp64": {"ms": tc64 * 1e3, "alg_bytes": bc64},
               "csr_spmv_fp32": {"ms": tc32 * 1e3, "alg_bytes": bc32}})
    finally:
        s.close()
    a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda:0")
    d = torch.empty_like(a)
    for _ in range(2):
        d.copy_(a)  # device to device: 1 GiB read + 1 GiB write
    torch.cuda.synchronize()


def far_field(solver=None, freq=20e9, limit=0.05, refine=4, nrep=5):
    """The far-field kernels (opem_far_field) at the size of the config-4 microstrip's radiation lid: its currents
    at ``freq`` split to ``limit`` wavelengths (synthetic values at the real positions and areas) on the pattern
    sphere refined ``refine`` times; FP64, bound by the vector FP64 rate (78.6 TFLOP/s: 53 flops per current-point
    pair plus one sincos, which is not counted)."""
    from . import pattern as pat
    own = solver is None
    if own:
        w = workloads.make("microstrip-c4")
        solver = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
    try:
        eta = np.sqrt(4e-7 * np.pi / 8.8541878176e-12)
        n = solver.ndof
        z = np.zeros(n, np.complex128)
        cur = pat.radiation_currents(solver, z, z, freq, limit)
        rng = np.random.default_rng(0)
        nc = cur["area"].size
        cur["J"] = rng.normal(size=(nc, 3)) + 1j * rng.normal(size=(nc, 3))
        cur["M"] = rng.normal(size=(nc, 3)) + 1j * rng.normal(size=(nc, 3))
        beta = 2 * np.pi * freq * 4e-7 * np.pi / eta
        xyz = solver.mesh.xyz
        center = 0.5 * (xyz.min(axis=0) + xyz.max(axis=0))
        sph = pat.Sphere(1e5 * 2 * np.pi / beta, center, refine)
        pat.far_field(solver, cur, center, beta, eta, sph.pts)
        ms = min(pat.far_field(solver, cur, center, beta, eta, sph.pts, timing=True)[1] for _ in range(nrep))
        pairs = nc * sph.pts.shape[0]
        d = dict(kernel="k_far_field_part + k_far_field_sum (FP64, current-point pairs staged through LDS)",
                 currents=int(nc), points=int(sph.pts.shape[0]), ms=ms, pairs_per_s=pairs / (ms * 1e-3),
                 tflops=pairs * 53 / (ms * 1e-3) / 1e12, peak_tflops=78.6,
                 frac=pairs * 53 / (ms * 1e-3) / 1e12 / 78.6)
        return d
    finally:
        if own:
            solver.close()


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "c3"
    if which == "lu_pmc":
        lu_pmc()
    elif which == "c4lu":
        c4lu(share=int(sys.argv[2]) if len(sys.argv) > 2 else 1)
    else:
        {"c2": c2, "c3": c3, "c4": c4, "c5": c5, "kernels": kernels, "far_field": lambda: _emit(far_field())}[which]()
