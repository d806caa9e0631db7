"""Benchmark of the per-frequency solve path on one MI355X.

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--workload waveguide-sweep] [--no-cpu]

One step = one batched frequency sweep: every frequency x every port excitation of the workload solved
to a true relative residual of 1e-10 (FP32 GMRES inside FP64 iterative refinement, all systems of the
sweep batched), followed by the H projection, port projections and S extraction.  The metric is
frequency points per second (a point counts when its full S matrix is done).  Beside it the plain FP64
GPU path (FP64 everywhere, one frequency at a time) runs on a subset of the same frequencies, and the
largest |S_mixed - S_fp64| over that subset is reported.  One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _pmc_traffic_ratio(kernel, shape):
    """HBM traffic / algorithmic bytes of a kernel at a shape, as measured by rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes and recorded in profiles/roofline_pmc.md ("traffic_ratio <kernel> <shape> <ratio>"); None if absent."""
    path = os.path.join(ROOT, "profiles", "roofline_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_ratio" and f[1] == kernel and f[2] == shape:
            return float(f[3])
    return None


def _cpu_baseline(work, its_per_system, restart):
    """Oracle (numpy/scipy FP64, one thread) on one frequency of the same workload, bounded in time.

    Times A(omega), the vertex-patch preconditioner build, a bounded number of GMRES(restart) iterations and
    the H projection for one excitation; a frequency point is charged R excitations at its_per_system
    iterations (the plain FP64 GPU path's mean: the same preconditioner and restart, no polynomial).
    """
    from threadpoolctl import threadpool_limits

    from oracle import assemble as oasm
    from oracle import solve as oslv
    from oracle.pipeline import Problem
    from openparem_amd.ports import provider_for

    m, p = work["mesh"], work["order"]
    f = float(work["freqs"][len(work["freqs"]) // 2])
    with threadpool_limits(limits=1):
        pr = Problem(m, p)                                      # one-time setup, untimed (as on the GPU)
        prov = provider_for(m)
        modes = {a: prov(m, a, f) for a in pr.ports}
        rs = {a: np.sqrt(np.pi * f * oasm.MU0 / v) for a, v in work["boundaries"].items()}
        t0 = time.perf_counter()
        A = oasm.system_matrix(pr.space, f, {a: modes[a]["gamma"] for a in pr.ports}, rs_of_attr=rs,
                               Ke=pr.Ke, Me=pr.Me)
        t_a = time.perf_counter() - t0
        drv = pr.ports[0]
        tr = pr.traces[drv]
        xm = tr.project(modes[drv]["E"](tr.pts), fixed_zero=np.intersect1d(tr.port_dofs, pr.pec))
        At, b = oslv.excite(A, pr.pec, tr.port_dofs, xm)
        dmask = np.zeros(pr.space.ndof, bool)
        dmask[pr.pec] = True
        dmask[tr.port_dofs] = True
        t0 = time.perf_counter()
        M = oslv.VertexPatchASM(At, pr.space, dmask)
        t_pc = time.perf_counter() - t0
        n_it = 10
        t0 = time.perf_counter()
        oslv.gmres(At, b, M=M, restart=restart, rtol=1e-30, maxit=n_it)
        t_it = (time.perf_counter() - t0) / n_it
        P, Q = oasm.h_matrices(pr.space, f, Me=pr.Me)
        t0 = time.perf_counter()
        oslv.h_field(P, Q, b)
        t_h = time.perf_counter() - t0
    R = len(pr.ports)
    per_point = t_a + R * (t_pc + its_per_system * t_it + t_h)
    return dict(value=1.0 / per_point, unit="freq-points/s", cores=1, kind="port",
                sample=(f"oracle (numpy/scipy FP64, 1 thread) at {f / 1e9:.3g} GHz on the same mesh: A(omega) "
                        f"{t_a:.2f}s, patch preconditioner {t_pc:.2f}s, {n_it} GMRES({restart}) iterations "
                        f"({t_it * 1e3:.0f} ms each), H projection {t_h:.2f}s; a point = {R} excitations x "
                        f"{its_per_system:.0f} iterations (the plain FP64 GPU path's mean)"))


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="waveguide-sweep")
    ap.add_argument("--fp64-points", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--restart", type=int, default=300)
    ap.add_argument("--coarse", type=int, default=0)
    ap.add_argument("--share", type=int, default=16, help="adjacent frequencies sharing one set of patch inverses")
    ap.add_argument("--poly", type=int, default=0, help="GMRES-polynomial preconditioner degree (0: off)")
    ap.add_argument("--arnoldi-step", type=int, default=3, help="Krylov vectors per Gram-Schmidt pass (1..4)")
    ap.add_argument("--warm-start", type=int, default=1,
                    help="1: even-indexed frequencies first, odd ones from their neighbours' interpolated solutions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the systems of a sweep are "
              f"independent and would be sharded across GPUs by frequency, which this build does not run)",
              file=sys.stderr)
        return 2

    from openparem_amd import workloads
    from openparem_amd.measure import assembly_c2
    from openparem_amd.solver import Solver

    work = workloads.make(args.workload)
    freqs = work["freqs"]
    s = Solver(work["mesh"], work["order"], boundaries=work["boundaries"])
    nports = len(s.ports)
    opts = dict(restart=args.restart, tol=1e-10, coarse=args.coarse, maxit=20000)

    for _ in range(args.warmup):
        s.sweep(freqs, precision="mixed", batch=len(freqs), orth_passes=1, precond_share=args.share,
                poly_degree=args.poly, arnoldi_step=args.arnoldi_step, warm_start=args.warm_start,
                **opts)
    _progress(f"{args.warmup} warm-up sweep(s) done ({s.ndof} DOF, {len(freqs)} frequencies)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = s.sweep(freqs, precision="mixed", batch=len(freqs), orth_passes=1, precond_share=args.share,
                    poly_degree=args.poly, arnoldi_step=args.arnoldi_step, warm_start=args.warm_start,
                    **opts)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    done = int(np.all(r["converged"].astype(bool), axis=1).sum())
    value = args.steps * done / elapsed

    _progress(f"{args.steps} timed sweep(s): {elapsed:.2f} s")
    # plain FP64 GPU path on an evenly spaced subset
    idx = np.unique(np.linspace(0, len(freqs) - 1, args.fp64_points).round().astype(int))
    t2 = time.perf_counter()
    r64 = s.sweep(freqs[idx], precision="fp64", batch=1, orth_passes=2, **opts)
    t3 = time.perf_counter()
    fp64_value = len(idx) / (t3 - t2)
    dS = float(np.abs(r["S"][idx] - r64["S"]).max())
    _progress(f"plain FP64 subset: {len(idx)} points in {t3 - t2:.2f} s")

    # Kernel rooflines at the batched shape of the sweep (all frequencies x excitations, FP32 inner solve):
    # the Gram-Schmidt projection (the largest share of the solve) with a basis of m/2 + 1 vectors (the mean
    # over a restart cycle) and the sweep's vectors per pass, the update, and the batched operator.
    S = len(freqs) * nports
    kmean = args.restart // 2 + 1
    nsb = max(1, args.arnoldi_step) if not args.poly else 1
    (t_mdot, t_mupd), (b_mdot, b_mupd) = s.krylov_bench("fp32", S=S, k=kmean, nrep=5, ns=nsb)
    t_spmv, b_spmv = s.spmv_bench("fp32", ngroups=len(freqs), R=nports, nrep=10)
    # The single-matrix complex CSR SpMV (explicit values, one vector) on the same pattern: the MatMult of a
    # one-system solve, HBM-bound, each launch on its own copy of the matrix.  Rows grouped as the pattern allows
    # (ND2: pairs of rows with one column list) and one row at a time.
    csr = {(p_, g_): s.csr_spmv_bench(p_, row_group=g_, nrep=50) for p_ in ("fp64", "fp32") for g_ in (0, 1)}
    csr_n, csr_nnz = s.ndof, s.nnz
    ratio_mdot = _pmc_traffic_ratio("k_mdot_s" if nsb > 1 else "k_mdot", f"S={S},k={kmean},ns={nsb},n={s.ndof}")
    its_mean = float(r["iters"].mean())
    s.close()
    _progress("kernel timings done")

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(work, float(r64["iters"].mean()), args.restart)
    if cpu is not None:
        _progress("CPU baseline done")
    asm = None if args.no_assembly else {f"nd{p_}": assembly_c2(p_, 1_000_000, HBM_PEAK_GBS) for p_ in (1, 2, 3)}

    line = {
        "metric": "freq-points/sec at |dS|<1e-6 (mixed+batched vs FP64 GPU); SpMV HBM GB/s",
        "value": value,
        "unit": "freq-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c-fp32 inner / c-fp64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: batched frequency sweep, {len(freqs)} frequencies "
                        f"{freqs[0] / 1e9:.3g}-{freqs[-1] / 1e9:.3g} GHz x {nports} port excitations = "
                        f"{len(freqs) * nports} systems on one pattern, ND{work['order']}, "
                        f"{s.ndof} DOF, nnz {s.nnz}",
            "order": work["order"], "ndof": s.ndof, "nnz": s.nnz, "systems": len(freqs) * nports,
            "precond": f"vertex-patch additive Schwarz (one inverse set per {args.share} adjacent frequencies)"
                       + (f", GMRES polynomial degree {args.poly}" if args.poly else "") + f", GMRES({args.restart})"
                       + (f", {args.arnoldi_step} Krylov vectors per Gram-Schmidt pass" if args.arnoldi_step > 1 else "")
                       + (", even-indexed frequencies first, odd ones from their neighbours' interpolated solutions"
                          if args.warm_start else ""),
        },
        "points_converged": done,
        "iters_mean": its_mean,
        "operator_applications_mean": its_mean * max(1, args.poly) + args.poly,
        "fp64_plain": {"value": fp64_value, "unit": "freq-points/s", "points": int(len(idx)),
                       "iters_mean": float(r64["iters"].mean())},
        "speedup_vs_fp64_plain": value / fp64_value,
        "max_abs_dS_vs_fp64": dS,
        "roofline": {"kernel": (f"k_mdot_s (Gram-Schmidt projection C = V^H [z_1..z_{nsb}], FP32 basis, {S} systems x "
                                f"{kmean} vectors)" if nsb > 1 else
                                f"k_mdot (Gram-Schmidt projection h = V^H w, FP32 basis, {S} systems x {kmean} vectors)"),
                     "bound": "hbm", "achieved": b_mdot / t_mdot / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_mdot / t_mdot / 1e9 / HBM_PEAK_GBS,
                     "traffic": (None if ratio_mdot is None else ratio_mdot * b_mdot),
                     "alg_bytes_per_launch": b_mdot, "launch_seconds": t_mdot},
        "kernels": {
            ("k_mupdate_s" if nsb > 1 else "k_mupdate"): {"GB/s": b_mupd / t_mupd / 1e9, "alg_bytes": b_mupd,
                                                          "ms": t_mupd * 1e3},
            "k_spmv_sweep_p (batched operator, one pass of Kt/Mr over each 32-system tile)": {
                "GB/s": b_spmv / t_spmv / 1e9, "alg_bytes": b_spmv, "ms": t_spmv * 1e3},
        },
        "spmv_csr": {
            "kernel": "k_csr_spmv (explicit complex CSR y = A x, one matrix and one vector, int32 indices)",
            "note": "frac: the bytes the kernel must move / time / 8 TB/s; canonical_frac: the plain CSR byte count "
                    "nnz (s_v + 4) + 4 (N + 1) + 2 N s_v / time / 8 TB/s (the grouped kernel reads each column "
                    "index once per row pair, so its own count is lower)",
            **{f"{p_}_{'grouped' if g_ == 0 else 'rows'}": {
                "GB/s": b_ / t_ / 1e9, "frac": b_ / t_ / 1e9 / HBM_PEAK_GBS, "alg_bytes": b_, "us": t_ * 1e6,
                "canonical_frac": (csr_nnz * ((16 if p_ == "fp64" else 8) + 4) + 4 * (csr_n + 1)
                                   + 2 * csr_n * (16 if p_ == "fp64" else 8)) / t_ / 1e9 / HBM_PEAK_GBS}
               for (p_, g_), (t_, b_) in csr.items()}},
        "cpu_baseline": cpu,
        "assembly_c2": asm,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
