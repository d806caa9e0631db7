"""Benchmark of the per-frequency solve path on one MI355X.

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--workload microstrip-c4] [--no-cpu] [--no-assembly]

One step = one batched frequency sweep of the workload (default: the shielded-microstrip box at 0.48 M ND2 DOF,
64 frequencies over 1-20 GHz x 2 port excitations = 128 systems on one pattern): every system solved to a true
relative residual of 1e-10 with FP32 GMRES inside FP64 iterative refinement, preconditioned by a sparse LU factor
(FP32, one per four adjacent frequencies), then the H projection, port projections and S extraction.  The plain
FP64 GPU path (FP64 LU factor of every frequency, one frequency at a time, FP64 GMRES) then solves every frequency
of the sweep; a frequency point counts towards the metric only when both paths converged and the mixed S is within
1e-6 of the FP64 S.  One JSON line is printed.

``--workload uneven-c5`` runs the uneven set instead (16 instances, ND1-3, 2-4 ports, 50 k - 1.5 M DOF, one after
another; see ``_uneven``).
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
    """HBM traffic of a kernel at a shape per unit of its algorithmic work (bytes per algorithmic byte, or per flop
    for an MFMA-bound kernel), as measured by rocprofv3 FETCH_SIZE/WRITE_SIZE passes and recorded in
    profiles/roofline_pmc.md ("traffic_ratio <kernel> <shape> <ratio>"); None if absent."""
    path = os.path.join(ROOT, "profiles", "roofline_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_ratio" and f[1] == kernel and f[2] == shape:
            return float(f[3])
    return None


def _cpu_baseline(s, work, fi):
    """CPU baseline: one frequency point of the workload (frequency index ``fi``, both excitations) solved on the host
    cores by oracle/cpu_lu.cpp, the same algorithm in C++ with std::thread (nested-dissection multifrontal LU in
    FP64, GMRES on the lifted, port-eliminated system) on the same matrix, with the Dirichlet values of the GPU's
    FP64 solution of that point; the CPU solution is compared with the GPU's.  value = 1 / (numeric factorization +
    port-root factorizations + solves) in freq-points/s; the per-mesh analysis (ordering, tree) is reported apart,
    as the GPU's timed sweeps exclude it too."""
    from openparem_amd.ports import provider_for
    from oracle import cpu_baseline as cb

    m = work["mesh"]
    f = float(work["freqs"][fi])
    mode_fn = provider_for(m)
    gam = {a: mode_fn(m, a, f)["gamma"] for a in s.ports}
    vals = s.system_matrix(f, gam)
    r = s.sweep(np.array([f]), precision="fp64", batch=1, restart=30, tol=1e-12, maxit=600, precond="lu",
                orth_passes=2, want_fields=True)
    xg = r["x"][0]
    sp_ = s.space()
    bd = sp_["bface_dofs"]
    pec_attrs = [a for a, k in m.bdr_kind.items() if k == "pec"]
    n = s.ndof
    excl = np.zeros(n, bool)
    excl[np.unique(bd[np.isin(m.bface_attr, pec_attrs)])] = True
    nport = np.zeros(n, np.int32)
    group = np.full(n, -1, np.int32)
    for e, a in enumerate(s.ports):
        d = np.unique(bd[m.bface_attr == a])
        nport[d] += 1
        group[d] = e
    excl |= nport > 1
    group[excl] = -1
    x, info = cb.port_solve(m.xyz, m.tets, work["order"], sp_["rowptr"], sp_["colind"], vals, excl, group, xg,
                            tol=1e-10, restart=30, maxit=200)
    dx = float(np.abs(x - xg).max() / np.abs(xg).max())
    t = info["factor_s"] + info["root_factor_s"] + info["solve_s"]
    return {"value": 1.0 / t, "unit": "freq-points/s", "cores": info["threads"], "kind": "port",
            "sample": f"one frequency point of the workload ({f / 1e9:.4g} GHz, {len(s.ports)} port excitations, "
                      f"{n} DOF): FP64 multifrontal LU factorization of A(w) (nested dissection, {info['flops']:.3g} "
                      f"real flops), the port-root factor of each excitation and FP64 GMRES(30) to 1e-10 with that "
                      f"factor, in C++ on {info['threads']} host threads (oracle/cpu_lu.cpp)",
            "cpu_model": cb.cpu_model(), "nproc": os.cpu_count(),
            "seconds": {"analysis": info["analysis_s"], "factor": info["factor_s"],
                        "port_root_factors": info["root_factor_s"], "solve": info["solve_s"]},
            "factor_gflops": info["flops"] / info["factor_s"] / 1e9, "gmres_iters": info["iters"],
            "rel_res": info["rel_res"], "max_rel_dx_vs_gpu_fp64": dx}


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_FP64_PEAK_TFLOPS = 78.6
MFMA_BF16_PEAK_TFLOPS = 2500.0
# the split-BF16 deep updates spend 12 BF16 MFMA multiply-adds (24 flops) per complex multiply-add (8 algorithmic
# flops): their ceiling in algorithmic flops is the BF16 peak x 8 / 24
SPLIT_BF16_PEAK_TFLOPS = MFMA_BF16_PEAK_TFLOPS * 8.0 / 24.0


def _lu_roofline(st, workload):
    """Rooflines of the sparse-LU kernels over a timed region (HIP events on the context stream, from opem_lu_stats):
    the trailing-update GEMMs (the 128-tile kernel of the deep FP32 updates and the 64-tile kernel of the rest)
    against the FP32 MFMA peak and the apply's factor products against HBM; the one that took longest is the
    roofline kernel.  Algorithmic flops are 8 m n k per complex update of an m x k by k x n block, whatever number of
    real MFMA products the kernel spends on it."""
    def mfma(name, key, peak=MFMA_FP32_PEAK_TFLOPS):
        d = {"kernel": name, "bound": "mfma", "peak": peak, "unit": "TFLOP/s",
             "seconds": st[key + "_s"], "launches": st[key + "_launches"], "flops": st[key + "_flops"],
             "achieved": st[key + "_flops"] / st[key + "_s"] / 1e12 if st[key + "_s"] else None}
        d["frac"] = d["achieved"] / peak if d["achieved"] else None
        return d
    if os.environ.get("OPEM_LU_BF16", "1") != "0":
        deep = mfma("k_dense_gemm_t128_bf16x3 (deep trailing updates of the FP32 fronts: 128x128 tiles, each real "
                    "product as three split-BF16 v_mfma_f32_16x16x32_bf16 products, four real products per complex "
                    "update; peak = the BF16 dense MFMA peak x 8 algorithmic / 24 BF16 flops)", "gemm128",
                    SPLIT_BF16_PEAK_TFLOPS)
        deep["frac_of_fp32_mfma_peak"] = (deep["achieved"] / MFMA_FP32_PEAK_TFLOPS) if deep["achieved"] else None
    else:
        deep = mfma("k_dense_gemm_t128 (deep trailing updates of the FP32 fronts: 128x128 tiles, three real "
                    "v_mfma_f32_16x16x4f32 products per complex update)", "gemm128")
    gemm = mfma("k_dense_gemm (the other trailing updates: 64x64 tiles, four real MFMA products per complex "
                "update)", "gemm")
    apply_ = {"kernel": "k_lu_apply_mfma (the preconditioner apply's products with the FP32 factor blocks on "
                        "v_mfma_f32_16x16x4f32, all four block kinds)",
              "bound": "hbm", "achieved": st["apply_bytes"] / st["apply_s"] / 1e9 if st["apply_s"] else None,
              "peak": HBM_PEAK_GBS, "unit": "GB/s", "seconds": st["apply_s"], "launches": st["apply_launches"],
              "bytes": st["apply_bytes"]}
    apply_["frac"] = apply_["achieved"] / HBM_PEAK_GBS if apply_["achieved"] else None
    dom = max((deep, gemm, apply_), key=lambda d: d["seconds"])
    kname = dom["kernel"].split()[0]
    ratio = _pmc_traffic_ratio(kname, workload)
    work_per_launch = (dom["bytes"] if dom is apply_ else dom["flops"]) / max(1.0, dom["launches"])
    roof = {"kernel": dom["kernel"], "bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"],
            "unit": dom["unit"], "frac": dom["frac"],
            # PMC traffic per launch: the apply's ratio to its algorithmic bytes, or the GEMM's bytes per flop
            "traffic": None if ratio is None or dom is gemm else ratio * work_per_launch,
            "per_launch": work_per_launch, "launch_seconds": dom["seconds"] / max(1.0, dom["launches"]),
            "launches": dom["launches"]}
    return roof, {deep["kernel"].split()[0]: deep, "k_dense_gemm": gemm, "k_lu_apply_mfma": apply_}


def _uneven(args):
    """The uneven set: 16 seeded guide, microstrip, tee and cross instances (ND1-3, 2-4 ports, 8-32 frequencies,
    50 k - 1.5 M DOF).  ``--concurrent K`` instances share the GPU at a time: K host threads, each with its own
    context (its own HIP stream and workspace, limited to 3/(4K) of the device memory), take instances largest first
    and, per instance, build the context (mesh upload, space) and sweep all its frequencies once in mixed precision
    (the first sweep of a context also runs the nested-dissection analysis and allocates the workspace).  value =
    passing points / wall time of that whole phase.  Then the plain FP64 path sweeps every instance the same way (its own timed phase).  A
    point passes when both converged and |dS| < 1e-6.  Lanes: with K > 1, lane 0 takes the large instances (those
    estimated at 7% or more of the set's work) one at a time with 55% of the device memory, then helps with the rest;
    lanes 1 .. K-1 share 20% and take the other instances largest first, so large factors never meet in memory."""
    import threading

    from openparem_amd import workloads
    from openparem_amd.solver import Solver, lu_stats

    keys = ("gemm_s", "gemm_flops", "gemm_launches", "panel_s", "panel_launches", "apply_s", "apply_bytes",
            "apply_launches", "factor_wall_s", "matrices_factored", "matrix_applies", "gemm128_s", "gemm128_flops",
            "gemm128_launches")
    inst = workloads.c5_instances()
    K = max(1, int(args.concurrent))
    dof_per_tet = {1: 1.2, 2: 5.3, 3: 13.5}
    cost = [(len(w["mesh"].tets) * dof_per_tet[w["order"]]) ** 1.4 * len(w["freqs"]) for w in inst]
    order = sorted(range(len(inst)), key=lambda i: -cost[i])  # largest first
    big = [i for i in order if K > 1 and cost[i] >= 0.07 * sum(cost)]
    small = [i for i in order if i not in big]

    def share_of(f):
        return 2 if len(f) > 1 and f[1] - f[0] <= 0.5e9 else 1

    def phase(fp64):
        out, errs = {}, []
        queues = {"big": list(big), "small": list(small)}
        lock = threading.Lock()

        def worker(lane):
            frac = 0.75 if K == 1 else (0.55 if lane == 0 else 0.2 / (K - 1))
            while True:
                with lock:
                    if errs:
                        return
                    q = queues["big"] if lane == 0 and queues["big"] else queues["small"]
                    if not q:
                        return
                    i = q.pop(0)
                w = inst[i]
                f = w["freqs"]
                try:
                    ta = time.perf_counter()
                    s = Solver(w["mesh"], w["order"], boundaries=w["boundaries"])
                    try:
                        opts = dict(restart=args.restart, tol=args.tol, maxit=600, precond="lu",
                                    mem_fraction=frac)
                        if fp64:
                            run = dict(precision="fp64", batch=1, orth_passes=2, **opts)
                        else:
                            run = dict(precision="mixed", precond_share=share_of(f), orth_passes=2, **opts)
                        lu_stats(s, reset=True)
                        t0 = time.perf_counter()
                        r = s.sweep(f, **run)
                        t1 = time.perf_counter()
                        st = lu_stats(s, reset=True)
                        out[i] = dict(r=r, sweep_s=t1 - t0, total_s=t1 - ta, ndof=s.ndof, ports=len(s.ports),
                                      stats=st)
                    finally:
                        s.close()
                    _progress(f"{'fp64' if fp64 else 'mixed'} instance {i} ({w['family']} ND{w['order']}, "
                              f"{out[i]['ndof']} DOF, {len(f)} freqs): sweep {t1 - t0:.2f} s, with set-up "
                              f"{t1 - ta:.2f} s")
                except Exception as e:  # noqa: BLE001 - reported after the phase
                    with lock:
                        errs.append((i, repr(e)))
                    return

        t0 = time.perf_counter()
        th = [threading.Thread(target=worker, args=(lane,)) for lane in range(K)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        if errs:
            raise RuntimeError(f"uneven set failed: {errs}")
        return out, wall

    mix, wall = phase(False)
    _progress(f"mixed phase: {wall:.2f} s wall with {K} concurrent instance(s)")
    p64, wall64 = phase(True)
    _progress(f"plain FP64 phase: {wall64:.2f} s wall")
    agg = {k: sum(mix[i]["stats"][k] for i in mix) for k in keys}
    per, tot_pass, tot_pts, tot_conv = [], 0, 0, 0
    for i, w in enumerate(inst):
        f = w["freqs"]
        r, r64 = mix[i]["r"], p64[i]["r"]
        dS = np.abs(r["S"] - r64["S"]).max(axis=(1, 2))
        conv = np.all(r["converged"].astype(bool), axis=1)
        ok = conv & np.all(r64["converged"].astype(bool), axis=1) & (dS < 1e-6)
        per.append({"instance": i, "family": w["family"], "order": w["order"], "ndof": mix[i]["ndof"],
                    "ports": mix[i]["ports"], "freqs": len(f), "band_GHz": [f[0] / 1e9, f[-1] / 1e9],
                    "share": share_of(f), "mixed_s": round(mix[i]["sweep_s"], 3),
                    "mixed_with_setup_s": round(mix[i]["total_s"], 3), "fp64_s": round(p64[i]["sweep_s"], 3),
                    "points_converged": int(conv.sum()), "points_passing": int(ok.sum()),
                    "iters_mean": float(r["iters"].mean()), "max_abs_dS_vs_fp64": float(dS.max())})
        tot_pass += int(ok.sum())
        tot_conv += int(conv.sum())
        tot_pts += len(f)
    roof, kern = _lu_roofline(agg, "uneven-c5")
    value = tot_pass / wall
    fp64_value = tot_pts / wall64
    line = {
        "metric": "freq-points/sec at |dS|<1e-6 (mixed+batched vs FP64 GPU); SpMV HBM GB/s",
        "value": value, "unit": "freq-points/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c-fp32 factor (deep updates as split-bf16 MFMA products) and inner GMRES / c-fp64 refinement", "data": "synthetic",
        "config": {"workload": f"uneven-c5: {len(per)} instances, {K} at a time on one GPU (one context and stream "
                               f"each), ND1-3, 2-4 ports, {tot_pts} frequency points, "
                               f"{min(p['ndof'] for p in per)}-{max(p['ndof'] for p in per)} DOF; wall time of "
                               f"the whole phase including each instance's set-up, analysis and allocation",
                   "concurrent": K,
                   "precond": "sparse LU per instance (FP32 factor, shared by 2 adjacent frequencies where they are "
                              f"at most 0.5 GHz apart), GMRES({args.restart}) inside FP64 refinement"},
        "points": tot_pts, "points_converged": tot_conv, "points_passing": tot_pass,
        "mixed_sweep_seconds_sum": sum(mix[i]["sweep_s"] for i in mix),
        "fp64_plain": {"value": fp64_value, "unit": "freq-points/s", "points": tot_pts, "wall_s": wall64},
        "speedup_vs_fp64_plain": value / fp64_value,
        "max_abs_dS_vs_fp64": max(p["max_abs_dS_vs_fp64"] for p in per),
        "roofline": roof, "kernels": kern,
        "instances": per, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="microstrip-c4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--restart", type=int, default=30)
    ap.add_argument("--share", type=int, default=4, help="adjacent frequencies sharing one FP32 LU factor")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--concurrent", type=int, default=4,
                    help="uneven-c5: instances sharing the GPU at a time (own context and stream each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the systems of a sweep are "
              f"independent and would be sharded across GPUs by frequency, which this build does not run)",
              file=sys.stderr)
        return 2
    # HIP-event timing of the sparse-LU kernels over the timed region (read by the library at context creation)
    os.environ["OPEM_LU_TIMING"] = "1"
    if args.workload == "uneven-c5":
        return _uneven(args)

    from openparem_amd import workloads
    from openparem_amd.measure import assembly_c2, far_field
    from openparem_amd.solver import Solver, lu_stats

    work = workloads.make(args.workload)
    freqs = work["freqs"]
    s = Solver(work["mesh"], work["order"], boundaries=work["boundaries"])
    nports = len(s.ports)
    opts = dict(restart=args.restart, tol=args.tol, maxit=600, precond="lu")
    # two classical Gram-Schmidt passes, the projections accumulated in FP64 (re-orthogonalisation)
    mixed = dict(precision="mixed", precond_share=args.share, orth_passes=2, **opts)

    for _ in range(args.warmup):
        s.sweep(freqs, **mixed)
    _progress(f"{args.warmup} warm-up sweep(s) done ({s.ndof} DOF, {len(freqs)} frequencies)")
    lu_stats(s, reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = s.sweep(freqs, **mixed)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    st = lu_stats(s, reset=True)
    _progress(f"{args.steps} timed sweep(s): {elapsed:.2f} s")

    # plain FP64 GPU path on every frequency: FP64 factor of each frequency, one frequency at a time
    t2 = time.perf_counter()
    r64 = s.sweep(freqs, precision="fp64", batch=1, orth_passes=2, **opts)
    t3 = time.perf_counter()
    fp64_value = len(freqs) / (t3 - t2)
    dS = np.abs(r["S"] - r64["S"]).max(axis=(1, 2))
    conv = np.all(r["converged"].astype(bool), axis=1)
    conv64 = np.all(r64["converged"].astype(bool), axis=1)
    passing = conv & conv64 & (dS < 1e-6)
    done = int(passing.sum())
    value = args.steps * done / elapsed
    _progress(f"plain FP64 path: {len(freqs)} points in {t3 - t2:.2f} s; {done} of {len(freqs)} points pass")

    # the single-matrix complex CSR SpMV (explicit values, one vector) on the same pattern: the MatMult of a
    # one-system solve, HBM-bound, each launch on its own copy of the matrix; rows grouped as the pattern allows
    # (ND2: pairs of rows with one column list) and one row at a time
    csr = {(p_, g_): s.csr_spmv_bench(p_, row_group=g_, nrep=50) for p_ in ("fp64", "fp32") for g_ in (0, 1)}
    csr_n, csr_nnz = s.ndof, s.nnz
    its_mean = float(r["iters"].mean())
    _progress("kernel timings done")

    # the far-field kernels at the size of this mesh's radiation lid (pattern sphere refined 4 times)
    ffd = far_field(s) if args.workload == "microstrip-c4" else None
    _progress("far-field kernels done")

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(s, work, len(freqs) // 2)
        _progress(f"CPU baseline done: {1.0 / cpu['value']:.1f} s per point on {cpu['cores']} threads, "
                  f"max |dx| vs GPU {cpu['max_rel_dx_vs_gpu_fp64']:.1e}")
    s.close()
    asm = None if args.no_assembly else {f"nd{p_}": assembly_c2(p_, 1_000_000, HBM_PEAK_GBS) for p_ in (1, 2, 3)}

    roof, kern = _lu_roofline(st, args.workload)

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
        "dtype": "c-fp32 factor (deep updates as split-bf16 MFMA products) and inner GMRES / c-fp64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: batched frequency sweep, {len(freqs)} frequencies "
                        f"{freqs[0] / 1e9:.3g}-{freqs[-1] / 1e9:.3g} GHz x {nports} port excitations = "
                        f"{len(freqs) * nports} systems on one pattern, ND{work['order']}, "
                        f"{s.ndof} DOF, nnz {s.nnz}",
            "order": work["order"], "ndof": s.ndof, "nnz": s.nnz, "systems": len(freqs) * nports,
            "precond": f"sparse LU (nested-dissection multifrontal, FP32 factor of the mean matrix of each "
                       f"{args.share} adjacent frequencies), GMRES({args.restart}) inside FP64 refinement",
        },
        "points": len(freqs),
        "points_converged": int(conv.sum()),
        "points_passing": done,
        "iters_mean": its_mean,
        "fp64_plain": {"value": fp64_value, "unit": "freq-points/s", "points": len(freqs),
                       "iters_mean": float(r64["iters"].mean()), "points_converged": int(conv64.sum())},
        "speedup_vs_fp64_plain": value / fp64_value if fp64_value else None,
        "max_abs_dS_vs_fp64": float(dS.max()),
        "roofline": roof,
        "kernels": {**kern,
                    "k_dense_panel": {"seconds": st["panel_s"], "launches": st["panel_launches"]},
                    "factorizations": {"matrices": st["matrices_factored"], "wall_seconds": st["factor_wall_s"]},
                    "matrix_applies": st["matrix_applies"]},
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
        "far_field": ffd,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
