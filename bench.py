"""Benchmark: SPAI columns/s on one MI355X.

Workload (the configuration the metric is quoted on): adaptive SPAI on the 2D
upwind convection-diffusion operator, 256 x 256 grid (n = 65,536), eps 0.4,
5 augmentation steps, 5 new indices per step, start pattern diag(A).  One
step = one complete build of M (every column of M) on a device-resident
copy of A; inputs are uploaded before the timed region.

Prints one JSON line.  `value` is the columns/s of the mixed-precision build
(FP32 QR + FP64 refinement); the FP64 build is timed beside it.
`--workload` selects another of the five workloads for profiling runs.
`roofline` refers to the column kernel (HIP events on the library's stream);
`cpu_baseline` times the CPU FP64 restatement of the reference on this host.
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

# workload name -> (generator in mspai_amd.gen, description)
WORKLOADS = {
    "laplace2d": ("config1", "static SPAI, 2D 5-point Laplacian {nx}x{nx} (n={n}), P=pattern(A)"),
    "convdiff2d": ("config2", "adaptive SPAI, 2D upwind convection-diffusion {nx}x{nx} (n={n}), eps=0.4, ns=5, "
                              "mn=5, J0=diag"),
    "random": ("config3", "adaptive SPAI, random nonsymmetric band matrix (n={n}), eps=0.3, ns=8, mn=4, J0=diag"),
    "probing3d": ("config4", "MSPAI probing, 3D 7-point Laplacian {nx}^3 (n={n}), rho=100, P=pattern(A)"),
    "varcoef3d": ("config5", "static SPAI, 3D variable-coefficient FV {nx}^3 (n={n}), P=pattern(A^2)"),
}


def _algorithmic_bytes(work, nnzM, n):
    """Gather + candidate union + rho reads + the CSC output of M."""
    return float(work[2] + work[3] + work[4] + 12.0 * nnzM + 4.0 * (n + 1))


def _traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"pmc_{name}.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def _fro_residual(A, M, B=None):
    """||A M - B||_F with B = I when no target is given."""
    import scipy.sparse as sps
    R = (A @ M).tocsc() - (B if B is not None else sps.identity(A.shape[1], format="csc"))
    return float(np.sqrt((R.multiply(R)).sum()))


def _column_rel_err(M, Mref):
    """max over columns of ||m_k - m_k^ref|| / ||m_k^ref|| (same pattern)."""
    d = (M - Mref).tocsc()
    num = np.sqrt(np.asarray(d.multiply(d).sum(axis=0)).ravel())
    den = np.sqrt(np.asarray(Mref.multiply(Mref).sum(axis=0)).ravel())
    return float(np.max(num / np.maximum(den, 1e-300)))


def _bwd_fwd(mspai_amd, gen, torch, prob=None, steps=4):
    """Backward/forward device time of the adjoint on a static workload (by
    default the MSPAI probing workload: 3D Laplacian 64^3, static pattern(A),
    rho = 100), both directions on the autotuned bin table."""
    name = None
    if prob is None:
        prob = gen.config4()
        name = "MSPAI probing, 3D Laplacian 64^3 (n=262144), static P(A), rho=100"
    b = mspai_amd.Builder(0)
    b.upload(prob.A, prob.row_pattern())
    b.set_pattern(prob.P)
    if prob.B is not None:
        b.set_target(prob.B)
    out = {"workload": name or prob.name}
    tune = b.autotune(repeats=1, max_impr_steps=0)
    out["bins"] = tune["name"]
    for label, precision in (("fp64", 0), ("mixed", 1)):
        fw, bw = [], []
        for it in range(steps):
            st = b.build(max_impr_steps=0, precision=precision, binning=2)
            gm = torch.ones(st["nnz"], dtype=torch.float64, device="cuda")
            gc = torch.empty(prob.A.nnz, dtype=torch.float64, device="cuda")
            gb = torch.empty(prob.B.nnz, dtype=torch.float64, device="cuda") if prob.B is not None else None
            torch.cuda.synchronize()
            sb = b.backward_device(gm.data_ptr(), gc.data_ptr(), gb.data_ptr() if gb is not None else None,
                                   precision=precision, binning=2)
            if it:
                fw.append(st["build_ms"])
                bw.append(sb["build_ms"])
        out[label] = {"fwd_ms": float(np.mean(fw)), "bwd_ms": float(np.mean(bw)),
                      "bwd_over_fwd": float(np.mean(bw) / np.mean(fw))}
    b.close()
    return out


BLOCK_WORKLOAD = ("block SPAI (-bs 3), coupled 3-field 2D upwind convection-diffusion 256x256 "
                  "(n=196608, 65536 block columns), eps=0.4, ns=5, mn=5, J0=block diagonal")


def _block_bench(mspai_amd, gen, steps, warmup, with_cpu):
    """Block SPAI build throughput on the coupled 3-field system, and its
    parity with the block oracle (whole matrix, bitwise)."""
    A = gen.block_convdiff2d(256, 3, seed=20)
    n = A.shape[0]
    opts = dict(epsilon=0.4, maxnew=5, max_impr_steps=5)
    b = mspai_amd.Builder(0)
    b.upload(A)
    b.set_block_size(3)
    nb = len(b.block_info()[0])
    for _ in range(warmup):
        b.build(**opts)
    dev, kern = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = b.build(**opts)
        dev.append(st["build_ms"])
        kern.append(st["solve_ms"])
    wall = (time.perf_counter() - t0) / steps
    M = b.download()
    info = b.column_info()
    b.close()
    out = {"workload": BLOCK_WORKLOAD, "block_columns_per_s": nb / wall, "scalar_columns_per_s": n / wall,
           "ms_per_build": 1000.0 * wall, "device_ms_per_build": float(np.mean(dev)),
           "kernel_ms": float(np.mean(kern)), "nnz_M": int(M.nnz), "fro_AM_minus_I": _fro_residual(A, M),
           "mean_solves_per_block_column": float(np.mean(info["nsolve"][:nb]))}
    if with_cpu:
        from oracle import block as OB
        cores = min(os.cpu_count() or 1, 16)
        Bm = OB.BlockMatrix(A, OB.constant_blocks(n, 3))
        t0 = time.perf_counter()
        reps = 0
        while True:
            ref = OB.spai_block(Bm, None, threads=cores, **opts)
            reps += 1
            if time.perf_counter() - t0 > 3.0 or reps >= 20:
                break
        cpu_t = (time.perf_counter() - t0) / reps
        Mref = OB.to_scalar(Bm, ref)
        out["cpu_baseline"] = {"value": nb / cpu_t, "unit": "block columns/s", "cores": cores, "kind": "port",
                               "sample": f"full workload x {reps} runs on {cores} threads"}
        out["parity"] = {"fp64_pattern_equal": bool(np.array_equal(Mref.indptr, M.indptr)
                                                    and np.array_equal(Mref.indices, M.indices)),
                         "fp64_values_bitwise": bool(np.array_equal(Mref.data.view(np.int64),
                                                                    M.data.view(np.int64)))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=None, help="grid size (default: the workload's full size)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["block"], default="convdiff2d",
                    help="block: only the block SPAI build (-bs 3) on the coupled 3-field system")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both",
                    help="timed paths; with both, value is the mixed path")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (scale for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the columns are independent and would be split "
              "across GPUs by column ranges, but only single-GPU runs are supported.",
              file=sys.stderr)
        sys.exit(2)

    try:  # torch first: one HIP runtime for torch and the library
        import torch
    except Exception:  # pragma: no cover
        torch = None
    import mspai_amd
    from mspai_amd import gen

    if args.workload == "block":
        blk = _block_bench(mspai_amd, gen, args.steps, args.warmup, not args.no_cpu)
        print(json.dumps({"metric": "block SPAI block columns/s", "value": blk["block_columns_per_s"],
                          "unit": "block columns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": blk["ms_per_build"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": BLOCK_WORKLOAD}, "block_spai": blk,
                          "cpu_baseline": blk.get("cpu_baseline")}))
        return

    make = getattr(gen, WORKLOADS[args.workload][0])
    if args.workload == "random":
        prob = make(n=args.nx) if args.nx else make()
    else:
        prob = make(args.nx) if args.nx else make()
    n = prob.n
    if args.workload == "random":
        nx = n
    else:
        nx = args.nx or (round(n ** 0.5) if args.workload in ("laplace2d", "convdiff2d") else round(n ** (1 / 3)))
    b = mspai_amd.Builder(0)
    b.upload(prob.A, prob.row_pattern())
    b.set_pattern(prob.P)
    if prob.B is not None:
        b.set_target(prob.B)
    base_opts = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                     use_mean=prob.use_mean)
    # one-time set-up outside the timed region: choose the size bins
    tune = b.autotune(repeats=2, **base_opts)

    def timed(precision):
        opts = dict(base_opts, precision=precision, binning=2)
        for _ in range(args.warmup):
            b.build(**opts)
        builds, solves, firsts = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = b.build(**opts)
            builds.append(st["build_ms"])
            solves.append(st["solve_ms"])
            firsts.append(st["first_bin_ms"])
        wall = time.perf_counter() - t0
        return {"ms_per_step": 1000.0 * wall / args.steps, "dev_ms": float(np.mean(builds)),
                "kern_ms": float(np.mean(solves)), "first_ms": float(np.mean(firsts)),
                "M": b.download(), "work": b.work_totals(),
                "bins": b.bin_counts(), "n_fallback": st["n_fallback"], "n_overflow": st["n_overflow"]}

    r64 = timed(0)
    tune_mixed = None
    rmx = None
    if args.precision in ("mixed", "both"):
        # the mixed kernels have their own register and LDS budgets: tune them apart
        tune_mixed = b.autotune(repeats=2, precision=1, **base_opts)
        rmx = timed(1)
    # the reference's LRU cache option (-cs): same bits, repeated A(I,J) blocks reuse factors
    cache_run = None
    if args.precision == "both" and args.workload == "convdiff2d":
        opts_c = dict(base_opts, precision=0, binning=2, cache=65536)
        for _ in range(args.warmup):
            b.build(**opts_c)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stc = b.build(**opts_c)
        cache_run = {"columns_per_s": n / ((time.perf_counter() - t0) / args.steps),
                     "cache_hits_per_build": stc["n_cache_hits"],
                     "pattern_equal_fp64": bool(np.array_equal(b.download().indices, r64["M"].indices))}
    b.close()
    bwd = None
    if torch is not None and args.precision == "both":
        if args.workload == "convdiff2d":
            bwd = _bwd_fwd(mspai_amd, gen, torch)
        elif args.workload in ("varcoef3d", "probing3d", "laplace2d"):
            bwd = _bwd_fwd(mspai_amd, gen, torch, prob, steps=3)
    block = None
    if args.precision == "both" and args.workload == "convdiff2d":
        block = _block_bench(mspai_amd, gen, max(3, args.steps // 2), args.warmup, not args.no_cpu)
    main_run = rmx if rmx is not None else r64
    M = main_run["M"]
    ms_per_step = main_run["ms_per_step"]
    work = main_run["work"]
    # the dominant kernel: the first size bin's launch when every column was
    # queued there (its own HIP events); else all column-kernel launches
    queued = main_run["bins"]["queued"]
    first_all = bool(queued) and queued[0] == n and not any(queued[1:])
    kern_ms = main_run["first_ms"] if first_all else main_run["kern_ms"]
    kern_name = (f"spai_column_kernel, first size bin (all {n} columns)" if first_all
                 else "spai_column_kernel (all size bins)")
    alg_bytes = _algorithmic_bytes(work, M.nnz, n)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    traffic = _traffic(f"{args.workload}_{nx}")
    fro_gpu = _fro_residual(prob.A, M, prob.B)
    M64 = r64["M"]
    mixed_check = None
    if rmx is not None:
        same_pat = bool(np.array_equal(M64.indptr, M.indptr) and np.array_equal(M64.indices, M.indices))
        rel = float(np.max(np.abs(M.data - M64.data) / np.maximum(np.abs(M64.data), 1e-300))) if same_pat else None
        fro64 = _fro_residual(prob.A, M64, prob.B)
        mixed_check = {"pattern_equal_fp64": same_pat,
                       "max_column_rel_err_vs_fp64": _column_rel_err(M, M64) if same_pat else None,
                       "max_entry_rel_err_vs_fp64": rel,
                       "fro_AM_minus_I_fp64": fro64,
                       "fro_rel_diff": abs(fro_gpu - fro64) / fro64,
                       "n_fp64_fallback": rmx["n_fallback"]}

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = min(os.cpu_count() or 1, 16)
        # the two large workloads: a stratified sample of 8,192 columns (seed 3)
        sampled = args.workload in ("random", "varcoef3d")
        cols = (np.sort(np.random.Generator(np.random.PCG64(3)).choice(n, size=8192, replace=False)).astype(np.int32)
                if sampled else None)
        ncols = cols.size if sampled else n
        t = time.perf_counter()
        reps = 0
        while True:
            ref = O.spai(prob, threads=cores, cols=cols)
            reps += 1
            if time.perf_counter() - t > 5.0 or reps >= 50:
                break
        cpu_t = (time.perf_counter() - t) / reps
        one = cols[:1024] if sampled else None
        t = time.perf_counter()
        O.spai(prob, threads=1, cols=one)
        cpu1 = (time.perf_counter() - t) / (one.size if sampled else n)
        if sampled:
            ok_pat = ok_val = True
            for tt, k in enumerate(cols):
                a0, a1 = M64.indptr[k], M64.indptr[k + 1]
                ok_pat &= bool(np.array_equal(M64.indices[a0:a1], ref["idx"][tt, :ref["len"][tt]]))
                ok_val &= bool(np.array_equal(M64.data[a0:a1].view(np.int64),
                                              ref["val"][tt, :ref["len"][tt]].view(np.int64)))
            parity = {"sampled_columns": int(cols.size), "fp64_pattern_equal": ok_pat, "fp64_values_bitwise": ok_val}
        else:
            Mref = O.to_csc(ref, n)
            parity = {"fp64_pattern_equal": bool(np.array_equal(Mref.indptr, M64.indptr)
                                                 and np.array_equal(Mref.indices, M64.indices)),
                      "fp64_values_bitwise": bool(np.array_equal(Mref.data.view(np.int64),
                                                                 M64.data.view(np.int64))),
                      "reported_pattern_equal": bool(np.array_equal(Mref.indptr, M.indptr)
                                                     and np.array_equal(Mref.indices, M.indices)),
                      "fro_AM_minus_I_oracle": _fro_residual(prob.A, Mref, prob.B)}
        what = (f"stratified sample of {ncols} of {n} columns (seed 3)" if sampled
                else f"full workload (n={n})")
        cpu = {"value": ncols / cpu_t, "unit": "columns/s", "cores": cores, "kind": "port",
               "sample": f"{what} x {reps} runs on {cores} threads; 1 thread: {1.0 / cpu1:.0f} columns/s",
               "single_thread_value": 1.0 / cpu1}

    value = n / (ms_per_step * 1e-3)
    out = {
        "metric": "SPAI columns/s, mixed vs FP64 GPU at matched ||AM-I||_F; bwd/fwd time",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 QR + f64 refinement" if rmx is not None else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload][1].format(nx=nx, n=n),
                   "n_columns": n, "nnz_A": int(prob.A.nnz), "nnz_M": int(M.nnz),
                   "precision": "mixed" if rmx is not None else "fp64",
                   "parallelism": "single GPU, persistent column queue"},
        "device_ms_per_build": main_run["dev_ms"],
        "fp64_columns_per_s": n / (r64["ms_per_step"] * 1e-3),
        "fp64_device_ms_per_build": r64["dev_ms"],
        "mixed_columns_per_s": (n / (rmx["ms_per_step"] * 1e-3)) if rmx is not None else None,
        "mixed_over_fp64": (r64["ms_per_step"] / rmx["ms_per_step"]) if rmx is not None else None,
        "mixed_check": mixed_check,
        "fro_AM_minus_I": fro_gpu,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kern_name, "kernel_ms": kern_ms, "all_bins_kernel_ms": main_run["kern_ms"],
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "qr_flop_per_launch": float(work[0] + work[1])},
        "bins": main_run["bins"],
        "autotune": tune,
        "autotune_mixed": tune_mixed,
        "fp64_with_cache": cache_run,
        "bwd_fwd": bwd,
        "block_spai": block,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
