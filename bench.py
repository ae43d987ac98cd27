"""Benchmark: SPAI columns/s on one MI355X.

Workload (the configuration the metric is quoted on): adaptive SPAI on the
random nonsymmetric band matrix, n = 262,144 (per row the diagonal plus 11
off-diagonal columns uniform within +-2000, values N(0,1), diagonal 1.2 x the
row's absolute sum; seed 20251202), eps 0.3, 8 augmentation steps, 4 new
indices per step, start pattern diag(A).  One step = one complete build of M
(every column) on a device-resident copy of A; inputs are uploaded before
the timed region.

Prints one JSON line.  `value` is the columns/s of the mixed-precision build
(FP32 QR + FP64 refinement); the FP64 build is timed beside it, and so are
the size-binned against the unbinned (generic kernel only) FP64 build, the
MSPAI probing workload (3D Laplacian 64^3, rho = 100) with its adjoint, and
the block SPAI build.  `roofline` refers to the dominant column-kernel launch
of the mixed build (HIP events on the library's stream around every launch);
`cpu_baseline` times the CPU FP64 restatement of the reference on this host.
`--workload` selects another workload as the main line (profiling runs).
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


def _traffic(name):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/pmc_<workload>.json) and the launch (and kernel) they
    were measured on, or (None, None, None)."""
    path = os.path.join(ROOT, "profiles", f"pmc_{name}.json")
    if not os.path.exists(path):
        return None, None, None
    try:
        d = json.load(open(path))
        return float(d["hbm_bytes_per_launch"]), d.get("launch"), d.get("kernel")
    except Exception:
        return None, None, None


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
        cores, _ = _cpu_cores()
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


def _cpu_cores():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most 16 (one GPU's share of the host on a shared GPU box)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return min(visible, 16), visible


def _problem(name, nx=None):
    from mspai_amd import gen
    make = getattr(gen, WORKLOADS[name][0])
    if name == "random":
        prob = make(n=nx) if nx else make()
        size = prob.n
    else:
        prob = make(nx) if nx else make()
        size = nx or (round(prob.n ** 0.5) if name in ("laplace2d", "convdiff2d") else round(prob.n ** (1 / 3)))
    return prob, size


def _builder(mspai_amd, prob):
    b = mspai_amd.Builder(0)
    b.upload(prob.A, prob.row_pattern())
    b.set_pattern(prob.P)
    if prob.B is not None:
        b.set_target(prob.B)
    return b


def _timed(b, opts, steps, warmup):
    """Wall time per build over `steps` builds after `warmup`, with the
    per-launch device times of every timed build averaged."""
    for _ in range(warmup):
        b.build(**opts)
    builds, launches = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = b.build(**opts)
        builds.append(st["build_ms"])
        launches.append(b.launch_stats())
    wall = time.perf_counter() - t0
    names = [x["name"] for x in launches[0]]
    per_launch = [{"name": nm, "ms": float(np.mean([ls[k]["ms"] for ls in launches]))} for k, nm in enumerate(names)]
    return {"ms_per_step": 1000.0 * wall / steps, "dev_ms": float(np.mean(builds)), "launches": per_launch,
            "n_fallback": st["n_fallback"], "n_overflow": st["n_overflow"], "bins": b.bin_counts()}


def _counted(b, opts):
    """One build with the work counters on: per-launch columns, algorithmic
    counts and near-margin decisions, M and the per-column info."""
    b.build(**dict(opts, counters=1))
    return {"launches": b.launch_stats(), "work": b.work_totals(), "M": b.download(), "info": b.column_info()}


def _dominant(timed, counted):
    """The launch with the largest device time, its algorithmic bytes
    (gather + candidate union + rho reads of the columns it finished, plus
    their CSC output 12 |J| + 4) and achieved GB/s."""
    k = int(np.argmax([x["ms"] for x in timed["launches"]]))
    c = counted["launches"][k]
    alg = float(c["work"][2] + c["work"][3] + c["work"][4] + 12.0 * c["nnz"] + 4.0 * c["columns"])
    ms = timed["launches"][k]["ms"]
    return {"kernel": timed["launches"][k]["name"], "kernel_ms": ms, "columns": int(c["columns"]),
            "algorithmic_bytes_per_launch": alg, "achieved": alg / (ms * 1e-3) / 1e9,
            "qr_flop_per_launch": float(c["work"][0] + c["work"][1])}


def _launch_table(timed, counted):
    return [{"name": t["name"], "ms": t["ms"], "columns": int(c["columns"]), "nnz": int(c["nnz"]),
             "near_margin": [int(v) for v in c["near"]]}
            for t, c in zip(timed["launches"], counted["launches"])]


def _probing_leg(mspai_amd, gen, torch, steps, warmup):
    """MSPAI probing on the 3D Laplacian 64^3 (n = 262,144), static
    pattern(A), rho = 100: FP64 and mixed builds on their autotuned bins,
    and the adjoint's backward/forward time."""
    prob = gen.config4()
    b = _builder(mspai_amd, prob)
    base = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=0, binning=2)
    out = {"workload": "MSPAI probing, 3D 7-point Laplacian 64^3 (n=262144), rho=100, P=pattern(A)"}
    for label, prec in (("fp64", 0), ("mixed", 1)):
        tune = b.autotune(repeats=1, precision=prec, max_impr_steps=0)
        t = _timed(b, dict(base, precision=prec), steps, warmup)
        out[label] = {"columns_per_s": prob.n / (t["ms_per_step"] * 1e-3), "ms_per_build": t["ms_per_step"],
                      "device_ms_per_build": t["dev_ms"], "bins": tune["name"], "n_fallback": t["n_fallback"]}
        if label == "fp64":
            M64 = b.download()
    M = b.download()
    same = bool(np.array_equal(M64.indptr, M.indptr) and np.array_equal(M64.indices, M.indices))
    out["mixed_over_fp64"] = out["mixed"]["columns_per_s"] / out["fp64"]["columns_per_s"]
    out["mixed_pattern_equal_fp64"] = same
    out["mixed_max_column_rel_err_vs_fp64"] = _column_rel_err(M, M64) if same else None
    b.close()
    if torch is not None:
        out["bwd_fwd"] = _bwd_fwd(mspai_amd, gen, torch, prob, steps=3)
    return out


def _cpu_leg(prob, M64, Mmx, sample):
    """The oracle (FP64, -O3) on this host: `sample` columns (None = all)
    on every usable core for about 10 s, plus its 1-thread rate; and the
    parity of the GPU builds on those columns."""
    from oracle import oracle as O
    cores, visible = _cpu_cores()
    n = prob.n
    cols = None
    if sample:
        cols = np.sort(np.random.Generator(np.random.PCG64(3)).choice(n, size=sample, replace=False)).astype(np.int32)
    ncols = cols.size if cols is not None else n
    t = time.perf_counter()
    reps = 0
    while True:
        ref = O.spai(prob, threads=cores, cols=cols)
        reps += 1
        if time.perf_counter() - t > 10.0 or reps >= 50:
            break
    cpu_t = (time.perf_counter() - t) / reps
    one = cols[:1024] if cols is not None else None
    t = time.perf_counter()
    O.spai(prob, threads=1, cols=one)
    cpu1 = (time.perf_counter() - t) / (one.size if one is not None else n)
    idx = cols if cols is not None else np.arange(n)
    pat = val = mpat = True
    rel = 0.0
    for tt, k in enumerate(idx):
        ln = ref["len"][tt]
        ri, rv = ref["idx"][tt, :ln], ref["val"][tt, :ln]
        a0, a1 = M64.indptr[k], M64.indptr[k + 1]
        pat &= bool(np.array_equal(M64.indices[a0:a1], ri))
        val &= bool(np.array_equal(M64.data[a0:a1].view(np.int64), rv.view(np.int64)))
        if Mmx is not None:
            b0, b1 = Mmx.indptr[k], Mmx.indptr[k + 1]
            same = bool(np.array_equal(Mmx.indices[b0:b1], ri))
            mpat &= same
            if same:
                rel = max(rel, float(np.linalg.norm(Mmx.data[b0:b1] - rv) / max(np.linalg.norm(rv), 1e-300)))
    what = f"stratified sample of {ncols} of {n} columns (seed 3)" if cols is not None else f"full workload (n={n})"
    cpu = {"value": ncols / cpu_t, "unit": "columns/s", "cores": cores, "visible_cores": visible, "kind": "port",
           "sample": f"{what} x {reps} runs on {cores} threads ({cpu_t:.2f} s per run); 1 thread: "
                     f"{1.0 / cpu1:.0f} columns/s",
           "single_thread_value": 1.0 / cpu1}
    parity = {"columns": int(ncols), "fp64_pattern_equal": pat, "fp64_values_bitwise": val,
              "near_margin_decisions_oracle": [int(v) for v in ref["near"].sum(axis=0)]}
    if Mmx is not None:
        parity.update({"mixed_pattern_equal": mpat, "mixed_max_column_rel_err_vs_oracle": rel})
    return cpu, parity


def _apply_leg(b, torch, A, M, reps=50):
    """The preconditioner application M x (and A x, M^T x) on the device,
    after the one-time CSR copy, timed with events on a stream of the leg's own.
    ms_per_call: `reps` products issued back to back from Python (host
    submission included).  ms_device: the same products queued behind a
    ~5 ms FP64 matrix product, so that all of them are queued before the first
    runs and the events see device time only.  Algorithmic bytes per product:
    12 per stored entry (value + index), 12 per output line (line pointer + y)
    and 8 per entry of x (read once)."""
    from mspai_amd.krylov import DeviceOperators
    ops = DeviceOperators(b)
    rng = np.random.Generator(np.random.PCG64(3))
    gate = torch.ones(6144, 6144, dtype=torch.float64, device="cuda")
    out = {}
    # a stream of its own: on the null stream (handle 0) mspai_apply runs on
    # the library's stream and waits for the product before it returns
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        _apply_products(ops, torch, A, M, rng, gate, reps, out)
    torch.cuda.synchronize()
    return out


def _apply_products(ops, torch, A, M, rng, gate, reps, out):
    for name, which, transpose, X in (("M_x", 1, False, M), ("A_x", 0, False, A), ("Mt_x", 1, True, M)):
        n_in = X.shape[0] if transpose else X.shape[1]
        n_out = X.shape[1] if transpose else X.shape[0]
        xh = rng.standard_normal(n_in)
        x = torch.tensor(xh, device="cuda")
        t0 = time.perf_counter()
        y = ops._apply(which, x, transpose)  # builds the CSR copy on first use
        torch.cuda.synchronize()
        first_ms = 1000.0 * (time.perf_counter() - t0)
        ref = (X.T @ xh) if transpose else (X @ xh)
        bitwise = bool(np.array_equal(y.cpu().numpy().view(np.int64), ref.view(np.int64)))
        for _ in range(3):
            ops._apply(which, x, transpose)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            ops._apply(which, x, transpose)
        ev1.record()
        ev1.synchronize()
        ms_call = ev0.elapsed_time(ev1) / reps
        gate.mm(gate)  # keeps the device busy while the products are queued
        ev0.record()
        for _ in range(reps):
            ops._apply(which, x, transpose)
        ev1.record()
        ev1.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        t0 = time.perf_counter()
        for _ in range(5):
            (X.T @ xh) if transpose else (X @ xh)
        host_ms = 1000.0 * (time.perf_counter() - t0) / 5
        alg = 12.0 * X.nnz + 12.0 * n_out + 8.0 * n_in
        out[name] = {"ms_device": ms, "ms_per_call": ms_call, "nnz": int(X.nnz), "algorithmic_bytes": alg,
                     "GB_per_s": alg / (ms * 1e-3) / 1e9,
                     "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "first_call_ms_with_csr_copy": first_ms,
                     "bitwise_equal_scipy": bitwise, "scipy_host_ms_1_thread": host_ms}


def _gmres_leg(b, torch, prob, maxiter=3000):
    """GMRES(30) to a relative residual of 1e-8 on A x = A x* (x* standard
    normal), without and with the built M on the right (the reference's
    KSPGMRES with the PCSHELL applying M, src/main.cpp:228-269): iterations,
    wall time, and the true relative residual checked on the host."""
    from mspai_amd.krylov import DeviceOperators, gmres
    ops = DeviceOperators(b)
    rng = np.random.Generator(np.random.PCG64(5))
    rhs = prob.A @ rng.standard_normal(prob.n)
    bt = torch.tensor(rhs, device="cuda")
    out = {"restart": 30, "rtol": 1e-8, "maxiter": maxiter, "orthogonalization": "classical Gram-Schmidt"}
    for name, M in (("none", None), ("spai_right", ops.M)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, it, hist = gmres(ops.A, bt, M, side="right", restart=30, rtol=1e-8, maxiter=maxiter)
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0)
        xh = x.cpu().numpy()
        out[name] = {"iterations": it, "wall_ms": ms, "ms_per_iteration": ms / max(it, 1),
                     "true_rel_residual": float(np.linalg.norm(rhs - prob.A @ xh) / np.linalg.norm(rhs)),
                     "converged": bool(hist[-1] <= 1e-8)}
    return out


def _setup_loop_leg(mspai_amd, gen, torch, restart=5, loops=20, its=5):
    """The loop of the reference's driver (src/main.cpp:256-269) on the
    256x256 convection-diffusion system: `loops` solves, each a
    preconditioner setup -- the full adaptive build when count % restart == 0,
    else the static re-solve on the pattern of the last full build
    (src/imspai.cpp:663-677, 850-854) -- then GMRES from x = 0 limited to
    `its` iterations at PETSc's default rtol 1e-5 (KSPSetTolerances(..., 5)).
    Wall times per setup kind and per solve (the first product with each new
    M includes its CSR copy)."""
    from mspai_amd.krylov import DeviceOperators, gmres
    prob = gen.config2()
    b = _builder(mspai_amd, prob)
    ops = DeviceOperators(b)
    bt = torch.tensor(prob.A @ np.ones(prob.n), device="cuda")
    full, resolve, solve, res = [], [], [], []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for count in range(loops):
        t0 = time.perf_counter()
        if count % restart == 0:
            b.set_pattern(prob.P)
            b.build(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                    precision=0, binning=1)
            full.append(1000.0 * (time.perf_counter() - t0))
        else:
            b.reuse_pattern()
            b.build(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=0, precision=0, binning=1)
            resolve.append(1000.0 * (time.perf_counter() - t0))
        t1 = time.perf_counter()
        _, it, hist = gmres(ops.A, bt, ops.M, side="right", restart=30, rtol=1e-5, maxiter=its)
        torch.cuda.synchronize()
        solve.append(1000.0 * (time.perf_counter() - t1))
        res.append(hist[-1])
    total = 1000.0 * (time.perf_counter() - t_all)
    nnz_m = int(b.download().nnz)
    b.close()
    return {"workload": WORKLOADS["convdiff2d"][1].format(nx=256, n=prob.n), "loops": loops, "restart": restart,
            "gmres_iterations_per_solve": its, "total_ms": total,
            "full_build_ms_mean": float(np.mean(full)), "resolve_ms_mean": float(np.mean(resolve)),
            "solve_ms_mean": float(np.mean(solve)), "last_rel_residual": float(res[-1]), "nnz_M": nnz_m}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=None, help="grid size / n (default: the workload's full size)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["block"], default="random",
                    help="block: only the block SPAI build (-bs 3) on the coupled 3-field system")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="only the main workload (profiling runs)")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both",
                    help="timed paths; with both, value is the mixed path")
    ap.add_argument("--profile", type=int, default=-1, metavar="TABLE",
                    help="profiling runs: only `steps` builds (after `warmup`) of one precision (--precision "
                         "fp64 or mixed) on bin table TABLE, no autotuning, legs or checks")
    ap.add_argument("--apply-only", action="store_true",
                    help="profiling runs: one FP64 build, then only the timed products M x, A x, M^T x")
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

    prob, size = _problem(args.workload, args.nx)
    n = prob.n
    if args.profile >= 0:
        # the dispatches of a rocprofv3 run are then exactly these builds
        prec = 1 if args.precision == "mixed" else 0
        b = _builder(mspai_amd, prob)
        name = b.set_bin_table(args.profile, precision=prec)
        opts = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                    use_mean=prob.use_mean, precision=prec, binning=2)
        r = _timed(b, opts, args.steps, args.warmup)
        b.close()
        print(json.dumps({"profile": args.workload, "n": n, "precision": args.precision, "table": name,
                          "builds": args.steps + args.warmup, "ms_per_build": r["ms_per_step"],
                          "launches": r["launches"]}))
        return
    print(f"[bench] {args.workload} n={n} nnz(A)={prob.A.nnz}", file=sys.stderr, flush=True)
    if args.apply_only:
        b = _builder(mspai_amd, prob)
        b.build(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                use_mean=prob.use_mean, precision=0, binning=2)
        print(json.dumps({"apply": _apply_leg(b, torch, prob.A, b.download())}))
        b.close()
        return
    free0 = torch.cuda.mem_get_info()[0] if torch is not None else None
    b = _builder(mspai_amd, prob)
    base = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                use_mean=prob.use_mean)
    # one-time set-up outside the timed regions: choose the size bins per precision
    tune = b.autotune(repeats=2, **base)
    r64 = _timed(b, dict(base, precision=0, binning=2), args.steps, args.warmup)
    c64 = _counted(b, dict(base, precision=0, binning=2))
    print(f"[bench] fp64 {r64['ms_per_step']:.3f} ms/build", file=sys.stderr, flush=True)
    tune_mixed = rmx = cmx = None
    if args.precision in ("mixed", "both"):
        tune_mixed = b.autotune(repeats=2, precision=1, **base)
        rmx = _timed(b, dict(base, precision=1, binning=2), args.steps, args.warmup)
        cmx = _counted(b, dict(base, precision=1, binning=2))
        print(f"[bench] mixed {rmx['ms_per_step']:.3f} ms/build", file=sys.stderr, flush=True)
    # device memory the builder holds after autotuning and the timed builds
    # (A, patterns, scratch, output, workspaces): its buffers only grow, so
    # this is its peak over every bin table tried
    hbm_held = (free0 - torch.cuda.mem_get_info()[0]) if free0 is not None else None
    unbinned = None
    if not args.no_legs and args.workload == "random":
        # the generic kernel alone (no size bins): the unbinned baseline
        r0 = _timed(b, dict(base, precision=0, binning=0), max(2, args.steps // 3), 1)
        unbinned = {"fp64_unbinned_ms_per_build": r0["ms_per_step"], "fp64_binned_ms_per_build": r64["ms_per_step"],
                    "binned_speedup": r0["ms_per_step"] / r64["ms_per_step"]}
    cache_run = gmres_run = None
    if not args.no_legs and args.workload in ("random", "convdiff2d"):
        # the reference's LRU cache option (-cs): same bits, repeated A(I,J) blocks reuse factors
        c2 = gen.config2() if args.workload == "random" else prob
        bc = _builder(mspai_amd, c2)
        o2 = dict(epsilon=c2.epsilon, maxnew=c2.maxnew, max_impr_steps=c2.max_impr_steps, precision=0, binning=1)
        plain = _timed(bc, o2, args.steps, args.warmup)
        Mp = bc.download()
        cached = _timed(bc, dict(o2, cache=65536), args.steps, args.warmup)
        stc = bc.build(**dict(o2, cache=65536))
        if torch is not None:
            gmres_run = _gmres_leg(bc, torch, c2)
            gmres_run["workload"] = WORKLOADS["convdiff2d"][1].format(nx=256, n=c2.n)
        cache_run = {"workload": WORKLOADS["convdiff2d"][1].format(nx=256, n=c2.n),
                     "columns_per_s_without": c2.n / (plain["ms_per_step"] * 1e-3),
                     "columns_per_s_with_cache": c2.n / (cached["ms_per_step"] * 1e-3),
                     "cache_hits_per_build": stc["n_cache_hits"],
                     "bitwise_equal": bool(np.array_equal(bc.download().data.view(np.int64), Mp.data.view(np.int64)))}
        bc.close()
    apply_run = setup_loop = None
    if not args.no_legs and torch is not None and args.workload == "random":
        apply_run = _apply_leg(b, torch, prob.A, b.download())  # M of the last build
        setup_loop = _setup_loop_leg(mspai_amd, gen, torch)
    b.close()
    probing = None
    if not args.no_legs and args.workload == "random" and args.precision == "both":
        probing = _probing_leg(mspai_amd, gen, torch, args.steps, args.warmup)
    bwd = None
    if not args.no_legs and torch is not None and args.precision == "both" and args.workload in (
            "varcoef3d", "probing3d", "laplace2d"):
        bwd = _bwd_fwd(mspai_amd, gen, torch, prob, steps=3)
    block = None
    if not args.no_legs and args.precision == "both" and args.workload == "random":
        block = _block_bench(mspai_amd, gen, max(3, args.steps // 2), args.warmup, not args.no_cpu)

    main_t, main_c = (rmx, cmx) if rmx is not None else (r64, c64)
    M = main_c["M"]
    M64 = c64["M"]
    dom = _dominant(main_t, main_c)
    traffic, traffic_launch, traffic_kernel = _traffic(f"{args.workload}_{size}")
    if traffic_launch != dom["kernel"]:
        traffic = None  # measured on another launch (another bin table won the autotuning)
    fro_gpu = _fro_residual(prob.A, M, prob.B)
    mixed_check = None
    if rmx is not None:
        same_pat = bool(np.array_equal(M64.indptr, M.indptr) and np.array_equal(M64.indices, M.indices))
        fro64 = _fro_residual(prob.A, M64, prob.B)
        mixed_check = {"pattern_equal_fp64": same_pat,
                       "max_column_rel_err_vs_fp64": _column_rel_err(M, M64) if same_pat else None,
                       "fro_AM_minus_I_fp64": fro64, "fro_rel_diff": abs(fro_gpu - fro64) / fro64,
                       "n_fp64_fallback": rmx["n_fallback"],
                       "fp64_fallback_fraction": rmx["n_fallback"] / n}
    cpu = parity = None
    if not args.no_cpu:
        sample = 8192 if args.workload in ("random", "varcoef3d") else None
        cpu, parity = _cpu_leg(prob, M64, M if rmx is not None else None, sample)

    value = n / (main_t["ms_per_step"] * 1e-3)
    near = np.sum([x["near"] for x in c64["launches"]], axis=0)
    out = {
        "metric": "SPAI columns/s, mixed vs FP64 GPU at matched ||AM-I||_F; bwd/fwd time",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_t["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 QR + f64 refinement" if rmx is not None else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload][1].format(nx=size, n=n),
                   "n_columns": n, "nnz_A": int(prob.A.nnz), "nnz_M": int(M.nnz),
                   "precision": "mixed" if rmx is not None else "fp64",
                   "parallelism": "single GPU, persistent column queue"},
        "device_ms_per_build": main_t["dev_ms"],
        "fp64_columns_per_s": n / (r64["ms_per_step"] * 1e-3),
        "fp64_ms_per_build": r64["ms_per_step"],
        "fp64_device_ms_per_build": r64["dev_ms"],
        "mixed_columns_per_s": (n / (rmx["ms_per_step"] * 1e-3)) if rmx is not None else None,
        "mixed_over_fp64": (r64["ms_per_step"] / rmx["ms_per_step"]) if rmx is not None else None,
        "mixed_check": mixed_check,
        "fro_AM_minus_I": fro_gpu,
        "device_memory_held_bytes": hbm_held,
        "near_margin_decisions_fp64": {"score_at_mean": int(near[0]), "maxnew_cut": int(near[1]),
                                       "residual_at_eps": int(near[2]),
                                       "margin": "at most 1e-12 relative"},
        "roofline": {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_launch": traffic_launch, "traffic_kernel": traffic_kernel,
                     **{k: v for k, v in dom.items() if k != "achieved"}},
        "launches": _launch_table(main_t, main_c),
        "launches_fp64": _launch_table(r64, c64),
        "bins": main_t["bins"],
        "autotune": tune,
        "autotune_mixed": tune_mixed,
        "binned_vs_unbinned": unbinned,
        "cache": cache_run,
        "apply": apply_run,
        "gmres": gmres_run,
        "setup_solve_loop": setup_loop,
        "probing3d": probing,
        "bwd_fwd": bwd,
        "block_spai": block,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
