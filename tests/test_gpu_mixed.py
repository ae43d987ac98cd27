"""Mixed precision (FP32 QR + FP64 refinement) against the FP64 oracle.

Bar: the sparsity pattern of M is identical; every column's values are within
1e-9 relative (2-norm over the column) of the oracle; the residual norms
within 1e-9 relative; solve counts equal.  Columns the guards send to the
FP64 kernels carry MSPAI_COL_FALLBACK and are bitwise equal to the oracle.
"""
import os

import numpy as np
import pytest

from mspai_amd import gen
from tests.util import csparse_problem

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
FALLBACK = 8
TOL = 1e-9

CASES = {
    "c1_laplace2d_64": lambda: gen.config1(64),
    "c2_convdiff_64": lambda: gen.config2(64),
    "c2_convdiff_256": lambda: gen.config2(256),
    "c3_random_16384": lambda: gen.config3(n=16384, bw=800),
    "c4_probing_16": lambda: gen.config4(16),
    "c5_varcoef_12": lambda: gen.config5(12),
    "cs_fs_183_1": lambda: csparse_problem("fs_183_1", eps=0.2, mn=5, ns=6),
    "cs_bcsstk01": lambda: csparse_problem("bcsstk01"),
}


def check_mixed(oracle, prob, M, info):
    ref = oracle.spai(prob, threads=THREADS)
    Mref = oracle.to_csc(ref, prob.n)
    assert np.array_equal(Mref.indptr, M.indptr) and np.array_equal(Mref.indices, M.indices), "pattern differs"
    assert np.array_equal(ref["nsolve"], info["nsolve"])
    assert np.array_equal(ref["status"], info["status"] & ~FALLBACK)
    d = M.data - Mref.data
    colerr = np.sqrt(np.add.reduceat(d * d, Mref.indptr[:-1])) if M.nnz else np.zeros(0)
    colnorm = np.sqrt(np.add.reduceat(Mref.data ** 2, Mref.indptr[:-1])) if M.nnz else np.zeros(0)
    rel = colerr / np.maximum(colnorm, 1e-300)
    assert rel.max(initial=0.0) <= TOL, rel.max()
    rr = np.abs(info["res_norm"] - ref["res_norm"]) / np.maximum(ref["res_norm"], 1e-300)
    assert rr.max(initial=0.0) <= TOL
    fb = (info["status"] & FALLBACK) != 0
    # columns rebuilt in FP64 are the oracle's bits
    for k in np.nonzero(fb)[0][:2000]:
        a, b = M.indptr[k], M.indptr[k + 1]
        assert np.array_equal(M.data[a:b].view(np.int64), Mref.data[a:b].view(np.int64))
    return rel.max(initial=0.0), fb.mean()


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_matches_oracle(oracle, gpu, name):
    prob = CASES[name]()
    print(f"[mixed {name}] n={prob.n}", flush=True)
    M, info, stats = gpu.spai(prob, precision=1)
    worst, fbfrac = check_mixed(oracle, prob, M, info)
    print(f"   worst column rel err {worst:.2e}, fallback fraction {fbfrac:.4f}, bins {stats['bins']}",
          flush=True)


def test_mixed_full_c3(oracle, gpu):
    prob = gen.config3()
    print(f"[mixed full {prob.name}] n={prob.n}", flush=True)
    M, info, stats = gpu.spai(prob, precision=1)
    worst, fbfrac = check_mixed(oracle, prob, M, info)
    print(f"   worst column rel err {worst:.2e}, fallback fraction {fbfrac:.4f}", flush=True)


def test_mixed_every_bin_table(oracle, gpu):
    """Each bin table with size bins, mixed: the bar above on a band with the
    bandwidth of the n = 262,144 workload (columns that outgrow a mixed bin go
    on in the next one from the step they outgrew); the bin-less table is
    refused for mixed builds."""
    prob = gen.config3(n=32768, bw=2000)
    b = gpu.Builder(0)
    try:
        b.upload(prob.A, prob.row_pattern())
        b.set_pattern(prob.P)
        opts = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps)
        for t in range(b.set_bin_table(-1)):
            try:
                name = b.set_bin_table(t, precision=1)
            except RuntimeError as e:
                assert "size bins" in str(e)
                continue
            st = b.build(binning=2, precision=1, **opts)
            print(f"[mixed table {t} {name}] {st['solve_ms']:.3f} ms", flush=True)
            worst, fbfrac = check_mixed(oracle, prob, b.download(), b.column_info())
            print(f"   worst column rel err {worst:.2e}, fallback fraction {fbfrac:.4f}", flush=True)
    finally:
        b.close()


def test_mixed_repeat_routes_fp64_columns_alike(oracle, gpu):
    """A mixed build whose FP64 remainder (guard fallbacks) is small sends it
    to the generic kernel in the next build instead of through the FP64 bins:
    the second build equals the first bit for bit, and both meet the bar."""
    prob = gen.config3(n=4096, bw=300)
    b = gpu.Builder(0)
    try:
        b.upload(prob.A, prob.row_pattern())
        b.set_pattern(prob.P)
        kw = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                  use_mean=prob.use_mean, precision=1)
        st1 = b.build(**kw)
        M1, info1 = b.download(), b.column_info()
        st2 = b.build(**kw)
        M2, info2 = b.download(), b.column_info()
    finally:
        b.close()
    print(f"[mixed repeat] n={prob.n} fallbacks {st1['n_fallback']} / {st2['n_fallback']}", flush=True)
    assert 0 < st1["n_fallback"] <= 4096 and st2["n_fallback"] == st1["n_fallback"]
    check_mixed(oracle, prob, M1, info1)
    assert np.array_equal(M1.indptr, M2.indptr) and np.array_equal(M1.indices, M2.indices)
    assert np.array_equal(M1.data.view(np.int64), M2.data.view(np.int64))
    assert np.array_equal(info1["res_norm"].view(np.int64), info2["res_norm"].view(np.int64))
    assert np.array_equal(info1["nsolve"], info2["nsolve"]) and np.array_equal(info1["status"], info2["status"])



def test_mixed_full_c4(oracle, gpu):
    """MSPAI probing on the full 3D Laplacian 64^3 (n = 262,144), every
    column against the oracle: identical pattern, <= 1e-9 per column."""
    prob = gen.config4()
    print(f"[mixed full {prob.name}] n={prob.n}", flush=True)
    M, info, stats = gpu.spai(prob, precision=1)
    worst, fbfrac = check_mixed(oracle, prob, M, info)
    print(f"   worst column rel err {worst:.2e}, fallback fraction {fbfrac:.4f}", flush=True)


def test_mixed_full_c5_sampled_columns(oracle, gpu):
    """n = 2,097,152 (static P(A^2)): the mixed build of every column, a
    stratified sample of 8,192 columns re-solved by the oracle: the same
    pattern, columns within 1e-9, residual norms within 1e-9."""
    prob = gen.config5()
    print(f"[mixed full {prob.name}] n={prob.n}", flush=True)
    M, info, stats = gpu.spai(prob, precision=1)
    assert np.array_equal(M.indptr, prob.P.indptr) and np.array_equal(M.indices, prob.P.indices)
    cols = np.sort(np.random.Generator(np.random.PCG64(3)).choice(prob.n, size=8192, replace=False)).astype(np.int32)
    ref = oracle.spai(prob, threads=THREADS, cols=cols)
    worst = 0.0
    for t, k in enumerate(cols):
        a, b = M.indptr[k], M.indptr[k + 1]
        rv = ref["val"][t, :ref["len"][t]]
        assert np.array_equal(M.indices[a:b], ref["idx"][t, :ref["len"][t]])
        worst = max(worst, float(np.linalg.norm(M.data[a:b] - rv) / max(np.linalg.norm(rv), 1e-300)))
    rr = np.abs(info["res_norm"][cols] - ref["res_norm"]) / np.maximum(ref["res_norm"], 1e-300)
    print(f"   worst sampled column rel err {worst:.2e}, residual {rr.max():.2e}, "
          f"fallbacks {stats['n_fallback']}", flush=True)
    assert worst <= TOL and rr.max() <= TOL
