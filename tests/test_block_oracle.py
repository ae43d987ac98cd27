"""Pins the block SPAI oracle (CPU only).

Checks, each chosen so that a plausible mistake fails it:
  * block detection (Find_Constant_Blocks / Find_Diagonal_Blocks) against
    hand-derived partitions;
  * Convert_To_Block_Matrix round trip: the blocks reassemble A exactly;
  * (A_j^T A_j)^-1 against numpy's inverse of the block column's Gram matrix;
  * block size 1 reduces to the scalar algorithm: same patterns and values as
    the scalar oracle;
  * a full start pattern gives the exact inverse (static solve on all blocks);
  * every rho of Compute_Rhos_Block against brute-force least squares
    min_X ||R - A(:,j) X||_F;
  * the per-block-column residual norm against ||A M(:,k) - E_k||_F from scipy.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from mspai_amd import gen


@pytest.fixture(scope="module")
def blk(oracle):
    from oracle import block as B
    return B


def _blockdiag(sizes, extra=()):
    n = int(sum(sizes))
    A = np.zeros((n, n))
    o = 0
    rng = np.random.Generator(np.random.PCG64(5))
    for s in sizes:
        A[o:o + s, o:o + s] = rng.standard_normal((s, s)) + 3 * np.eye(s)
        o += s
    for r, c in extra:
        A[r, c] = 0.5
    return sp.csc_matrix(A)


def test_constant_blocks(blk):
    assert blk.constant_blocks(10, 3).tolist() == [3, 3, 3, 1]
    assert blk.constant_blocks(9, 3).tolist() == [3, 3, 3]
    assert blk.constant_blocks(5, 1).tolist() == [1] * 5


def test_diagonal_blocks_hand_cases(blk):
    # dense diagonal blocks, off-diagonal entries away from the runs
    A = _blockdiag([3, 2, 1, 4], extra=[(9, 0), (0, 7), (6, 4)])
    assert blk.diagonal_blocks(A).tolist() == [3, 2, 1, 4]
    # the upper limit caps the first run
    assert blk.diagonal_blocks(A, upper_limit=2).tolist() == [2, 1, 2, 1, 2, 2]
    # tridiagonal: each run from a diagonal covers two rows
    T = sp.diags([np.ones(6), 2 * np.ones(7), np.ones(6)], [-1, 0, 1], format="csc")
    assert blk.diagonal_blocks(T).tolist() == [2, 2, 2, 1]
    # a zero diagonal entry ends a block at size one
    D = _blockdiag([2, 3]).toarray()
    D[2, 2] = 0.0
    assert blk.diagonal_blocks(sp.csc_matrix(D)).tolist() == [2, 1, 2]


def test_block_matrix_round_trip(blk):
    A = gen.block_convdiff2d(6, 3, seed=1)
    for sizes in (blk.constant_blocks(A.shape[0], 3), blk.constant_blocks(A.shape[0], 4),
                  blk.diagonal_blocks(A)):
        B = blk.BlockMatrix(A, sizes)
        D = np.zeros(A.shape)
        for j in range(B.nb):
            for e in range(B.colptr[j], B.colptr[j + 1]):
                i = B.rowind[e]
                h, w = B.bsize[i], B.bsize[j]
                D[B.start[i]:B.start[i] + h, B.start[j]:B.start[j] + w] = \
                    B.val[B.voff[e]:B.voff[e + 1]].reshape(w, h).T
        assert np.array_equal(D, A.toarray())
        # row structure is the transpose of the block pattern
        P = B.pattern()
        assert np.array_equal(B.rp_ptr, P.tocsr().indptr) and np.array_equal(B.rp_ind, P.tocsr().indices)


def test_square_inverse_matches_numpy(blk):
    A = gen.block_convdiff2d(5, 3, seed=2)
    B = blk.BlockMatrix(A, blk.constant_blocks(A.shape[0], 3))
    inv = blk.sq_inverse(B)
    o = 0
    Ad = A.toarray()
    for j in range(B.nb):
        w = B.bsize[j]
        cols = Ad[:, B.start[j]:B.start[j] + w]
        ref = np.linalg.inv(cols.T @ cols)
        got = inv[o:o + w * w].reshape(w, w).T
        assert np.allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
        o += w * w


@pytest.mark.parametrize("adaptive", [False, True])
def test_block_size_one_is_scalar_spai(blk, oracle, adaptive):
    A = gen.convdiff2d(12)
    n = A.shape[0]
    ns = 3 if adaptive else 0
    B = blk.BlockMatrix(A, np.ones(n, np.int32))
    out = blk.spai_block(B, None, epsilon=0.2, maxnew=3, max_impr_steps=ns)
    Mb = blk.to_scalar(B, out)
    prob = gen.Problem("c", A, gen.diag_pattern(n), None, epsilon=0.2, maxnew=3, max_impr_steps=ns)
    ref = oracle.spai(prob)
    Ms = oracle.to_csc(ref, n)
    Ms.eliminate_zeros()
    assert np.array_equal(Mb.indptr, Ms.indptr) and np.array_equal(Mb.indices, Ms.indices)
    # same J, same QR, same solves: identical values
    assert np.array_equal(Mb.data, Ms.data)
    assert np.allclose(out["res_norm"], ref["res_norm"], rtol=1e-14, atol=0)
    if adaptive:
        assert out["nsolve"].max() > 1


def test_full_pattern_gives_inverse(blk):
    A = gen.block_convdiff2d(3, 2, seed=3)
    for sizes in (blk.constant_blocks(A.shape[0], 2), blk.constant_blocks(A.shape[0], 4)):
        B = blk.BlockMatrix(A, sizes)
        full = sp.csc_matrix(np.ones((B.nb, B.nb)))
        out = blk.spai_block(B, full, max_impr_steps=0)
        M = blk.to_scalar(B, out).toarray()
        Ainv = np.linalg.inv(A.toarray())
        assert np.allclose(M, Ainv, rtol=1e-10, atol=1e-12 * np.abs(Ainv).max())
        assert out["res_norm"].max() < 1e-12


def test_rho_against_brute_force(blk):
    A = gen.block_convdiff2d(6, 3, seed=4)
    Ad = A.toarray()
    for sizes in (blk.constant_blocks(A.shape[0], 3), blk.constant_blocks(A.shape[0], 2)):
        B = blk.BlockMatrix(A, sizes)
        for col, J in ((7, [7]), (14, [8, 14, 20]), (B.nb - 1, [B.nb - 1])):
            rn, cand, rho = blk.rhos(B, col, J)
            assert len(cand) > 0
            sc = np.concatenate([np.arange(B.start[j], B.start[j + 1]) for j in J])
            E = np.zeros((A.shape[0], B.bsize[col]))
            E[B.start[col]:B.start[col + 1], :] = np.eye(B.bsize[col])
            X = np.linalg.lstsq(Ad[:, sc], E, rcond=None)[0]
            R = E - Ad[:, sc] @ X
            assert rn == pytest.approx(np.linalg.norm(R), rel=1e-10)
            for j, r in zip(cand, rho):
                Aj = Ad[:, B.start[j]:B.start[j + 1]]
                Y = np.linalg.lstsq(Aj, R, rcond=None)[0]
                ref = np.linalg.norm(R - Aj @ Y)
                assert r == pytest.approx(ref, rel=1e-7, abs=1e-9 * rn), (col, j)


@pytest.mark.parametrize("bs", [2, 3, 0])
def test_block_residuals_match_scipy(blk, bs):
    A = gen.block_convdiff2d(8, 3, seed=5)
    sizes = blk.diagonal_blocks(A) if bs == 0 else blk.constant_blocks(A.shape[0], bs)
    B = blk.BlockMatrix(A, sizes)
    out = blk.spai_block(B, None, epsilon=0.3, maxnew=3, max_impr_steps=3)
    M = blk.to_scalar(B, out)
    R = (A @ M - sp.identity(A.shape[0])).tocsc()
    for k in range(B.nb):
        cols = slice(B.start[k], B.start[k + 1])
        assert out["res_norm"][k] == pytest.approx(sp.linalg.norm(R[:, cols]), rel=1e-9, abs=1e-13)
    # augmentation ran and improved on the static block-diagonal start
    static = blk.to_scalar(B, blk.spai_block(B, None, max_impr_steps=0))
    assert sp.linalg.norm(A @ M - sp.identity(A.shape[0])) < sp.linalg.norm(A @ static - sp.identity(A.shape[0]))


def test_library_block_detection_matches_oracle(blk):
    """The library's host-side block partition (no device needed)."""
    import mspai_amd
    for A in (gen.block_convdiff2d(8, 3, seed=5), gen.convdiff2d(10), _blockdiag([3, 2, 1, 4], [(9, 0), (0, 7)])):
        for bs in (0, 1, 2, 3, 4):
            ref = blk.diagonal_blocks(A) if bs == 0 else blk.constant_blocks(A.shape[0], bs)
            assert np.array_equal(mspai_amd.block_sizes(A, bs), ref), bs


def test_zero_column_flags(blk, oracle):
    """A zero scalar column: its block column's solve reports the zero pivot
    (dtrtrs leaves Q^T E), and every block column that scores it as a
    candidate gets a NaN mean, so nothing is added there; M stays finite."""
    A = gen.block_convdiff2d(5, 3, seed=16).tolil()
    A[:, 7] = 0.0
    A = sp.csc_matrix(A)
    A.eliminate_zeros()
    B = blk.BlockMatrix(A, blk.constant_blocks(A.shape[0], 3))
    out = blk.spai_block(B, None, epsilon=0.3, maxnew=3, max_impr_steps=3)
    assert out["status"][2] & oracle.ST_SINGULAR
    nan_cols = np.nonzero(out["status"] & oracle.ST_NAN_RHO)[0]
    assert len(nan_cols) > 0
    # with one augmentation step block 2 can only be a candidate of columns
    # within two steps of it in the block graph (U is the union of the block
    # rows' patterns over I u {k}, I the neighbours of k)
    one = blk.spai_block(B, None, epsilon=0.3, maxnew=3, max_impr_steps=1)
    P = (B.pattern().toarray() != 0).astype(int)
    reach = (P + P @ P) != 0
    flagged = np.nonzero(one["status"] & oracle.ST_NAN_RHO)[0]
    assert len(flagged) > 0 and all(reach[2, c] for c in flagged)
    assert np.isfinite(blk.to_scalar(B, out).data).all()


def test_library_block_sizes_edges(blk):
    """Constant blocks larger than n, n = 1, and a diagonal without runs."""
    import mspai_amd
    assert mspai_amd.block_sizes(sp.identity(5, format="csc"), 7).tolist() == [5]
    assert mspai_amd.block_sizes(sp.identity(1, format="csc"), 3).tolist() == [1]
    assert mspai_amd.block_sizes(sp.identity(6, format="csc"), 0).tolist() == [1] * 6
    assert mspai_amd.block_sizes(sp.csc_matrix(np.ones((4, 4))), 0).tolist() == [4]
    with pytest.raises(mspai_amd.MspaiError):
        mspai_amd.block_sizes(sp.identity(4, format="csc"), -1)
