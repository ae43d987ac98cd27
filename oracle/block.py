"""Block SPAI oracle: block structure in numpy, column solves in liboracle.so.

TEST INFRASTRUCTURE ONLY: the checker for the GPU library's block path.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this.

Restates the reference's conversions around block SPAI (-bs):
  Find_Constant_Blocks        src/Matrix.imp:1345-1363
  Find_Diagonal_Blocks,
  Find_Diagonal_Block,
  Initial_Run_Length,
  Check_Next_Run              src/Matrix.imp:1381-1514
  Convert_To_Block_Matrix     src/Matrix.imp:1516-1745 (values and row structure)
  Scalar_Matrix               src/Matrix.imp (block -> scalar, zeros dropped)
The column solves are orc_spai_block (spai_block_oracle.cpp).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import oracle as _orc


def constant_blocks(n: int, bs: int) -> np.ndarray:
    """Find_Constant_Blocks: blocks of bs, the last one holds the remainder."""
    nb, odd = divmod(n, bs)
    sizes = [bs] * nb
    if odd:
        sizes.append(odd)
    return np.asarray(sizes, dtype=np.int32)


def _col_rows(A: sp.csc_matrix, j: int) -> np.ndarray:
    return A.indices[A.indptr[j]:A.indptr[j + 1]]


def _initial_run_length(A, i):
    rows = _col_rows(A, i)
    hit = np.nonzero(rows == i)[0]
    if len(hit) == 0:
        return 1
    bs, ii = 1, i + 1
    for r in rows[hit[0] + 1:]:
        if r == ii:
            bs += 1
            ii += 1
        else:
            return bs
    return bs


def _check_next_run(A, i, i_next, bs):
    rows = _col_rows(A, i_next)
    hit = np.nonzero(rows == i)[0]
    if len(hit) == 0:
        return 0
    new_bs, ii = 1, i + 1
    for r in rows[hit[0] + 1:]:
        if r == ii:
            new_bs += 1
            ii += 1
        else:
            break
    if new_bs > i_next - i:
        return bs if new_bs > bs else new_bs
    return 0


def _find_diagonal_block(A, i, limit):
    max_bs = _initial_run_length(A, i)
    bs = max_bs
    if max_bs == 1:
        return i
    if limit and max_bs > limit:
        bs = max_bs = limit
    n = A.shape[1]
    for k in range(max_bs):
        if k > 0 and i + k == n:
            return i + k - 1
        bs = _check_next_run(A, i, i + k, bs)
        if not bs:
            return i + k - 1
    return i + max_bs - 1


def diagonal_blocks(A, upper_limit: int = 1000000) -> np.ndarray:
    """Find_Diagonal_Blocks: variable blocks from the runs of consecutive rows
    that start at each block's first diagonal entry (-bs 0)."""
    A = sp.csc_matrix(A)
    A.sort_indices()
    sizes = []
    i, n = 0, A.shape[1]
    while i < n:
        iend = _find_diagonal_block(A, i, upper_limit)
        sizes.append(iend - i + 1)
        i = iend + 1
    return np.asarray(sizes, dtype=np.int32)


class BlockMatrix:
    """Convert_To_Block_Matrix: block CSC with column-major blocks (stored
    entries of a block that A does not hold are zeros) and the block row
    structure (block CSR pattern)."""

    def __init__(self, A, sizes):
        A = sp.csc_matrix(A)
        A.sort_indices()
        sizes = np.asarray(sizes, dtype=np.int32)
        n = A.shape[1]
        assert A.shape[0] == n and int(sizes.sum()) == n
        self.n, self.nb, self.bsize = n, len(sizes), sizes
        self.start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        bmap = np.repeat(np.arange(self.nb), sizes)
        C = A.tocoo()
        bi, bj = bmap[C.row], bmap[C.col]
        # stored blocks: distinct (block column, block row) pairs, column-major order
        key = bj.astype(np.int64) * self.nb + bi
        ukey, which = np.unique(key, return_inverse=True)
        ub_col, ub_row = ukey // self.nb, ukey % self.nb
        h, w = sizes[ub_row].astype(np.int64), sizes[ub_col].astype(np.int64)
        self.voff = np.concatenate([[0], np.cumsum(h * w)]).astype(np.int64)
        val = np.zeros(int(self.voff[-1]))
        pos = self.voff[which] + (C.col - self.start[bj]) * h[which] + (C.row - self.start[bi])
        val[pos] = C.data
        self.val = val
        self.rowind = ub_row.astype(np.int32)
        self.colptr = np.searchsorted(ub_col, np.arange(self.nb + 1)).astype(np.int32)
        pat = sp.csc_matrix((np.ones(len(self.rowind)), self.rowind, self.colptr), shape=(self.nb, self.nb))
        csr = pat.tocsr()
        csr.sort_indices()
        self.rp_ptr = np.asarray(csr.indptr, np.int32)
        self.rp_ind = np.asarray(csr.indices, np.int32)

    def pattern(self) -> sp.csc_matrix:
        """Block pattern as an nb x nb CSC pattern matrix."""
        return sp.csc_matrix((np.ones(len(self.rowind)), self.rowind, self.colptr), shape=(self.nb, self.nb))

    def struct(self, keep):
        arrs = [np.ascontiguousarray(self.bsize, np.int32), self.colptr, self.rowind, self.voff,
                np.ascontiguousarray(self.val, np.float64)]
        keep.extend(arrs)
        return _BCsc(self.nb, *[C.c_void_p(a.ctypes.data) for a in arrs])


class _BCsc(C.Structure):
    _fields_ = [("nb", C.c_int), ("bsize", C.c_void_p), ("colptr", C.c_void_p), ("rowind", C.c_void_p),
                ("voff", C.c_void_p), ("val", C.c_void_p)]


def block_pattern_start(B: BlockMatrix, P=None):
    """Start pattern over block columns: None = block diagonal."""
    if P is None:
        P = sp.identity(B.nb, format="csc")
    P = sp.csc_matrix(P)
    P.sort_indices()
    return np.asarray(P.indptr, np.int32), np.asarray(P.indices, np.int32)


def capacities(B: BlockMatrix, p_ptr, maxnew, ns):
    plen = int(np.max(np.diff(p_ptr))) if len(p_ptr) > 1 else 1
    capJ = min(B.nb, plen + maxnew * max(0, ns))
    maxbs = int(B.bsize.max())
    return capJ, capJ * maxbs * maxbs


def spai_block(B: BlockMatrix, P=None, epsilon=0.4, maxnew=5, max_impr_steps=5, use_mean=1, threads=1):
    """Block SPAI on the oracle.  Returns (per-block-column lists, info)."""
    lib = _orc.lib()
    p_ptr, p_ind = block_pattern_start(B, P)
    capJ, capX = capacities(B, p_ptr, maxnew, max_impr_steps)
    keep = []
    bs = B.struct(keep)
    o = _orc._Opts(epsilon, maxnew, max_impr_steps, use_mean, 0, 0, threads)
    nb = B.nb
    m_len = np.zeros(nb, np.int32)
    m_idx = np.zeros(nb * capJ, np.int32)
    m_val = np.zeros(nb * capX, np.float64)
    res = np.zeros(nb)
    ns_ = np.zeros(nb, np.int32)
    st = np.zeros(nb, np.int32)
    p = _orc._p
    rc = lib.orc_spai_block(C.byref(bs), p(B.rp_ptr), p(B.rp_ind), p(p_ptr), p(p_ind), C.byref(o),
                            C.c_int(capJ), C.c_longlong(capX), p(m_len), p(m_idx), p(m_val), p(res),
                            p(ns_), p(st))
    if rc != 0:
        raise RuntimeError(f"orc_spai_block failed ({rc})")
    out = {"len": m_len, "idx": m_idx.reshape(nb, capJ), "val": m_val.reshape(nb, capX),
           "res_norm": res, "nsolve": ns_, "status": st}
    return out


def to_scalar(B: BlockMatrix, out) -> sp.csc_matrix:
    """Scalar_Matrix: scalar CSC of the block result, exact zeros dropped."""
    rows, cols, vals = [], [], []
    for col in range(B.nb):
        bk = int(B.bsize[col])
        js = out["idx"][col, :out["len"][col]]
        if len(js) == 0:
            continue
        bj = B.bsize[js].astype(np.int64)
        nv = int((bj * bk).sum())
        v = out["val"][col, :nv]
        # block t holds bj x bk values, column-major: entry (jj, q) at q*bj + jj
        tb = np.repeat(np.arange(len(js)), bj * bk)
        local = np.arange(nv) - np.repeat(np.concatenate([[0], np.cumsum(bj * bk)[:-1]]), bj * bk)
        q, jj = local // bj[tb], local % bj[tb]
        keep = v != 0.0
        rows.append((B.start[js][tb] + jj)[keep])
        cols.append((B.start[col] + q)[keep])
        vals.append(v[keep])
    if rows:
        rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    M = sp.csc_matrix((vals, (rows, cols)), shape=(B.n, B.n))
    M.sort_indices()
    return M


def sq_inverse(B: BlockMatrix) -> np.ndarray:
    lib = _orc.lib()
    keep = []
    bs = B.struct(keep)
    out = np.zeros(int((B.bsize.astype(np.int64) ** 2).sum()))
    lib.orc_block_sq_inverse(C.byref(bs), _orc._p(out))
    return out


def rhos(B: BlockMatrix, col, J):
    """Residual norm of the static solve on J and the rho of every candidate."""
    lib = _orc.lib()
    keep = []
    bs = B.struct(keep)
    J = np.ascontiguousarray(sorted(J), np.int32)
    cap = B.nb
    cand = np.zeros(cap, np.int32)
    rho = np.zeros(cap)
    rn = C.c_double()
    nc = C.c_int()
    rc = lib.orc_block_rhos(C.byref(bs), _orc._p(B.rp_ptr), _orc._p(B.rp_ind), C.c_int(col), _orc._p(J),
                            C.c_int(len(J)), C.byref(rn), C.byref(nc), _orc._p(cand), _orc._p(rho),
                            C.c_int(cap))
    if rc != 0:
        raise RuntimeError(f"orc_block_rhos failed ({rc})")
    return rn.value, cand[:nc.value], rho[:nc.value]
