"""ctypes front end of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

NWORK = 8
ST_SINGULAR, ST_UNDERDET, ST_NAN_RHO, ST_K_NOT_IN_I = 1, 2, 4, 16


class _Opts(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("maxnew", C.c_int), ("max_impr_steps", C.c_int),
                ("use_mean", C.c_int), ("qr", C.c_int), ("cache", C.c_int), ("threads", C.c_int),
                ("up_ptr", C.c_void_p), ("up_ind", C.c_void_p)]


class _Csc(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("colptr", C.c_void_p),
                ("rowind", C.c_void_p), ("val", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.orc_spai.restype = C.c_int
        _lib.orc_ls_solve.restype = C.c_int
        _lib.orc_lane_dot.restype = C.c_double
        _lib.orc_lane_dot.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        _lib.orc_spai_backward.restype = C.c_int
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _csc_struct(M, keep):
    ip = np.ascontiguousarray(M.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(M.indices, dtype=np.int32)
    v = np.ascontiguousarray(M.data, dtype=np.float64)
    keep += [ip, ix, v]
    return _Csc(M.shape[0], M.shape[1], ip.ctypes.data, ix.ctypes.data, v.ctypes.data)


def capacity(prob) -> int:
    plen = np.diff(prob.P.indptr)
    return int(plen.max()) + prob.maxnew * max(0, prob.max_impr_steps)


def spai(prob, cols=None, threads: int = 1, qr: int = 0, cache: int = 0, epsilon=None,
         maxnew=None, max_impr_steps=None, use_mean=None, up=None):
    """Run the oracle on `prob` (a mspai_amd.gen.Problem); returns a dict with
    len, idx, val (padded to `cap`), res_norm, nsolve, status, work, cap, and
    near (per column: near-margin decisions at the mean, at the maxnew cut,
    at epsilon; fresh-QR path only).  `up`: an upper pattern (a sparse
    matrix with A's column count; its stored positions), which restricts
    every column's candidates to its column (-up)."""
    L = lib()
    keep = []
    A = _csc_struct(prob.A, keep)
    B = _csc_struct(prob.B, keep) if prob.B is not None else None
    rp_ptr, rp_ind = prob.row_pattern()
    p_ptr = np.ascontiguousarray(prob.P.indptr, dtype=np.int32)
    p_ind = np.ascontiguousarray(prob.P.indices, dtype=np.int32)
    o = _Opts(prob.epsilon if epsilon is None else epsilon,
              prob.maxnew if maxnew is None else maxnew,
              prob.max_impr_steps if max_impr_steps is None else max_impr_steps,
              prob.use_mean if use_mean is None else use_mean, qr, cache, threads)
    if up is not None:
        import scipy.sparse as sp
        U = sp.csc_matrix(up)
        U.sort_indices()
        if U.shape[1] != prob.A.shape[1]:
            raise ValueError("upper pattern needs A's column count")
        up_ptr = np.ascontiguousarray(U.indptr, dtype=np.int32)
        up_ind = np.ascontiguousarray(U.indices, dtype=np.int32)
        keep += [up_ptr, up_ind]
        o.up_ptr, o.up_ind = up_ptr.ctypes.data, up_ind.ctypes.data
    ns = o.max_impr_steps
    plen = np.diff(prob.P.indptr)
    cap = int(plen.max()) + o.maxnew * max(0, ns)
    if cols is None:
        ncl = prob.n
        colsarr = None
    else:
        colsarr = np.ascontiguousarray(cols, dtype=np.int32)
        ncl = colsarr.size
    m_len = np.zeros(ncl, np.int32)
    m_idx = np.full(ncl * cap, -1, np.int32)
    m_val = np.zeros(ncl * cap, np.float64)
    res = np.zeros(ncl, np.float64)
    nsolve = np.zeros(ncl, np.int32)
    status = np.zeros(ncl, np.int32)
    work = np.zeros(ncl * NWORK, np.float64)
    near = np.zeros(ncl * 3, np.int32)
    rc = L.orc_spai(C.byref(A), _p(rp_ptr), _p(rp_ind), _p(p_ptr), _p(p_ind),
                    C.byref(B) if B is not None else None, C.byref(o), _p(colsarr), C.c_int(ncl),
                    C.c_int(cap), _p(m_len), _p(m_idx), _p(m_val), _p(res), _p(nsolve), _p(status),
                    _p(work), _p(near))
    if rc != 0:
        raise RuntimeError(f"orc_spai failed: {rc}")
    return {"len": m_len, "idx": m_idx.reshape(ncl, cap), "val": m_val.reshape(ncl, cap),
            "res_norm": res, "nsolve": nsolve, "status": status,
            "work": work.reshape(ncl, NWORK), "near": near.reshape(ncl, 3), "cap": cap}


def to_csc(out, nrows):
    """Pack the padded per-column output into a scipy CSC matrix."""
    import scipy.sparse as sp
    ln = out["len"].astype(np.int64)
    ptr = np.concatenate([[0], np.cumsum(ln)]).astype(np.int32)
    mask = np.arange(out["cap"])[None, :] < ln[:, None]
    return sp.csc_matrix((out["val"][mask], out["idx"][mask], ptr), shape=(nrows, ln.size))


def ls_solve(A, b):
    """Dense least squares in the oracle's order. Returns (x, R-and-v, tau, status)."""
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m, n = A.shape
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    tau = np.zeros(max(1, min(m, n)), np.float64)
    bb = np.zeros(max(m, n), np.float64)
    bb[:m] = b
    st = lib().orc_ls_solve(C.c_int(m), C.c_int(n), _p(A), _p(tau), _p(bb))
    return bb[:n].copy(), A, tau, st


def lane_dot(x, y, r0=0, r1=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    if r1 is None:
        r1 = x.size
    return lib().orc_lane_dot(_p(x), _p(y), r0, r1)


def sort_rho(idx, rho):
    idx = np.ascontiguousarray(idx, dtype=np.int32).copy()
    rho = np.ascontiguousarray(rho, dtype=np.float64).copy()
    lib().orc_sort_rho(C.c_int(idx.size), _p(idx), _p(rho))
    return idx, rho


def backward(prob, M, gM, threads: int = 1):
    """CPU FP64 reference gradient of a static solve. Returns (gA, gB)."""
    keep = []
    A = _csc_struct(prob.A, keep)
    B = _csc_struct(prob.B, keep) if prob.B is not None else None
    mp = np.ascontiguousarray(M.indptr, dtype=np.int32)
    mi = np.ascontiguousarray(M.indices, dtype=np.int32)
    mv = np.ascontiguousarray(M.data, dtype=np.float64)
    g = np.ascontiguousarray(gM, dtype=np.float64)
    gA = np.zeros(prob.A.nnz, np.float64)
    gB = np.zeros(prob.B.nnz, np.float64) if prob.B is not None else None
    rc = lib().orc_spai_backward(C.byref(A), C.byref(B) if B is not None else None, _p(mp), _p(mi),
                                 _p(mv), _p(g), _p(gA), _p(gB) if gB is not None else None,
                                 C.c_int(threads))
    if rc != 0:
        raise RuntimeError(f"orc_spai_backward failed: {rc}")
    return gA, gB
