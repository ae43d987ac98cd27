"""mspai_amd: MI355X builder for SPAI/MSPAI preconditioners.

Python front end of the C ABI in include/mspai_gpu.h (libmspai_gpu.so).  The
library is the product; this module only moves arrays across the boundary.
There is no CPU fallback: if the library is missing or no gfx950 device is
visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmspai_gpu.so")
NWORK = 8

COL_SINGULAR, COL_UNDERDET, COL_NAN_RHO, COL_FALLBACK, COL_K_NOT_IN_I = 1, 2, 4, 8, 16

EXPORTED = ["mspai_default_opts", "mspai_ctx_create", "mspai_ctx_destroy", "mspai_last_error",
            "mspai_upload_csc", "mspai_update_values", "mspai_set_pattern", "mspai_set_upper_pattern",
            "mspai_reuse_pattern",
            "mspai_set_target",
            "mspai_build", "mspai_build_noalloc", "mspai_result_nnz", "mspai_download_csc",
            "mspai_download_column_info", "mspai_work_totals", "mspai_bin_counts", "mspai_launch_stats",
            "mspai_update_values_device", "mspai_update_target_device", "mspai_result_values_device",
            "mspai_backward", "mspai_autotune", "mspai_set_bin_table", "mspai_apply",
            "mspai_set_block_size", "mspai_block_info", "mspai_block_sizes"]


class LaunchStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("ms", C.c_double), ("columns", C.c_double),
                ("work", C.c_double * 8), ("near", C.c_double * 3), ("nnz", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("maxnew", C.c_int), ("max_impr_steps", C.c_int),
                ("use_mean", C.c_int), ("qr", C.c_int), ("cache", C.c_int),
                ("precision", C.c_int), ("binning", C.c_int), ("verbose", C.c_int),
                ("counters", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("solve_ms", C.c_double), ("nnz", C.c_longlong),
                ("ncols", C.c_int), ("n_fallback", C.c_int), ("n_overflow", C.c_int),
                ("n_cache_hits", C.c_int), ("first_bin_ms", C.c_double)]


class TuneReport(C.Structure):
    _fields_ = [("ncandidates", C.c_int), ("chosen", C.c_int), ("chosen_name", C.c_char * 32),
                ("candidate_ms", C.c_double * 32)]


_lib = None


def load():
    """Load libmspai_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(or python mspai_amd/build.py) first")
    lib = C.CDLL(LIB_PATH)
    for name in EXPORTED:
        getattr(lib, name)
    lib.mspai_last_error.restype = C.c_char_p
    lib.mspai_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    _lib = lib
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class MspaiError(RuntimeError):
    pass


class Builder:
    """One device context: upload C, set pattern/target, build, download M."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mspai_ctx_create(device, C.byref(h))
        if rc != 0:
            raise MspaiError(f"mspai_ctx_create failed ({rc}): no gfx950 device / HIP error")
        self.h = h
        self.shape = None

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.mspai_last_error(self.h)
            raise MspaiError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.mspai_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, A: sp.csc_matrix, row_pattern=None):
        A = sp.csc_matrix(A)
        A.sort_indices()
        ip, ix, v = _i32(A.indptr), _i32(A.indices), _f64(A.data)
        rp = ri = None
        if row_pattern is not None:
            rp, ri = _i32(row_pattern[0]), _i32(row_pattern[1])
        self._chk(self.lib.mspai_upload_csc(self.h, A.shape[0], A.shape[1], _p(ip), _p(ix), _p(v),
                                            _p(rp), _p(ri)), "mspai_upload_csc")
        self.shape = A.shape

    def update_values(self, vals):
        v = _f64(vals)
        self._chk(self.lib.mspai_update_values(self.h, _p(v)), "mspai_update_values")

    def set_pattern(self, P=None):
        if P is None:
            self._chk(self.lib.mspai_set_pattern(self.h, None, None), "mspai_set_pattern")
            return
        P = sp.csc_matrix(P)
        P.sort_indices()
        self._chk(self.lib.mspai_set_pattern(self.h, _p(_i32(P.indptr)), _p(_i32(P.indices))),
                  "mspai_set_pattern")

    def set_upper_pattern(self, UP=None):
        """Upper pattern (-up): column k's candidates are kept only where
        UP(:, k) has a stored entry.  None removes it."""
        if UP is None:
            self._chk(self.lib.mspai_set_upper_pattern(self.h, None, None), "mspai_set_upper_pattern")
            return
        U = sp.csc_matrix(UP)
        U.sort_indices()
        U.sum_duplicates()
        if U.shape[1] != self.shape[1]:
            raise ValueError("the upper pattern needs the matrix's column count")
        self._chk(self.lib.mspai_set_upper_pattern(self.h, _p(_i32(U.indptr)), _p(_i32(U.indices))),
                  "mspai_set_upper_pattern")

    def reuse_pattern(self):
        """Start pattern = the last built M's pattern (the reference's P_Memory
        re-solve between pattern recomputations)."""
        self._chk(self.lib.mspai_reuse_pattern(self.h), "mspai_reuse_pattern")

    def set_block_size(self, bs: int):
        """-bs: bs > 1 constant blocks, 0 variable blocks, 1 scalar SPAI.  In
        block mode set_pattern takes a pattern over block columns (None =
        block diagonal) and build() runs block SPAI; M stays scalar."""
        self._chk(self.lib.mspai_set_block_size(self.h, int(bs)), "mspai_set_block_size")

    def block_info(self):
        """(block sizes, block pattern of C as an nb x nb CSC pattern)."""
        nb = C.c_int()
        nnzb = C.c_longlong()
        self._chk(self.lib.mspai_block_info(self.h, C.byref(nb), C.byref(nnzb), None, None, None),
                  "mspai_block_info")
        sizes = np.zeros(nb.value, np.int32)
        ptr = np.zeros(nb.value + 1, np.int32)
        ind = np.zeros(max(1, nnzb.value), np.int32)
        self._chk(self.lib.mspai_block_info(self.h, None, None, _p(sizes), _p(ptr), _p(ind)), "mspai_block_info")
        P = sp.csc_matrix((np.ones(nnzb.value), ind[:nnzb.value], ptr), shape=(nb.value, nb.value))
        return sizes, P

    def set_target(self, B=None):
        if B is None:
            self._chk(self.lib.mspai_set_target(self.h, None, None, None), "mspai_set_target")
            return
        B = sp.csc_matrix(B)
        B.sort_indices()
        self._chk(self.lib.mspai_set_target(self.h, _p(_i32(B.indptr)), _p(_i32(B.indices)),
                                            _p(_f64(B.data))), "mspai_set_target")

    def build(self, epsilon=0.4, maxnew=5, max_impr_steps=5, use_mean=1, qr=0, cache=0,
              precision=0, binning=1, verbose=0, counters=0):
        """counters=1 stores the 8 work counters of every column (work_totals,
        column_info()["work"]); off by default, as in timed builds."""
        o = Opts(epsilon, maxnew, max_impr_steps, use_mean, qr, cache, precision, binning, verbose, counters)
        st = Stats()
        self._chk(self.lib.mspai_build(self.h, C.byref(o), C.byref(st)), "mspai_build")
        self.counters = bool(counters)
        return {"build_ms": st.build_ms, "solve_ms": st.solve_ms, "nnz": st.nnz,
                "ncols": st.ncols, "n_fallback": st.n_fallback, "n_overflow": st.n_overflow,
                "n_cache_hits": st.n_cache_hits, "first_bin_ms": st.first_bin_ms}

    def build_noalloc(self, colptr_ptr: int, rowind_ptr: int, val_ptr: int, max_nnz: int, epsilon=0.4, maxnew=5,
                      max_impr_steps=5, use_mean=1, qr=0, precision=0, binning=1):
        """Build M into caller-owned device arrays (colptr: ncols + 1 int32;
        rowind/val: max_nnz int32/float64, max_nnz >= ncols * (max |J0| +
        maxnew * max_impr_steps)) without any device allocation: the
        library's static scratch array holds the rest."""
        o = Opts(epsilon, maxnew, max_impr_steps, use_mean, qr, 0, precision, binning, 0, 0)
        st = Stats()
        self._chk(self.lib.mspai_build_noalloc(self.h, C.byref(o), C.c_void_p(colptr_ptr), C.c_void_p(rowind_ptr),
                                               C.c_void_p(val_ptr), C.c_longlong(max_nnz), C.byref(st)),
                  "mspai_build_noalloc")
        return {"build_ms": st.build_ms, "solve_ms": st.solve_ms, "nnz": st.nnz, "n_fallback": st.n_fallback}

    # ---- device-pointer entry points (addresses of device buffers, e.g. tensor storage)
    def update_values_device(self, ptr: int):
        self._chk(self.lib.mspai_update_values_device(self.h, C.c_void_p(ptr)), "mspai_update_values_device")

    def update_target_device(self, ptr: int):
        self._chk(self.lib.mspai_update_target_device(self.h, C.c_void_p(ptr)), "mspai_update_target_device")

    def apply_device(self, which: int, transpose: bool, x_ptr: int, y_ptr: int, stream_ptr: int = 0):
        """y = op(A) x (which 0) or op(M) x (which 1) on device vectors; with a
        stream handle the product is queued on that stream, else it runs on
        the context's stream and returns when y is written."""
        self._chk(self.lib.mspai_apply(self.h, int(which), int(bool(transpose)), C.c_void_p(x_ptr),
                                       C.c_void_p(y_ptr), C.c_void_p(stream_ptr) if stream_ptr else None),
                  "mspai_apply")

    def result_values_device(self, ptr: int):
        self._chk(self.lib.mspai_result_values_device(self.h, C.c_void_p(ptr)), "mspai_result_values_device")

    def backward_device(self, gm_ptr: int, gc_ptr: int, gb_ptr=None, precision=0, binning=1):
        o = Opts(0.0, 0, 0, 1, 0, 0, precision, binning, 0)
        st = Stats()
        self._chk(self.lib.mspai_backward(self.h, C.byref(o), C.c_void_p(gm_ptr), C.c_void_p(gc_ptr),
                                          C.c_void_p(gb_ptr) if gb_ptr else None, C.byref(st)),
                  "mspai_backward")
        return {"build_ms": st.build_ms, "solve_ms": st.solve_ms, "n_fallback": st.n_fallback,
                "n_overflow": st.n_overflow}

    def result_nnz(self) -> int:
        nnz = C.c_longlong()
        self._chk(self.lib.mspai_result_nnz(self.h, C.byref(nnz)), "mspai_result_nnz")
        return nnz.value

    def autotune(self, repeats=2, **build_opts):
        kw = dict(epsilon=0.4, maxnew=5, max_impr_steps=5, use_mean=1, qr=0, cache=0, precision=0)
        kw.update(build_opts)
        o = Opts(kw["epsilon"], kw["maxnew"], kw["max_impr_steps"], kw["use_mean"], kw["qr"], kw["cache"],
                 kw["precision"], 2, 0)
        r = TuneReport()
        self._chk(self.lib.mspai_autotune(self.h, C.byref(o), repeats, C.byref(r)), "mspai_autotune")
        return {"chosen": r.chosen, "name": r.chosen_name.decode(),
                "candidate_ms": list(r.candidate_ms[:r.ncandidates])}

    def set_bin_table(self, table, precision=0):
        """Pin the bin table of builds with binning=2 for `precision`; returns
        its name. table=-1 returns the number of tables instead."""
        name = C.create_string_buffer(32)
        cnt = C.c_int()
        self._chk(self.lib.mspai_set_bin_table(self.h, precision, table, name, C.byref(cnt)),
                  "mspai_set_bin_table")
        return cnt.value if table < 0 else name.value.decode()

    def work_totals(self):
        w = np.zeros(NWORK, np.float64)
        self._chk(self.lib.mspai_work_totals(self.h, _p(w)), "mspai_work_totals")
        return w

    def launch_stats(self):
        """Per column-kernel launch of the last build: name, device ms (HIP
        events), and with counters on the columns it finished, their work
        counter sums and near-margin decision counts (else -1)."""
        buf = (LaunchStat * 16)()
        n = C.c_int()
        self._chk(self.lib.mspai_launch_stats(self.h, 16, C.byref(n), buf), "mspai_launch_stats")
        return [{"name": buf[k].name.decode(), "ms": buf[k].ms, "columns": buf[k].columns,
                 "work": list(buf[k].work), "near": list(buf[k].near), "nnz": buf[k].nnz}
                for k in range(min(n.value, 16))]

    def bin_counts(self):
        nb = C.c_int()
        routed = np.zeros(8, np.int32)
        queued = np.zeros(8, np.int32)
        self._chk(self.lib.mspai_bin_counts(self.h, C.byref(nb), _p(routed), _p(queued)), "mspai_bin_counts")
        return {"nbins": nb.value, "routed": routed[:nb.value + 1].tolist(),
                "queued": queued[:nb.value + 1].tolist()}

    def download(self) -> sp.csc_matrix:
        nnz = C.c_longlong()
        self._chk(self.lib.mspai_result_nnz(self.h, C.byref(nnz)), "mspai_result_nnz")
        n = self.shape[1]
        ptr = np.zeros(n + 1, np.int32)
        ind = np.zeros(max(1, nnz.value), np.int32)
        val = np.zeros(max(1, nnz.value), np.float64)
        self._chk(self.lib.mspai_download_csc(self.h, _p(ptr), _p(ind), _p(val)), "mspai_download_csc")
        return sp.csc_matrix((val[:nnz.value], ind[:nnz.value], ptr), shape=(n, n))

    def column_info(self):
        n = self.shape[1]
        res = np.zeros(n, np.float64)
        ns = np.zeros(n, np.int32)
        stt = np.zeros(n, np.int32)
        wk = np.zeros(n * NWORK, np.float64) if getattr(self, "counters", False) else None
        self._chk(self.lib.mspai_download_column_info(self.h, _p(res), _p(ns), _p(stt), _p(wk)),
                  "mspai_download_column_info")
        return {"res_norm": res, "nsolve": ns, "status": stt,
                "work": wk.reshape(n, NWORK) if wk is not None else None}


def block_sizes(A, bs: int) -> np.ndarray:
    """Block partition of the square CSC matrix A (host only, no device):
    bs > 0 constant blocks, bs = 0 the variable blocks of the diagonal runs."""
    lib = load()
    A = sp.csc_matrix(A)
    A.sort_indices()
    n = A.shape[1]
    sizes = np.zeros(n, np.int32)
    nb = C.c_int()
    ip, ix = _i32(A.indptr), _i32(A.indices)
    rc = lib.mspai_block_sizes(C.c_int(n), _p(ip), _p(ix), C.c_int(bs), _p(sizes), C.byref(nb))
    if rc != 0:
        raise MspaiError(f"mspai_block_sizes failed ({rc})")
    return sizes[:nb.value].copy()


def spai(prob, device: int = 0, **opts):
    """Build M for a mspai_amd.gen.Problem on the GPU (work counters on
    unless counters=0 is passed).
    Returns (M as scipy CSC, per-column info, build stats)."""
    b = Builder(device)
    try:
        b.upload(prob.A, prob.row_pattern())
        b.set_pattern(prob.P)
        b.set_target(prob.B)
        kw = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                  use_mean=prob.use_mean, counters=1)
        kw.update(opts)
        stats = b.build(**kw)
        M = b.download()
        info = b.column_info()
        stats["work"] = b.work_totals() if kw["counters"] else None
        stats["bins"] = b.bin_counts()
    finally:
        b.close()
    return M, info, stats
