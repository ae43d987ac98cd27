"""PyTorch-ROCm harness for the adjoint of the SPAI build (test use).

`SpaiFunction` maps the values of C (and of the target B) to the values of M
with the pattern of the last build held fixed; its backward calls the
library's mspai_backward, which recomputes each column's QR and returns
dL/dC and dL/dB.  Tensors stay on the GPU; the library reads and writes their
storage through the device-pointer entry points.

torch is imported before the library is loaded so that one HIP runtime
serves both (a process that loads the library first and torch second ends up
with two runtimes on one device).

Also here: the MSPAI probing composition C = [A; rho e^T A],
B = [I; rho e^T] (the row layout of src/Read_mm_Matrix.cpp:532-667) as index
maps, the variable-coefficient operator A(kappa) of the FV workload, and the
loss L = ||A M 1 - 1||^2 + ||A M - I||_F^2.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

import mspai_amd as _m


class SpaiFunction(torch.autograd.Function):
    """m_vals = build(c_vals, b_vals).  The builder must already hold the
    pattern of C, the start pattern, and the target pattern (if any).  The
    backward uses the builder's last M, so forward/backward pairs must not be
    interleaved with other builds on the same builder."""

    @staticmethod
    def forward(ctx, c_vals, b_vals, builder, opts):
        torch.cuda.synchronize()
        builder.update_values_device(c_vals.data_ptr())
        if b_vals is not None:
            builder.update_target_device(b_vals.data_ptr())
        st = builder.build(**opts)
        m = torch.empty(st["nnz"], dtype=torch.float64, device=c_vals.device)
        builder.result_values_device(m.data_ptr())
        ctx.builder = builder
        ctx.opts = opts
        ctx.shapes = (c_vals.numel(), b_vals.numel() if b_vals is not None else 0)
        ctx.fwd_ms = st["build_ms"]
        return m

    @staticmethod
    def backward(ctx, gm):
        gm = gm.contiguous().to(torch.float64)
        nc, nb = ctx.shapes
        gc = torch.empty(nc, dtype=torch.float64, device=gm.device)
        gb = torch.empty(nb, dtype=torch.float64, device=gm.device) if nb else None
        torch.cuda.synchronize()
        ctx.builder.backward_device(gm.data_ptr(), gc.data_ptr(), gb.data_ptr() if gb is not None else None,
                                    precision=ctx.opts.get("precision", 0))
        return gc, gb, None, None


def spai_values(builder, c_vals, b_vals=None, **opts):
    return SpaiFunction.apply(c_vals, b_vals, builder, opts)


class ProbingMaps:
    """Index maps of C = [A; rho (e^T A)] and B = [I; rho e^T] built from A's
    pattern (e = ones), so that c_vals and b_vals are differentiable
    functions of A's values and rho."""

    def __init__(self, A: sp.csc_matrix, device="cuda"):
        A = sp.csc_matrix(A)
        A.sort_indices()
        n = A.shape[1]
        colsum = np.asarray(A.sum(axis=0)).ravel()
        probe_cols = np.nonzero(colsum)[0]
        self.n = n
        self.has_probe = np.zeros(n, bool)
        self.has_probe[probe_cols] = True
        lens = np.diff(A.indptr) + self.has_probe
        cptr = np.concatenate([[0], np.cumsum(lens)])
        a_col = np.repeat(np.arange(n), np.diff(A.indptr))
        a_pos = cptr[a_col] + (np.arange(A.nnz) - A.indptr[a_col])
        p_pos = cptr[probe_cols] + np.diff(A.indptr)[probe_cols]
        cind = np.empty(cptr[-1], np.int64)
        cind[a_pos] = A.indices
        cind[p_pos] = n
        self.C_pattern = sp.csc_matrix((np.ones(cptr[-1]), cind, cptr), shape=(n + 1, n))
        self.nnzC = int(cptr[-1])
        t = lambda x: torch.as_tensor(x, dtype=torch.int64, device=device)
        self.a_pos, self.p_pos, self.a_col, self.p_col = t(a_pos), t(p_pos), t(a_col), t(probe_cols)
        # B: column k holds (k, 1) then (n, rho)
        self.B_pattern = sp.csc_matrix((np.ones(2 * n), np.stack([np.arange(n), np.full(n, n)], 1).ravel(),
                                        np.arange(0, 2 * n + 1, 2)), shape=(n + 1, n))

    def c_values(self, a_vals, rho):
        colsum = torch.zeros(self.n, dtype=a_vals.dtype, device=a_vals.device).index_add(0, self.a_col, a_vals)
        c = torch.zeros(self.nnzC, dtype=a_vals.dtype, device=a_vals.device)
        c = c.index_put((self.a_pos,), a_vals)
        c = c.index_put((self.p_pos,), rho * colsum[self.p_col])
        return c

    def b_values(self, rho, dtype=torch.float64, device="cuda"):
        ones = torch.ones(self.n, dtype=dtype, device=device)
        return torch.stack([ones, rho * ones], dim=1).reshape(-1)


class ProductMaps:
    """Index triples of the product A M for fixed patterns of A and M:
    (A entry, M entry, output entry), so that the values of A M are a
    differentiable function of A's and M's values."""

    def __init__(self, A: sp.csc_matrix, M: sp.csc_matrix, device="cuda"):
        A = sp.csc_matrix(A)
        M = sp.csc_matrix(M)
        nrows = A.shape[0]
        mcol = np.repeat(np.arange(M.shape[1]), np.diff(M.indptr))
        j = M.indices.astype(np.int64)
        la = np.diff(A.indptr)[j]
        eM = np.repeat(np.arange(M.nnz), la)
        start = np.repeat(A.indptr[j], la)
        off = np.arange(eM.size) - np.repeat(np.cumsum(la) - la, la)
        eA = start + off
        i = A.indices[eA].astype(np.int64)
        k = mcol[eM].astype(np.int64)
        key = k * nrows + i
        uk, out = np.unique(key, return_inverse=True)
        self.nout = uk.size
        self.out_row = torch.as_tensor(uk % nrows, device=device)
        self.out_col = torch.as_tensor(uk // nrows, device=device)
        self.eA = torch.as_tensor(eA, device=device)
        self.eM = torch.as_tensor(eM, device=device)
        self.out = torch.as_tensor(out, device=device)
        self.nrows = nrows

    def values(self, a_vals, m_vals):
        prod = a_vals[self.eA] * m_vals[self.eM]
        return torch.zeros(self.nout, dtype=prod.dtype, device=prod.device).index_add(0, self.out, prod)


def loss_terms(pm: ProductMaps, a_vals, m_vals, n):
    """Per-term contributions of L = ||A M 1 - 1||^2 + ||A M - I||_F^2 (the
    terms are returned separately so finite differences can cancel the
    unchanged ones exactly)."""
    am = pm.values(a_vals, m_vals)
    diag = (pm.out_row == pm.out_col).to(am.dtype)
    t1 = (am - diag) ** 2
    y = torch.zeros(pm.nrows, dtype=am.dtype, device=am.device).index_add(0, pm.out_row, am)
    t2 = (y[:n] - 1.0) ** 2
    return t1, t2


def varcoef_values(kappa, fa, fb, nbnd, amap):
    """A's values (CSC order) from cell coefficients: off-diagonals
    -2 ka kb/(ka + kb) on interior faces, diagonal = sum of the cell's face
    coefficients + 2 kappa_c per Dirichlet face.  `amap` gives the CSC
    positions (diag_pos, fab_pos, fba_pos)."""
    ka, kb = kappa[fa], kappa[fb]
    t = 2.0 * ka * kb / (ka + kb)
    diag = 2.0 * kappa * nbnd
    diag = diag.index_add(0, fa, t).index_add(0, fb, t)
    n = kappa.numel()
    vals = torch.zeros(amap["nnz"], dtype=kappa.dtype, device=kappa.device)
    vals = vals.index_put((amap["diag_pos"],), diag)
    vals = vals.index_put((amap["fab_pos"],), -t)
    vals = vals.index_put((amap["fba_pos"],), -t)
    return vals


def varcoef_maps(A: sp.csc_matrix, fa, fb, device="cuda"):
    """CSC positions of the diagonal and of the (fa, fb) / (fb, fa) entries."""
    A = sp.csc_matrix(A)
    n = A.shape[1]

    def pos(rows, cols):
        out = np.empty(rows.size, np.int64)
        for lo in range(0, rows.size, 1 << 22):
            r, c = rows[lo:lo + (1 << 22)], cols[lo:lo + (1 << 22)]
            base = A.indptr[c]
            ln = A.indptr[c + 1] - base
            # rows are sorted within each column: search per column segment
            seg = np.zeros(r.size, np.int64)
            for d in range(int(ln.max())):
                hit = (d < ln) & (A.indices[np.minimum(base + d, A.nnz - 1)] == r)
                seg[hit] = d
            out[lo:lo + r.size] = base + seg
        return out
    idx = np.arange(n)
    t = lambda x: torch.as_tensor(x, device=device)
    return {"nnz": A.nnz, "diag_pos": t(pos(idx, idx)), "fab_pos": t(pos(fa, fb)), "fba_pos": t(pos(fb, fa))}
