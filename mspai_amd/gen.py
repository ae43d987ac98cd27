"""Synthetic inputs for the MSPAI column-solve workloads.

Every matrix is returned in compressed sparse column form with sorted row
indices per column, the layout of the reference's ``Compressed_Lines``
(src/Compressed_Lines.h:52-101): column ``c`` of the matrix is column ``c`` of
A, so the preconditioner satisfies A M ~ I (the ``-lp 0`` / MatrixMarket
semantics of src/imspai.cpp:694-697).

Workloads (all sizes are parameters; the defaults are the full sizes):

* ``laplace2d``     2D 5-point Laplacian, diagonal 4, off-diagonals -1,
                    Dirichlet boundary, lexicographic order (x fastest).
* ``convdiff2d``    2D upwind convection-diffusion, D=1, velocity
                    50 (cos t, sin t), t = 0.3, h = 1/(nx+1).
* ``random_band``   random nonsymmetric: per row the diagonal plus 11
                    distinct off-diagonal columns uniform within +-2000,
                    values N(0,1), diagonal 1.2 x the row's absolute sum.
* ``laplace3d``     3D 7-point Laplacian, diagonal 6, off-diagonals -1.
* ``varcoef3d``     3D 7-point finite-volume diffusion with cell
                    coefficients lognormal(0, 1), harmonic-mean faces and
                    Dirichlet faces 2 kappa_c to the ghost cell.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np
import scipy.sparse as sp


@dataclasses.dataclass
class Problem:
    """One SPAI build: C (= A, or A with probing rows), pattern, target."""

    name: str
    A: sp.csc_matrix                 # system matrix C: nrows x ncols
    P: sp.csc_matrix                 # start pattern J0 (values ignored)
    B: sp.csc_matrix | None = None   # target; None = identity
    epsilon: float = 0.4
    maxnew: int = 5
    max_impr_steps: int = 0
    use_mean: int = 1
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.A.shape[1]

    def row_pattern(self):
        """CSR pattern of C (indptr over rows, sorted column indices)."""
        R = self.A.tocsr()
        R.sort_indices()
        return R.indptr.astype(np.int32), R.indices.astype(np.int32)

    def digest(self) -> str:
        h = hashlib.sha256()
        for arr in (self.A.indptr, self.A.indices, self.A.data, self.P.indptr, self.P.indices):
            h.update(np.ascontiguousarray(arr).tobytes())
        if self.B is not None:
            for arr in (self.B.indptr, self.B.indices, self.B.data):
                h.update(np.ascontiguousarray(arr).tobytes())
        return h.hexdigest()


def _csc(rows, cols, vals, shape) -> sp.csc_matrix:
    A = sp.coo_matrix((np.asarray(vals, dtype=np.float64),
                       (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=shape).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def diag_pattern(n: int) -> sp.csc_matrix:
    """Start pattern {k} per column (Pattern::Diagonal_Pattern, src/Pattern.cpp:378-423)."""
    return _csc(np.arange(n), np.arange(n), np.ones(n), (n, n))


def pattern_of(A: sp.spmatrix) -> sp.csc_matrix:
    """Column pattern of A (Matrix::To_Pattern, src/Matrix.imp:285-356)."""
    A = sp.csc_matrix(A)
    P = sp.csc_matrix((np.ones(A.nnz), A.indices.copy(), A.indptr.copy()), shape=A.shape)
    P.sort_indices()
    P.indptr = P.indptr.astype(np.int32)
    P.indices = P.indices.astype(np.int32)
    return P


def pattern_power2(A: sp.spmatrix) -> sp.csc_matrix:
    """Exact pattern of A^2 (not the sparsified power of src/Matrix.imp:358-668)."""
    Pa = pattern_of(A)
    P2 = (Pa @ Pa).tocsc()
    return pattern_of(P2)


def laplace2d(nx: int = 64) -> sp.csc_matrix:
    n = nx * nx
    idx = np.arange(n)
    x, y = idx % nx, idx // nx
    rows, cols, vals = [idx], [idx], [np.full(n, 4.0)]
    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        ok = (x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < nx)
        rows.append(idx[ok])
        cols.append((idx + dx + dy * nx)[ok])
        vals.append(np.full(ok.sum(), -1.0))
    return _csc(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), (n, n))


def convdiff2d(nx: int = 256, theta: float = 0.3, speed: float = 50.0, D: float = 1.0) -> sp.csc_matrix:
    """Row i: D*(4u_i - sum neighbours)/h^2 + first-order upwind v.grad u / h."""
    n = nx * nx
    h = 1.0 / (nx + 1)
    vx, vy = speed * np.cos(theta), speed * np.sin(theta)
    assert vx > 0 and vy > 0
    idx = np.arange(n)
    x, y = idx % nx, idx // nx
    dh2 = D / (h * h)
    diag = 4.0 * dh2 + (vx + vy) / h
    coef = {(-1, 0): -dh2 - vx / h, (1, 0): -dh2, (0, -1): -dh2 - vy / h, (0, 1): -dh2}
    rows, cols, vals = [idx], [idx], [np.full(n, diag)]
    for (dx, dy), c in coef.items():
        ok = (x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < nx)
        rows.append(idx[ok])
        cols.append((idx + dx + dy * nx)[ok])
        vals.append(np.full(ok.sum(), c))
    return _csc(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), (n, n))


def random_band(n: int = 262144, k_off: int = 11, bw: int = 2000, seed: int = 20251202,
                dom: float = 1.2) -> sp.csc_matrix:
    """Row i: diagonal + k_off distinct columns uniform in [i-bw, i+bw]\\{i} (clipped)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    i = np.arange(n, dtype=np.int64)
    lo = np.maximum(0, i - bw)
    hi = np.minimum(n - 1, i + bw)
    width = hi - lo  # candidates excluding i
    if np.any(width < k_off):
        raise ValueError("bandwidth too small for k_off distinct columns")
    K = 4 * k_off
    t = rng.integers(0, width[:, None], size=(n, K))
    c = lo[:, None] + t
    c = c + (c >= i[:, None])
    order = np.argsort(c, axis=1, kind="stable")
    cs = np.take_along_axis(c, order, axis=1)
    dup_sorted = np.zeros_like(cs, dtype=bool)
    dup_sorted[:, 1:] = cs[:, 1:] == cs[:, :-1]
    dup = np.zeros_like(dup_sorted)
    np.put_along_axis(dup, order, dup_sorted, axis=1)
    keep = ~dup
    keep &= np.cumsum(keep, axis=1) <= k_off
    if np.any(keep.sum(axis=1) != k_off):
        raise RuntimeError("random_band: too few distinct draws; raise K")
    offc = c[keep].reshape(n, k_off)
    vals = rng.standard_normal((n, k_off))
    diag = dom * np.abs(vals).sum(axis=1)
    rows = np.concatenate([i, np.repeat(i, k_off)])
    cols = np.concatenate([i, offc.ravel()])
    v = np.concatenate([diag, vals.ravel()])
    return _csc(rows, cols, v, (n, n))


def laplace3d(nx: int = 64) -> sp.csc_matrix:
    n = nx ** 3
    idx = np.arange(n)
    x, y, z = idx % nx, (idx // nx) % nx, idx // (nx * nx)
    rows, cols, vals = [idx], [idx], [np.full(n, 6.0)]
    for dx, dy, dz in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        ok = ((x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < nx)
              & (z + dz >= 0) & (z + dz < nx))
        rows.append(idx[ok])
        cols.append((idx + dx + dy * nx + dz * nx * nx)[ok])
        vals.append(np.full(ok.sum(), -1.0))
    return _csc(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), (n, n))


def varcoef3d_kappa(nx: int = 128, seed: int = 2) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.standard_normal(nx ** 3))


def varcoef3d_structure(nx: int):
    """Index structure of the FV operator: for every matrix entry, which face
    (or which set of faces, for the diagonal) it is built from.

    Returns (rows, cols, face_a, face_b, kind) where for off-diagonal entries
    kind==0 and the value is -2 ka kb/(ka+kb); diagonal entries are the sum of
    their cell's interior faces plus 2 kappa_c per boundary face."""
    n = nx ** 3
    idx = np.arange(n)
    x, y, z = idx % nx, (idx // nx) % nx, idx // (nx * nx)
    ra, rb = [], []
    nbnd = np.zeros(n, dtype=np.int64)
    for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        ok = (x + dx < nx) & (y + dy < nx) & (z + dz < nx)
        a = idx[ok]
        b = a + dx + dy * nx + dz * nx * nx
        ra.append(a)
        rb.append(b)
    for dx, dy, dz in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        out = ((x + dx < 0) | (x + dx >= nx) | (y + dy < 0) | (y + dy >= nx)
               | (z + dz < 0) | (z + dz >= nx))
        nbnd += out
    return np.concatenate(ra), np.concatenate(rb), nbnd


def varcoef3d_values(kappa: np.ndarray, fa: np.ndarray, fb: np.ndarray, nbnd: np.ndarray):
    """Face coefficients and diagonal of the FV operator (float64)."""
    ka, kb = kappa[fa], kappa[fb]
    t = 2.0 * ka * kb / (ka + kb)
    diag = 2.0 * kappa * nbnd
    np.add.at(diag, fa, t)
    np.add.at(diag, fb, t)
    return t, diag


def varcoef3d(nx: int = 128, seed: int = 2) -> sp.csc_matrix:
    n = nx ** 3
    kappa = varcoef3d_kappa(nx, seed)
    fa, fb, nbnd = varcoef3d_structure(nx)
    t, diag = varcoef3d_values(kappa, fa, fb, nbnd)
    idx = np.arange(n)
    rows = np.concatenate([idx, fa, fb])
    cols = np.concatenate([idx, fb, fa])
    vals = np.concatenate([diag, -t, -t])
    return _csc(rows, cols, vals, (n, n))


def probing_system(A: sp.csc_matrix, rho: float, e: np.ndarray | None = None):
    """C = [A; rho e^T A], B = [I; rho e^T] with the probe row appended after
    the rows of A and its values scaled by rho, the layout
    Read_mm_Matrix::Data_To_Memory / Generate_Identity_Data builds
    (src/Read_mm_Matrix.cpp:532-667)."""
    n = A.shape[0]
    if e is None:
        e = np.ones(n)
    ceA = np.asarray(e @ A).ravel()
    nzc = np.nonzero(ceA)[0]
    Acoo = A.tocoo()
    C = _csc(np.concatenate([Acoo.row, np.full(nzc.size, n)]),
             np.concatenate([Acoo.col, nzc]),
             np.concatenate([Acoo.data, rho * ceA[nzc]]), (n + 1, n))
    nze = np.nonzero(e)[0]
    B = _csc(np.concatenate([np.arange(n), np.full(nze.size, n)]),
             np.concatenate([np.arange(n), nze]),
             np.concatenate([np.ones(n), rho * e[nze]]), (n + 1, n))
    return C, B


def block_convdiff2d(nx: int = 64, bs: int = 3, seed: int = 7, drop: float = 0.15) -> sp.csc_matrix:
    """Coupled system of bs fields on the 2D upwind convection-diffusion grid:
    node i couples to itself and its 4 neighbours through dense bs x bs blocks
    (the scalar operator's coefficient, scaled by its diagonal, times I, plus a
    random coupling).  A fraction `drop` of the off-diagonal coupling entries
    is zero, and each block's zeros are not stored, so blocks are partly empty
    as they are in the reference's converted matrices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = convdiff2d(nx).tocoo()
    scale = np.abs(S.data).max()
    k = S.nnz
    blk = (S.data / scale)[:, None, None] * np.eye(bs)[None] + 0.2 * rng.standard_normal((k, bs, bs)) / bs
    off = (S.row != S.col)[:, None, None] & (rng.random((k, bs, bs)) < drop)
    blk[off] = 0.0
    e, rr, cc = np.nonzero(blk)
    n = nx * nx * bs
    return _csc(S.row[e] * bs + rr, S.col[e] * bs + cc, blk[e, rr, cc], (n, n))


# ---------------------------------------------------------------- workloads

def config1(nx: int = 64) -> Problem:
    """Plain static SPAI, 2D Laplacian, pattern(A)."""
    A = laplace2d(nx)
    return Problem("laplace2d_static", A, pattern_of(A), None, epsilon=0.4, maxnew=5,
                   max_impr_steps=0, meta={"nx": nx})


def config2(nx: int = 256) -> Problem:
    """Adaptive SPAI on 2D upwind convection-diffusion: eps 0.4, ns 5, s 5, J0 diag."""
    A = convdiff2d(nx)
    return Problem("convdiff2d_adaptive", A, diag_pattern(A.shape[1]), None, epsilon=0.4,
                   maxnew=5, max_impr_steps=5, meta={"nx": nx})


def config3(n: int = 262144, bw: int = 2000, seed: int = 20251202) -> Problem:
    """Adaptive SPAI on the random banded matrix: eps 0.3, ns 8, s 4, J0 diag."""
    A = random_band(n, bw=bw, seed=seed)
    return Problem("random_band_adaptive", A, diag_pattern(n), None, epsilon=0.3, maxnew=4,
                   max_impr_steps=8, meta={"n": n, "bw": bw, "seed": seed})


def config4(nx: int = 64, rho: float = 100.0) -> Problem:
    """Static MSPAI probing on the 3D Laplacian: [A; rho e^T A] M ~ [I; rho e^T]."""
    A = laplace3d(nx)
    C, B = probing_system(A, rho)
    return Problem("laplace3d_probing", C, pattern_of(A), B, epsilon=0.4, maxnew=5,
                   max_impr_steps=0, meta={"nx": nx, "rho": rho})


def config5(nx: int = 128, seed: int = 2) -> Problem:
    """Static SPAI with the exact pattern of A^2 on the 3D variable-coefficient FV operator."""
    A = varcoef3d(nx, seed)
    return Problem("varcoef3d_pA2", A, pattern_power2(A), None, epsilon=0.4, maxnew=5,
                   max_impr_steps=0, meta={"nx": nx, "seed": seed})


CONFIGS = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}
