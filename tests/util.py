"""Shared helpers of the test suite (inputs and comparisons)."""
import os

import numpy as np
import scipy.sparse as sp

from mspai_amd import gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_csparse(name):
    """CSparse 'i j x' triplet file -> CSC; bcsstk01 stores a lower triangle."""
    t = np.loadtxt(os.path.join(GOLDEN, "csparse", name), ndmin=2)
    i, j, x = t[:, 0].astype(np.int64), t[:, 1].astype(np.int64), t[:, 2]
    shape = (int(i.max()) + 1, int(j.max()) + 1)
    if name == "bcsstk01":
        off = i != j
        i, j, x = np.concatenate([i, j[off]]), np.concatenate([j, i[off]]), np.concatenate([x, x[off]])
        shape = (max(shape), max(shape))
    return gen._csc(i, j, x, shape)


def csparse_problem(name, adaptive=True, eps=0.3, mn=4, ns=4):
    A = load_csparse(name)
    n = A.shape[1]
    if adaptive:
        P = gen._csc(np.arange(n), np.arange(n), np.ones(n), (n, n))
        return gen.Problem(f"csparse_{name}", A, P, None, epsilon=eps, maxnew=mn, max_impr_steps=ns)
    P = gen.pattern_of(A[:n, :n] if A.shape[0] >= n else A)
    return gen.Problem(f"csparse_{name}_static", A, P, None, epsilon=eps, maxnew=mn, max_impr_steps=0)


def column_lstsq(A, J, k, B=None):
    """numpy lstsq of min ||A(I,J) m - b(I)|| with I = rows of A(:,J)."""
    A = sp.csc_matrix(A)
    sub = A[:, J]
    I = np.unique(sub.indices)
    Ah = sub[I, :].toarray()
    if B is None:
        b = (I == k).astype(float)
    else:
        b = np.asarray(sp.csc_matrix(B)[:, k].toarray()).ravel()[I]
    x = np.linalg.lstsq(Ah, b, rcond=None)[0]
    return I, Ah, b, x


def same_csc_bitwise(M1, M2):
    return (np.array_equal(M1.indptr, M2.indptr) and np.array_equal(M1.indices, M2.indices)
            and np.array_equal(np.asarray(M1.data).view(np.int64), np.asarray(M2.data).view(np.int64)))


def same_pattern(M1, M2):
    return np.array_equal(M1.indptr, M2.indptr) and np.array_equal(M1.indices, M2.indices)


def ls_close(Ah, b, got, ref, c=50.0):
    """True when two least-squares solutions agree to within the problem's
    sensitivity: |dx|/|x| <= c eps (kappa + kappa^2 |r| / (|A| |x|))."""
    eps = np.finfo(float).eps
    if Ah.size == 0:
        return np.allclose(got, ref)
    s = np.linalg.svd(Ah, compute_uv=False)
    if s[-1] == 0:
        return False
    kappa = s[0] / s[-1]
    nx = max(np.linalg.norm(ref), 1e-300)
    r = np.linalg.norm(Ah @ ref - b)
    bound = c * eps * (kappa + kappa * kappa * r / (s[0] * nx)) + 1e-14
    return np.linalg.norm(got - ref) <= bound * nx
