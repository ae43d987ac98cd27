"""Pins the CPU FP64 oracle (CPU only).

The reference ships no golden vectors for the SPAI path, so the oracle is
checked against (a) numpy's LAPACK on committed fixtures, (b) closed forms,
(c) brute force on small inputs, (d) invariants.  Each check is chosen so a
plausible mistake (dropped term, sign, index, transposition) fails it.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from mspai_amd import gen
from tests.util import GOLDEN, column_lstsq, csparse_problem, ls_close, same_csc_bitwise


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "ls_lstsq.npz"))


def test_ls_solve_matches_numpy_lstsq_and_qr(oracle, fx):
    for k in range(int(fx["ncase"])):
        A, b, x_ref, R_ref = fx[f"A{k}"], fx[f"b{k}"], fx[f"x{k}"], fx[f"R{k}"]
        x, QR, tau, st = oracle.ls_solve(A, b)
        assert st == 0
        assert np.linalg.norm(x - x_ref) <= 1e-12 * max(1.0, np.linalg.norm(x_ref)), k
        n = A.shape[1]
        R = np.triu(QR[:n, :n])
        # LAPACK's sign convention: R matches entry for entry
        assert np.max(np.abs(R - R_ref[:n, :n])) <= 1e-12 * np.max(np.abs(R_ref)), k


def test_ls_solve_zero_pivot_leaves_qtb(oracle):
    rng = np.random.Generator(np.random.PCG64(3))
    A = rng.standard_normal((9, 4))
    A[:, 2] = 0.0
    b = rng.standard_normal(9)
    x, QR, tau, st = oracle.ls_solve(A, b)
    assert st & oracle.ST_SINGULAR
    Q, _ = np.linalg.qr(A, mode="complete")
    qtb = Q.T @ b
    # dtrtrs returns without solving: the first n entries are Q^T b
    assert np.allclose(x[:2], qtb[:2], rtol=1e-12, atol=1e-12)
    assert tau[2] == 0.0


def test_ls_solve_underdetermined_flags(oracle):
    A = np.arange(6.0).reshape(2, 3) + 1.0
    x, _, _, st = oracle.ls_solve(A, np.array([1.0, 0.0]))
    assert st & oracle.ST_UNDERDET
    assert np.all(x == 0.0)


def test_lane_dot_against_exact_sum(oracle):
    rng = np.random.Generator(np.random.PCG64(11))
    for m in (1, 2, 63, 64, 65, 200, 600):
        x, y = rng.standard_normal(m), rng.standard_normal(m)
        for r0 in (0, 1, m // 2):
            got = oracle.lane_dot(x, y, r0, m)
            exact = math.fsum((x[r0:] * y[r0:]).tolist())
            bound = 64 * np.finfo(float).eps * float(np.sum(np.abs(x[r0:] * y[r0:]))) + 1e-300
            assert abs(got - exact) <= bound
    # a single term is the rounded product
    assert oracle.lane_dot(np.array([3.0]), np.array([1.0 / 3.0])) == 3.0 * (1.0 / 3.0)


def test_identity_matrix_gives_identity(oracle):
    n = 50
    I = sp.identity(n, format="csc")
    p = gen.Problem("eye", gen._csc(np.arange(n), np.arange(n), np.ones(n), (n, n)),
                    gen.diag_pattern(n), None, epsilon=0.0, maxnew=3, max_impr_steps=3)
    out = oracle.spai(p)
    M = oracle.to_csc(out, n)
    assert same_csc_bitwise(M, sp.csc_matrix(I))
    assert np.all(out["res_norm"] == 0.0)
    assert np.all(out["nsolve"] == 1)


def test_diagonal_pattern_closed_form(oracle):
    """J = {k}: m_k = a_kk / ||a_k||^2 and ||r||^2 = 1 - a_kk^2/||a_k||^2."""
    p = gen.config2(16)
    p.max_impr_steps = 0
    out = oracle.spai(p)
    A = p.A
    d = A.diagonal()
    nrm2 = np.asarray(A.multiply(A).sum(axis=0)).ravel()
    assert np.allclose(out["val"][:, 0], d / nrm2, rtol=4e-16 * 8, atol=0)
    assert np.allclose(out["res_norm"] ** 2, 1 - d * d / nrm2, rtol=1e-13)


def test_translation_invariance_laplacian(oracle):
    """All interior columns of the 5-point Laplacian see the same A(I,J)."""
    nx = 12
    p = gen.config1(nx)
    out = oracle.spai(p)
    idx = np.arange(nx * nx)
    x, y = idx % nx, idx // nx
    interior = idx[(x >= 2) & (x < nx - 2) & (y >= 2) & (y < nx - 2)]
    vals = out["val"][interior, :5]
    assert np.all(vals.view(np.int64) == vals[0].view(np.int64))


@pytest.mark.parametrize("make", [lambda: gen.config1(16), lambda: gen.config5(6),
                                  lambda: csparse_problem("west0067", adaptive=False),
                                  lambda: csparse_problem("fs_183_1", adaptive=False),
                                  lambda: csparse_problem("bcsstk01", adaptive=False)])
def test_static_columns_match_numpy(oracle, make):
    p = make()
    out = oracle.spai(p)
    for k in range(p.n):
        J = out["idx"][k, :out["len"][k]]
        I, Ah, b, x = column_lstsq(p.A, J, k, p.B)
        got = out["val"][k, :out["len"][k]]
        assert ls_close(Ah, b, got, x), k
        assert abs(out["res_norm"][k] - np.linalg.norm(Ah @ got - b)) <= 1e-12 * max(1.0, out["res_norm"][k])


def _rho_bruteforce(A, r_ext, j):
    """rho_j^2 = min_alpha ||r - alpha a_j||^2 over all rows (scalar least squares)."""
    aj = A[:, j].toarray().ravel()
    alpha = np.linalg.lstsq(aj[:, None], r_ext, rcond=None)[0][0]
    return float(np.linalg.norm(r_ext - alpha * aj))


@pytest.mark.parametrize("make", [lambda: gen.config2(24), lambda: gen.config3(n=3000, bw=300),
                                  lambda: csparse_problem("bcsstk01"),
                                  lambda: csparse_problem("fs_183_1")])
def test_one_augmentation_step_bruteforce(oracle, make):
    """I-set, candidates, rho, the mean threshold and the top-mn selection of
    one augmentation step, recomputed independently with dense numpy."""
    _one_step_bruteforce(oracle, make())


def _random_upper_pattern(n, density, seed):
    """A nonsymmetric random pattern with the diagonal: column k allows the
    rows it stores (the candidates j of column k)."""
    U = sp.random(n, n, density=density, random_state=seed, format="csc")
    U = (U + sp.identity(n, format="csc")).tocsc()
    U.data[:] = 1.0
    U.sort_indices()
    return U


@pytest.mark.parametrize("make", [lambda: gen.config2(24), lambda: gen.config3(n=3000, bw=300)])
def test_one_augmentation_step_upper_pattern_bruteforce(oracle, make):
    """With an upper pattern (-up) the candidates of column k are those of
    the plain step that UP(:, k) holds (Set_Intersection,
    src/Spai_Sub.imp:1072-1094); rho, the mean over them and the selection
    recomputed with dense numpy.  UP is nonsymmetric, so intersecting with
    a row of UP instead of its column fails."""
    p = make()
    UP = _random_upper_pattern(p.n, 0.6 if p.n < 1000 else 0.02, 7)
    _one_step_bruteforce(oracle, p, UP)


def _one_step_bruteforce(oracle, p, UP=None):
    A = sp.csc_matrix(p.A)
    # structural pattern (stored entries, explicit zeros included), as the
    # reference's index sets are built from stored indices
    Ad = sp.csc_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape).toarray()
    n = p.n
    mn = p.maxnew
    one = oracle.spai(p, max_impr_steps=1, epsilon=0.0, up=UP)
    rng = np.random.Generator(np.random.PCG64(5))
    cols = rng.choice(n, size=min(n, 60), replace=False)
    checked = 0
    for k in cols:
        J0 = p.P.indices[p.P.indptr[k]:p.P.indptr[k + 1]]
        I, Ah, b, x = column_lstsq(A, J0, k)
        # brute-force I: rows where A(:, J0) has a nonzero
        assert np.array_equal(I, np.nonzero(np.any(Ad[:, J0] != 0, axis=1))[0])
        r = Ah @ x - b
        r_ext = np.zeros(A.shape[0])
        r_ext[I] = r
        L = np.union1d(I, [k])
        cand = np.setdiff1d(np.nonzero(np.any(Ad[L, :] != 0, axis=0))[0], J0)
        if UP is not None:
            cand = np.intersect1d(cand, UP.indices[UP.indptr[k]:UP.indptr[k + 1]])
        J1 = one["idx"][k, :one["len"][k]]
        added = np.setdiff1d(J1, J0)
        if cand.size == 0:
            assert added.size == 0
            continue
        rho = np.array([_rho_bruteforce(A, r_ext, j) for j in cand])
        mean = rho.mean()
        order = np.argsort(rho, kind="stable")
        q = int(np.sum(rho <= mean * (1 + 1e-12)))
        s = min(mn, q)
        if s < cand.size and s > 0:
            gap = rho[order[s]] - rho[order[s - 1]]
            if gap <= 1e-9 * max(rho[order[s]], 1e-300):
                continue  # tie at the cut: decided by std::sort's order, checked on the GPU side
        if np.any(np.abs(rho - mean) <= 1e-9 * mean):
            continue
        assert np.array_equal(np.sort(added), np.sort(cand[order[:s]])), k
        checked += 1
    assert checked >= 10


def test_upper_pattern_limits(oracle):
    """A full upper pattern changes nothing, bit for bit; an empty one leaves
    every column at its static solve on J0; a random one keeps every column
    of M inside J0 u UP(:, k)."""
    p = gen.config2(12)
    n = p.n
    ref = oracle.to_csc(oracle.spai(p), n)
    full = sp.csc_matrix(np.ones((n, n)))
    assert same_csc_bitwise(oracle.to_csc(oracle.spai(p, up=full), n), ref)
    empty = sp.csc_matrix((n, n))
    static = oracle.to_csc(oracle.spai(p, max_impr_steps=0), n)
    assert same_csc_bitwise(oracle.to_csc(oracle.spai(p, up=empty), n), static)
    UP = _random_upper_pattern(n, 0.3, 3)
    out = oracle.spai(p, up=UP)
    M = oracle.to_csc(out, n)
    assert not same_csc_bitwise(M, ref)
    P = sp.csc_matrix(p.P)
    for k in range(n):
        allowed = np.union1d(P.indices[P.indptr[k]:P.indptr[k + 1]], UP.indices[UP.indptr[k]:UP.indptr[k + 1]])
        assert np.all(np.isin(M.indices[M.indptr[k]:M.indptr[k + 1]], allowed)), k


@pytest.mark.parametrize("make", [lambda: gen.config2(32), lambda: gen.config3(n=4096, bw=400),
                                  lambda: csparse_problem("fs_183_1")])
def test_adaptive_result_is_least_squares_optimal(oracle, make):
    p = make()
    out = oracle.spai(p, threads=4)
    for k in range(0, p.n, max(1, p.n // 200)):
        J = out["idx"][k, :out["len"][k]]
        I, Ah, b, x = column_lstsq(p.A, J, k)
        got = out["val"][k, :out["len"][k]]
        assert ls_close(Ah, b, got, x), k
        rn = np.linalg.norm(Ah @ got - b)
        assert abs(out["res_norm"][k] - rn) <= 1e-12 * max(1.0, rn)


def _candidates_bruteforce(A, J, k):
    """J~ = (union of the row patterns of I u {k}) minus J, from dense boolean
    products (Augment_Sparsity, src/Spai_Sub.imp:1044-1069)."""
    D = (sp.csc_matrix(A).toarray() != 0)
    I = np.flatnonzero(D[:, J].any(axis=1))
    L = np.union1d(I, [k])
    U = np.flatnonzero(D[L, :].any(axis=0))
    return np.setdiff1d(U, J)


@pytest.mark.parametrize("make", [lambda: gen.config2(16), lambda: gen.config3(n=1024, bw=100),
                                  lambda: csparse_problem("fs_183_1")])
def test_stop_rule_is_norm_below_eps(oracle, make):
    """The augmentation loop stops at the first step with ||r||_2 <= eps (the
    norm, not its square), after ns augmentations, or when J~ is empty
    (src/Spai_Unrestrained.imp:78-129).  J after s steps is read from a run
    with ns = s; every step's solution is checked against numpy lstsq on that
    step's (I, J) and its ||r|| recomputed with numpy.  Columns with ||r|| in (eps, sqrt(eps)] at a step where the
    loop went on, and in (eps^2, eps] at the step where it stopped, are
    required: a rule on ||r||^2 <= eps would stop at the first kind, a rule on
    ||r|| <= eps^2 would go on at the second."""
    p = make()
    ns, eps = p.max_impr_steps, p.epsilon
    runs = [oracle.spai(p, max_impr_steps=s) for s in range(ns + 1)]
    full = runs[-1]
    went_on_below_sqrt = stopped_above_sq = 0
    for k in range(p.n):
        t = int(full["nsolve"][k]) - 1  # augmentation steps taken
        rns = []
        for s in range(t + 1):
            J = runs[s]["idx"][k, :runs[s]["len"][k]]
            I, Ah, b, x = column_lstsq(p.A, J, k)
            assert ls_close(Ah, b, runs[s]["val"][k, :J.size], x), (k, s)
            rns.append(float(np.linalg.norm(Ah @ runs[s]["val"][k, :J.size] - b)))
        for s in range(t):  # the loop went on: ||r|| above eps
            if abs(rns[s] - eps) > 1e-12:
                assert rns[s] > eps, (k, s, rns[s])
                went_on_below_sqrt += rns[s] <= math.sqrt(eps)
        rt = rns[t]
        assert abs(full["res_norm"][k] - rt) <= 1e-12 * max(1.0, rt)
        if abs(rt - eps) <= 1e-12:
            continue
        if rt <= eps:
            stopped_above_sq += rt > eps * eps
        else:  # out of steps, or no candidate left
            J = full["idx"][k, :full["len"][k]]
            assert t == ns or _candidates_bruteforce(p.A, J, k).size == 0, (k, t, rt)
    assert went_on_below_sqrt >= 1 and stopped_above_sq >= 1, (went_on_below_sqrt, stopped_above_sq)


@pytest.mark.parametrize("a,nsolve", [(0.3736, 1), (0.5774, 2)])
def test_stop_rule_hand_cases(oracle, a, nsolve):
    """A = [[1, 0], [a, 1]], J0 = {0}, eps = 0.4: column 0 has
    ||r|| = |a| / sqrt(1 + a^2) after the first solve, 0.35 in (eps^2, eps]
    for a = 0.3736 (stop: one solve) and 0.50 in (eps, sqrt(eps)] for
    a = 0.5774 (go on: J~ = {1}, a second solve)."""
    A = sp.csc_matrix(np.array([[1.0, 0.0], [a, 1.0]]))
    P = sp.csc_matrix(np.eye(2))
    p = gen.Problem("stop_rule", A, P, None, epsilon=0.4, maxnew=1, max_impr_steps=3)
    out = oracle.spai(p)
    rn0 = abs(a) / math.sqrt(1.0 + a * a)
    assert (rn0 <= 0.4) == (nsolve == 1)
    assert int(out["nsolve"][0]) == nsolve
    if nsolve == 1:
        assert abs(out["res_norm"][0] - rn0) <= 1e-15
    else:
        assert list(out["idx"][0, :out["len"][0]]) == [0, 1]


def test_threads_are_deterministic(oracle):
    p = gen.config3(n=4096, bw=400)
    a = oracle.spai(p, threads=1)
    b = oracle.spai(p, threads=5)
    assert same_csc_bitwise(oracle.to_csc(a, p.n), oracle.to_csc(b, p.n))


def test_cache_changes_nothing(oracle):
    """Spai_Caching reuses factors only on an exactly equal A-hat: bitwise
    identical output (src/Spai_Caching.imp:196-204)."""
    for p in (gen.config1(16), gen.config2(32)):
        a = oracle.spai(p)
        b = oracle.spai(p, cache=64)
        assert same_csc_bitwise(oracle.to_csc(a, p.n), oracle.to_csc(b, p.n))


@pytest.mark.parametrize("make", [lambda: gen.config2(32), lambda: gen.config3(n=4096, bw=400)])
def test_qr_update_path_is_least_squares_optimal(oracle, make):
    """-qr 2: appended Householder QR gives the least-squares solution of the
    augmented (I, J) (equality with a fresh QR up to rounding)."""
    p = make()
    upd = oracle.spai(p, qr=2)
    fresh = oracle.spai(p, qr=0)
    same = np.all(upd["len"] == fresh["len"]) and np.array_equal(upd["idx"], fresh["idx"])
    assert same or np.mean(np.all(upd["idx"] == fresh["idx"], axis=1)) > 0.99
    for k in range(0, p.n, max(1, p.n // 200)):
        J = upd["idx"][k, :upd["len"][k]]
        I, Ah, b, x = column_lstsq(p.A, J, k)
        got = upd["val"][k, :upd["len"][k]]
        assert ls_close(Ah, b, got, x), k


def test_probing_stacked_least_squares(oracle):
    """[A; rho e^T A] M ~ [I; rho e^T]: every column solves the stacked problem."""
    p = gen.config4(nx=5, rho=100.0)
    out = oracle.spai(p)
    for k in range(p.n):
        J = out["idx"][k, :out["len"][k]]
        I, Ah, b, x = column_lstsq(p.A, J, k, p.B)
        got = out["val"][k, :out["len"][k]]
        assert ls_close(Ah, b, got, x), k


def test_probing_limits(oracle):
    """rho -> 0 reduces to plain SPAI; rho -> inf enforces (e^T A) m_k = 1."""
    nx = 5
    A = gen.laplace3d(nx)
    plain = gen.Problem("plain", A, gen.pattern_of(A), None, max_impr_steps=0)
    ref = oracle.to_csc(oracle.spai(plain), A.shape[1])
    C, B = gen.probing_system(A, 1e-9)
    small = oracle.to_csc(oracle.spai(gen.Problem("p0", C, gen.pattern_of(A), B, max_impr_steps=0)), A.shape[1])
    assert np.max(np.abs(small.data - ref.data)) <= 1e-8 * np.max(np.abs(ref.data))
    C, B = gen.probing_system(A, 1e7)
    big = oracle.to_csc(oracle.spai(gen.Problem("pinf", C, gen.pattern_of(A), B, max_impr_steps=0)), A.shape[1])
    ceA = np.asarray(np.ones(A.shape[0]) @ A).ravel()
    cm = ceA @ big.toarray()  # (e^T A) m_k for every column k
    touches = np.asarray((abs(sp.csc_matrix(np.diag(ceA != 0).astype(float)) @ big)).sum(axis=0)).ravel() > 0
    assert np.max(np.abs(cm[touches] - 1.0)) <= 1e-6


def test_backward_matches_finite_differences(oracle):
    """dL/dA for L = sum_k g_k . m_k, by central differences (h = 1e-6)."""
    p = gen.config1(6)
    rng = np.random.Generator(np.random.PCG64(1))
    out = oracle.spai(p)
    M = oracle.to_csc(out, p.n)
    g = rng.standard_normal(M.nnz)
    gA, _ = oracle.backward(p, M, g)
    h = 1e-6
    for e in rng.choice(p.A.nnz, size=16, replace=False):
        vals = []
        for s in (+1, -1):
            q = gen.Problem("fd", p.A.copy(), p.P, None, max_impr_steps=0)
            q.A.data[e] += s * h
            Mq = oracle.to_csc(oracle.spai(q), p.n)
            vals.append(float(g @ Mq.data))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - gA[e]) <= 1e-6 * max(1.0, abs(fd)), e


def test_work_counters_static(oracle):
    p = gen.config1(10)
    out = oracle.spai(p)
    n = out["len"].astype(float)
    A = p.A
    m = np.array([np.unique(A[:, out["idx"][k, :out["len"][k]]].indices).size for k in range(p.n)], float)
    qr = 2 * m * n * n - (2.0 / 3.0) * n ** 3
    assert np.allclose(out["work"][:, 0], qr)
    assert np.allclose(out["work"][:, 6], m) and np.allclose(out["work"][:, 7], n)
