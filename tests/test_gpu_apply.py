"""The preconditioner application (mspai_apply) and the GMRES solve around it.

Products with A and M on the device equal scipy's CSC/CSR products bit for
bit (each entry of y sums its terms in increasing index order, multiply then
add).  GMRES(30) preconditioned with the built M converges on the
convection-diffusion operator in fewer iterations than without it, and the
returned x solves A x = b to the requested tolerance (checked on the host).
"""
import numpy as np
import pytest
import torch  # before the library: one HIP runtime per process

from mspai_amd import gen

pytestmark = pytest.mark.gpu


def _builder(gpu, prob, **opts):
    b = gpu.Builder(0)
    b.upload(prob.A, prob.row_pattern())
    b.set_pattern(prob.P)
    if prob.B is not None:
        b.set_target(prob.B)
    b.build(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps, **opts)
    return b


@pytest.mark.parametrize("make", [lambda: gen.config2(64), lambda: gen.config4(8), lambda: gen.config5(8)])
def test_products_match_scipy_bitwise(gpu, make):
    from mspai_amd.krylov import DeviceOperators
    prob = make()
    b = _builder(gpu, prob)
    M = b.download()
    ops = DeviceOperators(b)
    rng = np.random.Generator(np.random.PCG64(7))
    nr, nc = prob.A.shape
    for which, X in ((0, prob.A), (1, M)):
        for transpose in (False, True):
            n_in = X.shape[0] if transpose else X.shape[1]
            x = rng.standard_normal(n_in)
            y = ops._apply(which, torch.tensor(x, device="cuda"), transpose).cpu().numpy()
            ref = (X.T @ x) if transpose else (X @ x)
            print(f"[apply {prob.name} which={which} T={transpose}] n={y.size}", flush=True)
            assert np.array_equal(y.view(np.int64), ref.view(np.int64))
    # a second product after new values rebuilds the CSR copy of A
    vals = prob.A.data * 1.5
    b.update_values(vals)
    A2 = prob.A.copy()
    A2.data = vals
    x = rng.standard_normal(nc)
    y = ops.A(torch.tensor(x, device="cuda")).cpu().numpy()
    assert np.array_equal(y.view(np.int64), (A2 @ x).view(np.int64))
    b.close()


def _ragged(nrows, ncols, seed):
    """Lines of every length the staged product meets: empty lines, a whole
    block of empty lines, a row and a column longer than two LDS chunks (2048
    entries each), and lines that straddle chunk boundaries."""
    import scipy.sparse as sp
    rng = np.random.Generator(np.random.PCG64(seed))
    A = sp.random(nrows, ncols, density=0.002, random_state=seed, format="lil")
    A[7, :] = rng.standard_normal(ncols)
    A[:, 11] = rng.standard_normal((nrows, 1))
    A[:, 12] = rng.standard_normal((nrows, 1))
    A[512:1024, :] = 0
    A[:, 600:900] = 0
    A[nrows - 1, :] = 0
    A = sp.csc_matrix(A)
    A.sort_indices()
    return A


@pytest.mark.parametrize("shape", [(5000, 5000), (4500, 2600)])
def test_products_ragged_lines_bitwise(gpu, shape):
    from mspai_amd.krylov import DeviceOperators
    A = _ragged(*shape, seed=11)
    b = gpu.Builder(0)
    b.upload(A)
    ops = DeviceOperators(b)
    rng = np.random.Generator(np.random.PCG64(9))
    for transpose in (False, True):
        x = rng.standard_normal(A.shape[0] if transpose else A.shape[1])
        # negative-zero products keep the host's signs
        x[3] = -0.0
        print(f"[apply ragged {shape} T={transpose}] nnz={A.nnz}", flush=True)
        y = ops._apply(0, torch.tensor(x, device="cuda"), transpose).cpu().numpy()
        ref = (A.T @ x) if transpose else (A @ x)
        assert np.array_equal(y.view(np.int64), ref.view(np.int64))
    b.close()


def test_gmres_with_spai_converges_faster(gpu):
    from mspai_amd.krylov import DeviceOperators, gmres
    prob = gen.config2(128)
    b = _builder(gpu, prob)
    ops = DeviceOperators(b)
    n = prob.n
    rng = np.random.Generator(np.random.PCG64(5))
    xs = rng.standard_normal(n)
    rhs = prob.A @ xs
    bt = torch.tensor(rhs, device="cuda")
    _, it0, h0 = gmres(ops.A, bt, None, restart=30, rtol=1e-8, maxiter=3000)
    res = {}
    for side in ("right", "left"):
        x, it, h = gmres(ops.A, bt, ops.M, side=side, restart=30, rtol=1e-8, maxiter=3000)
        xh = x.cpu().numpy()
        rel = np.linalg.norm(rhs - prob.A @ xh) / np.linalg.norm(rhs)
        res[side] = (it, rel)
        print(f"[gmres {side}] iterations {it} (unpreconditioned {it0}), true rel residual {rel:.2e}", flush=True)
        assert it < it0
    assert res["right"][1] <= 1e-7
    b.close()
