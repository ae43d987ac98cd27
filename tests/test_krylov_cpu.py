"""The GMRES harness on the host (torch CPU tensors): both
orthogonalizations, both preconditioning sides and a restart shorter than
the iteration count reach the requested residual, and the solution solves the
system (checked against numpy).  A diagonal preconditioner that is exact for a
diagonal matrix converges in one iteration."""
import numpy as np
import pytest
import torch

from mspai_amd.krylov import gmres


def _system(n=150, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) * 1.5 / np.sqrt(n)
    b = rng.standard_normal(n)
    return A, b


@pytest.mark.parametrize("orthog", ["cgs", "mgs"])
@pytest.mark.parametrize("side", ["right", "left"])
def test_gmres_solves(orthog, side):
    A, b = _system()
    At, bt = torch.tensor(A), torch.tensor(b)
    Dinv = torch.tensor(1.0 / np.diag(A))
    x, it, hist = gmres(lambda v: At @ v, bt, lambda v: Dinv * v, side=side, restart=8, rtol=1e-10,
                        maxiter=400, orthog=orthog)
    assert it > 8  # at least one restart
    assert hist[-1] <= 1e-9
    assert np.allclose(x.numpy(), np.linalg.solve(A, b), rtol=0, atol=1e-8)


def test_gmres_exact_preconditioner_one_iteration():
    d = np.linspace(1.0, 50.0, 40)
    b = np.cos(np.arange(40.0))
    At, bt = torch.tensor(np.diag(d)), torch.tensor(b)
    x, it, _ = gmres(lambda v: At @ v, bt, lambda v: v / torch.tensor(d), rtol=1e-12, maxiter=50)
    assert it == 1
    assert np.allclose(x.numpy(), b / d, rtol=1e-13, atol=0)


def test_gmres_rejects_unknown_orthogonalization():
    with pytest.raises(ValueError):
        gmres(lambda v: v, torch.ones(3), orthog="householder")
