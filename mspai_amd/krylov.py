"""Restarted GMRES on the device with the built preconditioner.

The solve the reference wraps around the preconditioner (KSPGMRES with the
PCSHELL whose apply is MatMult(PM, x, y), src/main.cpp:228-269,
src/imspai.cpp:47-53): restarted GMRES(m) with modified Gram-Schmidt
Arnoldi and Givens rotations, preconditioned on the right (A M y = b,
x = M y) or on the left (M A x = M b).  Vectors are torch tensors on the GPU;
the products with A and M are the library's sparse kernels (mspai_apply),
queued on torch's current stream.
"""
from __future__ import annotations

import math

import torch


class DeviceOperators:
    """y = A x and y = M x through a Builder holding A and the last built M."""

    def __init__(self, builder):
        self.b = builder
        self.nrows, self.ncols = builder.shape

    def _apply(self, which, x, transpose=False):
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("x must be a contiguous float64 CUDA tensor")
        n_in = self.ncols if (which == 1 or not transpose) else self.nrows
        n_out = self.ncols if (which == 1 or transpose) else self.nrows
        if x.numel() != n_in:
            raise ValueError(f"x has {x.numel()} entries, the operator takes {n_in}")
        y = torch.empty(n_out, dtype=torch.float64, device=x.device)
        self.b.apply_device(which, transpose, x.data_ptr(), y.data_ptr(), torch.cuda.current_stream().cuda_stream)
        return y

    def A(self, x):
        return self._apply(0, x)

    def M(self, x):
        return self._apply(1, x)


def gmres(A, b, M=None, side="right", restart=30, rtol=1e-8, maxiter=1000, orthog="cgs"):
    """Solve A x = b; A and M are callables on vectors of b's device.  Returns
    (x, iterations, history of relative residual norms as GMRES sees them).

    orthog: "cgs" -- classical Gram-Schmidt without refinement, PETSc's
    default for KSPGMRES (KSPGMRESClassicalGramSchmidtOrthogonalization,
    refinement "never"): the projections on all basis vectors come from one
    product with the basis, and each iteration reads the device once; "mgs"
    -- modified Gram-Schmidt, one projection at a time.

    On a GPU whose current stream is the null stream the solve runs on a
    stream of its own (the caller's stream waits for it), so that the
    library's products are queued without a wait each (mspai_apply on the
    null stream waits for its product before it returns)."""
    if orthog not in ("cgs", "mgs"):
        raise ValueError("orthog must be 'cgs' or 'mgs'")
    if not b.is_cuda or torch.cuda.current_stream(b.device).cuda_stream != 0:
        return _gmres(A, b, M, side, restart, rtol, maxiter, orthog)
    cur = torch.cuda.current_stream(b.device)
    side_stream = torch.cuda.Stream(device=b.device)
    side_stream.wait_stream(cur)
    with torch.cuda.stream(side_stream):
        x, it, hist = _gmres(A, b, M, side, restart, rtol, maxiter, orthog)
    cur.wait_stream(side_stream)
    x.record_stream(cur)
    return x, it, hist


def _gmres(A, b, M, side, restart, rtol, maxiter, orthog):
    n = b.numel()
    x = torch.zeros_like(b)
    if M is not None and side == "left":
        rhs = M(b)
        op = lambda v: M(A(v))  # noqa: E731
    else:
        rhs = b
        op = (lambda v: A(M(v))) if M is not None else A
    bnorm = torch.linalg.norm(rhs).item()
    if bnorm == 0.0:
        return x, 0, [0.0]
    hist = []
    it = 0
    while it < maxiter:
        r = rhs - op(x) if (M is None or side == "left") else b - A(x)
        beta = torch.linalg.norm(r).item()
        hist.append(beta / bnorm)
        if beta / bnorm <= rtol:
            break
        m = min(restart, maxiter - it)
        V = torch.empty((m + 1, n), dtype=b.dtype, device=b.device)
        V[0] = r / beta
        H = [[0.0] * m for _ in range(m + 1)]
        cs, sn = [0.0] * m, [0.0] * m
        g = [0.0] * (m + 1)
        g[0] = beta
        k_used = 0
        for k in range(m):
            w = op(V[k])
            if orthog == "cgs":
                Vk = V[:k + 1]
                h = torch.mv(Vk, w)
                w = w - torch.mv(Vk.t(), h)
                hh = torch.cat((h, torch.linalg.norm(w).view(1))).tolist()  # the one read per iteration
                for j in range(k + 1):
                    H[j][k] = hh[j]
                hn = hh[k + 1]
            else:
                for j in range(k + 1):
                    h = torch.dot(w, V[j]).item()
                    H[j][k] = h
                    w = w - h * V[j]
                hn = torch.linalg.norm(w).item()
            H[k + 1][k] = hn
            if hn != 0.0:
                V[k + 1] = w / hn
            for j in range(k):  # previous rotations
                t = cs[j] * H[j][k] + sn[j] * H[j + 1][k]
                H[j + 1][k] = -sn[j] * H[j][k] + cs[j] * H[j + 1][k]
                H[j][k] = t
            den = math.hypot(H[k][k], H[k + 1][k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0.0 else (H[k][k] / den, H[k + 1][k] / den)
            H[k][k] = cs[k] * H[k][k] + sn[k] * H[k + 1][k]
            H[k + 1][k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            k_used = k + 1
            hist.append(abs(g[k + 1]) / bnorm)
            if abs(g[k + 1]) / bnorm <= rtol or hn == 0.0:
                break
        y = [0.0] * k_used  # back substitution H y = g
        for i in range(k_used - 1, -1, -1):
            y[i] = (g[i] - sum(H[i][j] * y[j] for j in range(i + 1, k_used))) / H[i][i]
        dx = torch.mv(V[:k_used].t(), torch.tensor(y, dtype=b.dtype, device=b.device)) if k_used else \
            torch.zeros_like(b)
        x = x + (M(dx) if (M is not None and side == "right") else dx)
        if hist[-1] <= rtol:
            r = b - A(x)
            hist.append(torch.linalg.norm(r).item() / torch.linalg.norm(b).item())
            break
    return x, it, hist
