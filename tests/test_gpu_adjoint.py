"""Adjoint of the SPAI build (mspai_backward) on the GPU.

Checks: the GPU gradient dL/dC, dL/dB against the CPU FP64 reference gradient
of the oracle; dL/dA and dL/drho of the MSPAI probing loss on the full 3D
Laplacian (64^3) against central finite differences (h = 1e-6, 64 random
entries); dL/dkappa of the variable-coefficient FV workload through A(kappa)
against central differences; and the backward/forward time ratio.
"""
import os

import numpy as np
import pytest
import torch  # before the library: one HIP runtime per process

from mspai_amd import gen

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _builder(gpu, prob):
    b = gpu.Builder(0)
    b.upload(prob.A, prob.row_pattern())
    b.set_pattern(prob.P)
    b.set_target(prob.B)
    return b


@pytest.mark.parametrize("make", [lambda: gen.config4(6), lambda: gen.config1(12), lambda: gen.config5(6)])
@pytest.mark.parametrize("precision", [0, 1])
def test_backward_matches_cpu_reference(oracle, gpu, make, precision):
    from mspai_amd.autograd import spai_values
    prob = make()
    b = _builder(gpu, prob)
    dev = "cuda"
    c = torch.tensor(prob.A.data, dtype=torch.float64, device=dev, requires_grad=True)
    bv = torch.tensor(prob.B.data, dtype=torch.float64, device=dev, requires_grad=True) if prob.B is not None else None
    m = spai_values(b, c, bv, max_impr_steps=0, precision=precision)
    g = torch.randn(m.numel(), dtype=torch.float64, device=dev, generator=torch.Generator(dev).manual_seed(4))
    (m * g).sum().backward()
    M = b.download()
    gA_ref, gB_ref = oracle.backward(prob, M, g.cpu().numpy(), threads=THREADS)
    gA = c.grad.cpu().numpy()
    err = np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref)
    print(f"[adjoint {prob.name} precision={precision}] rel err dL/dC {err:.2e}", flush=True)
    assert err <= (1e-11 if precision == 0 else 1e-9)
    if bv is not None:
        errb = np.linalg.norm(bv.grad.cpu().numpy() - gB_ref) / np.linalg.norm(gB_ref)
        assert errb <= (1e-11 if precision == 0 else 1e-9)
    b.close()


def test_probing_gradient_finite_differences(gpu):
    """L = ||A M 1 - 1||^2 + ||A M - I||_F^2 with M from
    [A; rho e^T A] M ~ [I; rho e^T], 3D Laplacian 64^3, rho = 100."""
    from mspai_amd.autograd import ProbingMaps, ProductMaps, loss_terms, spai_values
    nx, rho0 = 64, 100.0
    A = gen.laplace3d(nx)
    n = A.shape[1]
    maps = ProbingMaps(A)
    C0, B0 = gen.probing_system(A, rho0)
    assert np.array_equal(maps.C_pattern.indptr, C0.indptr) and np.array_equal(maps.C_pattern.indices, C0.indices)
    prob = gen.Problem("probe", C0, gen.pattern_of(A), B0, max_impr_steps=0)
    b = _builder(gpu, prob)
    dev = "cuda"
    a_vals = torch.tensor(A.data, dtype=torch.float64, device=dev)
    with torch.no_grad():
        m0 = spai_values(b, maps.c_values(a_vals, torch.tensor(rho0, dtype=torch.float64, device=dev)),
                         maps.b_values(torch.tensor(rho0, dtype=torch.float64, device=dev)), max_impr_steps=0)
    M0 = b.download()
    pm = ProductMaps(A, M0)
    # finite differences evaluate the loss on the host: the device sums use
    # atomics, whose order changes the last bits, and (p - q) / 2h would
    # amplify that to the size of the tolerance
    pm_host = ProductMaps(A, M0, device="cpu")
    maps_host = ProbingMaps(A, device="cpu")

    def terms(av, rv, precision=0):
        m = spai_values(b, maps.c_values(av, rv), maps.b_values(rv), max_impr_steps=0, precision=precision)
        return loss_terms(pm, av, m, n)

    def terms_host(av, rv, precision=0):
        ah, rh = av.cpu(), rv.cpu()  # C's probing row summed on the host too
        c = maps_host.c_values(ah, rh).to(dev)
        m = spai_values(b, c, maps.b_values(rv), max_impr_steps=0, precision=precision)
        return loss_terms(pm_host, ah, m.cpu(), n)

    for precision in (0, 1):
        av = a_vals.clone().requires_grad_(True)
        rv = torch.tensor(rho0, dtype=torch.float64, device=dev, requires_grad=True)
        t1, t2 = terms(av, rv, precision)
        (t1.sum() + t2.sum()).backward()
        ga, gr = av.grad.clone(), rv.grad.item()
        rng = np.random.Generator(np.random.PCG64(1))
        ents = rng.choice(A.nnz, size=64, replace=False)
        h = 1e-6
        worst = 0.0
        with torch.no_grad():
            for e in ents:
                ap, am_ = a_vals.clone(), a_vals.clone()
                ap[e] += h
                am_[e] -= h
                p1, p2 = terms_host(ap, torch.tensor(rho0, dtype=torch.float64, device=dev), precision)
                q1, q2 = terms_host(am_, torch.tensor(rho0, dtype=torch.float64, device=dev), precision)
                fd = ((p1 - q1).sum() + (p2 - q2).sum()).item() / (2 * h)
                ad = ga[e].item()
                rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-3)
                worst = max(worst, rel)
            p1, p2 = terms_host(a_vals, torch.tensor(rho0 + h, dtype=torch.float64, device=dev), precision)
            q1, q2 = terms_host(a_vals, torch.tensor(rho0 - h, dtype=torch.float64, device=dev), precision)
            fdr = ((p1 - q1).sum() + (p2 - q2).sum()).item() / (2 * h)
        relr = abs(fdr - gr) / max(abs(fdr), abs(gr), 1e-6)
        print(f"[probing FD precision={precision}] worst rel err dL/dA {worst:.2e}, dL/drho {gr:.6e} "
              f"fd {fdr:.6e} rel {relr:.2e}", flush=True)
        assert worst <= 1e-5
        assert relr <= 1e-5
    b.close()


def test_varcoef_kappa_gradient_finite_differences(gpu):
    """dL/dkappa through A(kappa) (harmonic-mean faces) on a 10^3 FV grid."""
    from mspai_amd.autograd import ProductMaps, loss_terms, spai_values, varcoef_maps, varcoef_values
    nx = 10
    prob = gen.config5(nx)
    A = prob.A
    n = A.shape[1]
    fa, fb, nbnd = gen.varcoef3d_structure(nx)
    amap = varcoef_maps(A, fa, fb)
    dev = "cuda"
    kappa0 = torch.tensor(gen.varcoef3d_kappa(nx, 2), dtype=torch.float64, device=dev)
    tfa, tfb = torch.as_tensor(fa, device=dev), torch.as_tensor(fb, device=dev)
    tnb = torch.as_tensor(nbnd, dtype=torch.float64, device=dev)
    av0 = varcoef_values(kappa0, tfa, tfb, tnb, amap)
    assert np.allclose(av0.cpu().numpy(), A.data, rtol=1e-14, atol=0)
    b = _builder(gpu, prob)
    with torch.no_grad():
        spai_values(b, av0, None, max_impr_steps=0)
    pm = ProductMaps(A, b.download())

    def terms(kap, precision=0):
        av = varcoef_values(kap, tfa, tfb, tnb, amap)
        m = spai_values(b, av, None, max_impr_steps=0, precision=precision)
        return loss_terms(pm, av, m, n)

    k = kappa0.clone().requires_grad_(True)
    t1, t2 = terms(k, precision=1)
    (t1.sum() + t2.sum()).backward()
    g = k.grad.clone()
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for c in np.random.Generator(np.random.PCG64(5)).choice(n, size=16, replace=False):
            kp, km = kappa0.clone(), kappa0.clone()
            kp[c] += h
            km[c] -= h
            p1, p2 = terms(kp)
            q1, q2 = terms(km)
            fd = ((p1 - q1).sum() + (p2 - q2).sum()).item() / (2 * h)
            worst = max(worst, abs(fd - g[c].item()) / max(abs(fd), abs(g[c].item()), 1e-3))
    print(f"[varcoef FD] worst rel err dL/dkappa {worst:.2e}", flush=True)
    assert worst <= 1e-5
    b.close()


@pytest.mark.parametrize("make", [lambda: gen.config4(), lambda: gen.config5()])
def test_backward_forward_time_ratio(gpu, make):
    """bwd/fwd device time on the full static workloads (reported)."""
    prob = make()
    b = _builder(gpu, prob)
    nnzm = None
    out = {}
    for precision in (0, 1):
        fw, bw = [], []
        for it in range(4):
            st = b.build(max_impr_steps=0, precision=precision)
            nnzm = st["nnz"]
            gm = torch.ones(nnzm, dtype=torch.float64, device="cuda")
            gc = torch.empty(prob.A.nnz, dtype=torch.float64, device="cuda")
            gb = torch.empty(prob.B.nnz, dtype=torch.float64, device="cuda") if prob.B is not None else None
            torch.cuda.synchronize()
            sb = b.backward_device(gm.data_ptr(), gc.data_ptr(), gb.data_ptr() if gb is not None else None,
                                   precision=precision)
            if it:
                fw.append(st["build_ms"])
                bw.append(sb["build_ms"])
        out[precision] = (np.mean(fw), np.mean(bw))
        print(f"[bwd/fwd {prob.name} precision={precision}] fwd {np.mean(fw):.3f} ms bwd {np.mean(bw):.3f} ms "
              f"ratio {np.mean(bw) / np.mean(fw):.2f}", flush=True)
    b.close()
    assert all(np.isfinite(v).all() for v in out.values())
