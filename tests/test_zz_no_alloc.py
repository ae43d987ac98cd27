"""mspai_build_noalloc: the build without device allocation (static scratch
array) equals mspai_build bit for bit, FP64 and mixed, on two calls in a row
on one context, the second with a smaller matrix; sizes beyond the scratch
array or an output array below the worst case fail before any launch."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch  # before the library: one HIP runtime for both

from mspai_amd import gen
from tests.util import same_csc_bitwise

pytestmark = pytest.mark.gpu


def _noalloc(gpu, b, prob, cap, **opts):
    n = prob.n
    colptr = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
    rowind = torch.full((n * cap,), -1, dtype=torch.int32, device="cuda")
    val = torch.full((n * cap,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = b.build_noalloc(colptr.data_ptr(), rowind.data_ptr(), val.data_ptr(), n * cap, **opts)
    ptr = colptr.cpu().numpy()
    nnz = int(ptr[-1])
    assert nnz == st["nnz"]
    return sp.csc_matrix((val[:nnz].cpu().numpy(), rowind[:nnz].cpu().numpy(), ptr), shape=(n, n))


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_build_twice(gpu, precision):
    b = gpu.Builder(0)
    try:
        for prob in (gen.config3(n=32768, bw=1000), gen.config2(48)):
            print(f"[noalloc precision={precision}] {prob.name} n={prob.n}", flush=True)
            opts = dict(epsilon=prob.epsilon, maxnew=prob.maxnew, max_impr_steps=prob.max_impr_steps,
                        use_mean=prob.use_mean, precision=precision)
            b.upload(prob.A, prob.row_pattern())
            b.set_pattern(prob.P)
            b.build(**opts)
            M = b.download()
            cap = int(np.diff(prob.P.indptr).max()) + prob.maxnew * prob.max_impr_steps
            Mn = _noalloc(gpu, b, prob, cap, **opts)
            assert same_csc_bitwise(M, Mn)
    finally:
        b.close()


def test_noalloc_rejects_sizes_beyond_the_scratch(gpu):
    """Both checks run on the host before any launch: nothing is written."""
    b = gpu.Builder(0)
    try:
        n = 4_000_000  # 4e6 columns x 33 output slots: far beyond the scratch array
        b.upload(sp.identity(n, format="csc"))
        b.set_pattern(None)
        colptr = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
        small = torch.zeros(16, dtype=torch.float64, device="cuda")
        with pytest.raises(gpu.MspaiError, match="scratch array"):
            b.build_noalloc(colptr.data_ptr(), small.data_ptr(), small.data_ptr(), n * 33, epsilon=0.3, maxnew=4,
                            max_impr_steps=8)
        prob = gen.config2(16)
        b.upload(prob.A, prob.row_pattern())
        b.set_pattern(prob.P)
        with pytest.raises(gpu.MspaiError, match="max_nnz"):
            b.build_noalloc(colptr.data_ptr(), small.data_ptr(), small.data_ptr(), 10, epsilon=0.4, maxnew=5,
                            max_impr_steps=5)
        torch.cuda.synchronize()
        assert int(colptr.abs().sum()) == 0 and float(small.abs().sum()) == 0.0
    finally:
        b.close()
