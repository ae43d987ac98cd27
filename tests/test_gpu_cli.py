"""The mspai_gpu command line on the GPU: a MatrixMarket file in, the
reference's precond.mtx layout out, byte for byte equal to the oracle's M
written in the layout of Write_Matrix_To_File (src/Matrix.cpp:250-342)."""
import os
import subprocess

import pytest

from mspai_amd import gen
from tests.mtx_format import format_mtx, write_input_mtx
from tests.util import load_csparse

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mspai_amd", "mspai_gpu")
THREADS = min(16, os.cpu_count() or 1)

CASES = {
    # name: (matrix, symmetric file, -pp, options)
    "west0067_adaptive": (lambda: load_csparse("west0067"), False, 1, dict(ep=0.4, mn=5, ns=5)),
    "bcsstk01_static_sym": (lambda: load_csparse("bcsstk01"), True, 2, dict(ep=0.4, mn=5, ns=0)),
    "convdiff_32_adaptive": (lambda: gen.config2(32).A, False, 1, dict(ep=0.4, mn=5, ns=5)),
    "convdiff_32_qr2": (lambda: gen.config2(32).A, False, 1, dict(ep=0.4, mn=5, ns=5, qr=2)),
    "laplace_16_pattern": (lambda: gen.config1(16).A, False, 2, dict(ep=0.4, mn=5, ns=0)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cli_writes_reference_layout(oracle, tmp_path, name):
    make, symmetric, pp, opts = CASES[name]
    A = make()
    src, out = tmp_path / "a.mtx", tmp_path / "precond.mtx"
    write_input_mtx(src, A, symmetric=symmetric)
    args = [CLI, str(src), "-pp", str(pp), "-out", str(out)]
    for k, v in opts.items():
        args += [f"-{k}", str(v)]
    print(f"[cli {name}] {' '.join(args[2:])}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = A.shape[1]
    P = gen.pattern_of(A) if pp == 2 else gen.diag_pattern(n)
    prob = gen.Problem(name, A.tocsc(), P, None,
                       epsilon=opts["ep"], maxnew=opts["mn"], max_impr_steps=opts["ns"])
    ref = oracle.spai(prob, threads=THREADS, qr=opts.get("qr", 0))
    assert out.read_text() == format_mtx(oracle.to_csc(ref, n))


BLOCK_CASES = {
    # name: (matrix, -bs, -pp, options)
    "block3_adaptive": (lambda: gen.block_convdiff2d(6, 3, seed=21), 3, 1, dict(ep=0.3, mn=3, ns=3)),
    "block0_pattern": (lambda: gen.block_convdiff2d(6, 3, seed=22), 0, 2, dict(ep=0.4, mn=5, ns=0)),
    "block2_hash": (lambda: gen.block_convdiff2d(5, 2, seed=23), 2, 2, dict(ep=0.4, mn=5, ns=5, ch=1)),
}


@pytest.mark.parametrize("name", sorted(BLOCK_CASES))
def test_cli_block_and_hash(oracle, tmp_path, name):
    """-bs (constant or variable blocks) and -ch 1 (static solve on the start
    pattern) through the command line, against the block oracle."""
    from oracle import block as OB
    make, bs, pp, opts = BLOCK_CASES[name]
    A = make()
    src, out = tmp_path / "a.mtx", tmp_path / "precond.mtx"
    write_input_mtx(src, A, symmetric=False)
    args = [CLI, str(src), "-bs", str(bs), "-pp", str(pp), "-out", str(out)]
    for k, v in opts.items():
        args += [f"-{k}", str(v)]
    print(f"[cli {name}] {' '.join(args[2:])}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sizes = OB.diagonal_blocks(A) if bs == 0 else OB.constant_blocks(A.shape[0], bs)
    B = OB.BlockMatrix(A, sizes)
    ns = 0 if opts.get("ch") else opts["ns"]
    ref = OB.spai_block(B, B.pattern() if pp == 2 else None, epsilon=opts["ep"], maxnew=opts["mn"],
                        max_impr_steps=ns, threads=THREADS)
    assert out.read_text() == format_mtx(OB.to_scalar(B, ref))

