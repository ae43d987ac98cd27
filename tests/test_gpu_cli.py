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


@pytest.mark.parametrize("up", ["2", "file"])
def test_cli_upper_pattern(oracle, tmp_path, up):
    """-up 2 (the pattern of A) and -up <file> (a MatrixMarket pattern)
    through the command line, against the oracle with the same upper
    pattern."""
    import numpy as np
    import scipy.sparse as sp
    A = gen.config2(32).A.tocsc()
    n = A.shape[1]
    src, out = tmp_path / "a.mtx", tmp_path / "precond.mtx"
    write_input_mtx(src, A, symmetric=False)
    if up == "file":
        UP = (sp.random(n, n, density=0.02, random_state=4, format="csc") + sp.identity(n, format="csc")).tocsc()
        UP.data[:] = 1.0
        write_input_mtx(tmp_path / "up.mtx", UP, symmetric=False)
        arg = str(tmp_path / "up.mtx")
    else:
        UP = gen.pattern_of(A)
        arg = "2"
    args = [CLI, str(src), "-pp", "1", "-up", arg, "-ep", "0.4", "-mn", "5", "-ns", "5", "-out", str(out)]
    print(f"[cli upper pattern {up}]", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    prob = gen.Problem("up", A, gen.diag_pattern(n), None, epsilon=0.4, maxnew=5, max_impr_steps=5)
    ref = oracle.spai(prob, threads=THREADS, up=UP)
    assert out.read_text() == format_mtx(oracle.to_csc(ref, n))
    assert not np.array_equal(oracle.to_csc(ref, n).indptr, oracle.to_csc(oracle.spai(prob, threads=THREADS), n).indptr)


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



PROBE_CASES = {
    # name: (system A, probe rows Ce (None: e^T A), Be (None: e^T), rho, target T or None, -pp)
    "laplace3d_6_e": (lambda: gen.laplace3d(6), None, None, 100.0, None, 2),
    "convdiff_12_random2_target": (lambda: gen.config2(12).A, "random", "random", 1.75, "random", 1),
}


@pytest.mark.parametrize("name", sorted(PROBE_CASES))
def test_cli_probing_against_oracle(oracle, tmp_path, name):
    """-Ce/-Be/-rho and a target file through the command line (Read_Matrix_File
    and Read_Target_File, src/Read_mm_Matrix.cpp:115-420; the static solve of
    MSPAI probing, src/Spai_Sub.imp): C = [A; rho Ce] and B = [T or I; rho Be],
    the start pattern taken from the rows of A only (Matrix::To_Pattern,
    src/Matrix.imp:285-356). The written M equals the oracle's M for the same C,
    B and pattern, byte for byte."""
    import numpy as np
    import scipy.sparse as sp
    make, ce, be, rho, tgt, pp = PROBE_CASES[name]
    A = sp.csc_matrix(make())
    n = A.shape[1]
    rng = np.random.RandomState(7)
    if ce is None:
        e = np.ones(n)
        Ce = sp.csr_matrix(np.asarray(e @ A).reshape(1, n))
        Be = sp.csr_matrix(e.reshape(1, n))
    else:
        Ce = sp.random(2, n, density=0.3, random_state=rng, format="csr")
        Be = sp.random(2, n, density=0.2, random_state=rng, format="csr")
    Ce.eliminate_zeros()
    Be.eliminate_zeros()
    T = None
    if tgt is not None:
        T = (sp.identity(n, format="csc") + sp.random(n, n, density=0.02, random_state=rng)).tocsc()
        T.eliminate_zeros()
    files = {}
    for key, M in (("a", A), ("ce", Ce), ("be", Be), ("t", T)):
        if M is not None:
            files[key] = str(tmp_path / f"{key}.mtx")
            write_input_mtx(files[key], M)
    out = tmp_path / "precond.mtx"
    args = [CLI, files["a"], f"-{pp}"]
    if T is not None:
        args.append(files["t"])
    args += ["-Ce", files["ce"], "-Be", files["be"], "-rho", repr(rho), "-ns", "0", "-ep", "0.4", "-mn", "5",
             "-out", str(out)]
    print(f"[cli probing {name}] {' '.join(args[2:])}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    C = sp.vstack([A, Ce.multiply(rho)]).tocsc()
    B = sp.vstack([T if T is not None else sp.identity(n), Be.multiply(rho)]).tocsc()
    C.sort_indices()
    B.sort_indices()
    P = gen.pattern_of(A) if pp == 2 else gen.diag_pattern(n)
    prob = gen.Problem(name, C, P, B, epsilon=0.4, maxnew=5, max_impr_steps=0)
    ref = oracle.spai(prob, threads=THREADS)
    assert out.read_text() == format_mtx(oracle.to_csc(ref, n))
