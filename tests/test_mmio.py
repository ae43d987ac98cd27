"""MatrixMarket input and the reference's preconditioner file layout (CPU)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from tests.mtx_format import format_mtx, write_input_mtx
from tests.util import load_csparse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def roundtrip(tmp_path_factory):
    exe = tmp_path_factory.mktemp("mmio") / "mmio_roundtrip"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mspai_amd", "host"),
                    os.path.join(ROOT, "tests", "helpers", "mmio_roundtrip.cpp"),
                    os.path.join(ROOT, "mspai_amd", "host", "mmio.cpp"), "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("name,symmetric", [("west0067", False), ("fs_183_1", False), ("bcsstk01", True),
                                            ("t1", False)])
def test_read_then_write_matches_reference_layout(roundtrip, tmp_path, name, symmetric):
    A = load_csparse(name)
    src, dst = tmp_path / "a.mtx", tmp_path / "out.mtx"
    write_input_mtx(src, A, symmetric=symmetric)
    subprocess.run([roundtrip, str(src), str(dst)], check=True)
    assert dst.read_text() == format_mtx(A)


def test_reader_rejects_bad_files(roundtrip, tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    r = subprocess.run([roundtrip, str(bad), str(tmp_path / "o.mtx")], capture_output=True, text=True)
    assert r.returncode == 1 and "fewer entries" in r.stderr
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    r = subprocess.run([roundtrip, str(bad), str(tmp_path / "o.mtx")], capture_output=True, text=True)
    assert r.returncode == 1


def test_negative_values_get_one_space():
    M = sp.csc_matrix(np.array([[-1.5, 0.0], [2.0, 3.25e-7]]))
    text = format_mtx(M)
    assert "1 1 -1.5000000000000e+00\n" in text and "2 1  2.0000000000000e+00\n" in text


CLI = os.path.join(ROOT, "mspai_amd", "mspai_gpu")


@pytest.mark.skipif(not os.path.exists(CLI), reason="mspai_gpu is built by __graft_entry__.build()")
def test_cli_option_errors_exit_1(tmp_path):
    """Option and input errors are reported before any device call."""
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    src = tmp_path / "a.mtx"
    write_input_mtx(src, sp.identity(4, format="csc"))
    for args, msg in ((["-xx", "1"], "unknown option -xx"), (["-pp", "3"], "-pp"), (["-ep"], "needs a value"),
                      (["-ch", "1", "-cs", "10"], "Caching and Hashing"),
                      (["-ch", "1", "-qr", "2"], "QR-Updates and Hashing"),
                      (["-ch", "2"], "Illegal hash size"), (["-bs", "-1"], "Illegal block size")):
        r = subprocess.run([CLI, str(src), *args], capture_output=True, text=True)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
    r = subprocess.run([CLI, str(tmp_path / "missing.mtx")], capture_output=True, text=True)
    assert r.returncode == 1 and "mspai_gpu:" in r.stderr


def _read_dump(path):
    with open(path) as f:
        m, n, nnz = map(int, f.readline().split())
        t = np.loadtxt(f, ndmin=2) if nnz else np.zeros((0, 3))
    return sp.csc_matrix((t[:, 2], (t[:, 0].astype(int), t[:, 1].astype(int))), shape=(m, n)), nnz


def _probe(roundtrip, tmp_path, A, Ce, Be, rho, T=None):
    files = {}
    for name, M in (("A", A), ("Ce", Ce), ("Be", Be), ("T", T)):
        if M is None:
            files[name] = "-"
            continue
        files[name] = str(tmp_path / f"{name}.mtx")
        write_input_mtx(files[name], M)
    c, b = tmp_path / "C.txt", tmp_path / "B.txt"
    r = subprocess.run([roundtrip, "--probe", files["A"], files["Ce"], files["Be"], repr(rho), files["T"],
                        str(c), str(b)], capture_output=True, text=True)
    if r.returncode:
        return None, None, r.stderr
    return _read_dump(c)[0], _read_dump(b)[0], ""


def _bitwise(X, Y):
    X, Y = sp.csc_matrix(X), sp.csc_matrix(Y)
    X.sort_indices()
    Y.sort_indices()
    return (X.shape == Y.shape and np.array_equal(X.indptr, Y.indptr) and np.array_equal(X.indices, Y.indices)
            and np.array_equal(X.data.view(np.int64), Y.data.view(np.int64)))


def test_probing_rows_layout(roundtrip, tmp_path):
    """-Ce/-Be with -rho (Read_Matrix_File / Read_Target_File,
    src/Read_mm_Matrix.cpp:115-420): C = [A; rho Ce] and B = [I; rho Be], the
    probe rows after the rows of A, each value multiplied by rho once; the
    MSPAI probing system of the 3D Laplacian with e^T as the probe equals the
    generator's [A; rho e^T A], [I; rho e^T] bit for bit."""
    from mspai_amd import gen
    A = gen.laplace3d(4)
    n = A.shape[0]
    e = np.ones(n)
    Ce = sp.csr_matrix(np.asarray(e @ A).reshape(1, n))
    Ce.eliminate_zeros()
    C, B, err = _probe(roundtrip, tmp_path, A, Ce, sp.csr_matrix(e.reshape(1, n)), 100.0)
    assert not err, err
    Cg, Bg = gen.probing_system(A, 100.0)
    assert _bitwise(C, Cg) and _bitwise(B, Bg)
    # two probe rows of random values, and a target file in place of I
    rng = np.random.Generator(np.random.PCG64(4))
    Ce2 = sp.random(2, n, density=0.3, random_state=np.random.RandomState(1), format="csr")
    Be2 = sp.random(2, n, density=0.2, random_state=np.random.RandomState(2), format="csr")
    T = sp.csc_matrix(sp.identity(n) * 2.0 + sp.random(n, n, density=0.05, random_state=np.random.RandomState(3)))
    rho = float(rng.uniform(0.5, 3.0))
    C, B, err = _probe(roundtrip, tmp_path, A, Ce2, Be2, rho, T)
    assert not err, err
    assert _bitwise(C, sp.vstack([A, Ce2.multiply(rho)]))
    assert _bitwise(B, sp.vstack([T, Be2.multiply(rho)]))
    # without probing files: C = A, B = T
    C, B, err = _probe(roundtrip, tmp_path, A, None, None, 1.0, T)
    assert not err and _bitwise(C, A) and _bitwise(B, T)


def test_probing_dimension_checks(roundtrip, tmp_path):
    from mspai_amd import gen
    A = gen.laplace3d(3)
    n = A.shape[0]
    ones = sp.csr_matrix(np.ones((1, n)))
    _, _, err = _probe(roundtrip, tmp_path, A, sp.csr_matrix(np.ones((1, n - 1))), ones, 1.0)
    assert "Mismatch in matrix and probing dimensions" in err
    _, _, err = _probe(roundtrip, tmp_path, A, ones, sp.csr_matrix(np.ones((2, n))), 1.0)
    assert "Ce and probing matrix Be" in err
    _, _, err = _probe(roundtrip, tmp_path, sp.csr_matrix(np.ones((n, n - 1))), None, None, 1.0)
    assert "must be square" in err
    _, _, err = _probe(roundtrip, tmp_path, A, None, None, 1.0, sp.identity(n + 1))
    assert "target matrix dimensions" in err


def test_skew_symmetric_file(roundtrip, tmp_path):
    """A skew-symmetric file stores one triangle; the other is its negative."""
    src, dst = tmp_path / "skew.mtx", tmp_path / "out.mtx"
    src.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2.0\n")
    subprocess.run([roundtrip, str(src), str(dst)], check=True)
    full = sp.csc_matrix(np.array([[0.0, -1.5, 0.0], [1.5, 0.0, 2.0], [0.0, -2.0, 0.0]]))
    assert dst.read_text() == format_mtx(full)
    src.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1.0\n")
    r = subprocess.run([roundtrip, str(src), str(dst)], capture_output=True, text=True)
    assert r.returncode == 1 and "diagonal" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="mspai_gpu is built by __graft_entry__.build()")
def test_cli_probing_option_errors(tmp_path):
    """The probing and target checks of PCSetFromOptions_SPAI
    (src/imspai.cpp:345-352, 367-395, 442-465), before any device call."""
    src, ce, be, pat = tmp_path / "a.mtx", tmp_path / "ce.mtx", tmp_path / "be.mtx", tmp_path / "p.mtx"
    write_input_mtx(src, sp.identity(4, format="csc"))
    write_input_mtx(ce, sp.csr_matrix(np.ones((1, 4))))
    write_input_mtx(be, sp.csr_matrix(np.ones((1, 4))))
    write_input_mtx(pat, sp.identity(4, format="csc"))
    for args, msg in ((["-Ce", str(ce), "-ns", "0"], "Only one probing condition"),
                      (["-Be", str(be), "-ns", "0"], "Only one probing condition"),
                      (["-Ce", str(ce), "-Be", str(be)], "only supported in static"),
                      (["-Ce", str(ce), "-Be", str(be), "-ns", "0", "-qr", "2"], "only supported in static"),
                      (["-1", str(src)], "only supported in static"),
                      (["-rho", "-1", "-ns", "0"], "Illegal parameter for rho"),
                      (["-schur", "2"], "Illegal value for using schur"),
                      (["-schur", "1", "-ns", "0"], "pattern filled where A has nnz"),
                      ([str(pat), "-schur", "1", "-ns", "0"], "No probing condition defined but schur"),
                      ([str(pat), "-schur", "1", "-ns", "0", "-Ce", str(ce), "-Be", str(be)], "not built"),
                      ([str(pat), str(src), "-schur", "1", "-ns", "0"], "Arbitrary target matrices")):
        r = subprocess.run([CLI, str(src), *args], capture_output=True, text=True)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
