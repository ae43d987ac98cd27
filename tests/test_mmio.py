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
