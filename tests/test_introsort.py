"""The GPU's emulation of libstdc++ std::sort must leave equal scores in the
same order as the host std::sort the oracle (and the reference) use: that
decides which candidates are kept on a tie at the maxnew cut
(src/Spai_Sub.imp:1106-1110, 1337-1342)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    out = tmp_path_factory.mktemp("introsort") / "libintrosort_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "mspai_amd", "csrc"),
                    os.path.join(ROOT, "tests", "helpers", "introsort_host.cpp"), "-o", str(out)],
                   check=True)
    lib = C.CDLL(str(out))

    def run(idx, rho):
        idx = np.ascontiguousarray(idx, np.int32).copy()
        rho = np.ascontiguousarray(rho, np.float64).copy()
        lib.emul_sort_rho(C.c_int(idx.size), C.c_void_p(idx.ctypes.data), C.c_void_p(rho.ctypes.data))
        return idx, rho
    return run


def _cases():
    rng = np.random.Generator(np.random.PCG64(2024))
    sizes = list(range(0, 40)) + [63, 64, 65, 100, 127, 128, 129, 500, 1000, 4096, 7000]
    for n in sizes:
        for kind in ("few", "many", "allequal", "sorted", "reversed", "organ", "saw"):
            if kind == "few":
                rho = rng.integers(0, 3, n).astype(float)
            elif kind == "many":
                rho = rng.integers(0, max(1, n // 2), n).astype(float) * 0.25
            elif kind == "allequal":
                rho = np.full(n, 0.7)
            elif kind == "sorted":
                rho = np.sort(rng.integers(0, 5, n)).astype(float)
            elif kind == "reversed":
                rho = np.sort(rng.integers(0, 5, n))[::-1].astype(float)
            elif kind == "organ":
                h = np.arange(n)
                rho = np.minimum(h, n - 1 - h).astype(float)
            else:
                rho = (np.arange(n) % 7).astype(float)
            yield n, kind, np.arange(n, dtype=np.int32) * 3 + 1, rho


def test_emulation_matches_std_sort(oracle, emul):
    count = 0
    for n, kind, idx, rho in _cases():
        a_idx, a_rho = oracle.sort_rho(idx, rho)
        b_idx, b_rho = emul(idx, rho)
        assert np.array_equal(a_idx, b_idx), (n, kind)
        assert np.array_equal(a_rho, b_rho), (n, kind)
        count += 1
    assert count > 300
