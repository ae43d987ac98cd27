"""The C ABI library loads and exports every function include/mspai_gpu.h
declares (no device needed; no compute calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mspai_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mspai_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import mspai_amd
    lib = mspai_amd.load()
    names = _declared()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(mspai_amd.EXPORTED)


def test_default_options_are_the_reference_defaults():
    import mspai_amd
    lib = mspai_amd.load()
    o = mspai_amd.Opts()
    lib.mspai_default_opts(C.byref(o))
    # src/Command_Line.cpp:192-226: ep 0.4, mn 5, ns 5, um 1, qr 0, cs 0
    assert (o.epsilon, o.maxnew, o.max_impr_steps, o.use_mean, o.qr, o.cache) == (0.4, 5, 5, 1, 0, 0)


def test_product_has_no_cpu_fallback():
    """Without a usable device the library refuses instead of computing on the host."""
    import mspai_amd
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(mspai_amd.MspaiError):
        mspai_amd.Builder(0)
