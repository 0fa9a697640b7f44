"""C-ABI checks that need no GPU: the library loads and exports every entry
point declared in include/batchpf.h; the host-side options defaults mirror
the reference (runpf.h:45-47)."""
import ctypes as C
import os
import re

import pytest

from mapol_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "batchpf.h")).read()
    return sorted(set(re.findall(r"\b(bpf_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(engine.EXPORTS)


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                    reason="libbatchpf.so not built")
def test_library_exports_every_symbol():
    L = engine.load_library()
    for name in _declared():
        assert hasattr(L, name), name


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                    reason="libbatchpf.so not built")
def test_default_options_follow_reference():
    L = engine.load_library()
    o = engine.BpfOptions()
    L.bpf_default_options(C.byref(o))
    assert (o.maxit_nr, o.maxit_fd, o.tol) == (10, 30, 1e-8)
    assert o.alg == engine.BPF_NR and o.precision == engine.BPF_FP64
