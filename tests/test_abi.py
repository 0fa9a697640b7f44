"""C-ABI checks that need no GPU: the library loads and exports every entry
point declared in include/batchpf.h; the host-side options defaults mirror
the reference (runpf.h:45-47)."""
import ctypes as C
import os
import re

import numpy as np
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


def _build_demo(tmp_path):
    import subprocess
    exe = str(tmp_path / "mirror_demo")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "mirror_demo.cpp"),
           "-L", os.path.dirname(engine.LIB_PATH), "-lbatchpf",
           f"-Wl,-rpath,{os.path.dirname(engine.LIB_PATH)}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                    reason="libbatchpf.so not built")
def test_cpp_mirror_compiles_and_links(tmp_path):
    assert os.path.exists(_build_demo(tmp_path))


@pytest.mark.gpu
def test_cpp_mirror_runs_like_reference(tmp_path):
    """The C++ mirror gives each particle the loss the oracle computes for
    the same controls (AVR set-point, OLTC tap)."""
    import subprocess
    import numpy as np
    import oracle
    from mapol_amd import synth
    out = subprocess.run([_build_demo(tmp_path)], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    losses = [float(l.split()[-1]) for l in out if l.startswith("particle")]
    # the same 3-bus case in oracle form (net injections as negative loads)
    case = synth.Case(
        bus_type=np.array([3, 2, 1], np.int32), Pd=np.array([0.0, -0.4, 0.9]),
        Qd=np.array([0.0, 0.0, 0.3]), Gs=np.zeros(3), Bs=np.zeros(3),
        Vm=np.array([1.02, 1.0, 1.0]), Va=np.zeros(3),
        br_from=np.array([0, 1, 0], np.int32), br_to=np.array([1, 2, 2], np.int32),
        R=np.full(3, 0.02), X=np.full(3, 0.08), B=np.full(3, 0.02),
        tap=np.array([0, 0, 0.98]), shift=np.zeros(3),
        br_status=np.ones(3, np.uint8), gen_bus=np.array([0, 1], np.int32),
        Pg=np.zeros(2), Qg=np.zeros(2), Qmax=np.full(2, np.inf),
        Qmin=np.full(2, -np.inf), VG=np.array([1.0, 1.01]),
        gen_status=np.ones(2, np.uint8))
    orc = oracle.Oracle(case)
    for k, got in enumerate(losses):
        vset = np.full(3, np.nan)
        vset[1] = 0.98 + 0.02 * k
        tap = np.full(3, np.nan)
        tap[2] = 0.95 + 0.03 * k
        r = orc.solve(vset=vset, tap=tap)
        assert r["converged"]
        assert got == pytest.approx(r["loss"], rel=1e-9)


@pytest.mark.gpu
def test_cpp_mirror_reads_matpower_case(tmp_path):
    """mapol::loadMatpowerFormat (the mirror of loadMatpowerFormat) feeds
    hybrid_runpf: the base-case loss of case118 equals the oracle's on the
    same file read by the Python reader."""
    import subprocess
    import numpy as np
    import oracle
    from mapol_amd import matpower
    from tests.test_matpower import write_case
    z = np.load(os.path.join(ROOT, "tests", "golden", "matpower_cases.npz"))
    path = str(tmp_path / "case118.m")
    write_case(path, float(z["case118/baseMVA"]), z["case118/bus"],
               z["case118/gen"], z["case118/branch"], style=1)
    out = subprocess.run([_build_demo(tmp_path), path], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "buses 118 branches 186 gens 54"
    losses = [float(l.split()[-1]) for l in out if l.startswith("particle")]
    r = oracle.Oracle(matpower.load(path)).solve()
    assert r["converged"]
    assert losses == pytest.approx([r["loss"]] * 2, rel=1e-9)


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                    reason="libbatchpf.so not built")
def test_tile_options_are_validated():
    """bpf_create refuses tile sizes the kernels cannot hold before it
    touches a device"""
    from mapol_amd import synth
    from mapol_amd.engine import Engine
    case, _ = synth.workload("pf118x200", T=2)
    for bad in (dict(tail_limit=65), dict(chunk_products=8192), dict(slot_chunk=32),
                dict(slot_chunk=1000), dict(ordering=3)):
        with pytest.raises(Exception, match="out of range|tile"):
            Engine(case, **bad)


@pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                    reason="libbatchpf.so not built")
def test_case_and_numbering_options_are_validated():
    """bpf_create refuses an unknown bus numbering and a branch or generator
    naming a bus outside the case, before it renumbers or touches a device"""
    from mapol_amd import synth
    from mapol_amd.engine import Engine
    case, _ = synth.workload("pf118x200", T=2)
    for b in (-1, 2):
        with pytest.raises(Exception, match="invalid options"):
            Engine(case, bus_numbering=b)
    bad = synth.Case(**{k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                        for k, v in case.__dict__.items()})
    bad.br_to[5] = case.nbus
    with pytest.raises(Exception, match="branch 5 has a bus out of range"):
        Engine(bad)
    bad.br_to[5] = case.br_to[5]
    bad.gen_bus[0] = -1
    with pytest.raises(Exception, match="generator bus out of range"):
        Engine(bad)
