"""Engine state across calls: graph replay reports no error and no fallback;
the adjoint in mixed precision uses its own FP64 factor storage, so a solve
after it replays the same captured graph, allocates nothing and gives the
same results; a second adjoint allocates nothing either."""
import numpy as np
import pytest

from mapol_amd import synth
from mapol_amd.engine import BPF_MIXED, Engine

pytestmark = pytest.mark.gpu


def test_replayed_solve_reports_no_error():
    case, cont = synth.workload("pf1354x200", T=32)
    e = Engine(case, precision=BPF_MIXED, refine_steps=1, enforce_qlim=True)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    for _ in range(3):
        e.solve()
        i = e.info()
        assert i["graph_used"] == 1 and i["graph_fallback"] == 0
        assert e.last_error() == ""
    assert e.info()["graph_captures"] == 1
    e.close()


def _device_free_bytes():
    """free device memory as the engine's HIP runtime reports it (the one
    libbatchpf.so links, already loaded in this process)"""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_mixed_solve_adjoint_solve_keeps_graph_and_buffers():
    case, cont = synth.workload("pf2869x100", T=24)
    e = Engine(case, precision=BPF_MIXED, refine_steps=1, enforce_qlim=True)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.solve()
    r1 = e.results()
    a1 = e.adjoint()
    free1 = _device_free_bytes()
    e.solve()
    r2 = e.results()
    a2 = e.adjoint()
    free2 = _device_free_bytes()
    assert free2 == free1, "a repeated solve + adjoint allocated device memory"
    i = e.info()
    assert i["graph_captures"] == 1 and i["graph_used"] == 1
    assert e.last_error() == ""
    for k in ("iterations", "verdict", "qlim_rounds"):
        np.testing.assert_array_equal(r1[k], r2[k])
    np.testing.assert_array_equal(r1["V"], r2["V"])
    for k in ("grad_P", "grad_Q", "grad_x", "objective"):
        np.testing.assert_array_equal(a1[k], a2[k])
    e.close()


@pytest.mark.parametrize("parts", [2, 4])
def test_parts_give_the_same_results(parts):
    """the batch run as concurrent parts (own streams and graphs) gives the
    single graph's results bit for bit, also with an instance count that
    the parts do not divide evenly"""
    case, cont = synth.workload("pf2869x100", T=23)
    e = Engine(case, precision=BPF_MIXED, refine_steps=1, enforce_qlim=True)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.solve()
    r1 = e.results()
    e.set_schedule(parts=parts)
    e.solve()
    r2 = e.results()
    i = e.info()
    assert i["graph_used"] == 1 and e.last_error() == ""
    assert i["graph_captures"] == 1 + parts
    for k in ("iterations", "verdict", "qlim_rounds", "loss"):
        np.testing.assert_array_equal(r1[k], r2[k])
    np.testing.assert_array_equal(r1["V"], r2["V"])
    e.close()
