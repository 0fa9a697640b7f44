"""The split factorization (bpf_set_factor_split): with 2, 4 or 8 workgroups
per instance on the subtrees below the cut, or chosen by the active-instance
count (the default), the Newton solves, the Newton step from given voltages
and the dense tail's inverse equal those of the one-workgroup factorization
bit for bit, in mixed precision and in FP64, with graph replay and the
batch run as parts side by side (the forward substitution kept in the
factorization throughout).  Each output keeps its product order and
lane groups (checked on the CPU by tests/cpp/split_check.cpp), so equality
is exact, not within a tolerance.  The workloads are grids the analysis
plans a split for (2869 and 9241 buses; at 1354 buses it saves too little
to be planned)."""
import numpy as np
import pytest

from mapol_amd import synth
from mapol_amd.engine import BPF_FP64, BPF_MIXED, Engine

pytestmark = pytest.mark.gpu


def _run(case, cont, V, wg, parts=1, **kw):
    e = Engine(case, enforce_qlim=True, **kw)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.set_factor_split(wg)
    # the forward substitution always in the factorization: whether it runs
    # there follows each part's active count and changes the rounding
    e.set_schedule(parts=parts, fused_forward=1 << 20)
    e.solve()
    r = e.results()
    dx = e.stage_newton_step(V)
    _, Si = e.stage_tail_inverse(1)
    e.close()
    return r["V"], r["iterations"], r["verdict"], dx, Si


@pytest.mark.parametrize("name,T", [("pf2869x100", 12), ("pf9241x200", 6)])
def test_split_factorization_bit_identical_mixed(name, T):
    print(f"[split] mixed, {name}, {T} instances", flush=True)
    case, cont = synth.workload(name, T=T)
    rng = np.random.default_rng(11)
    V = (1 + 0.02 * rng.standard_normal((T, case.nbus))) \
        * np.exp(0.05j * rng.standard_normal((T, case.nbus)))
    base = _run(case, cont, V, 1, precision=BPF_MIXED, refine_steps=1)
    for wg, parts in ((2, 1), (4, 1), (8, 1), (0, 1), (0, 4)):
        got = _run(case, cont, V, wg, parts, precision=BPF_MIXED, refine_steps=1)
        for a, b, what in zip(base, got, ("V", "iterations", "verdict", "dx", "tail inverse")):
            np.testing.assert_array_equal(a, b, err_msg=f"{what}, {wg} workgroups, {parts} parts")


def test_split_factorization_bit_identical_fp64():
    T = 8
    print(f"[split] FP64, pf2869x100, {T} instances", flush=True)
    case, cont = synth.workload("pf2869x100", T=T)
    rng = np.random.default_rng(12)
    V = (1 + 0.02 * rng.standard_normal((T, case.nbus))) \
        * np.exp(0.05j * rng.standard_normal((T, case.nbus)))
    base = _run(case, cont, V, 1, precision=BPF_FP64)
    for wg in (8, 0):
        got = _run(case, cont, V, wg, precision=BPF_FP64)
        for a, b, what in zip(base, got, ("V", "iterations", "verdict", "dx", "tail inverse")):
            np.testing.assert_array_equal(a, b, err_msg=f"{what}, {wg} workgroups")


def test_split_option_checked():
    case, cont = synth.workload("pf118x200", T=2)
    e = Engine(case)
    for bad in (-1, 3, 16):
        with pytest.raises(Exception, match="0 \\(adaptive\\), 1, 2, 4 or 8"):
            e.set_factor_split(bad)
    e.close()
