"""The forward substitution inside the factorization (bpf_set_fused_forward):
never (-1), while at most 8 instances of a part are active (so the mode
changes within a solve as instances finish), and always.  The Newton
solves reach the same iteration counts and verdicts, with voltages within
the convergence tolerance's rounding, and the Newton step from given
voltages agrees to rounding, in mixed precision and in FP64; the engine
reports the limit it ran with."""
import numpy as np
import pytest

from mapol_amd import synth
from mapol_amd.engine import BPF_FP64, BPF_MIXED, Engine

pytestmark = pytest.mark.gpu

LIMITS = (-1, 8, 1 << 20)


def _run(case, cont, V, limit, **kw):
    e = Engine(case, enforce_qlim=True, **kw)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.set_schedule(parts=2, fused_forward=limit)
    e.solve()
    r = e.results()
    used = e.info()["fused_forward"]
    dx = e.stage_newton_step(V)
    e.close()
    return r, dx, used


@pytest.mark.parametrize("precision", [BPF_MIXED, BPF_FP64])
def test_fused_forward_agrees(precision):
    T = 40
    print(f"[fused forward] precision {precision}, pf2869x100, {T} instances", flush=True)
    case, cont = synth.workload("pf2869x100", T=T)
    rng = np.random.default_rng(21)
    V = (1 + 0.02 * rng.standard_normal((T, case.nbus))) \
        * np.exp(0.05j * rng.standard_normal((T, case.nbus)))
    kw = dict(precision=precision, refine_steps=1 if precision == BPF_MIXED else 0)
    runs = {lim: _run(case, cont, V, lim, **kw) for lim in LIMITS}
    base, dx0, _ = runs[-1]
    conv = base["verdict"] == 0
    assert conv.sum() >= T - 2
    for lim, (r, dx, used) in runs.items():
        assert used == lim
        np.testing.assert_array_equal(r["iterations"], base["iterations"], err_msg=str(lim))
        np.testing.assert_array_equal(r["verdict"], base["verdict"], err_msg=str(lim))
        dv = np.abs(r["V"][conv] - base["V"][conv]).max()
        assert dv <= 1e-8, (lim, dv)
        tol = 1e-7 if precision == BPF_MIXED else 1e-10
        err = np.abs(dx - dx0).max() / max(1.0, np.abs(dx0).max())
        assert err <= tol, (lim, err)
