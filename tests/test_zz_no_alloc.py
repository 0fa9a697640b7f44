"""The allocation-free branch report (``bpf_get_branch_loading_noalloc``)
against the allocating one (``bpf_get_branch_loading``), called through the
C-ABI on the same solved engines: bit-identical loading, overload counts and
largest loading.  Two calls in a row, the second on a smaller input, so the
second reads a device scratch that still holds the first call's data; the
first needs two batches of instances (9241-bus grid, 16 049 branches: 31
instances fit the 4 MiB scratch)."""
import numpy as np
import pytest

from mapol_amd import synth
from mapol_amd.engine import BPF_MIXED, Engine

pytestmark = pytest.mark.gpu


def _solved(name, T):
    case, cont = synth.workload(name)
    e = Engine(case, precision=BPF_MIXED, refine_steps=1)
    e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
    e.solve()
    return case, e


def _rates(case, e, seed):
    Sf, St = e.branch_flows()
    base = np.maximum(np.abs(Sf), np.abs(St)).max(axis=0)
    rng = np.random.default_rng(seed)
    rate = np.maximum(base, 1e-3) * rng.uniform(0.7, 1.4, case.nbranch)
    rate[::13] = 0.0
    return rate


def _same(a, b):
    for k in ("overloads", "max_loading"):
        assert np.array_equal(a[k], b[k]), k
    if a["loading"] is not None and b["loading"] is not None:
        assert np.array_equal(a["loading"], b["loading"])


@pytest.mark.parametrize("per_branch", [True, False])
def test_noalloc_report_matches_allocating_report(per_branch):
    calls = [("pf9241x200", 40, 3), ("pf1354x200", 8, 4)]   # second smaller
    for name, T, seed in calls:
        print(f"[noalloc] {name} T={T} per_branch={per_branch}", flush=True)
        case, e = _solved(name, T)
        try:
            rate = _rates(case, e, seed)
            ref = e.branch_loading(rate)
            got = e.branch_loading_noalloc(rate, per_branch=per_branch)
            _same(got, ref)
            assert got["overloads"].sum() > 0, "ratings should overload some"
            # a second call on the same engine reuses the scratch as it is
            again = e.branch_loading_noalloc(rate, per_branch=per_branch)
            _same(again, ref)
        finally:
            e.close()


def test_noalloc_report_rejects_bad_ratings():
    case, e = _solved("pf118x200", 2)
    try:
        with pytest.raises(Exception, match="ratings"):
            e.branch_loading_noalloc(-np.ones(case.nbranch))
    finally:
        e.close()
