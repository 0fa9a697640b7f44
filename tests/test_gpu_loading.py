"""Branch flows and the per-contingency loading report
(``bpf_get_branch_flows``, ``bpf_get_branch_loading``) against the oracle's
branch flows at the oracle's own solution of every instance (the flows of
runpf.h:543-572; the oracle's flows are pinned by its power-balance test).
Ratings: U(0.8, 1.5) x the base case's larger end flow, so that N-1 outages
push some branches above 100 %."""
import numpy as np
import pytest

import oracle
from mapol_amd import synth
from mapol_amd.engine import BPF_MIXED, CONVERGED, Engine

pytestmark = pytest.mark.gpu


def _rates(case, seed=5):
    orc = oracle.Oracle(case)
    r = orc.solve()
    _, Sf, St = orc.objective(r["V"])
    base = np.maximum(np.abs(Sf), np.abs(St))
    rate = np.maximum(base, 1e-3) * np.random.default_rng(seed).uniform(0.8, 1.5, case.nbranch)
    rate[::17] = 0.0          # unrated branches report 0 %
    return rate


@pytest.mark.parametrize("name,T", [("pf118x200", 40), ("pf2869x100", 12)])
def test_flows_and_loading_match_oracle(name, T):
    print(f"[loading] {name} T={T}", flush=True)
    case, cont = synth.workload(name)
    rate = _rates(case)
    br_on, gen_on = cont.br_on[:T], cont.gen_on[:T]
    e = Engine(case, precision=BPF_MIXED, refine_steps=1)
    e.set_instances(br_on=br_on, gen_on=gen_on)
    e.solve()
    r = e.results()
    Sf, St = e.branch_flows()
    rep = e.branch_loading(rate)
    e.close()
    orc = oracle.Oracle(case)
    compared = overloaded = 0
    for t in range(T):
        inst = dict(br_on=br_on[t], gen_on=gen_on[t])
        ro = orc.solve(**inst)
        assert r["verdict"][t] == ro["verdict"], t
        if ro["verdict"] != CONVERGED:
            continue
        _, sf, st = orc.objective(ro["V"], **inst)
        scale = max(1.0, np.abs(sf).max())
        assert np.abs(Sf[t] - sf).max() <= 1e-7 * scale, t
        assert np.abs(St[t] - st).max() <= 1e-7 * scale, t
        # out-of-service branches carry no flow
        assert (Sf[t][br_on[t] == 0] == 0).all() and (St[t][br_on[t] == 0] == 0).all()
        lo = np.where(rate > 0, 100.0 * np.maximum(np.abs(sf), np.abs(st))
                      / np.where(rate > 0, rate, 1.0), 0.0)
        got = rep["loading"][t]
        assert np.abs(got - lo).max() <= 1e-6 * max(1.0, lo.max()), t
        assert (got[rate == 0] == 0).all()
        # counts: branches away from the 100 % boundary decide them
        clear = np.abs(lo - 100.0) > 1e-6
        assert rep["overloads"][t] >= (lo[clear] > 100).sum()
        assert rep["overloads"][t] <= (lo > 100 - 1e-6).sum()
        assert abs(rep["max_loading"][t] - lo.max()) <= 1e-6 * max(1.0, lo.max())
        compared += 1
        overloaded += rep["overloads"][t] > 0
    assert compared >= T // 2
    assert overloaded > 0, "the ratings should make some contingency overload"


def test_loading_rejects_bad_ratings():
    case, cont = synth.workload("pf118x200")
    e = Engine(case)
    e.set_instances(br_on=cont.br_on[:2], gen_on=cont.gen_on[:2])
    e.solve()
    with pytest.raises(Exception, match="ratings"):
        e.branch_loading(-np.ones(case.nbranch))
    with pytest.raises(Exception, match="shape"):
        e.branch_loading(np.ones(case.nbranch + 1))
    e.close()
