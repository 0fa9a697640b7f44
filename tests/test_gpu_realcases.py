"""The real PEGASE grids the reference ships (case1354/2869/9241pegase),
read through the C-ABI MATPOWER reader, solved by the HIP engine for N-1
contingencies (branch and generator outages) and compared with the oracle:
identical iteration counts and verdicts, |dV| <= 1e-8 pu (FP64 and mixed
precision, with and without reactive-limit switching).  Their topologies
(137 elimination-tree levels at 9241 buses, against 96 for the synthetic
grid) exercise other factorization schedules than the synthetic workloads."""
import os

import numpy as np
import pytest

import oracle
from mapol_amd import synth
from mapol_amd.engine import BPF_FP64, BPF_MIXED, CONVERGED, Engine, load_matpower
from tests.test_matpower import write_case

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    z = np.load(os.path.join(HERE, "golden", "matpower_cases.npz"))
    d = tmp_path_factory.mktemp("mp")
    out = {}
    for name in ("case1354pegase", "case2869pegase", "case9241pegase"):
        p = str(d / f"{name}.m")
        write_case(p, float(z[f"{name}/baseMVA"]), z[f"{name}/bus"],
                   z[f"{name}/gen"], z[f"{name}/branch"])
        out[name] = load_matpower(p)[0]
    return out


@pytest.mark.parametrize("qlim", [False, True])
@pytest.mark.parametrize("prec", [BPF_FP64, BPF_MIXED])
@pytest.mark.parametrize("name", ["case1354pegase", "case2869pegase", "case9241pegase"])
def test_real_grid_contingencies_match_oracle(cases, name, prec, qlim):
    print(f"[real grid] {name} precision={prec} qlim={qlim}", flush=True)
    case = cases[name]
    cont = synth.n1_contingencies(case, 12, seed=7, gen_frac=0.25)
    e = Engine(case, precision=prec, refine_steps=1, enforce_qlim=qlim)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.solve()
    r = e.results()
    e.close()
    orc = oracle.Oracle(case)
    opts = oracle.Oracle.opts(enforce_qlim=qlim)
    for t in range(cont.T):
        ro = orc.solve(opts, br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        assert r["iterations"][t] == ro["iterations"], t
        assert r["verdict"][t] == ro["verdict"], t
        assert r["qlim_rounds"][t] == ro["qlim_rounds"], t
        if ro["verdict"] == CONVERGED:
            assert np.abs(r["V"][t] - ro["V"]).max() <= 1e-8, t
