"""Parity of the HIP engine (through the C-ABI) against the oracle and the
golden vectors.  Bars: Ybus/mismatch/Jacobian to FP64 rounding; Newton
solutions with identical iteration counts and verdicts and |dV| <= 1e-8 pu
(FP64 and mixed precision)."""
import numpy as np
import pytest

import oracle
from mapol_amd import synth
from mapol_amd.engine import (BPF_FP64, BPF_MIXED, CONVERGED, MAX_ITER,
                              NONFINITE, SINGULAR, Engine)

pytestmark = pytest.mark.gpu

DV_TOL = 1e-8   # pu, the target bar for converged voltages


def _engine(name, T=None, **kw):
    case, cont = synth.workload(name, T=T)
    e = Engine(case, **kw)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    return case, cont, e


def test_ybus_matches_oracle():
    case, cont, e = _engine("pf118x200")
    e.solve()
    orc = oracle.Oracle(case)
    for t in (0, 17, 199):
        Yg = e.ybus(t)
        Yo = orc.ybus(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        np.testing.assert_array_equal(Yg.indptr, Yo.indptr)
        np.testing.assert_array_equal(Yg.indices, Yo.indices)
        scale = np.abs(Yo.data).max()
        assert np.abs(Yg.data - Yo.data).max() <= 1e-14 * scale
    e.close()


def test_mismatch_and_jacobian_match_oracle():
    case, cont, e = _engine("pf118x200", T=8)
    e.solve()
    orc = oracle.Oracle(case)
    rng = np.random.default_rng(5)
    n = case.nbus
    V = (1 + 0.03 * rng.normal(size=(8, n))) * np.exp(1j * 0.2 * rng.normal(size=(8, n)))
    F, nrm = e.stage_mismatch(V)
    for t in range(8):
        kw = dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        Fo, no = orc.mismatch(V[t], **kw)
        np.testing.assert_allclose(F[t, :len(Fo)], Fo, rtol=0, atol=1e-11)
        assert nrm[t] == pytest.approx(no, rel=1e-12, abs=1e-12)
    for t in (0, 5):
        Jg = e.stage_jacobian(V, t)
        Jo = orc.jacobian(V[t], br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        assert Jg.shape == Jo.shape
        d = (Jg - Jo).tocsr()
        assert abs(d).max() <= 1e-12 * abs(Jo).max()
        assert Jg.nnz == Jo.nnz
    e.close()


@pytest.mark.parametrize("precision,tol", [(BPF_FP64, 1e-11), (BPF_MIXED, 1e-9)])
def test_newton_step_matches_sparse_solve(precision, tol):
    import scipy.sparse.linalg as sla
    case, cont, e = _engine("pf1354x200", T=4, precision=precision,
                            refine_steps=2)
    e.solve()
    orc = oracle.Oracle(case)
    rng = np.random.default_rng(7)
    n = case.nbus
    V = case.Vm * np.exp(1j * case.Va) * (1 + 0.01 * rng.normal(size=(4, n)))
    dx = e.stage_newton_step(V)
    for t in range(4):
        kw = dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        J = orc.jacobian(V[t], **kw)
        F, _ = orc.mismatch(V[t], **kw)
        ref = sla.spsolve(J.tocsc(), F)
        got = dx[t, :len(ref)]
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())
    e.close()


@pytest.mark.parametrize("precision", [BPF_FP64, BPF_MIXED])
def test_nr_matches_golden_118x200(golden, precision):
    case, cont, e = _engine("pf118x200", precision=precision)
    e.solve()
    r = e.results()
    np.testing.assert_array_equal(r["iterations"], golden["pf118x200/nr/its"])
    np.testing.assert_array_equal(r["verdict"], golden["pf118x200/nr/verdict"])
    dv = np.abs(r["V"] - golden["pf118x200/nr/V"]).max()
    assert dv <= DV_TOL, dv
    np.testing.assert_allclose(r["loss"], golden["pf118x200/nr/loss"],
                               rtol=1e-9, atol=1e-10)
    e.close()


@pytest.mark.parametrize("precision", [BPF_FP64, BPF_MIXED])
def test_nr_matches_golden_14(golden, precision):
    case, cont, e = _engine("pf14", precision=precision)
    e.solve()
    r = e.results()
    assert r["iterations"][0] == golden["pf14/nr/its"][0]
    assert r["verdict"][0] == golden["pf14/nr/verdict"][0]
    assert np.abs(r["V"][0] - golden["pf14/nr/V"][0]).max() <= DV_TOL
    e.close()


@pytest.mark.parametrize("tag,qlim", [("nr", False), ("nrq", True)])
@pytest.mark.parametrize("precision", [BPF_FP64, BPF_MIXED])
def test_nr_matches_golden_1354(golden, tag, qlim, precision):
    idx = golden[f"pf1354x200/{tag}/idx"]
    case, cont = synth.workload("pf1354x200")
    e = Engine(case, precision=precision, enforce_qlim=qlim)
    e.set_instances(br_on=cont.br_on[idx], gen_on=cont.gen_on[idx])
    e.solve()
    r = e.results()
    np.testing.assert_array_equal(r["iterations"], golden[f"pf1354x200/{tag}/its"])
    np.testing.assert_array_equal(r["verdict"], golden[f"pf1354x200/{tag}/verdict"])
    np.testing.assert_array_equal(r["qlim_rounds"], golden[f"pf1354x200/{tag}/rounds"])
    ok = r["verdict"] == CONVERGED
    assert np.abs(r["V"][ok] - golden[f"pf1354x200/{tag}/V"][ok]).max() <= DV_TOL
    e.close()


def test_failure_verdicts_match_oracle():
    case, cont = synth.workload("pf118x200")
    br = synth.bridges(case)
    T = 3
    br_on = np.ones((T, case.nbranch), np.uint8)
    br_on[1, np.flatnonzero(br)[0]] = 0          # islands a bus
    e = Engine(case, maxit_nr=10)
    e.set_instances(br_on=br_on)
    e.solve()
    r = e.results()
    orc = oracle.Oracle(case)
    for t in range(T):
        ro = orc.solve(br_on=br_on[t])
        assert r["verdict"][t] == ro["verdict"]
        assert r["iterations"][t] == ro["iterations"]
    assert r["verdict"][1] == SINGULAR
    assert r["loss"][1] == np.finfo(np.float64).max
    e.close()
    e = Engine(case, maxit_nr=1)
    e.set_instances(T=2)
    e.solve()
    r = e.results()
    assert list(r["verdict"]) == [MAX_ITER, MAX_ITER]
    assert list(r["iterations"]) == [1, 1]
    e.close()
    c2, _ = synth.workload("pf118x200")
    c2.Pd = c2.Pd.copy()
    c2.Pd[5] = np.nan
    e = Engine(c2)
    e.set_instances(T=1)
    e.solve()
    assert e.results()["verdict"][0] == NONFINITE
    e.close()


@pytest.mark.parametrize("name,T", [("pf2869x100", 6), ("pf9241x200", 6)])
def test_nr_matches_oracle_large(name, T):
    case, cont = synth.workload(name)
    qlim = synth.WORKLOADS[name]["qlim"]
    orc = oracle.Oracle(case)
    for precision in (BPF_FP64, BPF_MIXED):
        e = Engine(case, precision=precision, enforce_qlim=qlim)
        e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        e.solve()
        r = e.results()
        for t in range(T):
            ro = orc.solve(oracle.Oracle.opts(enforce_qlim=qlim),
                           br_on=cont.br_on[t], gen_on=cont.gen_on[t])
            assert r["verdict"][t] == ro["verdict"], t
            assert r["iterations"][t] == ro["iterations"], t
            assert r["qlim_rounds"][t] == ro["qlim_rounds"], t
            if ro["converged"]:
                assert np.abs(r["V"][t] - ro["V"]).max() <= DV_TOL
                assert r["loss"][t] == pytest.approx(ro["loss"], rel=1e-8)
        e.close()


@pytest.mark.parametrize("alg,tag", [(1, "fdxb"), (2, "fdbx")])
def test_fdpf_matches_golden(golden, alg, tag):
    idx = golden[f"pf118x200/{tag}/idx"]
    case, cont = synth.workload("pf118x200")
    e = Engine(case, alg=alg)
    e.set_instances(br_on=cont.br_on[idx], gen_on=cont.gen_on[idx])
    e.solve()
    r = e.results()
    np.testing.assert_array_equal(r["iterations"], golden[f"pf118x200/{tag}/its"])
    np.testing.assert_array_equal(r["verdict"], golden[f"pf118x200/{tag}/verdict"])
    ok = r["verdict"] == CONVERGED
    assert np.abs(r["V"][ok] - golden[f"pf118x200/{tag}/V"][ok]).max() <= DV_TOL
    e.close()


def test_instances_can_be_set_again_with_another_count():
    """A handle re-used with a different number of instances (as the C++
    mirror's hybrid_runpf does on every call) gives what a fresh handle
    gives: the factor storage of the previous count must not survive."""
    case, cont = synth.workload("pf118x200")
    e = Engine(case, precision=BPF_MIXED, refine_steps=1)
    for T in (1, 16, 4, 16):
        print(f"[re-set instances] T={T}", flush=True)
        e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        e.solve()
        r = e.results()
        f = Engine(case, precision=BPF_MIXED, refine_steps=1)
        f.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        f.solve()
        rf = f.results()
        f.close()
        for k in ("iterations", "verdict", "loss"):
            np.testing.assert_array_equal(r[k], rf[k], err_msg=f"T={T} {k}")
        np.testing.assert_array_equal(r["V"], rf["V"])
    e.close()


@pytest.mark.parametrize("form,threads", [("slot", 1024), ("slot", 512), ("slot", 256),
                                          ("entry", 512), ("rows", 512)])
def test_sweep_forms_match_oracle(form, threads):
    """Every form of the triangular sweeps (slot form with 1024, 512 and 256
    lanes, entry-parallel, level rows; bpf_set_schedule) solves pf9241x200's
    first contingencies with the oracle's iterations and verdicts, FP64 and
    mixed precision."""
    print(f"[sweeps] form={form} threads={threads}", flush=True)
    sweep = {"slot": 1, "entry": 2, "rows": 3}[form]
    case, cont = synth.workload("pf9241x200")
    T = 4
    orc = oracle.Oracle(case)
    ref = [orc.solve(oracle.Oracle.opts(enforce_qlim=True), br_on=cont.br_on[t],
                     gen_on=cont.gen_on[t]) for t in range(T)]
    for precision in (BPF_FP64, BPF_MIXED):
        e = Engine(case, precision=precision, enforce_qlim=True)
        e.set_launch(solve=threads)
        e.set_schedule(sweep_form=sweep)
        e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        e.solve()
        r = e.results()
        e.close()
        for t in range(T):
            assert r["verdict"][t] == ref[t]["verdict"], t
            assert r["iterations"][t] == ref[t]["iterations"], t
            if ref[t]["converged"]:
                assert np.abs(r["V"][t] - ref[t]["V"]).max() <= DV_TOL


@pytest.mark.parametrize("lu,tail", [(1024, "mfma"), (1024, "valu"), (512, "valu"),
                                     (256, "valu")])
def test_factorization_workgroup_sizes_match_oracle(lu, tail):
    """Every workgroup size and dense-tail inverse of the factorization: with
    1024 lanes the 61-pivot tail is inverted on matrix cores (FP32, one
    32x32 tile per wave, pivot quads) or, with tail_inverse 2, by pivot
    pairs in 2x2 tiles of blocks; FP64 takes one pivot per step; 512 and 256
    lanes keep several entries per thread.  Same iterations, verdicts and
    voltages as the oracle."""
    print(f"[factor] lu threads={lu} tail={tail}", flush=True)
    case, cont = synth.workload("pf9241x200")
    T = 4
    orc = oracle.Oracle(case)
    ref = [orc.solve(oracle.Oracle.opts(enforce_qlim=True), br_on=cont.br_on[t],
                     gen_on=cont.gen_on[t]) for t in range(T)]
    for precision in (BPF_FP64, BPF_MIXED):
        e = Engine(case, precision=precision, enforce_qlim=True)
        e.set_launch(lu=lu)
        e.set_schedule(tail_inverse=2 if tail == "valu" else 1)
        e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        e.solve()
        r = e.results()
        e.close()
        for t in range(T):
            assert r["verdict"][t] == ref[t]["verdict"], t
            assert r["iterations"][t] == ref[t]["iterations"], t
            if ref[t]["converged"]:
                assert np.abs(r["V"][t] - ref[t]["V"]).max() <= DV_TOL


def test_graph_replay_matches_direct_launches():
    """The replayed graph (four Newton iterations per replay, the active
    count read once per replay; iterations past the last are empty) gives
    bit for bit what direct launches, one iteration at a time, give."""
    case, cont = synth.workload("pf1354x200")
    T = 24
    out = []
    for graph in (True, False):
        e = Engine(case, precision=BPF_MIXED, enforce_qlim=True, use_graph=graph)
        e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
        e.solve()
        assert e.info()["graph_used"] == (1 if graph else 0)
        out.append(e.results())
        e.close()
    a, b = out
    np.testing.assert_array_equal(a["iterations"], b["iterations"])
    np.testing.assert_array_equal(a["verdict"], b["verdict"])
    np.testing.assert_array_equal(a["V"], b["V"])


@pytest.mark.parametrize("tiles", [
    dict(ordering=1), dict(tail_limit=40), dict(tail_limit=47),
    dict(chunk_products=2048), dict(slot_chunk=1024), dict(slot_chunk=512),
    dict(ordering=1, tail_limit=48, chunk_products=3072, slot_chunk=1024)])
def test_tile_options_match_oracle(tiles):
    """every schedule the autotuner may choose (ordering, dense-tail size,
    chunk sizes, graph length, tail inverse, concurrent parts of the batch)
    gives the oracle's iterations, verdicts and voltages"""
    name, T = "pf9241x200", 6
    print(f"[tiles] {tiles}", flush=True)
    case, cont = synth.workload(name)
    e = Engine(case, precision=BPF_MIXED, enforce_qlim=True, **tiles)
    e.set_instances(br_on=cont.br_on[:T], gen_on=cont.gen_on[:T])
    for sched in (dict(), dict(newton_per_graph=2, tail_inverse=2, parts=2),
                  dict(newton_per_graph=4, tail_inverse=1, parts=4)):
        e.set_schedule(**sched)
        e.solve()
        r = e.results()
        orc = oracle.Oracle(case)
        for t in range(T):
            ro = orc.solve(oracle.Oracle.opts(enforce_qlim=True),
                           br_on=cont.br_on[t], gen_on=cont.gen_on[t])
            assert r["verdict"][t] == ro["verdict"], (t, sched)
            assert r["iterations"][t] == ro["iterations"], (t, sched)
            if ro["converged"]:
                assert np.abs(r["V"][t] - ro["V"]).max() <= DV_TOL
    info = e.info()
    if "tail_limit" in tiles:
        assert info["tail_n"] <= tiles["tail_limit"]
    e.close()

