"""Stage entry points of the C-ABI against the oracle:

* bpf_make_ybus (makeYbus.h:372-510) assembles every instance's Ybus without
  a solve: equal to the oracle's Ybus to FP64 rounding;
* bpf_stage_bmat (makeB.h:109, :392): B' and B'' of both fast-decoupled
  variants, equal to the oracle's (pinned by hand-built matrices on the CPU);
* bpf_stage_fd_halfstep (fdpf.h:480-560): one P and one Q half-step from
  perturbed voltages, equal to the same step written here with the oracle's
  mismatch and B matrices and a sparse direct solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from mapol_amd import synth
from mapol_amd.engine import BPF_FDBX, BPF_FDXB, Engine

pytestmark = pytest.mark.gpu


def _engine(name, T, **kw):
    case, cont = synth.workload(name, T=T)
    e = Engine(case, **kw)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    return case, cont, e


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200"])
def test_make_ybus_without_solve(name):
    case, cont, e = _engine(name, 6)
    with pytest.raises(Exception, match="Ybus"):
        e.ybus(0)
    e.make_ybus()
    orc = oracle.Oracle(case)
    for t in range(6):
        Yg = e.ybus(t)
        Yo = orc.ybus(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        np.testing.assert_array_equal(Yg.indptr, Yo.indptr)
        np.testing.assert_array_equal(Yg.indices, Yo.indices)
        assert np.abs(Yg.data - Yo.data).max() <= 1e-14 * np.abs(Yo.data).max()
    e.close()


@pytest.mark.parametrize("alg", [BPF_FDXB, BPF_FDBX])
@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200"])
def test_stage_bmat_matches_oracle(name, alg):
    case, cont, e = _engine(name, 5, alg=alg)
    orc = oracle.Oracle(case)
    for t in range(5):
        inst = dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        for which in (1, 2):
            Bg = e.stage_bmat(t, which)
            Bo = orc.bmat(alg, which, **inst)
            assert Bg.shape == Bo.shape, (t, which)
            d = abs(Bg - Bo)
            assert d.max() <= 1e-13 * abs(Bo).max(), (t, which)
    e.close()


@pytest.mark.parametrize("alg", [BPF_FDXB, BPF_FDBX])
def test_stage_fd_halfstep_matches_direct_step(alg):
    T = 4
    case, cont, e = _engine("pf1354x200", T, alg=alg)
    e.make_ybus()
    orc = oracle.Oracle(case)
    rng = np.random.default_rng(17)
    n = case.nbus
    V0 = np.stack([orc.initial_voltage(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
                   for t in range(T)])
    V = V0 * (1 + 0.01 * rng.normal(size=(T, n))) * np.exp(1j * 0.01 * rng.normal(size=(T, n)))
    slack = case.bus_type == 3
    V[:, slack] = V0[:, slack]
    for which in (1, 2):
        Vg, ng = e.stage_fd_halfstep(V, which)
        for t in range(T):
            inst = dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
            # bus types of the instance: a PV bus without an in-service
            # generator is PQ
            ty = case.bus_type.copy()
            for i in np.flatnonzero(ty == 2):
                on = (case.gen_bus == i) & (case.gen_status > 0) & (cont.gen_on[t] > 0)
                if not on.any():
                    ty[i] = 1
            pv, pq = np.flatnonzero(ty == 2), np.flatnonzero(ty == 1)
            F, _ = orc.mismatch(V[t], **inst)
            npv, npq = len(pv), len(pq)
            Vm, Va = np.abs(V[t]), np.arctan(V[t].imag / V[t].real)
            P = F[:npv + npq] / Vm[np.r_[pv, pq]]
            Q = F[npv + npq:npv + 2 * npq] / Vm[pq]
            nrm = max(np.abs(P).max(), np.abs(Q).max())
            assert ng[t] == pytest.approx(nrm, rel=1e-12)
            B = orc.bmat(alg, which, **inst).tocsc()
            Vn = V[t].copy()
            if which == 1:
                dth = spla.spsolve(B, P)
                idx = np.r_[pv, pq]
                Vn[idx] = Vm[idx] * np.exp(1j * (Va[idx] - dth))
            else:
                dvm = spla.spsolve(B, Q)
                Vn[pq] = (Vm[pq] - dvm) * np.exp(1j * Va[pq])
            assert np.abs(Vg[t] - Vn).max() <= 1e-12, (t, which)
    e.close()


def _superset_jacobian(case, J, ty):
    """the engine's 2x2-block Jacobian over the non-slack buses (bus order,
    scalar 2k + r: r = 0 P / angle, 1 Q / magnitude) from the oracle's
    reference-layout J: a PV bus keeps an identity Q row and a zero
    magnitude column"""
    import scipy.sparse as sp
    pv, pq = np.flatnonzero(ty == 2), np.flatnonzero(ty == 1)
    ns = np.r_[pv, pq]
    npv, npq = len(pv), len(pq)
    ang = {int(b): k for k, b in enumerate(ns)}
    mag = {int(b): npv + npq + k for k, b in enumerate(pq)}
    blk = {int(b): k for k, b in enumerate(np.sort(ns))}
    J = J.tocoo()
    rows, cols, vals = [], [], []
    inv_ang = {v: k for k, v in ang.items()}
    inv_mag = {v: k for k, v in mag.items()}
    for r, c, v in zip(J.row, J.col, J.data):
        if r < npv + npq:
            i, ri = inv_ang[r], 0
        else:
            i, ri = inv_mag[r], 1
        if c < npv + npq:
            j, cj = inv_ang[c], 0
        else:
            j, cj = inv_mag[c], 1
        rows.append(2 * blk[i] + ri)
        cols.append(2 * blk[j] + cj)
        vals.append(v)
    for b in pv:
        rows.append(2 * blk[int(b)] + 1)
        cols.append(2 * blk[int(b)] + 1)
        vals.append(1.0)
    n2 = 2 * len(ns)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n2, n2)), blk


@pytest.mark.parametrize("name,prec,tail", [
    ("pf1354x200", 0, 0), ("pf1354x200", 1, 1), ("pf1354x200", 1, 2),
    ("pf9241x200", 0, 0), ("pf9241x200", 1, 1), ("pf9241x200", 1, 2)])
def test_tail_inverse_matches_schur_complement(name, prec, tail):
    """the dense tail's inverse (matrix cores: tail 1, vector ALUs: tail 2;
    FP64 or FP32) against numpy's inverse of the Schur complement of the
    oracle's Jacobian; 9241 buses has an odd tail (61 pivots)"""
    import scipy.sparse.linalg as sla
    case, cont, e = _engine(name, 2, precision=prec, refine_steps=1)
    if tail:
        e.set_schedule(tail_inverse=tail)
    e.solve()
    r = e.results()
    rng = np.random.default_rng(5)
    V = r["V"] * (1 + 1e-3 * rng.normal(size=r["V"].shape))
    e.stage_newton_step(V)
    orc = oracle.Oracle(case)
    for t in range(2):
        bus, Si = e.stage_tail_inverse(t)
        s = len(bus)
        assert s >= 40 and (name != "pf9241x200" or s % 2 == 1)
        inst = dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
        J = orc.jacobian(V[t], **inst)
        Jb, blk = _superset_jacobian(case, J, r["bus_type"][t])
        tidx = np.array([2 * blk[int(b)] + q for b in bus for q in (0, 1)])
        nidx = np.setdiff1d(np.arange(Jb.shape[0]), tidx)
        Jtt = Jb[tidx][:, tidx].toarray()
        Jtn = Jb[tidx][:, nidx]
        Jnt = Jb[nidx][:, tidx].toarray()
        lu = sla.splu(Jb[nidx][:, nidx].tocsc())
        S = Jtt - Jtn @ lu.solve(Jnt)
        ref = np.linalg.inv(S)
        err = np.abs(Si - ref).max() / np.abs(ref).max()
        assert err <= (1e-9 if prec == 0 else 2e-3), (t, err)
    e.close()
