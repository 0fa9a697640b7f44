"""Host-side symbolic analysis of the engine (no GPU): the L+U block pattern,
the Crout level schedule and the sweep schedules exported by ``bpf_analyse``
are replayed in numpy exactly as the kernels execute them (phase A dot
products, inverted pivots, phase B scaling; forward/backward sweeps) on the
oracle's Jacobian, and the resulting Newton step must equal a sparse direct
solve of the same system."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from mapol_amd import engine, synth

HAS_A, DIAG, MASK = 1 << 30, 1 << 29, (1 << 29) - 1

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(engine.LIB_PATH),
                                reason="libbatchpf.so not built")


def _blocks(case, a, J, F):
    """Reference-layout J, F -> 2x2 blocks at L+U positions, block-order rhs
    (PV buses: identity Q row, zero magnitude column)."""
    nb = a["info"]["nblocks"]
    blk = np.full(case.nbus, -1)
    blk[a["bus_of_blk"]] = np.arange(nb)
    pv = np.flatnonzero(case.bus_type == 2)
    pq = np.flatnonzero(case.bus_type == 1)
    npv, npq = len(pv), len(pq)
    rowmap = [(blk[b], 0) for b in np.concatenate([pv, pq])] + \
             [(blk[b], 1) for b in pq]
    rp, ci = a["lu_rowptr"], a["lu_colind"]

    def pos(r, c):
        seg = ci[rp[r]:rp[r + 1]]
        k = np.searchsorted(seg, c)
        assert k < len(seg) and seg[k] == c, "J entry outside the L+U pattern"
        return rp[r] + k

    A = np.zeros((len(ci), 2, 2))
    J = J.tocoo()
    for r, c, v in zip(J.row, J.col, J.data):
        (br, cr), (bc, cc) = rowmap[r], rowmap[c]
        A[pos(br, bc), cr, cc] += v
    for b in pv:
        A[pos(blk[b], blk[b]), 1, 1] = 1.0
    rhs = np.zeros(2 * nb)
    for r, (br, cr) in enumerate(rowmap):
        rhs[2 * br + cr] = F[r]
    return A, rhs, rowmap


def _crout(a, A):
    A = A.copy()
    for l in range(a["info"]["lu_levels"]):
        for k in range(a["lvA"][l], a["lvA"][l + 1]):
            d = int(a["tA_dst"][k])
            p = d & MASK
            acc = A[p].copy() if d & HAS_A else np.zeros((2, 2))
            for q in range(a["tA_ptr"][k], a["tA_ptr"][k + 1]):
                acc -= A[a["pairL"][q]] @ A[a["pairU"][q]]
            A[p] = np.linalg.inv(acc) if d & DIAG else acc
        for k in range(a["lvB"][l], a["lvB"][l + 1]):
            A[a["tB_dst"][k]] = A[a["tB_dst"][k]] @ A[a["tB_diag"][k]]
    return A


def _sweeps(a, A, b):
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    x = b.reshape(-1, 2).copy()
    diag = np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i)
                     for i in range(len(rp) - 1)])
    for l in range(a["info"]["fw_levels"]):
        for i in a["fw_rows"][a["fw_ptr"][l]:a["fw_ptr"][l + 1]]:
            for p in range(rp[i], diag[i]):
                x[i] -= A[p] @ x[ci[p]]
    for l in range(a["info"]["bw_levels"]):
        for i in a["bw_rows"][a["bw_ptr"][l]:a["bw_ptr"][l + 1]]:
            for p in range(diag[i] + 1, rp[i + 1]):
                x[i] -= A[p] @ x[ci[p]]
            x[i] = A[diag[i]] @ x[i]
    return x.ravel()


@pytest.mark.parametrize("name", ["pf14", "pf118x200", "pf1354x200"])
def test_schedule_replay_gives_the_newton_step(name):
    case, cont = synth.workload(name)
    a = engine.analyse(case)
    orc = oracle.Oracle(case)
    V = orc.initial_voltage() * (1 + 0.01 * np.random.default_rng(0)
                                 .normal(size=case.nbus))
    J = orc.jacobian(V)
    F, _ = orc.mismatch(V)
    A, rhs, rowmap = _blocks(case, a, J, F)
    LU = _crout(a, A)
    x = _sweeps(a, LU, rhs)
    ref = sla.spsolve(J.tocsc(), F)
    got = np.array([x[2 * b + c] for b, c in rowmap])
    assert np.abs(got - ref).max() <= 1e-10 * max(1, np.abs(ref).max())
    # the PV magnitude unknowns of the superset system come out exactly zero
    pv_blocks = [r[0] for r in rowmap[:int((case.bus_type == 2).sum())]]
    assert np.all(x[2 * np.array(pv_blocks, int) + 1] == 0.0)


def test_schedule_invariants():
    case, _ = synth.workload("pf1354x200")
    a = engine.analyse(case)
    i = a["info"]
    nb, nnz = i["nblocks"], i["nnz_lu_blocks"]
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    # structurally symmetric pattern with a diagonal in every row
    pat = {(r, c) for r in range(nb) for c in ci[rp[r]:rp[r + 1]]}
    assert all((c, r) in pat for r, c in pat)
    assert all((r, r) in pat for r in range(nb))
    # every position is written exactly once by phase A
    dst = a["tA_dst"] & MASK
    assert np.array_equal(np.sort(dst), np.arange(nnz))
    # a pair only reads positions finished at an earlier level
    done_at = np.empty(nnz, int)
    for l in range(i["lu_levels"]):
        done_at[dst[a["lvA"][l]:a["lvA"][l + 1]]] = l
    for l in range(i["lu_levels"]):
        for k in range(a["lvA"][l], a["lvA"][l + 1]):
            q0, q1 = a["tA_ptr"][k], a["tA_ptr"][k + 1]
            assert np.all(done_at[a["pairL"][q0:q1]] < l)
            assert np.all(done_at[a["pairU"][q0:q1]] < l)
    # every block of the Jacobian pattern holds a Jacobian value
    assert i["nnz_jac_blocks"] == int((a["tA_dst"] & HAS_A > 0).sum())
