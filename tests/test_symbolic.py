"""Host-side symbolic analysis of the engine (no GPU): the L+U block pattern,
the Crout level schedule and the sweep schedules exported by ``bpf_analyse``
are replayed in numpy exactly as the kernels execute them (phase A dot
products, inverted pivots, phase B scaling, the dense tail's Gauss-Jordan
inverse; forward/backward sweeps over the non-tail rows and the tail rows
through that inverse) on the oracle's Jacobian: the resulting Newton step
must equal a sparse direct solve of the same system, and the transposed
(adjoint) schedule a dense solve of the transposed system."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from mapol_amd import engine, synth

HAS_A, DIAG, LCOL, MASK = 1 << 30, 1 << 29, 1 << 28, (1 << 28) - 1
NOT_LAST = 1 << 30      # slot row word: a segment before the row's last

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(engine.LIB_PATH),
                                reason="libbatchpf.so not built")


def _blocks(case, a, J, F):
    """Reference-layout J, F -> 2x2 blocks at L+U positions, block-order rhs
    (PV buses: identity Q row, zero magnitude column).  The right-hand-side
    blocks {b_P, 0; b_Q, 0} follow the factor's, as the Jacobian fill writes
    them for the factorization's forward substitution."""
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

    A = np.zeros((len(ci) + nb, 2, 2))
    J = J.tocoo()
    for r, c, v in zip(J.row, J.col, J.data):
        (br, cr), (bc, cc) = rowmap[r], rowmap[c]
        A[pos(br, bc), cr, cc] += v
    for b in pv:
        A[pos(blk[b], blk[b]), 1, 1] = 1.0
    rhs = np.zeros(2 * nb)
    for r, (br, cr) in enumerate(rowmap):
        rhs[2 * br + cr] = F[r]
    A[len(ci):, :, 0] = rhs.reshape(-1, 2)
    return A, rhs, rowmap


def _crout(a, A, c_start=0):
    """the kernel's factorization: chunks c_start.. of the levels below the dense tail
    (each product L*U of a chunk formed from finished blocks, then every
    output = its Jacobian value or init entry minus its products; a pivot
    stores its inverse, an L output is scaled by its pivot's inverse), then
    the tail's Schur complement assembled from its pair lists and inverted by
    block Gauss-Jordan in the kernel's pivot order.  Returns (LU, S^-1)."""
    A = A.copy()
    for c in range(c_start, a["info"]["chunks"]):
        q0, q1 = a["ch_pair"][c], a["ch_pair"][c + 1]
        prod = np.zeros((q1 - q0, 2, 2))
        for q in range(q0, q1):
            l, u = a["pairs"][q]
            prod[q - q0] = A[l] if u < 0 else -(A[l] @ A[u])
        for k in range(a["ch_task"][c], a["ch_task"][c + 1]):
            d, p0, p1, _ = a["tk_a"][k]
            p = d & MASK
            e = (A[p].copy() if d & HAS_A else np.zeros((2, 2))) \
                + prod[p0 - q0:p1 - q0].sum(0)
            if d & LCOL:
                _, v0, v1, _ = a["tk_b"][k]
                e = e @ np.linalg.inv(prod[v0 - q0:v1 - q0].sum(0))
            elif d & DIAG:
                e = np.linalg.inv(e)
            A[p] = e
    s = a["info"]["tail_n"]
    D = np.zeros((s, s, 2, 2))
    for k in range(a["info"]["tail_entries"]):
        dd = int(a["tl_dense"][k])
        acc = A[a["tl_pos"][k]].copy() if dd & HAS_A else np.zeros((2, 2))
        for q in range(a["tl_ptr"][k], a["tl_ptr"][k + 1]):
            acc -= A[a["tl_pairL"][q]] @ A[a["tl_pairU"][q]]
        D[divmod(dd & MASK, s)] = acc
    for k in range(s):
        P = np.linalg.inv(D[k, k])
        col, row = D[:, k].copy(), D[k, :].copy()
        for i in range(s):
            for j in range(s):
                if i == k:
                    D[i, j] = P if j == k else P @ row[j]
                elif j == k:
                    D[i, j] = -(col[i] @ P)
                else:
                    D[i, j] = D[i, j] - col[i] @ (P @ row[j])
    return A, D


def _fill_level0(a, A):
    """k_jacobian's elim0 fill: a level-0 pivot stored inverted, an L block
    of one times the pivot's inverse (from the Jacobian values)"""
    A = A.copy()
    J = A.copy()
    for r, (q, role) in enumerate(zip(a["jac_lu"], a["jac_l0"])):
        if role == -2:
            A[q] = np.linalg.inv(J[q])
        elif role >= 0:
            A[q] = J[q] @ np.linalg.inv(J[a["jac_lu"][role]])
    return A


def _diag(a):
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    return np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i)
                     for i in range(len(rp) - 1)])


def _mirror(a):
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    m = np.empty(len(ci), int)
    for r in range(len(rp) - 1):
        for p in range(rp[r], rp[r + 1]):
            c = ci[p]
            m[p] = rp[c] + np.searchsorted(ci[rp[c]:rp[c + 1]], r)
    return m


def _sweeps(a, A, Sinv, b, trans=False):
    """the kernels' sweeps: non-tail rows by the forward levels, the tail rows
    through S^-1 after subtracting their non-tail L entries, non-tail rows by
    the backward levels.  trans: the adjoint's transposed solve (U^T then
    L^T through the mirror map, S^-T)."""
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    x = b.reshape(-1, 2).copy()
    diag = _diag(a)
    mir = _mirror(a) if trans else None
    T = (lambda M: M.T) if trans else (lambda M: M)
    for l in range(a["info"]["fw_levels"]):
        for i in a["fw_rows"][a["fw_ptr"][l]:a["fw_ptr"][l + 1]]:
            for p in range(rp[i], diag[i]):
                x[i] -= (A[mir[p]].T if trans else A[p]) @ x[ci[p]]
            if trans:
                x[i] = A[diag[i]].T @ x[i]
    s = a["info"]["tail_n"]
    piv = a["tail_piv"]
    rs = np.zeros((s, 2))
    for k in range(s):
        rs[k] = x[piv[k]]
        for q in range(a["ts_ptr"][k], a["ts_ptr"][k + 1]):
            p = a["ts_pos"][q]
            rs[k] -= (A[mir[p]].T if trans else A[p]) @ x[a["ts_col"][q]]
    for k in range(s):
        x[piv[k]] = sum(T(Sinv[j, k] if trans else Sinv[k, j]) @ rs[j]
                        for j in range(s))
    for l in range(a["info"]["bw_levels"]):
        for i in a["bw_rows"][a["bw_ptr"][l]:a["bw_ptr"][l + 1]]:
            for p in range(diag[i] + 1, rp[i + 1]):
                x[i] -= (A[mir[p]].T if trans else A[p]) @ x[ci[p]]
            if not trans:
                x[i] = A[diag[i]] @ x[i]
    return x.ravel()


def _sweeps_ep(a, A, Sinv, b):
    """the entry-parallel sweeps: per chunk, all entry products L*x_col (or
    U*x_col) first, then every row of the chunk reduces its range; the tail
    rows between the forward and the backward chunks"""
    x = b.reshape(-1, 2).copy()
    nfw = a["info"]["sweep_fw_chunks"]
    ch, rows, ent = a["sw_chunk"], a["sw_row"], a["sw_ent"]

    def tail():
        s = a["info"]["tail_n"]
        piv = a["tail_piv"]
        rs = np.zeros((s, 2))
        for k in range(s):
            rs[k] = x[piv[k]]
            for q in range(a["ts_ptr"][k], a["ts_ptr"][k + 1]):
                rs[k] -= A[a["ts_pos"][q]] @ x[a["ts_col"][q]]
        for k in range(s):
            x[piv[k]] = sum(Sinv[k, j] @ rs[j] for j in range(s))

    for c in range(len(ch)):
        if c == nfw:
            tail()
        r0, r1, e0, e1 = ch[c]
        prod = np.array([A[p] @ x[col] for p, col in ent[e0:e1]]).reshape(-1, 2)
        for i, f0, f1, dpos in rows[r0:r1]:
            xi = x[i] - prod[f0 - e0:f1 - e0].sum(0)
            x[i] = A[dpos] @ xi if c >= nfw else xi
    if nfw == len(ch):
        tail()
    return x.ravel()


def _sweeps_slot(a, A, Sinv, b, skip_fw=False):
    """the slot-form sweeps: per chunk, every lane's entry product, then each
    row's lane group summed pairwise into its first lane (lane i adds lane
    i + 2^s for s < g), which updates x; the tail rows between the forward
    and the backward chunks.  skip_fw: b is the factorization's forward
    substitution (the tail rows' still their right-hand side), so the
    forward chunks are left out."""
    x = b.reshape(-1, 2).copy()
    nfw = a["info"]["slot_fw_chunks"]
    ch, sl = a["sl_chunk"], a["sl_slot"]
    LG = 27

    def tail():
        s = a["info"]["tail_n"]
        piv = a["tail_piv"]
        rs = np.zeros((s, 2))
        for k in range(s):
            rs[k] = x[piv[k]]
            for q in range(a["ts_ptr"][k], a["ts_ptr"][k + 1]):
                rs[k] -= A[a["ts_pos"][q]] @ x[a["ts_col"][q]]
        for k in range(s):
            x[piv[k]] = sum(Sinv[k, j] @ rs[j] for j in range(s))

    for c in range(nfw if skip_fw else 0, len(ch)):
        if c == nfw:
            tail()
        s0, s1 = ch[c, 0], ch[c, 1]
        colp = sl[s0:s1, 1] & ((1 << LG) - 1)
        prod = np.array([A[p] @ x[max(cp - 1, 0)] if cp > 0 else np.zeros(2)
                         for p, cp in zip(sl[s0:s1, 0], colp)])
        q = 0
        while q < s1 - s0:
            pos, w, lead, piv = sl[s0 + q]
            if lead < 0:        # padding to the chunk's 64-slot end
                break
            g = w >> LG
            row, last = lead & (NOT_LAST - 1), not lead & NOT_LAST
            v = prod[q:q + (1 << g)]
            while len(v) > 1:
                v = v[0::2] + v[1::2]
            xi = x[row] - v[0]
            x[row] = A[piv] @ xi if c >= nfw and last else xi
            q += 1 << g
    if nfw == len(ch):
        tail()
    return x.ravel()


def _sweeps_slot_t(a, A, Sinv, g):
    """the adjoint's transposed slot-form sweeps: the transposed slots hold
    mirror positions (blocks used transposed); the forward chunks (U^T)
    apply the transposed pivot inverse at a row's last segment; the tail
    through S^-T (its rows' non-tail entries through the mirror map)"""
    x = g.reshape(-1, 2).copy()
    nfw = a["slot_t_fw_chunks"]
    ch, sl = a["sl_chunk_t"], a["sl_slot_t"]
    mir = _mirror(a)
    LG = 27

    def tail():
        s = a["info"]["tail_n"]
        piv = a["tail_piv"]
        rs = np.zeros((s, 2))
        for k in range(s):
            rs[k] = x[piv[k]]
            for q in range(a["ts_ptr"][k], a["ts_ptr"][k + 1]):
                rs[k] -= A[mir[a["ts_pos"][q]]].T @ x[a["ts_col"][q]]
        for k in range(s):
            x[piv[k]] = sum(Sinv[j, k].T @ rs[j] for j in range(s))

    for c in range(len(ch)):
        if c == nfw:
            tail()
        s0, s1 = ch[c, 0], ch[c, 1]
        colp = sl[s0:s1, 1] & ((1 << LG) - 1)
        prod = np.array([A[p].T @ x[max(cp - 1, 0)] if cp > 0 else np.zeros(2)
                         for p, cp in zip(sl[s0:s1, 0], colp)])
        q = 0
        while q < s1 - s0:
            pos, w, lead, piv = sl[s0 + q]
            if lead < 0:        # padding to the chunk's 64-slot end
                break
            gl = w >> LG
            row, last = lead & (NOT_LAST - 1), not lead & NOT_LAST
            v = prod[q:q + (1 << gl)]
            while len(v) > 1:
                v = v[0::2] + v[1::2]
            xi = x[row] - v[0]
            x[row] = A[piv].T @ xi if c < nfw and last else xi
            q += 1 << gl
    if nfw == len(ch):
        tail()
    return x.ravel()


def _system(name, seed=0, case=None):
    if case is None:
        case, cont = synth.workload(name)
    a = engine.analyse(case)
    orc = oracle.Oracle(case)
    V = orc.initial_voltage() * (1 + 0.01 * np.random.default_rng(seed)
                                 .normal(size=case.nbus))
    J = orc.jacobian(V)
    F, _ = orc.mismatch(V)
    A, rhs, rowmap = _blocks(case, a, J, F)
    return case, a, J, F, A, rhs, rowmap


@pytest.mark.parametrize("name", ["pf14", "pf118x200", "pf1354x200"])
def test_schedule_replay_gives_the_newton_step(name):
    case, a, J, F, A, rhs, rowmap = _system(name)
    assert a["info"]["tail_n"] > 0
    LU, Sinv = _crout(a, A)
    x = _sweeps(a, LU, Sinv, rhs)
    ref = sla.spsolve(J.tocsc(), F)
    got = np.array([x[2 * b + c] for b, c in rowmap])
    assert np.abs(got - ref).max() <= 1e-10 * max(1, np.abs(ref).max())
    # the entry-parallel chunks give the same step
    xe = _sweeps_ep(a, LU, Sinv, rhs)
    assert np.abs(xe - x).max() <= 1e-12 * max(1, np.abs(x).max())
    # and so do the slot-form chunks
    assert a["info"]["slots"] > 0 or a["info"]["tail_n"] == a["info"]["nblocks"]
    xs = _sweeps_slot(a, LU, Sinv, rhs)
    assert np.abs(xs - x).max() <= 1e-12 * max(1, np.abs(x).max())
    # the PV magnitude unknowns of the superset system come out exactly zero
    pv_blocks = [r[0] for r in rowmap[:int((case.bus_type == 2).sum())]]
    assert np.all(x[2 * np.array(pv_blocks, int) + 1] == 0.0)


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200", "pf9241x200"])
def test_factorization_forms_the_forward_substitution(name):
    """The right-hand-side outputs of the factorization's chunks hold the
    forward substitution y = L^-1 b of every row below the dense tail
    (checked against a plain unit-lower sweep over the factor's L rows and
    against the kernels' forward chunks); a tail row's block keeps its b;
    the slot-form pass that starts from them, without its forward chunks,
    gives the Newton step of the full pass"""
    case, a, J, F, A, rhs, rowmap = _system(name)
    nb, nnz = a["info"]["nblocks"], a["info"]["nnz_lu_blocks"]
    LU, Sinv = _crout(a, A)
    y = LU[nnz:, :, 0]
    # the textbook forward substitution with the unit lower factor
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    diag = _diag(a)
    tail = set(int(k) for k in a["tail_piv"])
    yr = rhs.reshape(-1, 2).copy()
    for i in range(nb):
        if i in tail:
            continue
        for p in range(rp[i], diag[i]):
            yr[i] -= LU[p] @ yr[ci[p]]
    keep = np.array([i not in tail for i in range(nb)])
    scale = max(1.0, np.abs(yr).max())
    assert np.abs(y[keep] - yr[keep]).max() <= 1e-10 * scale
    assert np.array_equal(y[~keep], rhs.reshape(-1, 2)[~keep])
    assert np.all(LU[nnz:, :, 1] == 0.0)
    # the full slot pass and the one that starts from y
    x = _sweeps_slot(a, LU, Sinv, rhs)
    xf = _sweeps_slot(a, LU, Sinv, y.ravel(), skip_fw=True)
    assert np.abs(xf - x).max() <= 1e-10 * max(1.0, np.abs(x).max())
    ref = sla.spsolve(J.tocsc(), F)
    got = np.array([xf[2 * b + c] for b, c in rowmap])
    assert np.abs(got - ref).max() <= 1e-10 * max(1, np.abs(ref).max())


@pytest.mark.parametrize("name", ["pf14", "pf118x200", "pf1354x200", "pf9241x200"])
def test_fill_finishes_level_zero(name):
    """The Newton-step fill finishes level 0 of the elimination tree: the
    chunks from fill_level0_chunks on give the factors of the whole
    schedule, and the roles cover exactly the level-0 pivots and their L
    columns (U rows keep their Jacobian values)."""
    case, a, J, F, A, rhs, rowmap = _system(name, seed=3)
    c0 = a["info"]["fill_level0_chunks"]
    jl = a["jac_l0"]
    if a["info"]["tail_level"] == 0:
        # level 0 belongs to the dense tail: the fill leaves it alone
        assert c0 == 0 and (jl == -1).all()
        return
    assert c0 >= 1
    assert (jl >= -2).all() and (jl < len(jl)).all()
    # every L role names a pivot role, in the same column, below the pivot
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    piv = np.flatnonzero(jl == -2)
    assert (row[a["jac_lu"][piv]] == ci[a["jac_lu"][piv]]).all()
    for r in np.flatnonzero(jl >= 0):
        qp, q = a["jac_lu"][jl[r]], a["jac_lu"][r]
        assert jl[jl[r]] == -2 and ci[q] == ci[qp] and row[q] > row[qp]
    # the level-0 chunks' outputs are exactly the roles
    outs = {int(d) & MASK for d in a["tk_a"][:a["ch_task"][c0]][:, 0]}
    assert outs == set(int(q) for q in a["jac_lu"][jl != -1])
    LU, Sinv = _crout(a, A)
    LU0, Sinv0 = _crout(a, _fill_level0(a, A), c_start=c0)
    np.testing.assert_array_equal(LU0, LU)
    np.testing.assert_array_equal(Sinv0, Sinv)


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200", "pf9241x200"])
def test_tail_staging_reproduces_the_pair_lists(name):
    """The FP32 tail assembly reads its products from blocks staged in LDS:
    the staged positions are the distinct blocks of the pair lists, and each
    pair's packed indices name exactly its L and U blocks."""
    case, _ = synth.workload(name)
    a = engine.analyse(case)
    ns = a["info"]["tail_stage"]
    L, U = a["tl_pairL"], a["tl_pairU"]
    assert 0 < ns <= 5760
    stg = a["tl_stage"]
    np.testing.assert_array_equal(stg, np.unique(np.concatenate([L, U])))
    ps = a["tl_pairS"]
    np.testing.assert_array_equal(stg[ps & 0xFFFF], L)
    np.testing.assert_array_equal(stg[ps >> 16], U)


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200", "pf9241x200"])
def test_tail_entries_come_longest_first(name):
    """The tail entries are ordered by the length of their product lists,
    longest first (one lane per entry, a wave as slow as its longest list),
    and still cover every output of the dense tail once."""
    case, _ = synth.workload(name)
    a = engine.analyse(case)
    s = a["info"]["tail_n"]
    n = np.diff(a["tl_ptr"])
    assert (np.diff(n) <= 0).all()
    pos = a["tl_dense"] & MASK
    assert len(np.unique(pos)) == len(pos) and pos.max() < s * s
    # every diagonal block of the tail is an output
    assert set(range(0, s * s, s + 1)) <= set(pos.tolist())


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200"])
def test_transposed_replay_solves_the_adjoint_system(name):
    """k_adj_solve's schedule: (superset block matrix)^T lambda = g against a
    dense solve of the same block matrix"""
    case, a, J, F, A, rhs, rowmap = _system(name, seed=1)
    nb = a["info"]["nblocks"]
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    M = np.zeros((2 * nb, 2 * nb))
    for r in range(nb):
        for p in range(rp[r], rp[r + 1]):
            M[2 * r:2 * r + 2, 2 * ci[p]:2 * ci[p] + 2] = A[p]
    g = np.random.default_rng(2).normal(size=2 * nb)
    LU, Sinv = _crout(a, A)
    lam = _sweeps(a, LU, Sinv, g, trans=True)
    ref = np.linalg.solve(M.T, g)
    assert np.abs(lam - ref).max() <= 1e-10 * max(1, np.abs(ref).max())
    # and its slot form (k_adj_solve_slot), whose forward chunks also hold
    # the rows without L entries (they still apply their pivots)
    assert a["sl_chunk_t"].shape[0] > 0
    lam_s = _sweeps_slot_t(a, LU, Sinv, g)
    assert np.abs(lam_s - ref).max() <= 1e-10 * max(1, np.abs(ref).max())


def test_tail_inverse_is_the_schur_complement_inverse():
    """S^-1 from the Gauss-Jordan replay equals the inverse of
    A_SS - A_SN A_NN^-1 A_NS formed densely"""
    case, a, J, F, A, rhs, rowmap = _system("pf118x200", seed=3)
    nb = a["info"]["nblocks"]
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    M = np.zeros((2 * nb, 2 * nb))
    for r in range(nb):
        for p in range(rp[r], rp[r + 1]):
            M[2 * r:2 * r + 2, 2 * ci[p]:2 * ci[p] + 2] = A[p]
    S_ = np.concatenate([[2 * k, 2 * k + 1] for k in a["tail_piv"]])
    N_ = np.setdiff1d(np.arange(2 * nb), S_)
    schur = M[np.ix_(S_, S_)] - M[np.ix_(S_, N_)] @ np.linalg.solve(
        M[np.ix_(N_, N_)], M[np.ix_(N_, S_)])
    _, Sinv = _crout(a, A)
    s = a["info"]["tail_n"]
    got = Sinv.transpose(0, 2, 1, 3).reshape(2 * s, 2 * s)
    ref = np.linalg.inv(schur)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


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
    # every position is written at most once, by a chunk output; the blocks
    # of the dense tail are not written (the tail keeps its inverse) and a
    # U block without products keeps its Jacobian value
    # (positions nnz + k: the forward substitution's right-hand-side block
    # of pivot k, an output for every pivot below the tail with an L row)
    dst_all = a["tk_a"][:, 0] & MASK
    assert len(np.unique(dst_all)) == len(dst_all)
    tail = set(int(k) for k in a["tail_piv"])
    rhs_k = np.sort(dst_all[dst_all >= nnz] - nnz)
    has_l = [k for k in range(nb) if k not in tail and ci[rp[k]] < k]
    assert np.array_equal(rhs_k, has_l)
    dst = dst_all[dst_all < nnz]
    in_tail = np.array([(r in tail and ci[p] in tail)
                        for r in range(nb) for p in range(rp[r], rp[r + 1])])
    assert not np.any(in_tail[dst])
    rowof = np.repeat(np.arange(nb), np.diff(rp))
    unwritten = np.setdiff1d(np.flatnonzero(~in_tail), dst)
    # unwritten non-tail positions: U blocks (row < column) holding J values
    assert np.all(rowof[unwritten] < ci[unwritten])
    # a product only reads positions finished in an earlier chunk (or its
    # own pivot's block, as an init entry); an L output's pivot range lies
    # in its own chunk
    done_at = np.full(nnz + nb, 10 ** 9)
    done_at[unwritten] = -1          # final from the start
    done_at[nnz:] = -1               # b, for a pivot without an L row
    for c in range(i["chunks"]):
        done_at[dst_all[a["ch_task"][c]:a["ch_task"][c + 1]]] = c
    for c in range(i["chunks"]):
        q0, q1 = a["ch_pair"][c], a["ch_pair"][c + 1]
        assert q1 - q0 <= 4096
        for l, u in a["pairs"][q0:q1]:
            if u < 0:
                assert done_at[l] == c
            else:
                assert done_at[l] < c and done_at[u] < c
        for k in range(a["ch_task"][c], a["ch_task"][c + 1]):
            d, p0, p1, _ = a["tk_a"][k]
            assert q0 <= p0 <= p1 <= q1
            if d & LCOL:
                assert q0 <= a["tk_b"][k][1] <= a["tk_b"][k][2] <= q1
    # every block of the Jacobian pattern outside the tail holds a Jacobian
    # value: directly (HAS_A) or as an init entry
    n_init = int(((a["pairs"][:, 1] < 0) & (a["pairs"][:, 0] < nnz)).sum())
    assert n_init + int((a["tk_a"][:, 0] & HAS_A > 0).sum()) + len(unwritten) \
        + int((a["tl_dense"] & HAS_A > 0).sum()) == i["nnz_jac_blocks"]


def test_sweep_chunks_cover_the_levels():
    """the entry-parallel chunks hold every non-tail row once per sweep (the
    forward sweep only rows with L entries), in level order, with the row's
    own entries, at most 2048 per chunk"""
    case, _ = synth.workload("pf9241x200")
    a = engine.analyse(case)
    i = a["info"]
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    diag = _diag(a)
    nfw = i["sweep_fw_chunks"]
    ch, rows, ent = a["sw_chunk"], a["sw_row"], a["sw_ent"]
    assert np.all(ch[:, 3] - ch[:, 2] <= 2048)
    for lo, hi, order, fw in ((0, nfw, a["fw_rows"], True),
                              (nfw, len(ch), a["bw_rows"], False)):
        rws = rows[ch[lo, 0]:ch[hi - 1, 1]] if hi > lo else rows[:0]
        if fw:   # forward rows without L entries are left out (x = b)
            order = np.array([r for r in order if diag[r] > rp[r]])
        assert np.array_equal(rws[:, 0], order)
        for r, f0, f1, dpos in rws:
            p0, p1 = (rp[r], diag[r]) if fw else (diag[r] + 1, rp[r + 1])
            assert f1 - f0 == p1 - p0
            assert np.array_equal(ent[f0:f1, 0], np.arange(p0, p1))
            assert np.array_equal(ent[f0:f1, 1], ci[p0:p1])
            assert dpos == (-1 if fw else diag[r])


@pytest.mark.parametrize("name", ["pf1354x200", "pf9241x200"])
def test_slot_chunks_cover_the_levels(name):
    """the slot-form chunks hold every non-tail row once per sweep (the
    forward sweep only rows with L entries), in level order; a row's lane
    group is the power of two at least its entry count, starts at a multiple
    of its size within the chunk, holds the row's own entries in order and
    then padding slots; at most 2048 slots per chunk"""
    case, _ = synth.workload(name)
    a = engine.analyse(case)
    _check_slots(a, a["info"]["tail_n"])


def _check_slots(a, tail_n):
    i = a["info"]
    rp, ci = a["lu_rowptr"], a["lu_colind"]
    diag = _diag(a)
    nfw = i["slot_fw_chunks"]
    ch, sl = a["sl_chunk"], a["sl_slot"]
    assert i["slots"] == len(sl) > 0 and len(ch) == nfw + i["slot_bw_chunks"]
    assert np.all(ch[:, 1] - ch[:, 0] <= 2048) and np.all(ch[:, 1] > ch[:, 0])
    assert np.array_equal(ch[1:, 0], ch[:-1, 1]) and ch[0, 0] == 0 and ch[-1, 1] == len(sl)
    LG = 27
    long_rows = 0
    for lo, hi, order, fw in ((0, nfw, a["fw_rows"], True),
                              (nfw, len(ch), a["bw_rows"], False)):
        if fw:
            order = np.array([r for r in order if diag[r] > rp[r]])
        got = {}            # row -> its entries so far, in segment order
        done = []           # rows in the order their last segment comes
        for c in range(lo, hi):
            s0, s1 = ch[c, 0], ch[c, 1]
            q, gmax, rows_here = s0, 0, set()
            assert s0 % 64 == 0 and s1 % 64 == 0
            while q < s1:
                pos, w, lead, piv = sl[q]
                if lead < 0:
                    # padding to the chunk's 64-slot end: no entry, no row,
                    # fewer than 64 slots
                    pad = sl[q:s1]
                    assert s1 - q < 64 and np.all(pad[:, 1] == 0) and np.all(pad[:, 2] == -1)
                    assert np.all(pad[:, 0] == pad[:, 3]) and np.all(pad[:, 0] == sl[q - 1, 0])
                    break
                g = w >> LG
                gmax = max(gmax, g)
                row, last = lead & (NOT_LAST - 1), not lead & NOT_LAST
                assert lead >= 0 and (q - s0) % (1 << g) == 0 and q + (1 << g) <= s1
                assert row not in rows_here         # one segment of a row per chunk
                rows_here.add(row)
                grp = sl[q:q + (1 << g)]
                colp = grp[:, 1] & ((1 << LG) - 1)
                n = int((colp > 0).sum())
                assert np.all(colp[:n] > 0) and np.all(colp[n:] == 0)
                assert (1 << g) >= n and (g == 0 or (1 << (g - 1)) < n) and n <= 64
                assert np.all(grp[:, 1] >> LG == g) and np.all(grp[1:, 2] == -1)
                got.setdefault(row, ([], []))
                got[row][0].extend(grp[:n, 0])
                got[row][1].extend(colp[:n] - 1)
                assert piv == (diag[row] if not fw and last else pos)
                if last:
                    p0, p1 = (rp[row], diag[row]) if fw else (diag[row] + 1, rp[row + 1])
                    assert np.array_equal(got[row][0], np.arange(p0, p1))
                    assert np.array_equal(got[row][1], ci[p0:p1])
                    long_rows += p1 - p0 > 64
                    done.append(row)
                q += 1 << g
            assert ch[c, 2] == gmax
        # every row once, in level order (rows of one level may come in any
        # order of their last segments)
        assert sorted(done) == sorted(order.tolist())
        lv = {}
        for l, (b0, b1) in enumerate(zip(*(lambda p: (p[:-1], p[1:]))(a["fw_ptr" if fw else "bw_ptr"]))):
            for r in (a["fw_rows"] if fw else a["bw_rows"])[b0:b1]:
                lv[int(r)] = l
        levels = [lv[r] for r in done]
        assert levels == sorted(levels)
    return long_rows


def test_slot_form_splits_long_rows():
    """case9241pegase has rows of more than 64 entries below its dense tail:
    they come in segments of 64, one per chunk, the last one applying the
    pivot inverse; the slot-form replay gives the sparse direct solve's step"""
    import os
    from mapol_amd.engine import load_matpower
    from tests.test_matpower import write_case
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "matpower_cases.npz"))
    import tempfile
    d = tempfile.mkdtemp()
    path = os.path.join(d, "case9241pegase.m")
    write_case(path, float(z["case9241pegase/baseMVA"]), z["case9241pegase/bus"],
               z["case9241pegase/gen"], z["case9241pegase/branch"])
    case = load_matpower(path)[0]
    case, a, J, F, A, rhs, rowmap = _system(None, case=case)
    assert _check_slots(a, a["info"]["tail_n"]) > 0
    LU, Sinv = _crout(a, A)
    x = _sweeps(a, LU, Sinv, rhs)
    xs = _sweeps_slot(a, LU, Sinv, rhs)
    assert np.abs(xs - x).max() <= 1e-11 * max(1, np.abs(x).max())
    ref = sla.spsolve(J.tocsc(), F)
    got = np.array([x[2 * b + c] for b, c in rowmap])
    assert np.abs(got - ref).max() <= 1e-9 * max(1, np.abs(ref).max())


def _numbering_check(tmp_path_factory):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path_factory.mktemp("numbering") / "numbering_check")
    subprocess.run(["g++", "-std=c++17", "-O2",
                    os.path.join(root, "tests", "cpp", "numbering_check.cpp"),
                    os.path.join(root, "mapol_amd", "csrc", "symbolic.cpp"),
                    "-o", exe], check=True, capture_output=True)
    return exe


def _run_check(exe, n, m, types, f, t, ordering):
    import subprocess
    txt = f"{n} {m} {ordering}\n" + " ".join(map(str, types)) + "\n" + \
        "\n".join(f"{a} {b}" for a, b in zip(f, t)) + "\n"
    return subprocess.run([exe], input=txt, capture_output=True, text=True)


def test_engine_bus_numbering_keeps_the_analysis(tmp_path_factory):
    """the engine's own bus numbers (reverse Cuthill-McKee) with the
    elimination order taken from the caller's numbering give the caller's
    block pattern, schedules and dense tail, array for array
    (tests/cpp/numbering_check.cpp), and bring branch ends closer"""
    exe = _numbering_check(tmp_path_factory)
    cases = [("pf118x200", (0, 1, 2)), ("pf1354x200", (0, 1, 2)),
             ("pf9241x200", (0,))]
    for name, orders in cases:
        case, _ = synth.workload(name, T=2)
        for o in orders:
            r = _run_check(exe, case.nbus, case.nbranch, case.bus_type,
                           case.br_from, case.br_to, o)
            assert r.returncode == 0 and r.stdout.startswith("identical"), \
                (name, o, r.stdout, r.stderr)
            before, after = map(float, r.stdout.split()[-3::2])
            assert after < before / 4, (name, r.stdout)
    # several islands, an isolated bus, parallel branches, a PV-only island
    types = [3, 1, 2, 1, 1, 1, 1, 2, 2]
    f = [0, 1, 1, 2, 4, 7, 3]
    t = [1, 2, 2, 3, 5, 8, 0]
    for o in (0, 1, 2):
        r = _run_check(exe, len(types), len(f), types, f, t, o)
        assert r.returncode == 0 and r.stdout.startswith("identical"), (o, r.stdout)


def test_split_factorization_schedules(tmp_path_factory):
    """the split factorization's group schedules (tests/cpp/split_check.cpp):
    every output below the cut exactly once with the main schedule's
    products, order and lane groups; no group reads another group's outputs
    or its own before they are made; the dense tail's wave jobs cover every
    tail entry once"""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path_factory.mktemp("split") / "split_check")
    subprocess.run(["g++", "-std=c++17", "-O2",
                    os.path.join(root, "tests", "cpp", "split_check.cpp"),
                    os.path.join(root, "mapol_amd", "csrc", "symbolic.cpp"),
                    "-o", exe], check=True, capture_output=True)
    # (ordering, variants planned) per case
    for name, plans in [("pf118x200", ((0, 0),)),
                        ("pf1354x200", ((0, 0), (1, 0))),
                        ("pf2869x100", ((0, 2), (1, 3))),
                        ("pf9241x200", ((0, 3),))]:
        case, _ = synth.workload(name, T=2)
        for o, planned in plans:
            r = _run_check(exe, case.nbus, case.nbranch, case.bus_type,
                           case.br_from, case.br_to, o)
            assert r.returncode == 0, (name, o, r.stdout, r.stderr)
            assert r.stdout.strip().endswith(f"ok, {planned} variants"), (name, o, r.stdout)
