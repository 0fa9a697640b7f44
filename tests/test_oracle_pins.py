"""Independent pins of the oracle's initial voltage, Newton loop and fast-
decoupled path (CPU only).

Each check computes its answer by a route that shares no code with
oracle/pf_oracle.c, chosen so that a plausible slip in the oracle fails it:

* V0 (runpf.h:211-226): the closed form |V0| = set-point, arg V0 = table
  angle, with the set-point rule written out per bus (AVR control, else the
  last in-service generator's VG at a PV bus, else the table Vm).  Using Vm
  for a PV bus, the first generator's VG, an ignored control or degrees for
  radians all fail.
* Newton (newtonpf.h:365-523): the textbook matrix form of the polar
  Newton-Raphson power flow (complex derivatives dS/dVa = jV conj(diag(I) -
  Y diag(V)), dS/dVm = diag(V) conj(Y diag(V/|V|)) + conj(diag(I)) diag(V/|V|),
  as published with MATPOWER, Zimmerman et al., IEEE Trans. Power Syst.
  26(1), 2011), a Ybus assembled from the raw tables with scipy and a SuperLU
  solve, against the oracle's element-wise restatement, on the reference's
  own case files (src/datasets/*.m) and the synthetic grids: the same
  iteration counts and |dV| <= 1e-9.  A dropped Jacobian term or a wrong sign
  changes the iteration counts; a mismatch error moves the fixed point.
* B', B'' (makeB.h:311-501): written entry by entry from the pi-model of a
  4-bus case for both variants (XB: R = 0 in B'; BX: R = 0 in B''), with a
  tap on either side, a phase shifter between two non-slack buses (kept in
  B', zeroed in B''), line charging and shunts (dropped in B', kept in B'').
  Swapping XB and BX, keeping shunts or taps in B', or dropping the shift
  fails.
* FDPF (fdpf.h:450-596): the half-step loop written out in numpy over dense
  B'/B'' built from the published makeB rules, against the oracle: the same
  iteration counts and |dV| <= 1e-9 on the reference's case files and on
  synthetic contingencies.

The case files' own Vm/Va are not used as known answers: the mismatch at
them is 1e-2 pu (case14, rounded to three decimals) to 13-42 pu (the PEGASE
grids, whose phase-shift column the reference reads in degrees), so they are
not power-flow solutions of the files' data under the reference's model.
case14's stored solution is still checked at its printed precision class
(|dVm| <= 2e-3, |dVa| <= 0.02 deg), a coarse known answer.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from mapol_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
O = oracle.Oracle
PQ, PV, SLACK = 1, 2, 3


# ---------------------------------------------------------------------------
# the reference's case files, from the numeric tables (no loader under test)
# ---------------------------------------------------------------------------
def case_from_tables(z, name, honour_gen_status=True):
    """MATPOWER tables -> Case with the reference loader's units (per unit on
    baseMVA, Va in radians, 0-based buses in file order)."""
    base = float(z[f"{name}/baseMVA"])
    bus, gen, br = z[f"{name}/bus"], z[f"{name}/gen"], z[f"{name}/branch"]
    ids = {int(b): k for k, b in enumerate(bus[:, 0])}
    gb = np.array([ids[int(b)] for b in gen[:, 0]], np.int32)
    qmax = np.where(np.isinf(gen[:, 3]), np.inf, gen[:, 3] / base)
    qmin = np.where(np.isinf(gen[:, 4]), -np.inf, gen[:, 4] / base)
    gstat = (gen[:, 7] > 0) if honour_gen_status else np.ones(len(gen), bool)
    return synth.Case(
        bus_type=bus[:, 1].astype(np.int32), Pd=bus[:, 2] / base,
        Qd=bus[:, 3] / base, Gs=bus[:, 4] / base, Bs=bus[:, 5] / base,
        Vm=bus[:, 7].copy(), Va=np.deg2rad(bus[:, 8]),
        br_from=np.array([ids[int(b)] for b in br[:, 0]], np.int32),
        br_to=np.array([ids[int(b)] for b in br[:, 1]], np.int32),
        R=br[:, 2].copy(), X=br[:, 3].copy(), B=br[:, 4].copy(),
        tap=br[:, 8].copy(), shift=br[:, 9].copy(),
        br_status=(br[:, 10] > 0).astype(np.uint8), gen_bus=gb,
        Pg=gen[:, 1] / base, Qg=gen[:, 2] / base, Qmax=qmax, Qmin=qmin,
        VG=gen[:, 5].copy(), gen_status=gstat.astype(np.uint8))


@pytest.fixture(scope="module")
def tables():
    return np.load(os.path.join(HERE, "golden", "matpower_cases.npz"))


# ---------------------------------------------------------------------------
# numpy restatements (independent of the oracle's code)
# ---------------------------------------------------------------------------
def instance(case, br_on=None, gen_on=None, vset=None):
    """bus types, injections and voltage set-points of one instance"""
    n = case.nbus
    typ = case.bus_type.astype(int).copy()
    S = -(case.Pd + 1j * case.Qd)
    on_g = case.gen_status.astype(bool)
    if gen_on is not None:
        on_g = on_g & gen_on.astype(bool)
    ngen = np.zeros(n, int)
    v = case.Vm.astype(float).copy()
    for k in range(case.ngen):                 # later rows overwrite VG
        if not on_g[k]:
            continue
        b = case.gen_bus[k]
        ngen[b] += 1
        S[b] += case.Pg[k] + 1j * case.Qg[k]
        if case.bus_type[b] == PV:
            v[b] = case.VG[k]
    typ[(typ == PV) & (ngen == 0)] = PQ
    if vset is not None:
        v = np.where(np.isnan(vset), v, vset)
    on_b = case.br_status.astype(bool)
    if br_on is not None:
        on_b = on_b & br_on.astype(bool)
    return typ, S, v, on_b


def branch_admittances(R, X, B, tap, shift, on):
    """pi-model of each branch: (Yff, Yft, Ytf, Ytt), tap 0 meaning 1,
    shift in degrees (MATPOWER's branch model)"""
    st = on.astype(float)
    ys = st / (R + 1j * X)
    tau = np.where(tap != 0, tap, 1.0) * np.exp(1j * np.deg2rad(shift))
    ytt = ys + 1j * st * B / 2
    return ytt / (tau * np.conj(tau)), -ys / np.conj(tau), -ys / tau, ytt


def ybus_np(case, on, R=None, B=None, tap=None, shift=None, Bs=None):
    n = case.nbus
    f, t = case.br_from, case.br_to
    yff, yft, ytf, ytt = branch_admittances(
        case.R if R is None else R, case.X, case.B if B is None else B,
        case.tap if tap is None else tap, case.shift if shift is None else shift,
        on)
    Y = sp.csr_matrix((np.r_[yff, yft, ytf, ytt],
                       (np.r_[f, f, t, t], np.r_[f, t, f, t])), shape=(n, n))
    bs = case.Bs if Bs is None else Bs
    return (Y + sp.diags(case.Gs + 1j * bs)).tocsc()


def angle_ref(V):
    """the reference's angle, atan(Im/Re) (src/util/complexUtils.h:21)"""
    return np.arctan(V.imag / V.real)


def newton_matrix_form(case, br_on=None, gen_on=None, tol=1e-8, maxit=10):
    """polar Newton-Raphson in matrix form; returns (iterations, converged, V)"""
    typ, Sbus, v, on = instance(case, br_on, gen_on)
    Y = ybus_np(case, on)
    V = v * np.exp(1j * case.Va)
    pv = np.flatnonzero(typ == PV)
    pq = np.flatnonzero(typ == PQ)
    pvpq = np.r_[pv, pq]

    def F_of(V):
        mis = V * np.conj(Y @ V) - Sbus
        return np.r_[mis[pvpq].real, mis[pq].imag]

    F = F_of(V)
    it = 0
    conv = np.abs(F).max() < tol
    while not conv and it < maxit:
        it += 1
        Ib = Y @ V
        Vn = V / np.abs(V)
        dV = sp.diags(V)
        dS_dVm = dV @ np.conj(Y @ sp.diags(Vn)) + np.conj(sp.diags(Ib)) @ sp.diags(Vn)
        dS_dVa = 1j * dV @ np.conj(sp.diags(Ib) - Y @ dV)
        dS_dVm, dS_dVa = dS_dVm.tocsr(), dS_dVa.tocsr()
        J = sp.bmat([[dS_dVa[pvpq][:, pvpq].real, dS_dVm[pvpq][:, pq].real],
                     [dS_dVa[pq][:, pvpq].imag, dS_dVm[pq][:, pq].imag]]).tocsc()
        dx = spla.spsolve(J, F)
        Va, Vm = angle_ref(V), np.abs(V)
        Va[pvpq] -= dx[:len(pvpq)]
        Vm[pq] -= dx[len(pvpq):]
        V = Vm * np.exp(1j * Va)
        F = F_of(V)
        conv = np.abs(F).max() < tol
    return it, conv, V


def bmats_np(case, typ, on, alg, Bs):
    """B' over pv then pq and B'' over pq by the published makeB rules"""
    pv = np.flatnonzero(typ == PV)
    pq = np.flatnonzero(typ == PQ)
    pvpq = np.r_[pv, pq]
    z = np.zeros(case.nbranch)
    Yp = ybus_np(case, on, R=z if alg == oracle.FDXB else None, B=z, tap=z,
                 Bs=np.zeros(case.nbus))
    Ypp = ybus_np(case, on, R=z if alg == oracle.FDBX else None, shift=z, Bs=Bs)
    Bp = -Yp.imag.tocsr()[pvpq][:, pvpq].toarray()
    Bpp = -Ypp.imag.tocsr()[pq][:, pq].toarray()
    return Bp, Bpp, pv, pq


def fdpf_np(case, alg, br_on=None, gen_on=None, tol=1e-8, maxit=30):
    """the half-step loop of fdpf.h:450-596 with dense solves"""
    typ, Sbus, v, on = instance(case, br_on, gen_on)
    Y = ybus_np(case, on)
    V = v * np.exp(1j * case.Va)
    Bp, Bpp, pv, pq = bmats_np(case, typ, on, alg, case.Bs)
    pvpq = np.r_[pv, pq]

    def mis(V):
        m = (V * np.conj(Y @ V) - Sbus) / np.abs(V)
        P, Q = m[pvpq].real, m[pq].imag
        return P, Q, max(np.abs(P).max(initial=-1.0), np.abs(Q).max(initial=-1.0))

    P, Q, err = mis(V)
    conv = err < tol
    it = 0
    while not conv and it < maxit:
        it += 1
        dVa = np.linalg.solve(Bp, P)
        V[pvpq] = np.abs(V[pvpq]) * np.exp(1j * (angle_ref(V[pvpq]) - dVa))
        P, Q, err = mis(V)
        if err < tol:
            conv = True
            break
        dVm = np.linalg.solve(Bpp, Q)
        V[pq] = (np.abs(V[pq]) - dVm) * np.exp(1j * angle_ref(V[pq]))
        P, Q, err = mis(V)
        conv = err < tol
    return it, conv, V


# ---------------------------------------------------------------------------
# V0
# ---------------------------------------------------------------------------
def test_initial_voltage_closed_form(tables):
    """|V0| = set-point, arg V0 = table angle; set-point = AVR control, else
    the last in-service generator's VG at a PV bus, else Vm"""
    case = case_from_tables(tables, "case118")
    # a second generator at a PV bus with another VG: the later row wins
    k0 = int(np.flatnonzero(case.bus_type[case.gen_bus] == PV)[0])
    b0 = int(case.gen_bus[k0])
    for f in ("gen_bus", "Pg", "Qg", "Qmax", "Qmin", "VG", "gen_status"):
        a = getattr(case, f)
        setattr(case, f, np.r_[a, a[k0:k0 + 1]])
    case.VG[-1] = case.VG[k0] + 0.013
    case.Pg[-1] = 0.0
    assert np.abs(case.VG[case.bus_type[case.gen_bus] == PV] -
                  case.Vm[case.gen_bus[case.bus_type[case.gen_bus] == PV]]).max() > 1e-3
    rng = np.random.default_rng(5)
    gen_on = np.ones(case.ngen, np.uint8)
    pvgens = np.flatnonzero(case.bus_type[case.gen_bus] == PV)
    gen_on[rng.choice(pvgens, 6, replace=False)] = 0
    vset = np.full(case.nbus, np.nan)
    avr = rng.choice(np.flatnonzero(case.bus_type == PV), 5, replace=False)
    vset[avr] = rng.uniform(0.95, 1.05, size=5)
    for kw in (dict(), dict(gen_on=gen_on), dict(gen_on=gen_on, vset=vset)):
        V0 = O(case).initial_voltage(**kw)
        want = np.where(case.bus_type == SLACK, case.Vm, case.Vm.copy())
        on = case.gen_status.astype(bool)
        if "gen_on" in kw:
            on &= kw["gen_on"].astype(bool)
        for k in range(case.ngen):
            b = case.gen_bus[k]
            if on[k] and case.bus_type[b] == PV:
                want[b] = case.VG[k]
        if "vset" in kw:
            want = np.where(np.isnan(vset), want, vset)
        np.testing.assert_allclose(np.abs(V0), want, rtol=5e-16, atol=0)
        np.testing.assert_allclose(np.angle(V0), case.Va, rtol=0, atol=4e-16)
    assert np.abs(np.abs(O(case).initial_voltage())[b0] - case.VG[-1]) < 1e-15


# ---------------------------------------------------------------------------
# Newton on the reference's case files and on synthetic contingencies
# ---------------------------------------------------------------------------
def _compare_newton(case, cont_idx=(), cont=None):
    orc = O(case)
    runs = [dict()] + [dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
                       for t in cont_idx]
    for kw in runs:
        it, conv, V = newton_matrix_form(case, **kw)
        r = orc.solve(**kw)
        assert r["iterations"] == it
        assert r["converged"] == conv
        if conv:
            assert np.abs(r["V"] - V).max() <= 1e-9


@pytest.mark.parametrize("name", ["case14", "case57", "case118", "case300",
                                  "case1354pegase", "case2869pegase",
                                  "case9241pegase"])
def test_newton_matrix_form_on_case_files(tables, name):
    case = case_from_tables(tables, name)
    cont = synth.n1_contingencies(case, 3, seed=11, gen_frac=0.34)
    _compare_newton(case, range(3) if case.nbus <= 3000 else range(1), cont)


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200"])
def test_newton_matrix_form_on_synthetic(name):
    case, cont = synth.workload(name)
    _compare_newton(case, (0, 57, 133), cont)


def test_case14_stored_solution():
    """IEEE 14-bus: the stored Vm (3 decimals) / Va (2 decimals) within the
    precision class of the published solution"""
    z = np.load(os.path.join(HERE, "golden", "matpower_cases.npz"))
    case = case_from_tables(z, "case14")
    r = O(case).solve()
    bus = z["case14/bus"]
    assert r["converged"]
    assert np.abs(np.abs(r["V"]) - bus[:, 7]).max() <= 2e-3
    assert np.rad2deg(np.abs(np.angle(r["V"]) - np.deg2rad(bus[:, 8]))).max() <= 0.02


# ---------------------------------------------------------------------------
# B', B'' by hand
# ---------------------------------------------------------------------------
def _four_bus():
    return synth.Case(
        bus_type=np.array([3, 2, 1, 1], np.int32), Pd=np.array([0, .2, .5, .4]),
        Qd=np.array([0, .05, .15, .1]), Gs=np.array([0, 0, .01, .02]),
        Bs=np.array([0, .05, .1, .07]), Vm=np.array([1.02, 1.01, 1, 1]),
        Va=np.zeros(4),
        br_from=np.array([0, 2, 0, 2, 1], np.int32),
        br_to=np.array([1, 1, 2, 3, 3], np.int32),
        R=np.array([.02, .03, .01, .025, .015]),
        X=np.array([.1, .15, .08, .12, .09]),
        B=np.array([.04, .02, .03, .05, .01]),
        tap=np.array([0, .95, 0, 1.05, 0]),
        shift=np.array([0, 0, 0, 8.0, 0]),
        br_status=np.ones(5, np.uint8), gen_bus=np.array([0, 1], np.int32),
        Pg=np.array([0, .6]), Qg=np.zeros(2), Qmax=np.array([9, 9.]),
        Qmin=np.array([-9, -9.]), VG=np.array([1.02, 1.01]),
        gen_status=np.ones(2, np.uint8))


@pytest.mark.parametrize("alg", [oracle.FDXB, oracle.FDBX])
def test_bmatrices_by_hand(alg):
    c = _four_bus()
    R, X, B = c.R, c.X, c.B
    # -Im of a series admittance, with R kept or zeroed
    def y(k, zero_r):
        return 1 / X[k] if zero_r else X[k] / (R[k] ** 2 + X[k] ** 2)

    def g(k, zero_r):
        return 0.0 if zero_r else R[k] / (R[k] ** 2 + X[k] ** 2)

    zr_p, zr_pp = alg == oracle.FDXB, alg == oracle.FDBX
    # B' over (pv: 1; pq: 2, 3): taps 1, no charging, no shunts, shift of
    # branch 3 (2 -> 3, 8 deg) kept: -Im(-ys e^{+-j th}) = b cos th -+ g sin th
    th = np.deg2rad(8.0)
    Bp = np.zeros((3, 3))
    Bp[0, 0] = y(0, zr_p) + y(1, zr_p) + y(4, zr_p)
    Bp[1, 1] = y(1, zr_p) + y(2, zr_p) + y(3, zr_p)
    Bp[2, 2] = y(3, zr_p) + y(4, zr_p)
    Bp[0, 1] = Bp[1, 0] = -y(1, zr_p)
    Bp[0, 2] = Bp[2, 0] = -y(4, zr_p)
    # from bus 2 (row 1) to bus 3 (column 2): Yft = -ys / conj(tau)
    Bp[1, 2] = -(y(3, zr_p) * np.cos(th) - g(3, zr_p) * np.sin(th))
    Bp[2, 1] = -(y(3, zr_p) * np.cos(th) + g(3, zr_p) * np.sin(th))
    # B'' over (pq: 2, 3): shift zeroed, taps, charging and shunts kept
    Bpp = np.zeros((2, 2))
    Bpp[0, 0] = ((y(1, zr_pp) - B[1] / 2) / 0.95 ** 2      # from side of 2->1
                 + y(2, zr_pp) - B[2] / 2                    # to side of 0->2
                 + (y(3, zr_pp) - B[3] / 2) / 1.05 ** 2      # from side of 2->3
                 - c.Bs[2])
    Bpp[1, 1] = (y(3, zr_pp) - B[3] / 2 + y(4, zr_pp) - B[4] / 2 - c.Bs[3])
    Bpp[0, 1] = Bpp[1, 0] = -y(3, zr_pp) / 1.05
    orc = O(c)
    np.testing.assert_allclose(orc.bmat(alg, 1).toarray(), Bp, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(orc.bmat(alg, 2).toarray(), Bpp, rtol=1e-14, atol=1e-12)
    # an outage removes the branch from both; a generator outage moves its
    # bus from the P-only rows of B' to the pq set of B''
    on = np.ones(5, np.uint8)
    on[4] = 0
    Bp2 = orc.bmat(alg, 1, br_on=on).toarray()
    assert Bp2[0, 2] == 0 and Bp2[2, 0] == 0
    assert Bp2[0, 0] == pytest.approx(Bp[0, 0] - y(4, zr_p), rel=1e-14)
    assert orc.bmat(alg, 2, gen_on=np.array([1, 0], np.uint8)).shape == (3, 3)


# ---------------------------------------------------------------------------
# FDPF loop
# ---------------------------------------------------------------------------
def _compare_fdpf(case, runs, alg):
    orc = O(case)
    for kw in runs:
        it, conv, V = fdpf_np(case, alg, **kw)
        r = orc.solve(O.opts(alg=alg), **kw)
        assert r["iterations"] == it, kw
        assert r["converged"] == conv
        if conv:
            assert np.abs(r["V"] - V).max() <= 1e-9


@pytest.mark.parametrize("alg", [oracle.FDXB, oracle.FDBX])
@pytest.mark.parametrize("name", ["case14", "case118", "case300"])
def test_fdpf_loop_on_case_files(tables, name, alg):
    case = case_from_tables(tables, name)
    cont = synth.n1_contingencies(case, 2, seed=3, gen_frac=0.5)
    _compare_fdpf(case, [dict()] + [dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
                                    for t in range(2)], alg)


@pytest.mark.parametrize("alg", [oracle.FDXB, oracle.FDBX])
def test_fdpf_loop_on_synthetic(alg):
    case, cont = synth.workload("pf118x200")
    _compare_fdpf(case, [dict(br_on=cont.br_on[t], gen_on=cont.gen_on[t])
                         for t in (0, 41, 99, 170)], alg)
