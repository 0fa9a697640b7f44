"""The engine's own bus numbering (reverse Cuthill-McKee, bpf_options
bus_numbering = 1, the default) stays inside the engine: against an engine
that keeps the caller's numbering (bus_numbering = 0), on the same case,
outages and controls (AVR set-points and shunts, which name buses, and OLTC
taps), every per-bus input and output of the interface agrees:

* solve results (voltages, bus types after reactive-limit switching,
  iterations, verdicts, losses) and branch flows;
* the Ybus pattern (rows by bus, columns ascending) and its values, bit for
  bit (each entry adds the same branch terms in the same order);
* the stage entries in the reference's layouts (mismatch, Jacobian, Newton
  step, the dense tail's buses and inverse, the adjoint's gradients);
* fast-decoupled B' and B'' and one P and Q half-step.

Both engines factor with the same elimination order (taken from the caller's
numbering), so they differ only in the order of the terms each bus's
current adds: results agree to rounding, iterations exactly.  The oracle
tests pin either against the reference's algorithm."""
import numpy as np
import pytest

from mapol_amd import synth
from mapol_amd.engine import BPF_FDXB, BPF_MIXED, Engine

pytestmark = pytest.mark.gpu

AVR, OLTC, SHC = 1, 2, 3


def _controls(case, T, seed):
    """one AVR (a PV bus), one SHC (a PQ bus) and one OLTC (a tapped branch)
    control per instance, NaN (keep the default) in every third instance"""
    rng = np.random.default_rng(seed)
    pv = np.flatnonzero(case.bus_type == 2)
    pq = np.flatnonzero(case.bus_type == 1)
    tapped = np.flatnonzero(np.asarray(case.tap) != 0)
    ctl = [(AVR, int(pv[len(pv) // 3])), (SHC, int(pq[len(pq) // 2]))]
    if len(tapped):
        ctl.append((OLTC, int(tapped[0])))
    val = np.empty((T, len(ctl)))
    val[:, 0] = rng.uniform(0.98, 1.04, T)
    val[:, 1] = rng.uniform(0.0, 0.15, T)
    if len(tapped):
        val[:, 2] = rng.uniform(0.95, 1.05, T)
    val[::3, :] = np.nan
    return ctl, val


def _pair(case, cont, ctl, val, **kw):
    out = []
    for b in (0, 1):
        e = Engine(case, bus_numbering=b, **kw)
        e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on, ctl=ctl, ctl_val=val)
        out.append(e)
    return out


def _close(a, b, rel):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    scale = max(float(np.abs(b).max()), 1e-300)
    assert float(np.abs(a - b).max()) <= rel * scale


@pytest.mark.parametrize("name", ["pf118x200", "pf1354x200"])
def test_newton_interface_in_callers_numbering(name):
    T = 8
    print(f"[numbering] Newton, {name}, {T} instances", flush=True)
    case, cont = synth.workload(name, T=T)
    ctl, val = _controls(case, T, 5)
    e0, e1 = _pair(case, cont, ctl, val, precision=BPF_MIXED, refine_steps=1,
                   enforce_qlim=True)
    for e in (e0, e1):
        e.solve()
    r0, r1 = e0.results(), e1.results()
    np.testing.assert_array_equal(r0["verdict"], r1["verdict"])
    np.testing.assert_array_equal(r0["iterations"], r1["iterations"])
    np.testing.assert_array_equal(r0["bus_type"], r1["bus_type"])
    assert np.abs(r0["V"] - r1["V"]).max() <= 1e-9
    _close(r1["loss"], r0["loss"], 1e-9)
    for a, b in zip(e0.branch_flows(), e1.branch_flows()):
        _close(b, a, 1e-8)
    for t in (0, T - 1):
        Y0, Y1 = e0.ybus(t), e1.ybus(t)
        np.testing.assert_array_equal(Y0.indptr, Y1.indptr)
        np.testing.assert_array_equal(Y0.indices, Y1.indices)
        np.testing.assert_array_equal(Y0.data, Y1.data)
    # stage entries from perturbed voltages
    rng = np.random.default_rng(7)
    V = r0["V"] * (1 + 0.01 * rng.standard_normal(r0["V"].shape)) \
        * np.exp(0.01j * rng.standard_normal(r0["V"].shape))
    F0, n0 = e0.stage_mismatch(V)
    F1, n1 = e1.stage_mismatch(V)
    _close(F1, F0, 1e-12)
    _close(n1, n0, 1e-12)
    for t in (0, 3):
        J0, J1 = e0.stage_jacobian(V, t), e1.stage_jacobian(V, t)
        assert J0.shape == J1.shape
        assert abs(J0 - J1).max() <= 1e-12 * abs(J0).max()
    dx0, dx1 = e0.stage_newton_step(V), e1.stage_newton_step(V)
    _close(dx1, dx0, 1e-6)
    for t in (0, 5):
        b0, S0 = e0.stage_tail_inverse(t)
        b1, S1 = e1.stage_tail_inverse(t)
        np.testing.assert_array_equal(b0, b1)
        _close(S1, S0, 1e-4)
    # the adjoint at the solution
    for e in (e0, e1):
        e.solve()
    rate = np.full(case.nbranch, 1.0)
    g0, g1 = e0.adjoint(rate, 0.5), e1.adjoint(rate, 0.5)
    for k in ("grad_P", "grad_Q", "grad_x", "objective"):
        _close(g1[k], g0[k], 1e-8)
    e0.close()
    e1.close()


def test_fast_decoupled_interface_in_callers_numbering():
    T = 6
    print(f"[numbering] fast decoupled, pf1354x200, {T} instances", flush=True)
    case, cont = synth.workload("pf1354x200", T=T)
    ctl, val = _controls(case, T, 9)
    e0, e1 = _pair(case, cont, ctl, val, alg=BPF_FDXB)
    for e in (e0, e1):
        e.solve()
    r0, r1 = e0.results(), e1.results()
    np.testing.assert_array_equal(r0["verdict"], r1["verdict"])
    np.testing.assert_array_equal(r0["iterations"], r1["iterations"])
    assert np.abs(r0["V"] - r1["V"]).max() <= 1e-9
    for t in (0, T - 1):
        for which in (1, 2):
            B0, B1 = e0.stage_bmat(t, which), e1.stage_bmat(t, which)
            assert B0.shape == B1.shape
            assert abs(B0 - B1).max() <= 1e-14 * abs(B0).max()
    e0.make_ybus()
    e1.make_ybus()
    rng = np.random.default_rng(3)
    V = r0["V"] * (1 + 0.01 * rng.standard_normal(r0["V"].shape))
    for which in (1, 2):
        V0, n0 = e0.stage_fd_halfstep(V, which)
        V1, n1 = e1.stage_fd_halfstep(V, which)
        assert np.abs(V0 - V1).max() <= 1e-11
        _close(n1, n0, 1e-12)
    e0.close()
    e1.close()

