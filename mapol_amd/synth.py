"""Seeded synthetic transmission grids and contingency sets.

The reference solves MATPOWER cases read by ``loadMatpowerFormat``
(src/files/loadMatpowerFormat.h:14-209).  This module produces cases of the
same shape (bus / branch / generator tables, per-unit on a 100 MVA base) from a
seed, so that every workload can be regenerated bit-for-bit on any machine.

Grid recipe (all draws from ``numpy.random.Generator(PCG64(seed))`` in the
order written here):

1. ``n`` uniform points in the unit square.
2. Candidate edges: the 8 nearest neighbours of every point (``cKDTree``).
3. A random spanning tree: Kruskal over the candidate edges with i.i.d.
   U(0,1) weights (a random *local* tree; components left over are joined by
   their closest pair of points).
4. Chords: the remaining candidate edges in order of increasing length until
   the branch count reaches ``m``.
5. Slack = the bus nearest the centre; ``npv`` PV buses drawn without
   replacement from the rest; all other buses PQ.
6. Branches: r ~ U(0.01, 0.08), x = r * U(2, 5), b ~ U(0, 0.05); a fraction
   ``tap_frac`` of branches carries an off-nominal tap ~ U(tap_lo, tap_hi),
   the others tap 0 (MATPOWER's "no transformer", i.e. ratio 1).
7. Loads on every bus: Pd ~ U(0, 0.5) * load_scale pu, Qd = 0.3 Pd.
8. One generator per PV bus plus one at the slack.  Each generator owns the
   buses electrically closest to it (Dijkstra over x); a PV generator
   supplies its cell's load plus a load-proportional share of the
   DC-estimated losses, so the slack's net injection stays near zero.
   Set-points |V| ~ U(1.0, 1.05) (slack too).
9. ``start="flat"``: the bus table holds a flat start (|V| = set-point on
   generator buses, 1.0 elsewhere, all angles 0).
   ``start="solved"``: the generator solves the base case once with a
   distributed slack (every PV generator picks up losses in proportion to
   its output; own small Newton solver below), writes the adjusted
   generation back and stores the solved |V|, angle in the bus table -- the
   way MATPOWER case files carry a solved operating point.  The reference
   builds every instance's initial voltage from that table
   (src/powerflow/runpf.h:211-226).
10. Reactive limits (only when ``qlim`` is set), centred on each generator's
    base-case output: Qmax = Qg + P_g * U(0, 0.5), Qmin = Qg + P_g *
    U(-0.3, 0); otherwise +-inf.  The base case is then within limits and
    PV->PQ switching is caused by the contingencies.

Why not a flat start with generation at 105 % of load for the large grids:
at n >= 1354 the single slack must then absorb ~5 % of system load through
2-4 branches (DC angles above 600 degrees), and even with locally balanced
generation the I^2 R losses of the R/X = 0.2-0.5 branches must travel from
the one slack across 20-50 hops; Newton diverges from flat start at every
load level tested down to 5 %.  The three large workloads therefore use
``start="solved"`` and a load scale of 0.2, the largest tested value at
which the 9241-bus grid keeps a solvable base case.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np

PQ, PV, SLACK = 1, 2, 3

INF = float("inf")


@dataclass
class Case:
    """Bus/branch/generator tables in per unit (angles in radians).

    Field meaning follows the reference's ``Bus``/``Branch``/``Generator``
    (src/powersystem/Bus.h:17-30, Branch.h:81-93, Generator.h:140-146),
    except that loads and generation are kept apart (``Pd``/``Qd`` vs
    ``Pg``/``Qg``) so that generator outages can be expressed; the net
    injection the reference stores in ``Bus.P`` is ``sum(Pg on bus) - Pd``
    (loadMatpowerFormat.h:88,126).
    """

    bus_type: np.ndarray          # int32 [n]
    Pd: np.ndarray                # float64 [n]
    Qd: np.ndarray
    Gs: np.ndarray
    Bs: np.ndarray
    Vm: np.ndarray                # initial magnitude (flat start: 1.0)
    Va: np.ndarray                # initial angle, radians
    br_from: np.ndarray           # int32 [m]
    br_to: np.ndarray
    R: np.ndarray
    X: np.ndarray
    B: np.ndarray
    tap: np.ndarray               # 0 => no transformer
    shift: np.ndarray             # degrees
    br_status: np.ndarray         # uint8
    gen_bus: np.ndarray           # int32 [g]
    Pg: np.ndarray
    Qg: np.ndarray
    Qmax: np.ndarray
    Qmin: np.ndarray
    VG: np.ndarray
    gen_status: np.ndarray        # uint8
    baseMVA: float = 100.0
    name: str = ""
    rate: np.ndarray | None = None  # branch MVA rating (pu), optional
    xy: np.ndarray | None = field(default=None, repr=False)

    @property
    def nbus(self) -> int:
        return int(self.bus_type.shape[0])

    @property
    def nbranch(self) -> int:
        return int(self.br_from.shape[0])

    @property
    def ngen(self) -> int:
        return int(self.gen_bus.shape[0])

    def counts(self):
        t = self.bus_type
        return dict(n=self.nbus, m=self.nbranch, g=self.ngen,
                    npv=int((t == PV).sum()), npq=int((t == PQ).sum()),
                    nslack=int((t == SLACK).sum()))


class _DSU:
    def __init__(self, n):
        self.p = np.arange(n)

    def find(self, a):
        p = self.p
        root = a
        while p[root] != root:
            root = p[root]
        while p[a] != root:
            p[a], a = root, p[a]
        return root

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra == rb:
            return False
        if ra < rb:
            self.p[rb] = ra
        else:
            self.p[ra] = rb
        return True


def _topology(rng, n, m, knn=8):
    from scipy.spatial import cKDTree

    pts = rng.uniform(size=(n, 2))
    k = min(knn, n - 1)
    tree = cKDTree(pts)
    _, nb = tree.query(pts, k=k + 1)
    pairs = set()
    for i in range(n):
        for j in nb[i, 1:]:
            a, b = (i, int(j)) if i < j else (int(j), i)
            pairs.add((a, b))
    cand = np.array(sorted(pairs), dtype=np.int64)
    length = np.hypot(*(pts[cand[:, 0]] - pts[cand[:, 1]]).T)
    w = rng.uniform(size=len(cand))
    dsu = _DSU(n)
    in_tree = np.zeros(len(cand), dtype=bool)
    for e in np.argsort(w, kind="stable"):
        if dsu.union(cand[e, 0], cand[e, 1]):
            in_tree[e] = True
    edges = [tuple(cand[e]) for e in np.flatnonzero(in_tree)]
    # join leftover components through their closest pair of points
    roots = sorted({dsu.find(i) for i in range(n)})
    while len(roots) > 1:
        comp = np.array([dsu.find(i) for i in range(n)])
        a_idx = np.flatnonzero(comp == roots[0])
        b_idx = np.flatnonzero(comp != roots[0])
        d, j = cKDTree(pts[b_idx]).query(pts[a_idx], k=1)
        s = int(np.argmin(d))
        a, b = int(a_idx[s]), int(b_idx[j[s]])
        dsu.union(a, b)
        edges.append((min(a, b), max(a, b)))
        roots = sorted({dsu.find(i) for i in range(n)})
    if len(edges) > m:
        raise ValueError(f"spanning tree needs {len(edges)} branches > m={m}")
    used = set(edges)
    for e in np.argsort(length, kind="stable"):
        if len(edges) >= m:
            break
        pe = (int(cand[e, 0]), int(cand[e, 1]))
        if pe not in used:
            used.add(pe)
            edges.append(pe)
    if len(edges) < m:
        raise ValueError("not enough candidate chords; raise knn")
    # order branches by (from, to) like a case file sorted by bus number
    edges.sort()
    e = np.array(edges, dtype=np.int64)
    return pts, e[:, 0], e[:, 1]


def _dispatch(n, f, t, R, X, Pd, slack, pv):
    """Active generation of [slack, *pv] (slack entry 0).

    Every generator bus owns the buses electrically closest to it (Dijkstra
    over branch reactance).  Each PV generator supplies the load of its cell
    plus a share of the estimated losses; the slack supplies its own cell, so
    its net injection stays near zero.  Losses are estimated from the DC flows
    of the cell-balanced dispatch as sum(r * f^2)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import dijkstra
    import scipy.sparse.linalg as sla

    W = sp.coo_matrix((np.concatenate([X, X]),
                       (np.concatenate([f, t]), np.concatenate([t, f]))),
                      shape=(n, n)).tocsr()
    src = np.concatenate([[slack], pv])
    _, _, owner = dijkstra(W, directed=False, indices=src, min_only=True,
                           return_predecessors=True)
    pos = {int(b): k for k, b in enumerate(src)}
    cell = np.zeros(len(src))
    for i in range(n):
        cell[pos[int(owner[i])]] += Pd[i]
    Pg = cell.copy()
    # DC losses of the cell-balanced dispatch
    b = 1.0 / X
    B = sp.coo_matrix((np.concatenate([b, b, -b, -b]),
                       (np.concatenate([f, t, f, t]),
                        np.concatenate([f, t, t, f]))), shape=(n, n)).tocsr()
    P = -Pd.copy()
    np.add.at(P, src, Pg)
    keep = np.flatnonzero(np.arange(n) != slack)
    th = np.zeros(n)
    th[keep] = sla.spsolve(B[keep][:, keep].tocsc(), P[keep])
    loss = float((R * ((th[f] - th[t]) * b) ** 2).sum())
    share = cell[1:] / cell[1:].sum()
    Pg[1:] += loss * share
    Pg[0] = 0.0
    return Pg


def _ybus_np(case: "Case"):
    """Ybus from the branch pi-model (same model as makeYbus.h:63-80)."""
    import scipy.sparse as sp
    n, f, t = case.nbus, case.br_from, case.br_to
    Ys = case.br_status / (case.R + 1j * case.X)
    Bc = case.br_status * 1j * case.B
    tap = np.where(case.tap != 0, case.tap, 1.0) * np.exp(
        1j * np.pi / 180 * case.shift)
    Ytt = Ys + Bc / 2
    Yff = Ytt / (tap * np.conj(tap))
    Yft = -Ys / np.conj(tap)
    Ytf = -Ys / tap
    Y = sp.coo_matrix((np.concatenate([Yff, Yft, Ytf, Ytt]),
                       (np.concatenate([f, f, t, t]),
                        np.concatenate([f, t, f, t]))), shape=(n, n)).tocsr()
    return (Y + sp.diags(case.Gs + 1j * case.Bs)).tocsr()


def solve_base_distributed(case: "Case", tol=1e-12, maxit=40):
    """Base-case AC power flow with a distributed slack.

    Unknowns: angles of non-slack buses, |V| of PQ buses and one scalar
    ``eps``; every PV generator's output becomes ``Pg + eps * Pg/sum(Pg)``
    and the slack bus keeps its own net injection.  Polar Newton with the
    complex-derivative Jacobian, bordered by the ``eps`` column and the slack
    row.  Returns ``(V, eps, iterations)`` or ``(None, eps, maxit)``."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    n = case.nbus
    Y = _ybus_np(case)
    s = int(np.flatnonzero(case.bus_type == SLACK)[0])
    pv = np.flatnonzero(case.bus_type == PV)
    pq = np.flatnonzero(case.bus_type == PQ)
    pvpq = np.concatenate([pv, pq])
    npv, npq = len(pv), len(pq)
    Sinj = -(case.Pd + 1j * case.Qd).astype(complex)
    np.add.at(Sinj, case.gen_bus, case.Pg + 1j * case.Qg)
    w = np.zeros(n)
    np.add.at(w, case.gen_bus[1:], case.Pg[1:])
    w /= w.sum()
    Vm = np.ones(n)
    Vm[case.gen_bus] = case.VG
    Va = np.zeros(n)
    eps = 0.0
    Psl = Sinj[s].real
    for it in range(maxit):
        V = Vm * np.exp(1j * Va)
        I = Y @ V
        S = V * np.conj(I)
        mis = S - Sinj - eps * w
        F = np.concatenate([mis[pvpq].real, mis[pq].imag, [S[s].real - Psl]])
        if not np.all(np.isfinite(F)):
            break
        if np.abs(F).max() < tol:
            return V, eps, it
        dVa = (sp.diags(1j * V) @ np.conj(sp.diags(I) - Y @ sp.diags(V))).tocsr()
        Vn = V / np.abs(V)
        dVm = (sp.diags(V) @ np.conj(Y @ sp.diags(Vn))
               + sp.diags(np.conj(I) * Vn)).tocsr()
        J = sp.bmat([
            [dVa[pvpq][:, pvpq].real, dVm[pvpq][:, pq].real,
             sp.csr_matrix(-w[pvpq][:, None])],
            [dVa[pq][:, pvpq].imag, dVm[pq][:, pq].imag,
             sp.csr_matrix((npq, 1))],
            [dVa[s][:, pvpq].real, dVm[s][:, pq].real, None]]).tocsc()
        dx = sla.spsolve(J, F)
        Va[pvpq] -= dx[:npv + npq]
        Vm[pq] -= dx[npv + npq:npv + 2 * npq]
        eps -= dx[-1]
    return None, eps, maxit


def make_grid(n: int, npv: int, m: int, seed: int, *, tap_frac=0.10,
              tap_lo=0.95, tap_hi=1.05, qlim=False, load_scale=1.0,
              start="flat", name="") -> Case:
    """Build a connected synthetic grid (recipe in the module docstring)."""
    if npv + 1 >= n:
        raise ValueError("need at least one PQ bus")
    rng = np.random.Generator(np.random.PCG64(seed))
    pts, f, t = _topology(rng, n, m)
    slack = int(np.argmin(np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5)))
    others = np.array([i for i in range(n) if i != slack])
    pv = np.sort(rng.choice(others, size=npv, replace=False))
    bus_type = np.full(n, PQ, dtype=np.int32)
    bus_type[pv] = PV
    bus_type[slack] = SLACK

    R = rng.uniform(0.01, 0.08, size=m)
    X = R * rng.uniform(2.0, 5.0, size=m)
    B = rng.uniform(0.0, 0.05, size=m)
    ntap = int(round(tap_frac * m))
    tap = np.zeros(m)
    tap_idx = rng.choice(m, size=ntap, replace=False)
    tap[tap_idx] = rng.uniform(tap_lo, tap_hi, size=ntap)

    Pd = rng.uniform(0.0, 0.5, size=n) * load_scale
    Qd = 0.3 * Pd
    gen_bus = np.concatenate([[slack], pv]).astype(np.int32)
    g = len(gen_bus)
    Pg = _dispatch(n, f, t, R, X, Pd, slack, pv)
    VG = rng.uniform(1.0, 1.05, size=g)
    cmax = rng.uniform(0.0, 0.5, size=g)
    cmin = rng.uniform(-0.3, 0.0, size=g)
    Vm = np.ones(n)
    Vm[gen_bus] = VG                    # bus Vm column carries set-points
    case = Case(
        bus_type=bus_type, Pd=Pd, Qd=Qd, Gs=np.zeros(n), Bs=np.zeros(n),
        Vm=Vm, Va=np.zeros(n),
        br_from=f.astype(np.int32), br_to=t.astype(np.int32), R=R, X=X, B=B,
        tap=tap, shift=np.zeros(m), br_status=np.ones(m, dtype=np.uint8),
        gen_bus=gen_bus, Pg=Pg, Qg=np.zeros(g), Qmax=np.full(g, INF),
        Qmin=np.full(g, -INF), VG=VG,
        gen_status=np.ones(g, dtype=np.uint8), name=name or f"synth{n}",
        xy=pts)
    if start == "solved":
        V, eps, it = solve_base_distributed(case)
        if V is None:
            raise RuntimeError(f"{case.name}: base case has no solution")
        case.Pg = case.Pg.copy()
        case.Pg[1:] += eps * case.Pg[1:] / case.Pg[1:].sum()
        case.Vm = np.abs(V)
        case.Va = np.angle(V)
        S = V * np.conj(_ybus_np(case) @ V)
        case.Qg = S.imag[gen_bus] + Qd[gen_bus]
        case.Pg[0] = S.real[slack] + Pd[slack]
    elif start != "flat":
        raise ValueError(start)
    if qlim:
        case.Qmax = case.Qg + case.Pg * cmax
        case.Qmin = case.Qg + case.Pg * cmin
        case.Qmax[0], case.Qmin[0] = INF, -INF   # the slack never switches
    return case


def bridges(case: Case) -> np.ndarray:
    """Boolean mask of branches whose removal disconnects the grid
    (iterative Tarjan low-link over the multigraph; parallel branches are
    never bridges)."""
    n, m = case.nbus, case.nbranch
    adj = [[] for _ in range(n)]
    for e in range(m):
        if case.br_status[e]:
            a, b = int(case.br_from[e]), int(case.br_to[e])
            adj[a].append((b, e))
            adj[b].append((a, e))
    disc = np.full(n, -1)
    low = np.zeros(n, dtype=np.int64)
    is_bridge = np.zeros(m, dtype=bool)
    timer = 0
    for root in range(n):
        if disc[root] >= 0:
            continue
        stack = [(root, -1, 0)]
        disc[root] = low[root] = timer
        timer += 1
        while stack:
            v, pe, it = stack[-1]
            if it < len(adj[v]):
                stack[-1] = (v, pe, it + 1)
                w, e = adj[v][it]
                if e == pe:
                    continue
                if disc[w] < 0:
                    disc[w] = low[w] = timer
                    timer += 1
                    stack.append((w, e, 0))
                else:
                    low[v] = min(low[v], disc[w])
            else:
                stack.pop()
                if stack:
                    u = stack[-1][0]
                    low[u] = min(low[u], low[v])
                    if low[v] > disc[u]:
                        is_bridge[pe] = True
    return is_bridge


@dataclass
class Contingencies:
    """Per-instance outage masks (1 = in service)."""

    br_on: np.ndarray    # uint8 [T, m]
    gen_on: np.ndarray   # uint8 [T, g]
    label: list

    @property
    def T(self) -> int:
        return int(self.br_on.shape[0])


def n1_contingencies(case: Case, T: int, seed: int, gen_frac: float = 0.0,
                     include_base: bool = False) -> Contingencies:
    """``T`` single outages: non-bridge branch outages, and a fraction
    ``gen_frac`` of PV-generator outages (the slack generator is never
    taken out).  Candidates are drawn without replacement and cycled when
    ``T`` exceeds their number."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m, g = case.nbranch, case.ngen
    ok = np.flatnonzero(~bridges(case) & (case.br_status > 0))
    n_gen = int(round(gen_frac * T))
    n_br = T - n_gen - (1 if include_base else 0)
    br_pick = rng.permutation(ok)
    br_list = [int(br_pick[i % len(br_pick)]) for i in range(n_br)]
    slack_bus = np.flatnonzero(case.bus_type == SLACK)
    gen_ok = np.flatnonzero(~np.isin(case.gen_bus, slack_bus)
                            & (case.gen_status > 0))
    gen_pick = rng.permutation(gen_ok) if len(gen_ok) else gen_ok
    gen_list = [int(gen_pick[i % len(gen_pick)]) for i in range(n_gen)]
    items = [("branch", b) for b in br_list] + [("gen", k) for k in gen_list]
    order = rng.permutation(len(items))
    items = [items[i] for i in order]
    if include_base:
        items = [("base", -1)] + items
    br_on = np.ones((T, m), dtype=np.uint8)
    gen_on = np.ones((T, g), dtype=np.uint8)
    for t, (kind, idx) in enumerate(items):
        if kind == "branch":
            br_on[t, idx] = 0
        elif kind == "gen":
            gen_on[t, idx] = 0
    return Contingencies(br_on=br_on, gen_on=gen_on,
                         label=[f"{k}:{i}" for k, i in items])


# ---------------------------------------------------------------------------
# The five workloads.  Sizes and distributions are fixed here once.
# ---------------------------------------------------------------------------
SEED0 = 20261202

WORKLOADS = {
    # 14 buses: 1 slack, 4 PV, 9 PQ, 20 branches; one instance
    "pf14": dict(n=14, npv=4, m=20, T=1, gen_frac=0.0, qlim=False,
                 tap_frac=0.10, tap_lo=0.95, tap_hi=1.05),
    # 118 buses: 53 PV, 64 PQ, 186 branches; 200 N-1 branch outages
    "pf118x200": dict(n=118, npv=53, m=186, T=200, gen_frac=0.0, qlim=False,
                      tap_frac=0.10, tap_lo=0.95, tap_hi=1.05),
    # 1354 buses: 260 PV, 1991 branches; 200 mixed branch/generator outages
    "pf1354x200": dict(n=1354, npv=260, m=1991, T=200, gen_frac=0.25,
                       qlim=True, tap_frac=0.10, tap_lo=0.95, tap_hi=1.05,
                       load_scale=0.2, start="solved"),
    # 2869 buses: 510 PV, 4582 branches; 100 N-1 branch outages
    "pf2869x100": dict(n=2869, npv=510, m=4582, T=100, gen_frac=0.0,
                       qlim=False, tap_frac=0.10, tap_lo=0.95, tap_hi=1.05,
                       load_scale=0.2, start="solved"),
    # 9241 buses: 1445 PV, 16049 branches, 20 % taps in U(0.9, 1.1);
    # 200 mixed branch/generator outages
    "pf9241x200": dict(n=9241, npv=1445, m=16049, T=200, gen_frac=0.25,
                       qlim=True, tap_frac=0.20, tap_lo=0.90, tap_hi=1.10,
                       load_scale=0.2, start="solved"),
}
WORKLOAD_ORDER = ["pf14", "pf118x200", "pf1354x200", "pf2869x100",
                  "pf9241x200"]


def workload(name: str, T: int | None = None):
    """Return ``(case, contingencies)`` for a named workload."""
    w = WORKLOADS[name]
    idx = WORKLOAD_ORDER.index(name)
    case = make_grid(w["n"], w["npv"], w["m"], SEED0 + idx,
                     tap_frac=w["tap_frac"], tap_lo=w["tap_lo"],
                     tap_hi=w["tap_hi"], qlim=w["qlim"],
                     load_scale=w.get("load_scale", 1.0),
                     start=w.get("start", "flat"), name=name)
    TT = w["T"] if T is None else T
    if TT == 1 and w["T"] == 1:
        cont = Contingencies(br_on=np.ones((1, case.nbranch), np.uint8),
                             gen_on=np.ones((1, case.ngen), np.uint8),
                             label=["base"])
    else:
        cont = n1_contingencies(case, TT, SEED0 + 100 + idx,
                                gen_frac=w["gen_frac"])
    return case, cont
