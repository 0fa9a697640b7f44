"""The autotuner's search and cache logic, on the CPU, with scripted engines
whose solve and kernel times are known: it must pick the cheapest candidate
of every knob, persist the entry, reuse it without measuring, keep other
keys of the cache, skip candidates the analysis refuses, and apply launch
shapes and schedule to an engine."""
import json

import numpy as np
import pytest

from mapol_amd import synth, tune


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeEngine:
    """solve() advances the clock by a cost that depends on the tiles,
    launch shapes and schedule; kernel_time reports per-family costs"""

    def __init__(self, clock, tiles, log):
        self.clock, self.tiles, self.log = clock, tiles, log
        self.launch = {"lu": 1024, "solve": 1024, "mismatch": 256}
        self.sched = {"newton_per_graph": 4, "tail_inverse": 1, "parts": 1,
                      "fused_forward": 32}
        self.timing = False
        self.kt = {}
        self.closed = False

    def cost_ms(self):
        t, s = self.tiles, self.sched
        c = 10.0
        c += 0.4 if t["ordering"] != 1 else 0           # best 1
        c += abs(t["tail_limit"] - 48) * 0.01          # best 48
        c += 0.5 if t["chunk_products"] != 3072 else 0   # best 3072
        c += 0.3 if t["slot_chunk"] != 2048 else 0       # best 2048
        c += {2: 0.2, 4: 0.1, 8: 0.0}[s["newton_per_graph"]]  # best 8
        c += 0.05 if s["tail_inverse"] == 1 else 0       # best 2 (VALU)
        c += {1: 0.3, 2: 0.0, 4: 0.1}[s["parts"]]           # best 2
        c += 0.0 if s["fused_forward"] == 8 else 0.05      # best 8
        return c

    def family_us(self):
        best = {"lu_factor": 512, "lu_solve": 1024, "mismatch": 256}
        fam = {"lu_factor": "lu", "lu_solve": "solve", "mismatch": "mismatch"}
        return {k: 100.0 + abs(self.launch[fam[k]] - best[k]) / 8 for k in best}

    def solve(self):
        self.clock.t += self.cost_ms() * 1e-3
        self.log.append(("solve", dict(self.tiles)))
        if self.timing:
            for k, us in self.family_us().items():
                n, ms = self.kt.get(k, (0, 0.0))
                self.kt[k] = (n + 1, ms + us * 1e-3)

    def set_launch(self, lu=0, solve=0, mismatch=0):
        for k, v in (("lu", lu), ("solve", solve), ("mismatch", mismatch)):
            if v:
                self.launch[k] = v

    def set_schedule(self, newton_per_graph=0, tail_inverse=0, sweep_form=0,
                     parts=0, fused_forward=None):
        if fused_forward is not None:
            self.sched["fused_forward"] = fused_forward
        if newton_per_graph:
            self.sched["newton_per_graph"] = newton_per_graph
        if tail_inverse:
            self.sched["tail_inverse"] = tail_inverse
        if parts:
            self.sched["parts"] = parts

    def set_timing(self, on):
        self.timing = bool(on)
        self.kt = {}

    def kernel_time(self, name):
        return self.kt.get(name, (0, 0.0))

    def close(self):
        self.closed = True


@pytest.fixture
def clock(monkeypatch):
    c = Clock()
    monkeypatch.setattr(tune.time, "perf_counter", c)
    return c


def _factory(clock, made, log, refuse=None):
    def make(**tiles):
        if refuse and refuse(tiles):
            raise RuntimeError("a pivot's factorization products exceed one LDS chunk")
        e = FakeEngine(clock, tiles, log)
        made.append(e)
        return e
    return make


def test_signature_separates_workloads():
    case, _ = synth.workload("pf118x200", T=2)
    a = tune.signature(case, 200, 1, 1, True)
    assert a == tune.signature(case, 200, 1, 1, True)
    assert a != tune.signature(case, 100, 1, 1, True)
    assert a != tune.signature(case, 200, 0, 0, True)
    other = synth.Case(**{k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                          for k, v in case.__dict__.items()})
    other.br_to = other.br_to.copy()
    other.br_to[[0, 1]] = other.br_to[[1, 0]]
    assert a != tune.signature(other, 200, 1, 1, True)


def test_tune_picks_cheapest_and_caches(tmp_path, clock):
    path = str(tmp_path / "tuned.json")
    with open(path, "w") as fh:
        json.dump({"other-key": {"tiles": {}}}, fh)
    made, log = [], []
    got = tune.tune(_factory(clock, made, log), "k1", cache_path=path, reps=3)
    assert got["cached"] is False
    assert got["tiles"] == {"ordering": 1, "tail_limit": 48, "chunk_products": 3072,
                            "slot_chunk": 2048}
    assert got["launch"] == {"lu": 512, "solve": 1024, "mismatch": 256}
    assert got["schedule"] == {"newton_per_graph": 8, "tail_inverse": 2, "parts": 2,
                               "fused_forward": 8}
    assert all(e.closed for e in made)
    assert got["ms_per_solve"] == pytest.approx(10.0, abs=1e-9)
    cache = json.load(open(path))
    assert set(cache) == {"other-key", "k1"}
    assert cache["k1"]["tiles"] == got["tiles"]
    # a second call measures nothing and creates no engine
    made2 = []
    again = tune.tune(_factory(clock, made2, []), "k1", cache_path=path)
    assert again["cached"] is True and made2 == []
    assert again["launch"] == got["launch"] and again["schedule"] == got["schedule"]


def test_refused_tiles_are_skipped(tmp_path, clock):
    made, log = [], []
    refuse = lambda t: t["chunk_products"] == 3072   # noqa: E731
    got = tune.tune(_factory(clock, made, log, refuse), "k2",
                    cache_path=str(tmp_path / "c.json"), reps=1)
    assert got["tiles"]["chunk_products"] in (4096, 2048)
    errs = [t for t in got["trials"] if "error" in t]
    assert errs and all(t["tiles"]["chunk_products"] == 3072 for t in errs)


def test_apply_sets_launch_and_schedule(clock):
    e = FakeEngine(clock, {}, [])
    tune.apply(e, {"launch": {"lu": 256, "solve": 512, "mismatch": 1024},
                   "schedule": {"newton_per_graph": 2, "tail_inverse": 2, "parts": 4,
                                "fused_forward": -1}})
    assert e.launch == {"lu": 256, "solve": 512, "mismatch": 1024}
    assert e.sched == {"newton_per_graph": 2, "tail_inverse": 2, "parts": 4,
                       "fused_forward": -1}
