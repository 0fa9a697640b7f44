"""Autotuner of the engine's launch shapes.

The per-instance kernels (LU factorization, triangular solves with
refinement, mismatch + verdicts) run one workgroup per contingency; their
best workgroup size depends on the elimination-tree shape (level widths) and
on the instance count.  The tuner times each kernel family on the live
engine with HIP events (``bpf_get_kernel_time``) for every candidate size,
keeps the fastest per family, applies it with ``bpf_set_launch`` and can
persist the choice in a JSON cache keyed by (buses, L+U blocks, instances,
precision).
"""
from __future__ import annotations

import json
import os

CANDIDATES = (256, 512, 1024)
FAMILIES = {"lu": "lu_factor", "solve": "lu_solve", "mismatch": "mismatch"}


def key_of(engine) -> str:
    i = engine.info()
    return (f"n{i['nbus']}-lu{i['nnz_lu_blocks']}-T{i['instances']}-"
            f"p{engine.options.precision}-r{engine.options.refine_steps}")


def _avg_us(engine, kernel):
    n, ms = engine.kernel_time(kernel)
    return 1e3 * ms / n if n else float("inf")


def autotune(engine, candidates=CANDIDATES, reps=2, cache_path=None,
             verbose=False):
    """Pick workgroup sizes for the per-instance kernels; returns a dict with
    the chosen sizes and the measured average kernel times (us)."""
    key = key_of(engine)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if key in cache:
            c = cache[key]
            engine.set_launch(c["lu"], c["solve"], c["mismatch"])
            return dict(c, cached=True)
    engine.solve()                                   # warm-up
    best = {}
    timings = {}
    current = {"lu": 0, "solve": 0, "mismatch": 0}
    for fam, kernel in FAMILIES.items():
        res = {}
        for c in candidates:
            args = dict(current)
            args[fam] = c
            engine.set_launch(args["lu"], args["solve"], args["mismatch"])
            engine.solve()                           # graph re-capture
            engine.set_timing(True)
            for _ in range(reps):
                engine.solve()
            res[c] = _avg_us(engine, kernel)
            engine.set_timing(False)
            if verbose:
                print(f"[tune] {kernel} {c:5d} threads: {res[c]:9.1f} us")
        best[fam] = min(res, key=res.get)
        current[fam] = best[fam]
        timings[kernel] = {str(c): round(v, 2) for c, v in res.items()}
    engine.set_launch(best["lu"], best["solve"], best["mismatch"])
    out = dict(best, timings_us=timings, key=key, cached=False)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                cache = json.load(fh)
        cache[key] = {k: out[k] for k in ("lu", "solve", "mismatch",
                                          "timings_us")}
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
    return out
