"""Autotuner of the engine's tile sizes, schedule and launch shapes.

What it picks, per case topology, instance count and precision:

* the bus ordering and the tiles of the symbolic schedules
  (``bpf_options``; each candidate is a new engine, since they shape the
  analysis): minimum degree or its least-height equivalent, pivots of the
  dense tail
  (``tail_limit``), block products per factorization chunk
  (``chunk_products``) and lane slots per slot-form sweep chunk
  (``slot_chunk``);
* launch shapes (``bpf_set_launch``): workgroup sizes of the factorization,
  the sweeps and the mismatch, timed per kernel family with HIP events;
* the run-time schedule (``bpf_set_schedule``): Newton iterations per
  captured graph, the dense-tail inverse on matrix cores or vector ALUs and
  the number of concurrent parts the batch runs as.

Tiles and schedule are judged by the wall time of whole batched solves (they
move time between kernels), the launch shapes by their own kernel's time.
The search is coordinate descent in the order above, starting from the
engine's defaults.  Results persist in a JSON cache keyed by
:func:`signature`; a cached entry is applied without measuring.
"""
from __future__ import annotations

import json
import os
import time
import zlib

TILE_CANDIDATES = {
    "ordering": (0, 1),            # minimum degree; its least-height reorder
    "tail_limit": (64, 56, 48, 40),
    "chunk_products": (4096, 3072, 2048),
    "slot_chunk": (2048, 1024, 512),
}
SCHEDULE_CANDIDATES = {
    "newton_per_graph": (4, 2, 8),
    "tail_inverse": (1, 2),        # matrix cores, vector ALUs
    "parts": (1, 2, 4),            # concurrent parts of the batch
    # the forward substitution in the factorization while at most this
    # many instances of a part are active (-1 never, 1 << 20 always)
    "fused_forward": (32, -1, 8, 64, 1 << 20),
}
CANDIDATES = (256, 512, 1024)      # workgroup sizes
FAMILIES = {"lu": "lu_factor", "solve": "lu_solve", "mismatch": "mismatch"}
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                             "tuned.json")


def signature(case, T, precision, refine, qlim) -> str:
    """cache key: sizes, a CRC of the topology and bus types, the instance
    count and the numerical settings"""
    import numpy as np
    h = zlib.crc32(np.ascontiguousarray(case.br_from, np.int32).tobytes())
    h = zlib.crc32(np.ascontiguousarray(case.br_to, np.int32).tobytes(), h)
    h = zlib.crc32(np.ascontiguousarray(case.bus_type, np.int32).tobytes(), h)
    return (f"n{case.nbus}-m{case.nbranch}-{h:08x}-T{int(T)}-p{int(precision)}-"
            f"r{int(refine)}-q{int(bool(qlim))}")


def load_cache(path):
    if path and os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def save_cache(path, key, entry):
    cache = load_cache(path)
    cache[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _solve_ms(engine, reps):
    """median wall time of ``reps`` solves (bpf_solve returns after the
    device has finished) after one untimed solve"""
    engine.solve()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        engine.solve()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return 1e3 * ts[len(ts) // 2]


def _avg_us(engine, kernel):
    n, ms = engine.kernel_time(kernel)
    return 1e3 * ms / n if n else float("inf")


def autotune(engine, candidates=CANDIDATES, reps=2, verbose=False):
    """Launch shapes only, on a live engine: for each kernel family the
    workgroup size with the least average kernel time; applied and returned
    with the timings (us)."""
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
    return dict(best, timings_us=timings)


def apply(engine, entry):
    """Apply the launch shapes and schedule of a tuning entry to an engine
    (the tiles are engine options: pass ``entry["tiles"]`` when creating
    it)."""
    la, sc = entry["launch"], entry["schedule"]
    engine.set_launch(la["lu"], la["solve"], la["mismatch"])
    engine.set_schedule(sc["newton_per_graph"], sc["tail_inverse"], 0,
                        sc.get("parts", 1), sc.get("fused_forward"))


def tune(make_engine, key, cache_path=DEFAULT_CACHE, reps=3, verbose=False,
         tile_candidates=TILE_CANDIDATES, schedule_candidates=SCHEDULE_CANDIDATES,
         launch_candidates=CANDIDATES):
    """Tune (or look up) everything for one workload.

    ``make_engine(**tiles)`` returns an engine with its instances set, built
    with the given tile options; ``key`` is the cache key
    (:func:`signature`).  Returns the entry: ``tiles``, ``launch``,
    ``schedule``, ``ms_per_solve`` (best whole-solve time seen), ``trials``
    (every measurement) and ``cached``."""
    cache = load_cache(cache_path)
    if key in cache:
        return dict(cache[key], cached=True)
    trials = []

    def log(msg):
        if verbose:
            print(f"[tune] {msg}", flush=True)

    # 1. tiles: one engine per candidate, whole solves
    tiles = {k: v[0] for k, v in tile_candidates.items()}
    best_ms = None
    for name, cands in tile_candidates.items():
        res = {}
        for c in cands:
            t = dict(tiles, **{name: c})
            try:
                e = make_engine(**t)
            except Exception as ex:            # a schedule the analysis refuses
                trials.append(dict(stage="tiles", tiles=t, error=str(ex)))
                log(f"tiles {t}: refused ({ex})")
                continue
            try:
                res[c] = _solve_ms(e, reps)
            finally:
                e.close()
            trials.append(dict(stage="tiles", tiles=t, ms=round(res[c], 3)))
            log(f"tiles {t}: {res[c]:.3f} ms")
        if res:
            tiles[name] = min(res, key=res.get)
            best_ms = res[tiles[name]]
    # 2. launch shapes and 3. schedule on the engine with the chosen tiles
    e = make_engine(**tiles)
    try:
        got = autotune(e, candidates=launch_candidates, reps=max(1, reps - 1),
                       verbose=verbose)
        launch = {k: got[k] for k in FAMILIES}
        trials.append(dict(stage="launch", timings_us=got["timings_us"]))
        schedule = {k: v[0] for k, v in schedule_candidates.items()}
        for name, cands in schedule_candidates.items():
            res = {}
            for c in cands:
                s = dict(schedule, **{name: c})
                e.set_schedule(s["newton_per_graph"], s["tail_inverse"], 0,
                               s.get("parts", 1), s.get("fused_forward"))
                res[c] = _solve_ms(e, reps)
                trials.append(dict(stage="schedule", schedule=s, ms=round(res[c], 3)))
                log(f"schedule {s}: {res[c]:.3f} ms")
            schedule[name] = min(res, key=res.get)
            best_ms = res[schedule[name]]
    finally:
        e.close()
    entry = dict(tiles=tiles, launch=launch, schedule=schedule,
                 ms_per_solve=round(best_ms, 3) if best_ms is not None else None,
                 trials=trials)
    if cache_path:
        save_cache(cache_path, key, entry)
    return dict(entry, cached=False)
