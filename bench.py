"""Batched contingency power-flow benchmark (one MI355X).

A step is one batched solve of the workload: Ybus assembly with outage
masks, Newton-Raphson for every contingency (with reactive-limit switching
when the workload enables it) and losses, inputs already resident in HBM.
The headline is mixed precision (FP32 LU + FP64 refinement) at 9241 buses x
200 contingencies; the plain FP64 GPU Newton and the CPU oracle run on the
same inputs and are reported beside it.

Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("contingency PF solves/s, 9241-bus×200, mixed vs FP64 NR at "
          "1e-8 pu mismatch")
HBM_PEAK_GBS = 8000.0
KERNELS = ["ybus", "init", "mismatch", "jacobian", "lu_factor", "lu_solve",
           "residual", "update", "loss"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(info, case, T, its_total, s, refine):
    """Algorithmic HBM bytes per kernel over one solve: the compulsory bytes
    of one instance-launch (every array a kernel must read or write, once)
    times the launches.

    ``its_total`` = sum over instances of Newton iterations (instance-
    iterations); ``s`` = bytes per factor scalar (4 mixed, 8 FP64).  With
    the forward substitution in the factorization (``info["fused_forward"]``
    >= 0, an active-instance limit, so per launch either form may run) each
    kernel counts the smaller of its two forms, a lower bound: the fill and
    the factorization without the right-hand-side blocks, the first solve
    without the forward chunks (every strictly lower block counted out of
    its factor read, the few the tail rows still read included) and with
    the right-hand-side blocks in place of the mismatch."""
    n, m = case.nbus, case.nbranch
    nb, nY = info["nblocks"], info["nnz_ybus"]
    nJ, nLU = info["nnz_jac_blocks"], info["nnz_lu_blocks"]
    fused = info.get("fused_forward", -1) >= 0
    # L blocks strictly below the diagonal: (nLU - nb) / 2 of the pattern
    nL = (nLU - nb) // 2
    first = (4 * s * (nLU - nL) + 4 * s * nb + 16 * nb) if fused \
        else (4 * s * nLU + 32 * nb)
    per_inst = {
        "ybus": 16 * nY + 9 * m + 8 * n,
        "mismatch": 16 * nY + 48 * n + 16 * nb + n,
        "jacobian": 16 * nY + 32 * n + n + 4 * s * nJ,
        "lu_factor": 4 * s * nJ + 4 * s * nLU,
        # every pass reads the factor and its right-hand side and writes dx;
        # the refinement passes also read dx (they accumulate)
        "lu_solve": first + refine * (4 * s * nLU + 32 * nb) + refine * 16 * nb,
        "residual": refine * (16 * nY + 33 * n + 48 * nb),
        "update": 48 * nb + n,
        "loss": 16 * n + 41 * m,
    }
    launches_inst = {
        "ybus": T, "loss": T, "mismatch": T + its_total,
        "jacobian": its_total, "lu_factor": its_total, "lu_solve": its_total,
        "residual": its_total,
        "update": its_total,
    }
    return {k: per_inst[k] * launches_inst[k] for k in per_inst}


def run_gpu(case, cont, precision, refine, qlim, warmup, steps, timing=True,
            tune=True, graph=True, fixed=None, cache=None):
    import torch
    from mapol_amd import tune as tuner
    from mapol_amd.engine import Engine

    def make(**tiles):
        e = Engine(case, precision=precision, refine_steps=refine,
                   enforce_qlim=qlim, use_graph=graph, **tiles)
        e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
        return e

    entry = None
    if fixed:
        e = make()
        e.set_launch(*fixed)
        launch = dict(zip(("lu", "solve", "mismatch"), fixed))
    elif tune:
        # tiles, launch shapes and schedule from the autotuner (its JSON cache
        # when it holds this workload; tuning is outside the timed steps)
        entry = tuner.tune(make, tuner.signature(case, cont.T, precision, refine, qlim),
                           cache_path=cache or tuner.DEFAULT_CACHE)
        e = make(**entry["tiles"])
        tuner.apply(e, entry)
        launch = entry["launch"]
        log(f"[bench] tuning ({'cached' if entry['cached'] else 'measured'}): "
            f"tiles {entry['tiles']} schedule {entry['schedule']}")
    else:
        e = make()
        launch = None
    if launch:
        log(f"[bench] launch shapes: lu {launch['lu']} solve {launch['solve']} "
            f"mismatch {launch['mismatch']}")
    for _ in range(warmup):
        e.solve()
    torch.cuda.synchronize()
    # the measured steps: each Newton iteration replayed as one hipGraph
    t0 = time.perf_counter()
    for _ in range(steps):
        e.solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    r = e.results()
    info = e.info()
    message = e.last_error()
    # per-kernel durations: the same steps again with HIP events around
    # direct launches on the engine's stream (events inside replayed graphs
    # do not time reliably)
    kt, dt_ev = {}, None
    if timing:
        e.set_timing(True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            e.solve()
        torch.cuda.synchronize()
        dt_ev = time.perf_counter() - t1
        kt = {k: e.kernel_time(k) for k in KERNELS}
        e.set_timing(False)
    e.close()
    return dict(seconds=dt, seconds_timed=dt_ev, kt=kt, res=r, info=info,
                launch=launch, message=message,
                tuning=(None if entry is None else
                        {k: entry[k] for k in ("tiles", "launch", "schedule", "cached")}))


def copy_calibration(mib):
    """One plain device-to-device copy of ``mib`` MiB, for scaling the HBM
    byte counters of a profiler run (FETCH_SIZE / WRITE_SIZE of this copy
    should read ``mib`` MiB each)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copied {mib} MiB ({n * 4} bytes)")


def cpu_info():
    """host CPU as the box reports it: logical CPUs and the model name"""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(case, cont, qlim, threads, runs=5, warmup=2, single_s=2.5):
    """The oracle (FP64 restatement of the reference's CPU Newton with its own
    sparse LU, symbolic analysis once per case) on the host, as the
    reference's exec modes run it (src/pso/run.h:86-95): all contingencies on
    ``threads`` threads (one contingency per thread), and one core solving a
    sample of the contingencies in sequence.  Each is the median of ``runs``
    timed runs after ``warmup`` untimed ones (run.sh:6 repeats 5 times).
    Returns the all-thread results of every contingency (parity) and the
    figures."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import oracle

    orc = oracle.Oracle(case)
    opts = oracle.Oracle.opts(enforce_qlim=qlim)

    def one(t):
        return t, orc.solve(opts, br_on=cont.br_on[t], gen_on=cont.gen_on[t])

    def all_threads():
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            out = dict(ex.map(one, range(cont.T)))
        return out, time.perf_counter() - t0

    out = None
    times = []
    for k in range(warmup + runs):
        res, dt = all_threads()
        if out is None:
            out = res
        if k >= warmup:
            times.append(dt)
    multi = cont.T / float(np.median(times))
    # one core, contingencies in sequence: a sample stratified by the Newton
    # iteration counts of the all-thread run (the middle contingency of each
    # of ns equal strata in order of iterations), as many as one core solves
    # in about single_s seconds; its rate is scaled to the whole set by
    # iterations, since a few hard contingencies (non-convergent ones run to
    # the iteration limit) carry a share of the time no small sample matches
    its = np.array([out[t]["iterations"] for t in range(cont.T)], float)
    rank = np.argsort(its, kind="stable")

    def stratified(k):
        return rank[((np.arange(k) + 0.5) * cont.T / k).astype(int)]

    t0 = time.perf_counter()
    probe = stratified(4)
    for t in probe:
        one(int(t))
    per = (time.perf_counter() - t0) / len(probe)
    ns = int(max(4, min(cont.T, round(single_s / max(per, 1e-6)))))
    sample = stratified(ns)
    st = []
    for k in range(warmup + runs):
        t0 = time.perf_counter()
        for t in sample:
            one(int(t))
        if k >= warmup:
            st.append(time.perf_counter() - t0)
    sample_its = float(its[sample].sum())
    single_raw = ns / float(np.median(st))
    single = single_raw * (sample_its / ns) / max(float(its.mean()), 1e-9)
    return out, dict(multi=multi, multi_runs=times, single=single, single_n=ns,
                     single_runs=st, single_raw=single_raw,
                     single_its=sample_its / ns, all_its=float(its.mean()))


REAL_GRIDS = {
    # the reference's own 9241-bus grid (src/datasets/case9241pegase.m, its
    # tables kept as a fixture), 200 N-1 branch and generator outages
    "case9241pegase": dict(T=200, seed=9241, gen_frac=0.25, qlim=True),
}


def load_workload(name):
    """(case, contingencies, reactive limits on, data description)"""
    from mapol_amd import synth
    if name not in REAL_GRIDS:
        case, cont = synth.workload(name)
        return (case, cont, bool(synth.WORKLOADS[name]["qlim"]),
                "synthetic (seeded generator, mapol_amd/synth.py)")
    import tempfile
    import numpy as np
    from mapol_amd import matpower
    from mapol_amd.engine import load_matpower
    w = REAL_GRIDS[name]
    z = np.load(os.path.join(ROOT, "tests", "golden", "matpower_cases.npz"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, name + ".m")
        matpower.write(path, float(z[f"{name}/baseMVA"]), z[f"{name}/bus"],
                       z[f"{name}/gen"], z[f"{name}/branch"])
        case = load_matpower(path)[0]
    cont = synth.n1_contingencies(case, w["T"], w["seed"], gen_frac=w["gen_frac"])
    return (case, cont, w["qlim"],
            f"the reference's {name}.m (bpf_load_matpower), {w['T']} N-1 outages")


def rates_from_base(case):
    """branch ratings U(0.8, 1.5) x the larger end flow of the base case, as
    the 2869-bus adjoint workload defines them (solved on the GPU)"""
    import numpy as np
    from mapol_amd.engine import Engine
    e = Engine(case)
    e.set_instances(T=1)
    e.solve()
    Sf, St = e.branch_flows()
    e.close()
    base = np.maximum(np.abs(Sf[0]), np.abs(St[0]))
    rng = np.random.default_rng(3)
    return np.maximum(base, 1e-3) * rng.uniform(0.8, 1.5, case.nbranch)


def adjoint_leg(case, cont, qlim, refine, tuning, reps=3):
    """The adjoint of losses + branch-flow penalty for every contingency after
    the mixed-precision solve (FP64 factors of its own); device time by HIP
    events around its kernels, and the whole call (host copies of the
    gradients included)."""
    import torch
    rate = rates_from_base(case)
    e = _tuned_engine(case, qlim, refine, tuning)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.solve()
    e.adjoint(rate=rate, w=1.0)            # warm-up: allocates its buffers
    e.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        e.adjoint(rate=rate, w=1.0)
    wall = (time.perf_counter() - t0) / reps
    n, ms = e.kernel_time("adjoint")
    e.set_timing(False)
    captures = e.info()["graph_captures"]
    e.solve()                              # the solve after it: same graph
    info = e.info()
    e.close()
    return dict(device_ms=round(ms / max(n, 1), 3), call_ms=round(wall * 1e3, 3),
                contingencies=cont.T,
                solve_after_adjoint_recaptured=bool(info["graph_captures"] != captures))


def fd_leg(name, steps, warmup):
    """Fast decoupled XB and BX (B' and B'' factored once in FP64) and the
    mixed-precision Newton on the 1354-bus workload: solves/s of each."""
    import torch
    from mapol_amd import synth
    from mapol_amd.engine import BPF_FDBX, BPF_FDXB, BPF_MIXED, BPF_NR, CONVERGED, Engine
    case, cont = synth.workload(name)
    qlim = bool(synth.WORKLOADS[name]["qlim"])
    out = {}
    for label, alg, prec in (("nr_mixed", BPF_NR, BPF_MIXED), ("fdxb", BPF_FDXB, 0),
                             ("fdbx", BPF_FDBX, 0)):
        # reactive-limit switching is a Newton feature
        e = Engine(case, alg=alg, precision=prec, refine_steps=1,
                   enforce_qlim=qlim and alg == BPF_NR)
        e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
        for _ in range(warmup):
            e.solve()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r = e.results()
        e.close()
        out[label] = dict(value=round(cont.T * steps / dt, 2),
                          ms_per_step=round(dt / steps * 1e3, 3),
                          converged=int((r["verdict"] == CONVERGED).sum()),
                          max_iterations=int(r["iterations"].max()))
    return out


def _tuned_engine(case, qlim, refine, tuning):
    from mapol_amd import tune as tuner
    from mapol_amd.engine import BPF_MIXED, Engine
    e = Engine(case, precision=BPF_MIXED, refine_steps=refine, enforce_qlim=qlim,
               **(tuning["tiles"] if tuning else {}))
    if tuning:
        tuner.apply(e, tuning)
    return e


def setup_leg(case, cont, qlim, refine, tuning, steps):
    """A new batch every step: bpf_set_instances (upload of the outage masks,
    per-instance arrays built on the device) followed by the solve."""
    import torch
    e = _tuned_engine(case, qlim, refine, tuning)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    e.solve()
    torch.cuda.synchronize()
    t_set = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
        t_set += time.perf_counter() - a
        e.solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    captures = e.info()["graph_captures"]
    e.close()
    mask_bytes = int(cont.br_on.nbytes + cont.gen_on.nbytes)
    return dict(value=round(cont.T * steps / dt, 2), ms_per_step=round(dt / steps * 1e3, 3),
                set_instances_ms=round(t_set / steps * 1e3, 3),
                h2d_bytes_per_batch=mask_bytes, graph_captures=captures)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="pf9241x200")
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--launch", default=None, metavar="LU,SOLVE,MISMATCH",
                    help="fixed workgroup sizes for the mixed-precision run "
                         "(no autotuning), e.g. the ones a tuned run chose")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each kernel directly instead of replaying "
                         "the captured Newton-iteration graph")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain copy of MIB MiB (profiler "
                         "counter calibration) and exit")
    ap.add_argument("--tune-cache", default=None, metavar="PATH",
                    help="autotuner cache (default mapol_amd/tuned.json)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the other configuration legs (set-up + solve, "
                         "adjoint, fast decoupled at 1354 buses)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: contingencies are independent "
              "and would shard across GPUs with no exchange during the Newton "
              "iterations; runs on several GPUs are not supported.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import numpy as np
    import torch
    from mapol_amd import synth
    from mapol_amd.engine import BPF_FP64, BPF_MIXED, CONVERGED

    torch.cuda.set_device(0)
    case, cont, qlim, data_desc = load_workload(args.workload)
    T = cont.T
    log(f"[bench] workload {args.workload}: {case.counts()} T={T} qlim={qlim}")

    fixed = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    mix = run_gpu(case, cont, BPF_MIXED, args.refine, qlim, args.warmup,
                  args.steps, tune=not args.no_tune, graph=not args.no_graph,
                  fixed=fixed, cache=args.tune_cache)
    log(f"[bench] mixed: {mix['seconds'] / args.steps * 1e3:.2f} ms/step")
    fp = None
    if not args.no_fp64:
        fp = run_gpu(case, cont, BPF_FP64, 0, qlim, args.warmup, args.steps,
                     tune=not args.no_tune, graph=not args.no_graph,
                     cache=args.tune_cache)
        log(f"[bench] fp64: {fp['seconds'] / args.steps * 1e3:.2f} ms/step")

    value = T * args.steps / mix["seconds"]
    rm = mix["res"]
    its_total = int(rm["iterations"].sum())
    info = mix["info"]
    # roofline of the dominant kernel (by time) of the mixed run
    bytes_step = algorithmic_bytes(info, case, T, its_total, 4, args.refine)
    kt = mix["kt"]
    kernels = {}
    for k in bytes_step:
        nl, ms = kt.get(k, (0, 0.0))
        if nl == 0 or ms <= 0:
            continue
        per_launch_ms = ms / nl
        launches_per_step = nl / args.steps
        bytes_launch = bytes_step[k] / launches_per_step
        gbs = bytes_launch / (per_launch_ms * 1e-3) / 1e9
        kernels[k] = dict(launches=nl, avg_us=round(per_launch_ms * 1e3, 2),
                          ms_per_step=round(ms / args.steps, 3),
                          bytes_per_launch=int(bytes_launch),
                          GBps=round(gbs, 1))
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.workload}.txt")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])
    roof = dict(bound="hbm", kernel=dom, achieved=kernels[dom]["GBps"],
                peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(kernels[dom]["GBps"] / HBM_PEAK_GBS, 4),
                traffic=traffic)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mix["seconds"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 LU+solve, f64 mismatch/refinement/update",
        "data": data_desc,
        "config": {"workload": args.workload, "buses": case.nbus,
                   "branches": case.nbranch, "contingencies": T,
                   "precision": "mixed", "refine_steps": args.refine,
                   "tol": 1e-8, "enforce_qlim": qlim,
                   "launch": ({k: mix["launch"][k] for k in ("lu", "solve", "mismatch")}
                              if mix["launch"] else None),
                   "tuning": mix["tuning"]},
        "roofline": roof,
        "kernels": kernels,
        "newton_instance_iterations": its_total,
        "graph_replayed": bool(info["graph_used"]),
        "graph_fallback": bool(info["graph_fallback"]),
        "engine_message": mix["message"],
        "ms_per_step_kernel_timing": (round(mix["seconds_timed"] / args.steps * 1e3, 3)
                                      if mix["seconds_timed"] else None),
        "converged": int((rm["verdict"] == CONVERGED).sum()),
    }
    if fp is not None:
        rf = fp["res"]
        fp_value = T * args.steps / fp["seconds"]
        ok = (rm["verdict"] == CONVERGED) & (rf["verdict"] == CONVERGED)
        out["fp64_gpu"] = {
            "value": round(fp_value, 2),
            "ms_per_step": round(fp["seconds"] / args.steps * 1e3, 3),
            "lu_factor_ms_per_step": round(fp["kt"]["lu_factor"][1] / args.steps, 3),
            "lu_solve_ms_per_step": round(fp["kt"]["lu_solve"][1] / args.steps, 3),
        }
        out["speedup_vs_fp64"] = round(value / fp_value, 3)
        out["mixed_vs_fp64"] = {
            "iterations_equal": int((rm["iterations"] == rf["iterations"]).sum()),
            "verdicts_equal": int((rm["verdict"] == rf["verdict"]).sum()),
            "max_dV": float(np.abs(rm["V"][ok] - rf["V"][ok]).max()) if ok.any() else None,
        }
    if not args.no_legs:
        log("[bench] leg: set-up + solve per batch")
        out["setup_plus_solve"] = setup_leg(case, cont, qlim, args.refine,
                                            mix["tuning"], args.steps)
        if args.workload == "pf9241x200":
            log("[bench] leg: adjoint for every contingency")
            out["adjoint"] = adjoint_leg(case, cont, qlim, args.refine, mix["tuning"])
            log("[bench] leg: pf1354x200 Newton (mixed), FDXB, FDBX")
            out["pf1354x200"] = fd_leg("pf1354x200", args.steps, args.warmup)
    if not args.no_cpu:
        # the box's CPU share for one GPU is 16 threads, whatever the host has
        nproc, model = cpu_info()
        threads = args.cpu_threads or min(16, nproc)
        log(f"[bench] cpu baseline: {threads} threads, then 1 core")
        res, cb = cpu_baseline(case, cont, qlim, threads)
        nsolve = len(res)
        its_eq = sum(int(rm["iterations"][t] == r["iterations"]) for t, r in res.items())
        ver_eq = sum(int(rm["verdict"][t] == r["verdict"]) for t, r in res.items())
        dv = max((float(np.abs(rm["V"][t] - r["V"]).max())
                  for t, r in res.items()
                  if r["converged"] and rm["verdict"][t] == CONVERGED),
                 default=None)
        out["cpu_baseline"] = {
            "value": round(cb["multi"], 3), "unit": "solves/s",
            "cores": threads, "kind": "port",
            "sample": f"all {nsolve} contingencies of {args.workload} "
                      f"(oracle/pf_oracle.c, FP64, own sparse LU), one "
                      f"contingency per thread, median of "
                      f"{len(cb['multi_runs'])} runs after 2 warm-up runs "
                      f"({', '.join(f'{x:.2f}' for x in cb['multi_runs'])} s)",
            "single_core": {
                "value": round(cb["single"], 3), "cores": 1,
                "sample": f"{cb['single_n']} contingencies in sequence, one "
                          f"from each of {cb['single_n']} equal strata of the "
                          f"{nsolve} ranked by Newton iterations "
                          f"({cb['single_its']:.2f} iterations per contingency "
                          f"against {cb['all_its']:.2f} over all), median of "
                          f"{len(cb['single_runs'])} runs after 2 warm-up runs "
                          f"({', '.join(f'{x:.2f}' for x in cb['single_runs'])} s); "
                          f"value scaled to the whole set by iterations",
                "sample_rate": round(cb["single_raw"], 3)},
            "nproc": nproc, "cpu_model": model,
        }
        out["speedup_vs_cpu"] = round(value / cb["multi"], 2)
        out["speedup_vs_cpu_single_core"] = round(value / cb["single"], 2)
        out["parity_vs_oracle"] = {"instances": nsolve,
                                   "iterations_equal": its_eq,
                                   "verdicts_equal": ver_eq, "max_dV": dv}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
