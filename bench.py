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
    iterations); ``s`` = bytes per factor scalar (4 mixed, 8 FP64)."""
    n, m = case.nbus, case.nbranch
    nb, nY = info["nblocks"], info["nnz_ybus"]
    nJ, nLU = info["nnz_jac_blocks"], info["nnz_lu_blocks"]
    per_inst = {
        "ybus": 16 * nY + 9 * m + 8 * n,
        "mismatch": 16 * nY + 48 * n + 16 * nb + n,
        "jacobian": 16 * nY + 32 * n + n + 4 * s * nJ,
        "lu_factor": 4 * s * nJ + 4 * s * nLU,
        # every pass reads the factor and its right-hand side and writes dx;
        # the refinement passes also read dx (they accumulate)
        "lu_solve": (1 + refine) * (4 * s * nLU + 32 * nb) + refine * 16 * nb,
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
            tune=True, graph=True, fixed=None):
    import torch
    from mapol_amd.engine import Engine
    from mapol_amd.tune import autotune

    e = Engine(case, precision=precision, refine_steps=refine,
               enforce_qlim=qlim, use_graph=graph)
    e.set_instances(br_on=cont.br_on, gen_on=cont.gen_on)
    if fixed:
        e.set_launch(*fixed)
        launch = dict(zip(("lu", "solve", "mismatch"), fixed))
    else:
        launch = autotune(e) if tune else None
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
                launch=launch)


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


def cpu_baseline(case, cont, qlim, threads, limit_s=30.0):
    """Oracle (own sparse LU, FP64) over the contingencies on host threads;
    stops starting new instances after ``limit_s`` seconds of wall time."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle

    orc = oracle.Oracle(case)
    opts = oracle.Oracle.opts(enforce_qlim=qlim)

    def one(t):
        return t, orc.solve(opts, br_on=cont.br_on[t], gen_on=cont.gen_on[t])

    t0 = time.perf_counter()
    out = {}
    with ThreadPoolExecutor(max_workers=threads) as ex:
        futs = []
        for t in range(cont.T):
            futs.append(ex.submit(one, t))
        for f in futs:
            t, r = f.result()
            out[t] = r
    dt = time.perf_counter() - t0
    return out, dt


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
    case, cont = synth.workload(args.workload)
    qlim = bool(synth.WORKLOADS[args.workload]["qlim"])
    T = cont.T
    log(f"[bench] workload {args.workload}: {case.counts()} T={T} qlim={qlim}")

    fixed = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    mix = run_gpu(case, cont, BPF_MIXED, args.refine, qlim, args.warmup,
                  args.steps, tune=not args.no_tune, graph=not args.no_graph,
                  fixed=fixed)
    log(f"[bench] mixed: {mix['seconds'] / args.steps * 1e3:.2f} ms/step")
    fp = None
    if not args.no_fp64:
        fp = run_gpu(case, cont, BPF_FP64, 0, qlim, args.warmup, args.steps,
                     tune=not args.no_tune, graph=not args.no_graph)
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
        "data": "synthetic (seeded generator, mapol_amd/synth.py)",
        "config": {"workload": args.workload, "buses": case.nbus,
                   "branches": case.nbranch, "contingencies": T,
                   "precision": "mixed", "refine_steps": args.refine,
                   "tol": 1e-8, "enforce_qlim": qlim,
                   "launch": ({k: mix["launch"][k] for k in ("lu", "solve", "mismatch")}
                              if mix["launch"] else None)},
        "roofline": roof,
        "kernels": kernels,
        "newton_instance_iterations": its_total,
        "graph_replayed": bool(info["graph_used"]),
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
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        res, dt = cpu_baseline(case, cont, qlim, threads)
        nsolve = len(res)
        its_eq = sum(int(rm["iterations"][t] == r["iterations"]) for t, r in res.items())
        ver_eq = sum(int(rm["verdict"][t] == r["verdict"]) for t, r in res.items())
        dv = max((float(np.abs(rm["V"][t] - r["V"]).max())
                  for t, r in res.items()
                  if r["converged"] and rm["verdict"][t] == CONVERGED),
                 default=None)
        out["cpu_baseline"] = {
            "value": round(nsolve / dt, 3), "unit": "solves/s",
            "cores": threads, "kind": "port",
            "sample": f"all {nsolve} contingencies of {args.workload} "
                      f"(oracle/pf_oracle.c, FP64, own sparse LU), one "
                      f"contingency per thread, {dt:.1f} s wall",
        }
        out["speedup_vs_cpu"] = round(value / (nsolve / dt), 2)
        out["parity_vs_oracle"] = {"instances": nsolve,
                                   "iterations_equal": its_eq,
                                   "verdicts_equal": ver_eq, "max_dV": dv}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
