#!/usr/bin/env python3
"""Benchmark: RHS evaluations per second of the device-resident earthquake-cycle
loop on the 8-fault, 8,192-element dense network, mixed precision (FP32
stiffness storage, FP64 accumulation / state / solves) against the FP64 plain
form, with the event-time error between the two.

A step is one accepted time step of main() (scripts/Functions_RSFDFN3DMain_D.jl:
146-322): one coupling product over all elements plus one or more local-solve
passes.  `value` = RHS evaluations / s of the mixed-precision run over the
timed steps, inputs resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quake-dfn_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "RHS evals/sec + sim-years/hour, mixed vs FP64 on GPU at stated event-time error"
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(ROOT, "profiles", "coupling_fault_network_8192.md")


def measured_traffic():
    """HBM bytes per FP32 coupling launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated); None if absent."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch_fp32:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(qd, cfg):
    from qdfn import configs
    seg = cfg["seg"]
    ops = {}
    for name, prec in (("fp64", qd.PREC_FP64), ("fp32", qd.PREC_FP32)):
        t0 = time.perf_counter()
        ops[name] = qd.DenseOperator.from_rect_segments(seg, cfg["mu"], cfg["nu"], prec)
        log(f"[bench] assembled + built {name} operator n={seg.shape[0]} in {time.perf_counter() - t0:.2f}s")
    ks, _ = ops["fp64"].kself()
    ksum, knsum = ops["fp64"].matvec(np.ones(seg.shape[0]), loading_form=False)
    prob = configs.problem_from_segments(cfg, ks, ksum, knsum)
    return ops, prob


def run_timed(qd, op, prob, steps, warmup, graph_attempts, kernel_steps):
    """Warm up, then the timed region: `steps` accepted steps replayed from the
    captured attempt graph (wall clock around a synchronous call), followed by
    `kernel_steps` steps launched directly with HIP events around every
    coupling kernel on the run's stream (per-kernel durations)."""
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0,
                        trace_capacity=steps + warmup + kernel_steps)
    if warmup:
        sim.run(warmup, graph_attempts=graph_attempts)
    s0 = sim.stats()
    t0 = time.perf_counter()
    sim.run(steps, graph_attempts=graph_attempts)
    el = time.perf_counter() - t0
    s1 = sim.stats()
    t1 = time.perf_counter()
    sim.run(kernel_steps, timed=True)
    el_k = time.perf_counter() - t1
    s2 = sim.stats()
    tr = sim.trace()
    sim.close()
    rhs = s1["rhs_evals"] - s0["rhs_evals"]
    return dict(elapsed=el, rhs=rhs, retries=s1["retry_passes"] - s0["retry_passes"],
                sim_seconds=s1["t"] - s0["t"], coupling_ms=s2["coupling_ms_avg"],
                coupling_timed=s2["coupling_timed"], trace=tr, status=s2["status"],
                direct_evals_per_s=(s2["rhs_evals"] - s1["rhs_evals"]) / el_k)


def cpu_baseline(ops, prob, n_steps):
    """The oracle's FP64 main() loop (C restatement, OpenMP rows) on the same
    stiffness, for a bounded number of steps."""
    from oracle import oracle as orc
    n = ops["fp64"].n
    # the matrices come from the GPU assembly (entries agree with the oracle's
    # own assembly to 1e-10; assembling 8192^2 pairs on the CPU is not the
    # loop being timed)
    from qdfn import configs  # noqa: F401
    import qdfn
    seg = prob["_seg"]
    KS, KN = qdfn.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, prob["mu"], prob["_nu"])
    p = dict(prob)
    t0 = time.perf_counter()
    st, res = orc.sim_dense(p, KS, KN, total_step=n_steps, record_step=0, trace=False)
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=res["rhs_evals"] / el, elapsed=el, steps=n_steps, status=st, cores=threads)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--graph-attempts", type=int, default=32)
    ap.add_argument("--kernel-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU layout is "
              "design-only: see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import qdfn
    from qdfn import configs
    info = qdfn.device_info()
    log(f"[bench] device {info}")
    cfg = configs.fault_network()
    ops, prob = build_problem(qdfn, cfg)
    prob["_seg"], prob["_nu"] = cfg["seg"], cfg["nu"]
    n = cfg["seg"].shape[0]

    res = {}
    for name in ("fp64", "fp32"):
        log(f"[bench] {name}: warmup {args.warmup} steps, timed {args.steps} steps")
        r = run_timed(qdfn, ops[name], prob, args.steps, args.warmup, args.graph_attempts, args.kernel_steps)
        if r["status"] != 0:
            raise SystemExit(f"{name} run failed with status {r['status']}")
        r["evals_per_s"] = r["rhs"] / r["elapsed"]
        r["sim_years_per_hour"] = r["sim_seconds"] / configs.SEC_PER_YEAR / (r["elapsed"] / 3600.0)
        log(f"[bench] {name}: {r['rhs']} RHS evals in {r['elapsed']:.3f}s = {r['evals_per_s']:.1f}/s, "
            f"coupling {r['coupling_ms']*1e3:.1f} us avg over {r['coupling_timed']}, "
            f"sim {r['sim_seconds']/configs.SEC_PER_YEAR:.4f} yr, retries {r['retries']}, "
            f"direct-launch {r['direct_evals_per_s']:.1f}/s")
        res[name] = r

    # event-time error between mixed and FP64 over the whole run
    def onset(r):
        t = r["trace"]
        return configs.first_event_time(t["told"], t["dt"], t["vmax"])
    t64, t32 = onset(res["fp64"]), onset(res["fp32"])
    ev_err = abs(t32 - t64) / t64 if (t64 and t32) else None
    # divergence of the step sequences (simulated time reached) as a second measure
    tt64 = res["fp64"]["trace"]["told"]
    tt32 = res["fp32"]["trace"]["told"]
    m = min(tt64.size, tt32.size)
    time_dev = float(np.max(np.abs(tt32[:m] - tt64[:m]) / np.maximum(tt64[:m], 1.0))) if m else None

    mixed = res["fp32"]
    bytes_launch = ops["fp32"].bytes_per_product()
    achieved = bytes_launch / (mixed["coupling_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle FP64 loop, {args.cpu_steps} steps")
        c = cpu_baseline(ops, prob, args.cpu_steps)
        cpu = {"value": round(c["value"], 3), "unit": "RHS evals/s", "cores": c["cores"], "kind": "port",
               "sample": f"{c['steps']} accepted steps of the same 8,192-element network, FP64 C restatement "
                         f"of main() with OpenMP over rows ({c['elapsed']:.1f}s)"}
        log(f"[bench] CPU baseline {c['value']:.2f} RHS evals/s on {c['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(mixed["evals_per_s"], 2),
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["elapsed"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 8-fault network, back-slip loading)",
        "config": {"workload": "fault_network_8192", "n_elements": n, "storage": "fp32 stiffness, fp64 state",
                   "steps_timed": args.steps},
        "fp64": {"rhs_evals_per_s": round(res["fp64"]["evals_per_s"], 2),
                 "sim_years_per_hour": round(res["fp64"]["sim_years_per_hour"], 4),
                 "coupling_us": round(res["fp64"]["coupling_ms"] * 1e3, 2)},
        "mixed": {"rhs_evals_per_s": round(mixed["evals_per_s"], 2),
                  "sim_years_per_hour": round(mixed["sim_years_per_hour"], 4),
                  "coupling_us": round(mixed["coupling_ms"] * 1e3, 2)},
        "speedup_mixed_vs_fp64": round(mixed["evals_per_s"] / res["fp64"]["evals_per_s"], 3),
        "event_time_rel_err": ev_err,
        "sim_time_rel_dev": time_dev,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
