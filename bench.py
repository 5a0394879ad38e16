#!/usr/bin/env python3
"""Benchmark: RHS evaluations per second of the device-resident earthquake-cycle
loop, mixed precision against the FP64 plain form, at a stated event-time error.

Headline workload: the rough strike-slip fault of 20,000 triangles (100 x 100
cells of 50 m, sinusoidal roughness), H-matrix operator (cluster leaves of 64,
ACA tolerance 1e-6) stored in FP32 with FP64 accumulation and FP64 state,
against the dense FP64 operator ("plain" form, main() of
scripts/Functions_RSFDFN3DMain_D.jl).  Both loops are main_H/main semantics on
the device; each run is loaded by back-slip tables built from its own operator
(UniformBackslip.jl).  `bf16_far_h` reports the same H build with its low-rank
factors stored in BF16 (dense blocks FP32), with its own operator error.
Secondary workload: the 8-fault, 8,192-element dense network, FP32 storage
against FP64.

A step is one accepted time step (one coupling product plus one or more
local-solve passes).  `value` = RHS evaluations / s of the mixed-precision H run
over the timed steps (replayed from a captured HIP graph), inputs resident in
HBM.  Prints one JSON line on stdout; progress goes to stderr.
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
PROFILE = os.path.join(ROOT, "profiles", "hmat_phase2_rough_fault_20000.md")


def measured_traffic():
    """HBM bytes per launch of the phase-2 H kernel from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated); None if absent."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_timed(qd, op, prob, steps, warmup, graph_attempts, kernel_steps):
    """Warm up, then the timed region: `steps` accepted steps replayed from the
    captured attempt graph (wall clock around a synchronous call), followed by
    `kernel_steps` steps launched directly with HIP events around the coupling
    kernel (the phase-2 kernel for an H operator)."""
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0,
                        trace_capacity=steps + warmup + kernel_steps)
    if warmup:
        sim.run(warmup, graph_attempts=graph_attempts)
    s0 = sim.stats()
    t0 = time.perf_counter()
    sim.run(steps, graph_attempts=graph_attempts)
    el = time.perf_counter() - t0
    s1 = sim.stats()
    if kernel_steps:
        sim.run(kernel_steps, timed=True)
    s2 = sim.stats()
    tr = sim.trace()
    sim.close()
    rhs = s1["rhs_evals"] - s0["rhs_evals"]
    return dict(elapsed=el, rhs=rhs, retries=s1["retry_passes"] - s0["retry_passes"],
                sim_seconds=s1["t"] - s0["t"], coupling_ms=s2["coupling_ms_avg"],
                coupling_timed=s2["coupling_timed"], trace=tr, status=s2["status"],
                evals_per_s=rhs / el)


def summarize(name, r, sec_per_year):
    r["sim_years_per_hour"] = r["sim_seconds"] / sec_per_year / (r["elapsed"] / 3600.0)
    log(f"[bench] {name}: {r['rhs']} RHS evals in {r['elapsed']:.3f}s = {r['evals_per_s']:.1f}/s, "
        f"coupling {r['coupling_ms']*1e3:.1f} us over {r['coupling_timed']}, "
        f"{r['sim_years_per_hour']:.3f} sim-yr/h, retries {r['retries']}")


def divergence(tr, ref, rel=1e-3):
    """First step at which the simulated time of a run departs from the FP64
    plain run by more than `rel` (relative), with the plain run's time there."""
    a, b = np.asarray(tr["told"]), np.asarray(ref["told"])
    m = min(a.size, b.size)
    d = np.abs(a[:m] - b[:m]) > rel * np.maximum(np.abs(b[:m]), 1.0)
    if not d.any():
        return None
    k = int(np.argmax(d))
    return {"step": k, "sim_time_s": float(b[k])}


def _build_summary(h):
    b = h.build_info()
    keys = ("entries_evaluated", "parts_sampled", "blocks_fallback", "steps", "ms_build", "ms_evaluation",
            "ms_sampling", "ms_full_pivot", "ms_recompression", "ms_layout")
    out = {k: (round(b[k], 1) if isinstance(b[k], float) else b[k]) for k in keys}
    out["stored_entries"] = h.info()["stored_entries"]
    return out


def rough_fault(qd, args):
    from qdfn import configs
    m = configs.rough_fault()
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    m = dict(m, P1=P1, P2=P2, P3=P3)
    ctr = (P1 + P2 + P3) / 3
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=args.crit,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [a[p] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    t0 = time.perf_counter()
    d64 = qd.DenseOperator.from_triangles(*geo, m["mu"], lam, qd.PREC_FP64)
    log(f"[bench] rough fault n={n}: dense FP64 operator in {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    h32 = qd.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=args.aca_tol,
                               prec=qd.PREC_FP32)
    info = h32.info()
    rk = h32.ranks()
    log(f"[bench] H FP32 operator in {time.perf_counter() - t0:.2f}s: {info['blocks']} blocks, "
        f"{int(st['admissible'].sum())} admissible, {info['stored_entries'] / 1e6:.1f} M entries, "
        f"{info['bytes_per_product'] / 1e6:.1f} MB per product")
    builds = {"full": _build_summary(h32)}
    if not args.no_sampled_build:
        # the same operator from sampled rows and columns (no full assembly)
        hs = qd.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=args.aca_tol,
                                  prec=qd.PREC_FP32, sampled=True)
        builds["sampled"] = _build_summary(hs)
        xr = np.random.default_rng(6).standard_normal(n)
        a, b, c = h32.matvec(xr, loading_form=False), hs.matvec(xr, loading_form=False), d64.matvec(xr, False)
        builds["product_rel_err_vs_dense_fp64"] = {
            k: [float(np.linalg.norm(y[w] - c[w]) / np.linalg.norm(c[w])) for w in (0, 1)]
            for k, y in (("full", a), ("sampled", b))}
        hs.close()
        log(f"[bench] H build full {builds['full']['ms_build']:.0f} ms ({builds['full']['entries_evaluated']} "
            f"pairs), sampled {builds['sampled']['ms_build']:.0f} ms ({builds['sampled']['entries_evaluated']} "
            f"pairs); product errors {builds['product_rel_err_vs_dense_fp64']}")
    hb = None
    if not args.no_bf16_far:
        # the same build with the low-rank factors stored in BF16
        t0 = time.perf_counter()
        hb = qd.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=args.aca_tol,
                                  prec=qd.PREC_BF16_FAR)
        log(f"[bench] H BF16-factor operator in {time.perf_counter() - t0:.2f}s: "
            f"{hb.info()['bytes_per_product'] / 1e6:.1f} MB per product")
    d32 = qd.DenseOperator.from_triangles(*geo, m["mu"], lam, qd.PREC_FP32)
    ks, _ = d64.kself()
    ksum, knsum = d64.matvec(np.ones(n), loading_form=False)
    prob = configs.problem_from_triangles(m, ks, ksum, knsum, order=p)
    prob["ext_at_told"] = 1  # main_H reads the external tables at TOld; the plain runs use the same timing
    hsum, hnsum = h32.matvec(np.ones(n), loading_form=False)
    prob_h = dict(prob, ext=configs.backslip_tables(hsum, hnsum))
    # operator accuracy against the dense FP64 product
    def accuracy(h):
        rng = np.random.default_rng(5)
        acc = {}
        for label, x in (("random", rng.standard_normal(n)), ("uniform", np.ones(n))):
            a = d64.matvec(x, loading_form=True)
            b = h.matvec(x, loading_form=True)
            acc[label] = {"tl_rel": float(np.abs(a[0] - b[0]).max() / np.abs(a[0]).max()),
                          "sigma_rel": float(np.abs(a[1] - b[1]).max() / np.abs(a[1]).max())}
        return acc

    acc = accuracy(h32)
    ph1, ph2, tot = h32.time_phases(20)
    runs = [("fp64_dense", d64, prob), ("mixed_h", h32, prob_h), ("fp32_dense", d32, prob)]
    bf = None
    if hb is not None:
        bsum, bnsum = hb.matvec(np.ones(n), loading_form=False)
        runs.append(("bf16_far_h", hb, dict(prob, ext=configs.backslip_tables(bsum, bnsum))))
        bi = hb.info()
        bf = {"operator_rel_err": accuracy(hb), "bytes_per_product": bi["bytes_per_product"],
              "phase_bytes": [bi["phase1_bytes"], bi["phase2_bytes"]],
              "h_product_us": [round(v * 1e3, 2) for v in hb.time_phases(20)]}
    res = {}
    for name, op, pr in runs:
        log(f"[bench] {name}: warmup {args.warmup}, timed {args.steps} steps")
        r = run_timed(qd, op, pr, args.steps, args.warmup, args.graph_attempts, args.kernel_steps)
        if r["status"] != 0:
            raise SystemExit(f"{name} run failed with status {r['status']}")
        summarize(name, r, configs.SEC_PER_YEAR)
        r["first_event"] = configs.first_event_time(r["trace"]["told"], r["trace"]["dt"], r["trace"]["vmax"])
        res[name] = r
    if hb is not None:
        r = res["bf16_far_h"]
        bf.update(rhs_evals_per_s=round(r["evals_per_s"], 2), sim_years_per_hour=round(r["sim_years_per_hour"], 4),
                  first_event_s=r["first_event"], divergence_from_fp64=divergence(r["trace"], res["fp64_dense"]["trace"]))
        bf["h_product_us"] = dict(zip(("phase1", "phase2", "total"), bf["h_product_us"]))
        hb.close()
    out = dict(n=n, info=info, st=st, rk=rk, acc=acc, phases=(ph1, ph2, tot), res=res, h32=h32, prob_h=prob_h,
               builds=builds, bf16_far=bf)
    d64.close()
    d32.close()
    return out


def network(qd, args):
    from qdfn import configs
    cfg = configs.fault_network()
    seg = cfg["seg"]
    ops = {name: qd.DenseOperator.from_rect_segments(seg, cfg["mu"], cfg["nu"], prec)
           for name, prec in (("fp64", qd.PREC_FP64), ("fp32", qd.PREC_FP32))}
    ks, _ = ops["fp64"].kself()
    ksum, knsum = ops["fp64"].matvec(np.ones(seg.shape[0]), loading_form=False)
    prob = configs.problem_from_segments(cfg, ks, ksum, knsum)
    res = {}
    for name in ("fp64", "fp32"):
        r = run_timed(qd, ops[name], prob, args.network_steps, args.warmup, args.graph_attempts, 0)
        if r["status"] != 0:
            raise SystemExit(f"network {name} run failed with status {r['status']}")
        summarize("network " + name, r, configs.SEC_PER_YEAR)
        res[name] = r
    for o in ops.values():
        o.close()
    t = [configs.first_event_time(res[k]["trace"]["told"], res[k]["trace"]["dt"], res[k]["trace"]["vmax"])
         for k in ("fp64", "fp32")]
    err = abs(t[1] - t[0]) / t[0] if (t[0] and t[1]) else None
    return {"workload": "fault_network_8192", "fp64_rhs_evals_per_s": round(res["fp64"]["evals_per_s"], 2),
            "mixed_rhs_evals_per_s": round(res["fp32"]["evals_per_s"], 2),
            "speedup_mixed_vs_fp64": round(res["fp32"]["evals_per_s"] / res["fp64"]["evals_per_s"], 3),
            "event_time_rel_err": err, "steps_timed": args.network_steps}


def ensemble(qd, args):
    """256 independent planar faults of 500-4,000 elements (FP32 operators, one
    per distinct geometry) advanced together by qdfn_ensemble_run for up to
    20 simulated years each (step cap args.ensemble_steps per member)."""
    from qdfn import configs
    from qdfn.ensemble import problem_for_member
    t0 = time.perf_counter()
    members = configs.ensemble_members()
    e = qd.Ensemble.from_members(members, problem_for_member, max_rhs=0, record_step=10)
    info = e.info()
    log(f"[bench] ensemble: {len(members)} members, {sum(m['seg'].shape[0] for m in members)} elements, "
        f"{info['groups']} operator groups (max_rhs {info['max_rhs']}, autotuned), built in "
        f"{time.perf_counter() - t0:.1f}s")
    t_end = 20 * configs.SEC_PER_YEAR
    s0 = e.stats()
    t0 = time.perf_counter()
    e.run(args.ensemble_steps, t_end=t_end, graph_attempts=16)
    el = time.perf_counter() - t0
    s1 = e.stats()
    rhs = sum(b["rhs_evals"] - a["rhs_evals"] for a, b in zip(s0, s1))
    steps = sum(b["steps"] - a["steps"] for a, b in zip(s0, s1))
    years = sum(b["t"] - a["t"] for a, b in zip(s0, s1)) / configs.SEC_PER_YEAR
    done = sum(1 for b in s1 if b["t"] >= t_end)
    log(f"[bench] ensemble: {rhs} RHS evals ({steps} steps) in {el:.2f}s = {rhs / el:.0f}/s; "
        f"{done}/{len(members)} members reached 20 years")
    out = {"workload": "ensemble_256", "members": len(members),
           "elements": int(sum(m["seg"].shape[0] for m in members)), "operator_groups": info["groups"],
           "max_rhs": info["max_rhs"], "matrix_bytes_per_sweep": info["bytes_per_sweep"],
           "rhs_evals_per_s": round(rhs / el, 1), "steps": steps, "elapsed_s": round(el, 3),
           "sim_years_per_hour": round(years / (el / 3600.0), 1),
           "members_at_20_years": done, "step_cap_per_member": args.ensemble_steps, "dtype": "f64 state, fp32 operator"}
    e.close()
    return out


def adjoint_case(qd, args):
    """64 x 64 elements of 50 m (4,096), per-element a, b, Dc; loss = squared
    slip misfit at 32 seeded elements (+ squared first-event-time misfit when
    the window holds the onset) over a window of accepted steps; gradient by
    qdfn_sim_adjoint, checked against central differences on 16 random
    parameters.  Windows: args.adjoint_steps steps (through the first event;
    the rupture's sensitivity grows exponentially, so differences are taken on
    4 parameters only, to show it), the onset plus 38 steps (the first HighV
    solves) and the interseismic steps before the onset."""
    from qdfn import adjoint, configs
    m = configs.adjoint_fault()
    seg = m["seg"]
    op = qd.DenseOperator.from_rect_segments(seg, m["mu"], m["nu"], qd.PREC_FP64)
    ks, _ = op.kself()
    ksum, knsum = op.matvec(np.ones(op.n), loading_form=False)
    prob = configs.problem_from_segments(m, ks, ksum, knsum)
    rng = np.random.default_rng(41)
    obs = np.sort(rng.choice(op.n, 32, replace=False))
    out = {"workload": "adjoint_fault_4096", "observations": 32}
    K = args.adjoint_steps
    windows = [("full_window", K, 4)]
    while windows:
        name, K, n_fd = windows.pop(0)
        sim = qd.Simulation(op, prob, record_step=1, record_capacity=K)
        t0 = time.perf_counter()
        sim.run(K)
        fwd_s = time.perf_counter() - t0
        h = sim.history()
        dK = h["History_Disp"][-1]
        target = dK[obs] * 1.1 + 1e-3
        _, t_first = adjoint.misfit(h, 0, obs, target)
        if t_first is not None:
            t_target = t_first * 1.05
            w_time = float(np.sum((dK[obs] - target) ** 2)) / (t_first - t_target) ** 2
        else:
            t_target, w_time = 0.0, 0.0
        t0 = time.perf_counter()
        L, g, info = sim.adjoint(prob["v_i"], prob["theta_i"], obs, target, w_time, t_target)
        bwd_s = time.perf_counter() - t0
        sim.close()
        rng2 = np.random.default_rng(43)
        params = [(("a", "b", "dc")[q % 3], int(rng2.integers(op.n))) for q in range(16)][:n_fd]
        chk = adjoint.fd_check(op, prob, K, obs, target, g, params, w_time, t_target)
        errs = np.array([c["rel_err"] for c in chk])
        smooth = np.array([c["smoothness"] for c in chk])
        log(f"[bench] adjoint ({name}): {K} steps forward {fwd_s * 1e3:.1f} ms, backward "
            f"{info['ms_backward']:.1f} ms (GPU), onset step {info['onset_step']}, retried steps "
            f"{info['retried_steps']}, max |grad| {max(np.abs(v).max() for v in g.values()):.3e}; central "
            f"differences on {n_fd} parameters: rel. error median {np.median(errs):.2e} max {errs.max():.2e} "
            f"(one-sided differences disagree by {np.median(smooth):.1e} at the best step)")
        out[name] = {"steps": K, "loss": L, "time_term": w_time > 0.0, "forward_ms": round(fwd_s * 1e3, 2),
                     "backward_ms": round(info["ms_backward"], 2), "backward_wall_ms": round(bwd_s * 1e3, 2),
                     "onset_step": info["onset_step"], "retried_steps": info["retried_steps"],
                     "grad_max_abs": float(max(np.abs(v).max() for v in g.values())),
                     "fd_params": n_fd, "fd_rel_err_median": float(np.median(errs)),
                     "fd_rel_err_max": float(errs.max()), "fd_one_sided_disagreement_median": float(np.median(smooth)),
                     "fd": [{k: (float(f"{v:.9g}") if isinstance(v, float) else v) for k, v in c.items()}
                            for c in chk]}
        if name == "full_window" and info["onset_step"] > 2:
            # through the onset and the first 30-40 steps of the event (HighV solves)
            windows.append(("onset_window", min(K, int(info["onset_step"]) + 38), 16))
            windows.append(("interseismic", int(info["onset_step"]) - 2, 16))
    op.close()
    return out


def cpu_baseline(h32, st, prob_h, n_steps, threads):
    """The oracle's FP64 main_H (C restatement of HmatSolver_Pararllel and of the
    loop; products and local solves over OpenMP threads, outputs independent of
    the count) on the same stored blocks, for a bounded number of steps, at each
    thread count in `threads`."""
    from oracle import oracle as orc
    HS, HN = h32.blocks(st["range_SR"])
    out = {}
    prev = orc.set_threads(1)
    for th in threads:
        orc.set_threads(th)
        t0 = time.perf_counter()
        status, r = orc.sim_hmat(prob_h, HS, HN, total_step=n_steps, record_step=0, trace=False)
        el = time.perf_counter() - t0
        out[th] = dict(value=r["rhs_evals"] / el, elapsed=el, steps=n_steps, status=status)
    orc.set_threads(prev)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000, help="timed accepted steps per run (rough fault)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--graph-attempts", type=int, default=32)
    ap.add_argument("--kernel-steps", type=int, default=300, help="direct-launch steps with per-kernel events")
    ap.add_argument("--crit", type=float, default=1.0, help="DistDiamRatioCrit of the cluster tree")
    ap.add_argument("--aca-tol", type=float, default=1e-6)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--network-steps", type=int, default=6000)
    ap.add_argument("--no-network", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=20000, help="accepted-step cap per ensemble member")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-sampled-build", action="store_true")
    ap.add_argument("--no-bf16-far", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU layout is "
              "design-only: see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import qdfn
    log(f"[bench] device {qdfn.device_info()}")
    rf = rough_fault(qdfn, args)
    res = rf["res"]
    mixed, plain = res["mixed_h"], res["fp64_dense"]
    t64, th, t32 = plain["first_event"], mixed["first_event"], res["fp32_dense"]["first_event"]
    ev_err = abs(th - t64) / t64 if (th and t64) else None
    ev_ref = abs(t32 - t64) / t64 if (t32 and t64) else None
    ph1, ph2, tot = rf["phases"]
    bytes2 = rf["info"]["phase2_bytes"]
    achieved = bytes2 / (mixed["coupling_ms"] * 1e-3) / 1e9
    net = None if args.no_network else network(qdfn, args)
    ens = None if args.no_ensemble else ensemble(qdfn, args)
    adj = None if args.no_adjoint else adjoint_case(qdfn, args)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle FP64 main_H, {args.cpu_steps} steps")
        # all of the host's share of cores (16 per GPU on a shared box), and one
        nall = min(int(os.environ.get("OMP_NUM_THREADS", "0") or os.cpu_count() or 1), os.cpu_count() or 1, 16)
        cs = cpu_baseline(rf["h32"], rf["st"], rf["prob_h"], args.cpu_steps, sorted({1, nall}))
        c, c1 = cs[nall], cs[1]
        cpu = {"value": round(c["value"], 4), "unit": "RHS evals/s", "cores": nall, "kind": "port",
               "sample": f"{c['steps']} accepted steps of the 20,000-triangle H-matrix run, FP64 C restatement "
                         f"of main_H / HmatSolver on the same blocks, {nall} OpenMP threads "
                         f"({c['elapsed']:.1f}s; one thread: {c1['value']:.3f} RHS evals/s, "
                         f"{c1['elapsed']:.1f}s)",
               "single_thread": {"value": round(c1["value"], 4), "cores": 1}}
        log(f"[bench] CPU baseline {c['value']:.3f} RHS evals/s on {nall} threads, {c1['value']:.3f} on one")
    rf["h32"].close()
    rk = rf["rk"]
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
        "data": "synthetic (seeded rough fault, back-slip loading)",
        "config": {"workload": "rough_fault_20000", "n_elements": rf["n"], "operator": "H-matrix, FP32 storage",
                   "leaf": 64, "dist_diam_ratio": args.crit, "aca_tol": args.aca_tol, "rank_cap": 1024,
                   "plain": "dense FP64", "steps_timed": args.steps},
        "fp64_plain": {"rhs_evals_per_s": round(plain["evals_per_s"], 2),
                       "sim_years_per_hour": round(plain["sim_years_per_hour"], 4),
                       "coupling_us": round(plain["coupling_ms"] * 1e3, 2)},
        "mixed": {"rhs_evals_per_s": round(mixed["evals_per_s"], 2),
                  "sim_years_per_hour": round(mixed["sim_years_per_hour"], 4),
                  "phase2_kernel_us": round(mixed["coupling_ms"] * 1e3, 2),
                  "h_product_us": {"phase1": round(ph1 * 1e3, 2), "phase2": round(ph2 * 1e3, 2),
                                   "total": round(tot * 1e3, 2)},
                  "stored_entries": rf["info"]["stored_entries"],
                  "bytes_per_product": rf["info"]["bytes_per_product"],
                  "mean_rank_shear": round(float(rk["rank_S"][rk["kind_S"] == 1].mean()), 2),
                  "mean_rank_normal": round(float(rk["rank_N"][rk["kind_N"] == 1].mean()), 2)},
        "fp32_dense": {"rhs_evals_per_s": round(res["fp32_dense"]["evals_per_s"], 2)},
        "speedup_mixed_vs_fp64": round(mixed["evals_per_s"] / plain["evals_per_s"], 3),
        "event_time_rel_err": ev_err,
        "event_time_rel_err_fp32_dense": ev_ref,
        "first_event_s": {"fp64_plain": t64, "mixed_h": th, "fp32_dense": t32},
        "divergence_from_fp64": {"mixed_h": divergence(mixed["trace"], plain["trace"]),
                                 "fp32_dense": divergence(res["fp32_dense"]["trace"], plain["trace"])},
        "operator_rel_err": rf["acc"],
        "bf16_far_h": rf["bf16_far"],
        "h_build": rf["builds"],
        "network_8192": net,
        "ensemble_256": ens,
        "adjoint_4096": adj,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
