#!/usr/bin/env python3
"""Benchmark: RHS evaluations per second and simulated years per hour of the
device-resident earthquake-cycle loop, mixed precision against the FP64 plain
form, at a stated event-time error.

Headline workload: the rough strike-slip fault of 20,000 triangles (100 x 100
cells of 50 m, three seeded sinusoidal roughness modes), H-matrix operator
with cluster leaves of 64, admissibility eta = 2 read through the reference's
criterion as DistDiamRatioCrit 1.5, ACA tolerance 1e-6; storage FP32 for dense
blocks and leading low-rank components, FP16 unit vectors for the trailing
components of each block within its error budget; FP64 accumulation, state,
local solves and time.  Plain form: the dense FP64 operator (main() of
scripts/Functions_RSFDFN3DMain_D.jl).  Both loops have main_H / main semantics
on the device; each run is loaded by back-slip tables built from its own
operator (UniformBackslip.jl).  Reported beside: the same H structure in FP32
storage, the reference's default DistDiamRatioCrit 1.0, and (--bf16-far)
BF16 low-rank factors.

A step is one accepted time step (one coupling product plus one or more
local-solve passes).  `value` = RHS evaluations / s of the headline run over
exactly --steps accepted steps replayed from a captured HIP graph, inputs
resident in HBM.  The accuracy half of the metric does not depend on --steps:
fixed-duration runs to the first event (or the configuration's duration)
give the first-event time of the mixed run, of the FP64 run and of FP64 runs
whose initial state is perturbed at rounding level (the spread that measures
the workload's conditioning).  Prints one JSON line on stdout; progress on
stderr.
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
PROFILE = os.path.join(ROOT, "profiles", "hmat_phase2_mixed_rough_fault_20000.md")
YEAR = 3.15576e7
CHECKPOINT_EVERY = 50  # adjoint workload: forward checkpoints, segments recomputed in the backward sweep


def measured_traffic(key="traffic_bytes_per_launch"):
    """HBM bytes per launch of a headline product kernel (phase 2 by default,
    key traffic_phase1_bytes_per_launch for phase 1) from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated); None if absent."""
    try:
        for line in open(PROFILE):
            if line.startswith(key + ":"):
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
    sim.close()
    rhs = s1["rhs_evals"] - s0["rhs_evals"]
    return dict(elapsed=el, rhs=rhs, retries=s1["retry_passes"] - s0["retry_passes"],
                sim_seconds=s1["t"] - s0["t"], coupling_ms=s2["coupling_ms_avg"],
                coupling_timed=s2["coupling_timed"], status=s2["status"], evals_per_s=rhs / el)


def summarize(name, r):
    r["sim_years_per_hour"] = r["sim_seconds"] / YEAR / (r["elapsed"] / 3600.0)
    log(f"[bench] {name}: {r['rhs']} RHS evals in {r['elapsed']:.3f}s = {r['evals_per_s']:.1f}/s, "
        f"coupling {r['coupling_ms']*1e3:.1f} us over {r['coupling_timed']}, "
        f"{r['sim_years_per_hour']:.3f} sim-yr/h, retries {r['retries']}")


def onsets(told, dt, vmax, threshold=1e-3, rearm=1e-5):
    """Event onsets: accepted steps whose max V reaches `threshold` after the
    run has been below `rearm` since the last onset (so the oscillations of
    one rupture around the threshold count once), placed by linear
    interpolation in log V (time axis = end of step)."""
    from qdfn import configs
    v = np.asarray(vmax)
    out = []
    armed = True
    for k in range(v.size):
        if armed and v[k] >= threshold:
            if k == 0:
                out.append(float(told[0] + dt[0]))
            else:
                out.append(configs.first_event_time(told[k - 1:k + 1], dt[k - 1:k + 1], v[k - 1:k + 1], threshold))
            armed = False
        elif not armed and v[k] < rearm:
            armed = True
    return out


def first_event_peak(vmax, threshold=1e-3, rearm=1e-5):
    """Peak V of the first event: the largest max-V of the accepted steps from
    its onset (max V >= threshold) until max V falls back below `rearm`;
    (peak, ended) with ended False when the run stops inside the event, None
    without an event."""
    v = np.asarray(vmax)
    on = np.nonzero(v >= threshold)[0]
    if on.size == 0:
        return None, False
    k = int(on[0])
    below = np.nonzero(v[k:] < rearm)[0]
    end = k + int(below[0]) if below.size else v.size
    return float(v[k:end].max()), bool(below.size)


def run_window(qd, op, prob, years, stop_at_event, max_steps, chunk=2000, graph_attempts=32):
    """One run from t = 0 to `years` simulated years (or, with stop_at_event,
    through its first event: until max V is back below 1e-5 m/s), max_steps at
    most, launched in chunks of accepted steps replayed from the captured
    attempt graph."""
    from qdfn import configs
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=max_steps)
    t_end = years * YEAR
    ev = None
    wall = 0.0
    while True:
        s = sim.stats()
        left = max_steps - s["steps"]
        if left <= 0 or s["t"] >= t_end:
            break
        t0 = time.perf_counter()
        sim.run_until(t_end, min(chunk, left), graph_attempts=graph_attempts)
        wall += time.perf_counter() - t0
        if stop_at_event:
            tr = sim.trace()
            ev = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
            if ev is not None and first_event_peak(tr["vmax"])[1]:
                break
    s = sim.stats()
    tr = sim.trace()
    sim.close()
    if ev is None:
        ev = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
    peak, ended = first_event_peak(tr["vmax"])
    return dict(first_event_s=ev, peak_v=peak, event_ended=ended, steps=s["steps"], sim_seconds=s["t"], wall_s=wall,
                rhs=s["rhs_evals"],
                retries=s["retry_passes"], onsets=onsets(tr["told"], tr["dt"], tr["vmax"]), status=s["status"],
                sim_years_per_hour=(s["t"] / YEAR) / (wall / 3600.0) if wall > 0 else None,
                rhs_evals_per_s=s["rhs_evals"] / wall if wall > 0 else None)


def event_accuracy(qd, ops, probs, years, n_pert, max_steps):
    """First-event times (onset of V >= 1e-3 m/s) within `years` of the FP64
    plain run, of n_pert FP64 plain runs whose initial theta is multiplied by
    1 + 1e-15 xi (xi standard normal, seeds 1..n_pert: rounding-level
    perturbations of the reference's own path) and of each listed mixed run."""
    base = run_window(qd, ops["fp64_dense"], probs["fp64_dense"], years, True, max_steps)
    log(f"[bench] event: fp64_dense {fmt_ev(base['first_event_s'])}, peak V {base['peak_v']}, in {base['steps']} "
        f"steps, {base['wall_s']:.1f}s")
    pert, pert_peak = [], []
    for seed in range(1, n_pert + 1):
        p = dict(probs["fp64_dense"])
        xi = np.random.default_rng(seed).standard_normal(p["theta_i"].size)
        p["theta_i"] = p["theta_i"] * (1.0 + 1e-15 * xi)
        r = run_window(qd, ops["fp64_dense"], p, years, True, max_steps)
        pert.append(r["first_event_s"])
        pert_peak.append(r["peak_v"])
        log(f"[bench] event: fp64_dense perturbed (seed {seed}) {fmt_ev(r['first_event_s'])}, peak V {r['peak_v']}")
    out = {"window_years": years, "fp64_plain_s": base["first_event_s"], "fp64_perturbed_s": pert,
           "peak_v_definition": "largest max-V of the accepted steps from the first onset (1e-3 m/s) until max V "
                                "is back below 1e-5 m/s",
           "fp64_plain_peak_v": base["peak_v"], "fp64_perturbed_peak_v": pert_peak}
    pv = [x for x in pert_peak + [base["peak_v"]] if x is not None]
    out["fp64_peak_v_rel_spread"] = (max(pv) - min(pv)) / base["peak_v"] if pv and base["peak_v"] else None
    evs = [t for t in pert + [base["first_event_s"]] if t is not None]
    out["fp64_spread"] = {"min_s": min(evs) if evs else None, "max_s": max(evs) if evs else None,
                          "runs_without_event": sum(1 for t in pert + [base["first_event_s"]] if t is None),
                          "runs": n_pert + 1,
                          "rel_spread": (max(evs) - min(evs)) / base["first_event_s"]
                          if evs and base["first_event_s"] else None}
    for name in ops:
        if name == "fp64_dense":
            continue
        r = run_window(qd, ops[name], probs[name], years, True, max_steps)
        t, t64 = r["first_event_s"], base["first_event_s"]
        entry = {"first_event_s": t, "peak_v": r["peak_v"],
                 "peak_v_rel_err": abs(r["peak_v"] - base["peak_v"]) / base["peak_v"]
                 if r["peak_v"] is not None and base["peak_v"] else None}
        if t is not None and t64 is not None:
            entry["rel_err"] = abs(t - t64) / t64
        else:
            entry["rel_err"] = None
            entry["status"] = (f"no event within {years} years" if t is None else "FP64 run has no event") + \
                (f" (FP64: {t64 / YEAR:.3f} years)" if t64 else "")
        lo, hi = out["fp64_spread"]["min_s"], out["fp64_spread"]["max_s"]
        none_seen = out["fp64_spread"]["runs_without_event"] > 0
        entry["within_fp64_perturbation_range"] = bool(
            (t is None and none_seen) or (t is not None and lo is not None and lo <= t <= (np.inf if none_seen else hi)))
        out[name] = entry
        log(f"[bench] event: {name} {fmt_ev(t)}; rel. error {entry['rel_err']}; within the FP64 spread "
            f"{entry['within_fp64_perturbation_range']}; peak V rel. error {entry['peak_v_rel_err']}")
    return out, base


def fmt_ev(t):
    return "no event" if t is None else f"{t / YEAR:.4f} years"


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
    t0 = time.perf_counter()
    structs = {c: qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=c,
                                    arrange_point=(0.0, 0.0, 0.0)) for c in sorted({args.crit, 1.0})}
    st = structs[args.crit]
    p = st["perm"]
    geo = [a[p] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    d64 = qd.DenseOperator.from_triangles(*geo, m["mu"], lam, qd.PREC_FP64)
    log(f"[bench] rough fault n={n}: structures {time.perf_counter() - t0:.2f}s (dense FP64 operator included)")

    def build(prec, crit=args.crit, sampled=False):
        s = structs[crit]
        g = [a[s["perm"]] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
        t0 = time.perf_counter()
        h = qd.HMatrix.build_tri(*g, m["mu"], lam, s["range_SR"], s["admissible"], aca_rtol=args.aca_tol,
                                 prec=prec, sampled=sampled)
        info = h.info()
        log(f"[bench] H (crit {crit}, prec {prec}{', sampled' if sampled else ''}) in {time.perf_counter() - t0:.2f}s: "
            f"{info['blocks']} blocks, {int(s['admissible'].sum())} admissible, {info['stored_entries'] / 1e6:.1f} M "
            f"entries, {info['bytes_per_product'] / 1e6:.1f} MB per product")
        return h, s

    hm, _ = build(qd.PREC_MIXED)
    h32, _ = build(qd.PREC_FP32)
    hm1, st1 = build(qd.PREC_MIXED, crit=1.0) if args.crit != 1.0 else (None, None)
    builds = {"full": _build_summary(hm)}
    xr = np.random.default_rng(6).standard_normal(n)
    ref_prod = d64.matvec(xr, False)

    def prod_err(h, x=xr, c=ref_prod):
        y = h.matvec(x, loading_form=False)
        return [float(np.linalg.norm(y[w] - c[w]) / np.linalg.norm(c[w])) for w in (0, 1)]

    builds["product_rel_err_vs_dense_fp64"] = {"full": prod_err(hm)}
    if not args.no_sampled_build:
        hs, _ = build(qd.PREC_MIXED, sampled=True)
        builds["sampled"] = _build_summary(hs)
        builds["product_rel_err_vs_dense_fp64"]["sampled"] = prod_err(hs)
        hs.close()
    hb = build(qd.PREC_BF16_FAR)[0] if args.bf16_far else None
    ks, _ = d64.kself()
    ksum, knsum = d64.matvec(np.ones(n), loading_form=False)
    prob = configs.problem_from_triangles(m, ks, ksum, knsum, order=p)
    prob["ext_at_told"] = 1  # main_H reads the external tables at TOld; the plain run uses the same timing

    def own_loading(h, order=p, fric=m):
        hsum, hnsum = h.matvec(np.ones(n), loading_form=False)
        kk, _ = h.kself()
        pr = configs.problem_from_triangles(fric, kk, hsum, hnsum, order=order)
        pr["ext_at_told"] = 1
        return pr

    probs = {"fp64_dense": prob, "mixed_h": own_loading(hm), "fp32_h": own_loading(h32)}
    # the same mesh and operators with the single fault's friction on a central
    # velocity-weakening patch: a first event that is a well-conditioned
    # function of the operator (the accuracy runs)
    mp = configs.rough_fault_patch(nx=100, nz=100)
    mp = dict(mp, P1=P1, P2=P2, P3=P3)
    probs_patch = {"fp64_dense": configs.problem_from_triangles(mp, ks, ksum, knsum, order=p),
                   "mixed_h": own_loading(hm, fric=mp), "fp32_h": own_loading(h32, fric=mp)}
    probs_patch["fp64_dense"]["ext_at_told"] = 1
    ops = {"fp64_dense": d64, "mixed_h": hm, "fp32_h": h32}
    if hm1 is not None:
        probs["mixed_h_crit1"] = own_loading(hm1, st1["perm"])
        ops["mixed_h_crit1"] = hm1
    if hb is not None:
        probs["bf16_far_h"] = own_loading(hb)
        ops["bf16_far_h"] = hb

    # operator accuracy against the dense FP64 product (same element order)
    def accuracy(h):
        rng = np.random.default_rng(5)
        acc = {}
        for label, x in (("random", rng.standard_normal(n)), ("uniform", np.ones(n))):
            a = d64.matvec(x, loading_form=True)
            b = h.matvec(x, loading_form=True)
            acc[label] = {"tl_rel": float(np.abs(a[0] - b[0]).max() / np.abs(a[0]).max()),
                          "sigma_rel": float(np.abs(a[1] - b[1]).max() / np.abs(a[1]).max())}
        return acc

    acc = {"mixed_h": accuracy(hm), "fp32_h": accuracy(h32)}
    if hb is not None:
        acc["bf16_far_h"] = accuracy(hb)
    res = {}
    for name in ops:
        log(f"[bench] {name}: warmup {args.warmup}, timed {args.steps} steps")
        r = run_timed(qd, ops[name], probs[name], args.steps, args.warmup, args.graph_attempts, args.kernel_steps)
        if r["status"] != 0:
            raise SystemExit(f"{name} run failed with status {r['status']}")
        summarize(name, r)
        res[name] = r
    phases = {k: [round(v * 1e3, 2) for v in ops[k].time_phases(20)] for k in ops if k != "fp64_dense"}
    hinfo = {k: dict(ops[k].info(), bytes_by_part=ops[k].bytes_by_part()) for k in ops if k != "fp64_dense"}
    out = dict(n=n, st=st, probs=probs, ops=ops, res=res, acc=acc, phases=phases, hinfo=hinfo, builds=builds,
               d64=d64, hm=hm, d64_bytes=d64.bytes_per_product(), probs_patch=probs_patch)
    return out


def network(qd, args):
    """Config 1: 8 rectangular faults, 8,192 elements, dense FP64 against
    FP32 storage and against a mixed-storage H operator (DistDiamRatioCrit
    1.0, leaves of 64, ACA args.aca_tol; elements in the cluster tree's
    order); first events within 100 years and their FP64 perturbation spread
    (as for the rough fault)."""
    from qdfn import configs
    cfg = configs.fault_network()
    seg = cfg["seg"]
    ops = {name: qd.DenseOperator.from_rect_segments(seg, cfg["mu"], cfg["nu"], prec)
           for name, prec in (("fp64_dense", qd.PREC_FP64), ("fp32_dense", qd.PREC_FP32))}
    ks, _ = ops["fp64_dense"].kself()
    ksum, knsum = ops["fp64_dense"].matvec(np.ones(seg.shape[0]), loading_form=False)
    prob = configs.problem_from_segments(cfg, ks, ksum, knsum)
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    segs = np.ascontiguousarray(seg[p])
    t0 = time.perf_counter()
    ops["mixed_h"] = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], st["range_SR"], st["admissible"],
                                           aca_rtol=args.aca_tol, prec=qd.PREC_MIXED)
    h_build_s = time.perf_counter() - t0
    h_bytes = ops["mixed_h"].info()["bytes_per_product"]
    log(f"[bench] network mixed H built in {h_build_s:.2f}s: {h_bytes / 1e6:.1f} MB per product")
    probs = {"fp64_dense": prob, "fp32_dense": prob,
             "mixed_h": configs.problem_from_segments(dict(cfg, seg=segs), ks[p], ksum[p], knsum[p])}
    res = {}
    for name in ops:
        r = run_timed(qd, ops[name], probs[name], args.network_steps, args.warmup, args.graph_attempts, 0)
        if r["status"] != 0:
            raise SystemExit(f"network {name} run failed with status {r['status']}")
        summarize("network " + name, r)
        res[name] = r
    acc, base = event_accuracy(qd, ops, probs, 100.0, args.perturbations, 60000)
    # the configuration's 100 years on the mixed H operator, through its events
    full = run_window(qd, ops["mixed_h"], probs["mixed_h"], 100.0, False, args.network_cycle_steps, chunk=50000)
    log(f"[bench] network mixed_h over 100 years: {full['sim_seconds'] / YEAR:.1f} years in {full['steps']} steps, "
        f"{full['wall_s']:.2f}s, {full['sim_years_per_hour']:.0f} sim-yr/h, {len(full['onsets'])} events")
    out = {"workload": "fault_network_8192", "fp64_rhs_evals_per_s": round(res["fp64_dense"]["evals_per_s"], 2),
           "mixed_rhs_evals_per_s": round(res["fp32_dense"]["evals_per_s"], 2),
           "speedup_mixed_vs_fp64": round(res["fp32_dense"]["evals_per_s"] / res["fp64_dense"]["evals_per_s"], 3),
           "mixed_h_rhs_evals_per_s": round(res["mixed_h"]["evals_per_s"], 2),
           "speedup_mixed_h_vs_fp64": round(res["mixed_h"]["evals_per_s"] / res["fp64_dense"]["evals_per_s"], 3),
           "mixed_h_bytes_per_product": int(h_bytes), "mixed_h_build_s": round(h_build_s, 2),
           "steps_timed": args.network_steps, "coupling_us": {k: round(res[k]["coupling_ms"] * 1e3, 2) for k in res},
           "event_time": acc,
           "mixed_h_over_duration": {"years": full["sim_seconds"] / YEAR, "target_years": 100.0, "steps": full["steps"],
                                  "step_cap": args.network_cycle_steps, "wall_s": round(full["wall_s"], 3),
                                  "rhs_evals_per_s": round(full["rhs_evals_per_s"], 1),
                                  "sim_years_per_hour": round(full["sim_years_per_hour"], 1),
                                  "events": len(full["onsets"]), "retry_passes": full["retries"]},
           "tolerance": "first-event time of FP32 storage and of the mixed H operator within 5e-5 of the FP64 "
                        "run's (test_gpu_network.py; measured 2.85e-5 and 2.4e-5; FP64 runs perturbed at 1e-15 "
                        "agree to 1e-9)"}
    for o in ops.values():
        o.close()
    return out


def single_fault(qd, args):
    """Config 0: the 512-element single fault, dense FP64, run toward 200
    simulated years (at most args.single_steps accepted steps): RHS evals/s
    and sim-years/hour over whole cycles (first to last onset in the window),
    event count and recurrence."""
    from qdfn import configs
    cfg = configs.single_fault()
    op = qd.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], qd.PREC_FP64)
    ks, _ = op.kself()
    ksum, knsum = op.matvec(np.ones(op.n), loading_form=False)
    prob = configs.problem_from_segments(cfg, ks, ksum, knsum)
    r = run_window(qd, op, prob, 200.0, False, args.single_steps, chunk=100_000)
    op.close()
    ev = r["onsets"]
    log(f"[bench] single fault: {r['sim_seconds'] / YEAR:.1f} years in {r['steps']} steps, {r['wall_s']:.2f}s = "
        f"{r['rhs_evals_per_s']:.0f} RHS evals/s, {r['sim_years_per_hour']:.0f} sim-yr/h, {len(ev)} events")
    return {"workload": "single_fault_512", "years": r["sim_seconds"] / YEAR, "target_years": 200.0,
            "steps": r["steps"], "step_cap": args.single_steps, "rhs_evals": r["rhs"], "retry_passes": r["retries"],
            "wall_s": round(r["wall_s"], 3), "rhs_evals_per_s": round(r["rhs_evals_per_s"], 1),
            "sim_years_per_hour": round(r["sim_years_per_hour"], 1), "events": len(ev),
            "mean_recurrence_years": float(np.mean(np.diff(ev)) / YEAR) if len(ev) > 1 else None,
            "first_event_years": ev[0] / YEAR if ev else None, "dtype": "f64"}


def ensemble(qd, args):
    """256 independent planar faults of 500-4,000 elements (FP32 operators, one
    per distinct geometry) advanced together by qdfn_ensemble_run to 20
    simulated years each (step cap args.ensemble_steps per member)."""
    from qdfn import configs
    from qdfn.ensemble import problem_for_member
    t0 = time.perf_counter()
    members = configs.ensemble_members()
    e = qd.Ensemble.from_members(members, problem_for_member, max_rhs=0, record_step=10)
    info = e.info()
    log(f"[bench] ensemble: {len(members)} members, {sum(m['seg'].shape[0] for m in members)} elements, "
        f"{info['groups']} operator groups (max_rhs {info['max_rhs']}, autotuned), built in "
        f"{time.perf_counter() - t0:.1f}s")
    t_end = 20 * YEAR
    s0 = e.stats()
    t0 = time.perf_counter()
    e.run(args.ensemble_steps, t_end=t_end, graph_attempts=16)
    el = time.perf_counter() - t0
    s1 = e.stats()
    rhs = sum(b["rhs_evals"] - a["rhs_evals"] for a, b in zip(s0, s1))
    steps = sum(b["steps"] - a["steps"] for a, b in zip(s0, s1))
    years = sum(b["t"] - a["t"] for a, b in zip(s0, s1)) / YEAR
    done = sum(1 for b in s1 if b["t"] >= t_end)
    log(f"[bench] ensemble: {rhs} RHS evals ({steps} steps) in {el:.2f}s = {rhs / el:.0f}/s; "
        f"{done}/{len(members)} members reached 20 years")
    out = {"workload": "ensemble_256", "members": len(members),
           "elements": int(sum(m["seg"].shape[0] for m in members)), "operator_groups": info["groups"],
           "max_rhs": info["max_rhs"], "matrix_bytes_per_sweep": info["bytes_per_sweep"],
           "rhs_evals_per_s": round(rhs / el, 1), "steps": steps, "elapsed_s": round(el, 3),
           "sim_years_per_hour": round(years / (el / 3600.0), 1), "members_at_20_years": done,
           "steps_per_member": {"min": int(min(b["steps"] for b in s1)), "max": int(max(b["steps"] for b in s1))},
           "step_cap_per_member": args.ensemble_steps, "dtype": "f64 state, fp32 operator"}
    e.close()
    return out


def adjoint_case(qd, args):
    """64 x 64 elements of 50 m (4,096), per-element a, b, Dc; loss = squared
    slip misfit at 32 seeded elements (+ squared first-event-time misfit when
    the window holds the onset) over a window of accepted steps; gradient by
    qdfn_sim_adjoint, checked against central differences on random
    parameters, and by the dot-product identity against the forward-mode
    (dual-number) oracle along random directions.  Windows: args.adjoint_steps
    steps through the first event (its adjoint growth reported), the onset plus
    38 steps (the first HighV solves), the onset plus 3 steps and the
    interseismic steps before the onset."""
    from qdfn import adjoint, configs
    m = configs.adjoint_fault()
    seg = m["seg"]
    op = qd.DenseOperator.from_rect_segments(seg, m["mu"], m["nu"], qd.PREC_FP64)
    KS = KN = None
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
        # forward with a checkpoint every 50 steps; the backward sweep
        # recomputes each 50-step segment from its checkpoint
        sim = qd.Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=K)
        t0 = time.perf_counter()
        sim.run_checkpointed(K, CHECKPOINT_EVERY)
        fwd_s = time.perf_counter() - t0
        dK = sim.state()["Disp"]
        target = dK[obs] * 1.1 + 1e-3
        tr = sim.trace()
        t_first = adjoint._onset(tr["told"], tr["dt"], tr["vmax"], 1e-3)
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
        dual = None
        if name in ("interseismic", "onset_plus_3") and not args.no_dual:
            # dot-product identity against the forward-mode (dual-number) oracle
            from oracle import oracle_dual as od
            if KS is None:
                KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, m["mu"], m["nu"])
            dual = []
            for sd in (1, 2):
                r3 = np.random.default_rng(sd)
                delta = {k: prob[k] * r3.standard_normal(op.n) for k in ("a", "b", "dc")}
                adj_dd = sum(float(np.dot(g[k], delta[k])) for k in delta)
                t0 = time.perf_counter()
                st_d, rd = od.sim_dense(prob, KS, KN, total_step=K, record_step=1, seed=delta)
                Dd, dDd = rd["hist_disp"][-1], rd["d_hist_disp"][-1]
                dd = float(np.sum(2 * (Dd[obs] - target) * dDd[obs]))
                if w_time > 0:
                    v, dv = rd["trace_vmax"], rd["d_trace_vmax"]
                    te, dte = rd["trace_told"] + rd["trace_dt"], rd["d_trace_told"] + rd["d_trace_dt"]
                    k = int(np.nonzero(v >= 1e-3)[0][0])
                    l0, l1, lt = np.log10(v[k - 1]), np.log10(v[k]), np.log10(1e-3)
                    dl0, dl1 = dv[k - 1] / (v[k - 1] * np.log(10)), dv[k] / (v[k] * np.log(10))
                    f = (lt - l0) / (l1 - l0)
                    df = (-dl0 * (l1 - l0) - (lt - l0) * (dl1 - dl0)) / (l1 - l0) ** 2
                    tf = te[k - 1] + f * (te[k] - te[k - 1])
                    dd += 2 * w_time * (tf - t_target) * (dte[k - 1] + df * (te[k] - te[k - 1]) +
                                                          f * (dte[k] - dte[k - 1]))
                dual.append({"adjoint": adj_dd, "dual_forward": dd, "rel_err": abs(adj_dd - dd) / abs(dd),
                             "cpu_s": round(time.perf_counter() - t0, 2)})
            log(f"[bench] adjoint ({name}): dot-product check against the dual-number oracle: rel. errors "
                + ", ".join(f"{d_['rel_err']:.2e}" for d_ in dual))
        out[name] = {"steps": K, "checkpoint_every": CHECKPOINT_EVERY, "loss": L, "time_term": w_time > 0.0,
                     "forward_ms": round(fwd_s * 1e3, 2),
                     "lambda_growth": info["lambda_growth"], "growth_step": info["growth_step"],
                     "verifiable": info["growth_step"] < 0, "dual_dot_product": dual,
                     "backward_ms": round(info["ms_backward"], 2), "backward_wall_ms": round(bwd_s * 1e3, 2),
                     "onset_step": info["onset_step"], "retried_steps": info["retried_steps"],
                     "grad_max_abs": float(max(np.abs(v).max() for v in g.values())),
                     "fd_params": n_fd, "fd_rel_err_median": float(np.median(errs)),
                     "fd_rel_err_max": float(errs.max()), "fd_one_sided_disagreement_median": float(np.median(smooth)),
                     "fd": [{k: (float(f"{v:.9g}") if isinstance(v, float) else v) for k, v in c.items()}
                            for c in chk]}
        if name == "full_window" and info["onset_step"] > 2:
            windows.append(("onset_window", min(K, int(info["onset_step"]) + 38), 16))
            # the device and the CPU oracle follow one discrete path only a few
            # rupture steps past the onset (their secant solves stop at 1e-9 and
            # the rupture amplifies ulp differences), so the dual-number check
            # runs there and on the interseismic window
            windows.append(("onset_plus_3", min(K, int(info["onset_step"]) + 3), 4))
            windows.append(("interseismic", int(info["onset_step"]) - 2, 16))
    op.close()
    return out


def cpu_baseline(h, st, prob, n_steps, threads):
    """The oracle's FP64 main_H (C restatement of HmatSolver_Pararllel and of the
    loop; products and local solves over OpenMP threads, outputs independent of
    the count) on the headline operator's stored blocks, for a bounded number
    of steps, at each thread count in `threads`."""
    from oracle import oracle as orc
    HS, HN = h.blocks(st["range_SR"], stored=True)
    out = {}
    prev = orc.set_threads(1)
    for th in threads:
        orc.set_threads(th)
        t0 = time.perf_counter()
        status, r = orc.sim_hmat(prob, HS, HN, total_step=n_steps, record_step=0, trace=False)
        el = time.perf_counter() - t0
        out[th] = dict(value=r["rhs_evals"] / el, elapsed=el, steps=n_steps, status=status)
    orc.set_threads(prev)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000, help="timed accepted steps per run (rough fault)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--graph-attempts", type=int, default=32)
    ap.add_argument("--kernel-steps", type=int, default=300, help="direct-launch steps with per-kernel events")
    ap.add_argument("--crit", type=float, default=1.5, help="DistDiamRatioCrit of the cluster tree (eta = 2)")
    ap.add_argument("--aca-tol", type=float, default=1e-6)
    ap.add_argument("--event-years", type=float, default=50.0, help="accuracy runs: first event within this window")
    ap.add_argument("--perturbations", type=int, default=4, help="rounding-level perturbed FP64 runs")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--network-steps", type=int, default=6000)
    ap.add_argument("--no-network", action="store_true")
    ap.add_argument("--network-cycle-steps", type=int, default=3500000, help="step cap of the 100-year network run")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--single-steps", type=int, default=2500000, help="step cap of the 200-year single-fault run")
    ap.add_argument("--ensemble-steps", type=int, default=200000, help="accepted-step cap per ensemble member")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-dual", action="store_true", help="skip the adjoint's dual-number oracle check")
    ap.add_argument("--no-sampled-build", action="store_true")
    ap.add_argument("--bf16-far", action="store_true", help="also build and run BF16 low-rank factors")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU layout, "
              "receiver-row shards with an all-gather of the slip per step, is a design only)", file=sys.stderr)
        sys.exit(2)

    import qdfn
    log(f"[bench] device {qdfn.device_info()}")
    rf = rough_fault(qdfn, args)
    res = rf["res"]
    mixed, plain = res["mixed_h"], res["fp64_dense"]
    hb = rf["hinfo"]["mixed_h"]["bytes_by_part"]
    t2 = mixed["coupling_ms"] * 1e-3  # phase-2 kernel, HIP events on the run's stream
    achieved = hb["phase2_entries"] / t2 / 1e9
    t1 = rf["phases"]["mixed_h"][0] * 1e-6  # phase-1 kernel (us), HIP events around it (standalone product)
    achieved1 = hb["phase1_entries"] / t1 / 1e9
    # accuracy: fixed-duration runs, independent of --steps
    ev_ops = {k: rf["ops"][k] for k in rf["ops"] if k in ("fp64_dense", "mixed_h", "fp32_h", "mixed_h_crit1")}
    acc, base = event_accuracy(qdfn, ev_ops, rf["probs"], args.event_years, args.perturbations, 200000)
    # the same operators with the patch friction (well-conditioned first event)
    pops = {k: rf["ops"][k] for k in rf["probs_patch"]}
    acc_patch, _ = event_accuracy(qdfn, pops, rf["probs_patch"], args.event_years, args.perturbations, 200000)
    # the headline run over the configuration's duration (50 years), through its events
    full = run_window(qdfn, rf["ops"]["mixed_h"], rf["probs"]["mixed_h"], args.event_years, False, 2_000_000,
                      chunk=20000)
    log(f"[bench] mixed_h over {args.event_years} years: {full['steps']} steps in {full['wall_s']:.2f}s, "
        f"{full['sim_years_per_hour']:.1f} sim-yr/h, {len(full['onsets'])} events")
    net = None if args.no_network else network(qdfn, args)
    single = None if args.no_single else single_fault(qdfn, args)
    ens = None if args.no_ensemble else ensemble(qdfn, args)
    adj = None if args.no_adjoint else adjoint_case(qdfn, args)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle FP64 main_H on the headline operator's blocks, {args.cpu_steps} steps")
        # all of the host's share of cores (16 per GPU on a shared box), and one
        nall = min(int(os.environ.get("OMP_NUM_THREADS", "0") or os.cpu_count() or 1), os.cpu_count() or 1, 16)
        cs = cpu_baseline(rf["hm"], rf["st"], rf["probs"]["mixed_h"], args.cpu_steps, sorted({1, nall}))
        c, c1 = cs[nall], cs[1]
        cpu = {"value": round(c["value"], 4), "unit": "RHS evals/s", "cores": nall, "kind": "port",
               "sample": f"{c['steps']} accepted steps of the 20,000-triangle H-matrix run, FP64 C restatement "
                         f"of main_H / HmatSolver on the headline operator's stored blocks, {nall} OpenMP threads "
                         f"({c['elapsed']:.1f}s; one thread: {c1['value']:.3f} RHS evals/s, {c1['elapsed']:.1f}s)",
               "single_thread": {"value": round(c1["value"], 4), "cores": 1}}
        log(f"[bench] CPU baseline {c['value']:.3f} RHS evals/s on {nall} threads, {c1['value']:.3f} on one")
    hi = rf["hinfo"]["mixed_h"]
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
        "dtype": "f64 state/solve/accumulate; operator storage fp32 + fp16 tail (H), plain f64",
        "data": "synthetic (seeded rough fault, back-slip loading)",
        "config": {"workload": "rough_fault_20000", "n_elements": rf["n"],
                   "operator": "H-matrix: FP32 dense blocks and leading low-rank components, FP16 trailing "
                               "components within each block's ACA budget (admissible blocks and the "
                               "near-field blocks off the diagonal compressed, diagonal blocks dense), FP64 "
                               "accumulation",
                   "leaf": 64, "admissibility_eta": 2, "dist_diam_ratio": args.crit, "aca_tol": args.aca_tol,
                   "rank_cap": 1024, "plain": "dense FP64", "steps_timed": args.steps},
        "fp64_plain": {"rhs_evals_per_s": round(plain["evals_per_s"], 2),
                       "sim_years_per_hour": round(plain["sim_years_per_hour"], 4),
                       "coupling_us": round(plain["coupling_ms"] * 1e3, 2),
                       "coupling_hbm_frac": round(rf["d64_bytes"] / (plain["coupling_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                                  4)},
        "mixed": {"rhs_evals_per_s": round(mixed["evals_per_s"], 2),
                  "sim_years_per_hour": round(mixed["sim_years_per_hour"], 4),
                  "phase2_kernel_us": round(mixed["coupling_ms"] * 1e3, 2),
                  "h_product_us": dict(zip(("phase1", "phase2", "total"), rf["phases"]["mixed_h"])),
                  "stored_entries": hi["stored_entries"], "bytes_per_product": hi["bytes_per_product"],
                  "bytes_by_part": hb},
        "speedup_mixed_vs_fp64": round(mixed["evals_per_s"] / plain["evals_per_s"], 3),
        "compared": {k: {"rhs_evals_per_s": round(res[k]["evals_per_s"], 2),
                         "bytes_per_product": rf["hinfo"][k]["bytes_per_product"] if k in rf["hinfo"] else None,
                         "h_product_us": rf["phases"].get(k)} for k in res},
        "event_time_rel_err": acc_patch["mixed_h"]["rel_err"],
        "event_time_rel_err_basis": "first onset (max V >= 1e-3 m/s) of the mixed run against the FP64 dense run on "
                                    "the headline mesh and operators with the single fault's friction on a central "
                                    "3 x 3 km velocity-weakening patch (event_time_patch, with the spread of FP64 "
                                    "runs perturbed at 1e-15); the configuration's element-random friction gives "
                                    "an ill-conditioned first event (event_time: its FP64 runs perturbed at 1e-15 "
                                    "scatter by the spread shown, runs without an event included)",
        "peak_v_rel_err": acc_patch["mixed_h"]["peak_v_rel_err"],
        "event_time_patch": acc_patch,
        "event_time": acc,
        "mixed_over_duration": {"years": full["sim_seconds"] / YEAR, "steps": full["steps"],
                                "wall_s": round(full["wall_s"], 3),
                                "rhs_evals_per_s": round(full["rhs_evals_per_s"], 1),
                                "sim_years_per_hour": round(full["sim_years_per_hour"], 2),
                                "events": len(full["onsets"]), "retry_passes": full["retries"]},
        "fp64_to_first_event": {"years": base["sim_seconds"] / YEAR, "steps": base["steps"],
                                "wall_s": round(base["wall_s"], 3), "rhs_evals_per_s": round(base["rhs_evals_per_s"], 1),
                                "sim_years_per_hour": round(base["sim_years_per_hour"], 2)},
        "operator_rel_err": rf["acc"],
        "h_build": rf["builds"],
        "network_8192": net,
        "single_fault_512": single,
        "ensemble_256": ens,
        "adjoint_4096": adj,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "kernel": "k_hunit<float, fp16_t, false> (H product phase 2)",
                     "bytes": "stored panel entries only", "bytes_per_launch": hb["phase2_entries"],
                     "index_and_gather_bytes": hb["phase2_index"] + hb["phase2_gather"],
                     "phase1": {"kernel": "k_vpanel_mixed<false> (H product phase 1)", "us": round(t1 * 1e6, 2),
                                "achieved": round(achieved1, 1), "frac": round(achieved1 / HBM_PEAK_GBS, 4),
                                "bytes_per_launch": hb["phase1_entries"], "bytes": "stored V entries only",
                                "traffic": measured_traffic("traffic_phase1_bytes_per_launch")}},
        "cpu_baseline": cpu,
    }
    for o in rf["ops"].values():
        o.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
