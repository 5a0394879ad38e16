"""Run the earthquake-cycle loop alone on the 20,000-triangle rough fault (or
the 8-fault, 8,192-element rectangular network), for kernel traces of a whole
attempt (product phases, local solves, decision).

    python -m qdfn.profile_loop [--steps 1000] [--warmup 100] [--aca-tol 1e-6]
                                [--prec fp32|mixed|mixed8|bf16|fp64] [--crit C]
                                [--repeat 1] [--workload rough_fault|network]

Builds the H operator (storage --prec: FP32; FP32 with FP16 trailing low-rank
components; BF16 low-rank factors; FP64 — leaves of 64, DistDiamRatioCrit
--crit, 1.5 for the rough fault and 1.0 for the network unless given) and the
run's arrays the way the benchmark does, replays the captured
attempt graph for `steps` accepted steps and prints the RHS evaluations per
second (with --repeat R: R such runs one after another, each timed, and the
rates of every run).
"""
import argparse
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--aca-tol", type=float, default=1e-6)
    ap.add_argument("--prec", choices=("fp32", "mixed", "mixed8", "bf16", "fp64"), default="fp32")
    ap.add_argument("--crit", type=float, default=None)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--workload", choices=("rough_fault", "network"), default="rough_fault")
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    prec = {"fp32": qdfn.PREC_FP32, "mixed": qdfn.PREC_MIXED, "mixed8": qdfn.PREC_MIXED8,
            "bf16": qdfn.PREC_BF16_FAR, "fp64": qdfn.PREC_FP64}[a.prec]
    if a.workload == "network":
        a.crit = 1.0 if a.crit is None else a.crit
        cfg = configs.fault_network()
        seg = cfg["seg"]
        n = seg.shape[0]
        st = qdfn.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=64, max_levels=40,
                                 dist_diam_ratio=a.crit, arrange_point=(0.0, 0.0, 0.0))
        segs = np.ascontiguousarray(seg[st["perm"]])
        h = qdfn.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], st["range_SR"], st["admissible"],
                                    aca_rtol=a.aca_tol, prec=prec)
        ks, _ = h.kself()
        hs, hn = h.matvec(np.ones(n), loading_form=False)
        prob = configs.problem_from_segments(dict(cfg, seg=segs), ks, hs, hn)
        return run(a, h, prob, n)
    a.crit = 1.5 if a.crit is None else a.crit
    m = configs.rough_fault()
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qdfn.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    m = dict(m, P1=P1, P2=P2, P3=P3)
    ctr = (P1 + P2 + P3) / 3
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    st = qdfn.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=a.crit,
                             arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [x[p] for x in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    h = qdfn.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=a.aca_tol,
                               prec=prec)
    ks, _ = h.kself()
    hs, hn = h.matvec(np.ones(n), loading_form=False)
    prob = configs.problem_from_triangles(m, ks, hs, hn, order=p)
    prob["ext_at_told"] = 1
    return run(a, h, prob, n)


def run(a, h, prob, n):
    import qdfn
    sim = qdfn.Simulation(h, prob, record_step=0, record_capacity=0,
                          trace_capacity=a.steps * a.repeat + a.warmup)
    sim.run(a.warmup)
    rates = []
    for _ in range(a.repeat):
        s0 = sim.stats()
        t0 = time.perf_counter()
        sim.run(a.steps)
        el = time.perf_counter() - t0
        s1 = sim.stats()
        rates.append((s1["rhs_evals"] - s0["rhs_evals"]) / el)
    if a.repeat > 1:
        print("rates of the runs (RHS evals/s): " + " ".join(f"{r:.0f}" for r in rates), flush=True)
    sim.close()
    info = h.info()
    ph = h.time_phases(20)
    h.close()
    print(f"loop {a.workload} n={n} prec {a.prec} crit {a.crit}: {info['bytes_per_product'] / 1e6:.1f} MB per product, "
          f"{s1['rhs_evals'] - s0['rhs_evals']} RHS evals in {el:.3f} s = "
          f"{(s1['rhs_evals'] - s0['rhs_evals']) / el:.1f}/s, status {s1['status']}; product phases (ms): "
          f"phase 1 {ph[0]:.4f}, phase 2 {ph[1]:.4f}, total {ph[2]:.4f}", flush=True)
    return 0 if s1["status"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
