"""Run the earthquake-cycle loop alone on the 20,000-triangle rough fault, for
kernel traces of a whole attempt (product phases, local solves, decision).

    python -m qdfn.profile_loop [--steps 1000] [--warmup 100] [--aca-tol 1e-6] [--bf16-far]

Builds the H operator (FP32 storage, or BF16 low-rank factors with --bf16-far;
leaves of 64, DistDiamRatioCrit 1.0) and the run's arrays the way the
benchmark does, replays the captured attempt graph for `steps` accepted steps
and prints the RHS evaluations per second.
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
    ap.add_argument("--bf16-far", action="store_true")
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    m = configs.rough_fault()
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qdfn.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    m = dict(m, P1=P1, P2=P2, P3=P3)
    ctr = (P1 + P2 + P3) / 3
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    st = qdfn.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                             arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [x[p] for x in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    h = qdfn.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=a.aca_tol,
                               prec=qdfn.PREC_BF16_FAR if a.bf16_far else qdfn.PREC_FP32)
    ks, _ = h.kself()
    hs, hn = h.matvec(np.ones(n), loading_form=False)
    prob = configs.problem_from_triangles(m, ks, hs, hn, order=p)
    prob["ext_at_told"] = 1
    sim = qdfn.Simulation(h, prob, record_step=0, record_capacity=0, trace_capacity=a.steps + a.warmup)
    sim.run(a.warmup)
    s0 = sim.stats()
    t0 = time.perf_counter()
    sim.run(a.steps)
    el = time.perf_counter() - t0
    s1 = sim.stats()
    sim.close()
    h.close()
    print(f"loop n={n}: {s1['rhs_evals'] - s0['rhs_evals']} RHS evals in {el:.3f} s = "
          f"{(s1['rhs_evals'] - s0['rhs_evals']) / el:.1f}/s, status {s1['status']}", flush=True)
    return 0 if s1["status"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
