"""Run the batched ensemble alone, for kernel traces of its attempts.

    python -m qdfn.profile_ensemble [--steps 2000] [--members 256] [--max-rhs 0]

Builds the 256 planar faults of 500-4,000 elements (FP32 operators, one per
distinct geometry, members of one geometry read as one multi-right-hand-side
product), advances every member `steps` accepted steps together
(qdfn_ensemble_run, attempts replayed from a captured HIP graph) and prints
the aggregate RHS evaluations per second and the matrix bytes of one sweep.
"""
import argparse
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--max-rhs", type=int, default=0)
    ap.add_argument("--split-mb", type=float, default=None, help="MB of matrix per coupling workgroup "
                    "(default: qdfn.ensemble.SPLIT_MB; 0: each operator's single-run split count)")
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    from qdfn.ensemble import problem_for_member
    members = configs.ensemble_members(m=a.members)
    from qdfn.ensemble import SPLIT_MB
    split_mb = SPLIT_MB if a.split_mb is None else a.split_mb
    e = qdfn.Ensemble.from_members(members, problem_for_member, max_rhs=a.max_rhs, record_step=0,
                                   split_mb=split_mb or None)
    info = e.info()
    s0 = e.stats()
    t0 = time.perf_counter()
    e.run(a.steps, graph_attempts=16)
    el = time.perf_counter() - t0
    s1 = e.stats()
    rhs = sum(b["rhs_evals"] - x["rhs_evals"] for x, b in zip(s0, s1))
    print(f"ensemble: {len(members)} members, {sum(m['seg'].shape[0] for m in members)} elements, "
          f"{info['groups']} operator groups, max_rhs {info['max_rhs']}, {info['bytes_per_sweep'] / 1e6:.1f} MB "
          f"per sweep; {rhs} RHS evals in {el:.3f} s = {rhs / el:.0f}/s", flush=True)
    e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
