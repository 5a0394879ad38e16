"""Sampled H-matrix construction on a rough fault too large for the full build,
then the product and a short run of the loop on the result.

    python -m qdfn.profile_sampled [--cells 316] [--rtol 1e-4] [--rows 48] [--steps 300]

A `cells` x `cells` rough fault (two triangles per 50 m cell) is clustered
(leaves of 64, DistDiamRatioCrit 1.0) and its H operator is built from sampled
rows and columns (qdfn_hmat_build_tri_sampled, FP32 storage).  The product is
checked on `rows` random receivers against their rows of the stiffness
assembled in full (qdfn_tde_stiffness_tri on those receivers only), and the
loop runs `steps` accepted steps.  Set QDFN_HMAT_VERBOSE for build progress.
"""
import argparse
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=316)
    ap.add_argument("--rtol", type=float, default=1e-4)
    ap.add_argument("--rows", type=int, default=48)
    ap.add_argument("--steps", type=int, default=300)
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    m = configs.rough_fault(nx=a.cells, nz=a.cells)
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qdfn.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    m = dict(m, P1=P1, P2=P2, P3=P3)
    ctr = (P1 + P2 + P3) / 3
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    print(f"mesh n={n}: full K_S and K_N in FP64 would take {2 * 8 * n * n / 1e9:.0f} GB", flush=True)
    st = qdfn.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=60, dist_diam_ratio=1.0,
                             arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [x[p] for x in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    print(f"structure: {len(st['admissible'])} blocks, {int(st['admissible'].sum())} admissible", flush=True)
    t0 = time.perf_counter()
    h = qdfn.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=a.rtol,
                               prec=qdfn.PREC_FP32, sampled=True)
    print(f"sampled build: {time.perf_counter() - t0:.1f} s wall, {h.build_info()}", flush=True)
    info = h.info()
    t1, t2, tot = h.time_phases(5)
    print(f"stored {info['stored_entries'] / 1e9:.3f} G entries, {info['bytes_per_product'] / 1e9:.2f} GB per "
          f"product; product {tot * 1e3:.0f} us (phase 1 {t1 * 1e3:.0f} us, phase 2 {t2 * 1e3:.0f} us)", flush=True)
    rng = np.random.default_rng(4)
    rows = np.sort(rng.choice(n, a.rows, replace=False))
    x = rng.standard_normal(n)
    yS, yN = h.matvec(x, loading_form=False)
    KSr, KNr = qdfn.StiffnessMatrix_ByParts_Calculation_Tri(geo[0], geo[1], geo[2], geo[3], geo[4][rows], m["mu"],
                                                           lam, geo[5][rows], geo[6][rows])
    rs, rn = KSr @ x, KNr @ x
    scale = a.rtol * np.linalg.norm(KSr) * np.linalg.norm(x)
    print(f"rows: |y_S - K_S x| max {np.abs(yS[rows] - rs).max():.3e} (max |K_S x| {np.abs(rs).max():.3e}), "
          f"|y_N - K_N x| max {np.abs(yN[rows] - rn).max():.3e} (max |K_N x| {np.abs(rn).max():.3e}); "
          f"rtol ||K_S rows||_F ||x|| = {scale:.3e}", flush=True)
    ks, _ = h.kself()
    hs, hn = h.matvec(np.ones(n), loading_form=False)
    prob = configs.problem_from_triangles(m, ks, hs, hn, order=p)
    prob["ext_at_told"] = 1
    sim = qdfn.Simulation(h, prob, record_step=0, record_capacity=0, trace_capacity=a.steps + 100)
    sim.run(50)
    s0 = sim.stats()
    t0 = time.perf_counter()
    sim.run(a.steps)
    el = time.perf_counter() - t0
    s1 = sim.stats()
    sim.close()
    h.close()
    ev = s1["rhs_evals"] - s0["rhs_evals"]
    print(f"loop: {ev} RHS evals in {el:.2f} s = {ev / el:.1f}/s, status {s1['status']}", flush=True)
    return 0 if s1["status"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
