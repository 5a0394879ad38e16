"""GPU: members of the batched-ensemble workload (256 planar faults of
500-4,000 elements, seeded friction) advanced together by qdfn_ensemble_run
track the oracle's main() (scripts/Functions_RSFDFN3DMain_D.jl:2-324, one run
per input as QUAKEDFN_RunARG.jl:247-262 starts it).  The members are the
smallest, a middle-sized and the largest of the workload's 256; the oracle
runs each on the member's operator as the ensemble stores it (E and K_N
rounded to FP32), so both integrate the same equations: step times to 1e-9,
step sizes to 1e-7 and max V to 1e-6, up to the first event or the
predictability horizon of the member, whichever comes first.  The members'
element-scale random friction makes their creep unstable: rounding-level
differences (GPU and CPU sum the coupling in different orders) grow, so the
horizon is measured on the oracle itself, as the first step at which an
oracle run with the initial state perturbed by a relative 1e-15 departs from
the unperturbed one by 1e-10 in simulated time; it must span at least 50
steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
STEPS = 300


def test_ensemble_members_track_oracle(qd, orc):
    from qdfn import configs
    from qdfn.ensemble import problem_for_member
    allm = configs.ensemble_members()
    sizes = np.array([m["seg"].shape[0] for m in allm])
    order = np.argsort(sizes, kind="stable")
    pick = [allm[order[0]], allm[order[len(order) // 2]], allm[order[-1]]]
    assert sizes[order[0]] < 700 and sizes[order[-1]] > 3800
    e = qd.Ensemble.from_members(pick, problem_for_member, max_rhs=0, trace_capacity=STEPS)
    try:
        probs = [problem_for_member(m, s.op) for m, s in zip(pick, e.sims)]
        e.run(STEPS)
        for m, s, prob in zip(pick, e.sims, probs):
            tr, st = s.trace(), s.stats()
            assert st["status"] == 0 and st["steps"] == STEPS
            seg = m["seg"]
            KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, m["mu"], m["nu"])
            d = np.diag(KS).copy()
            E32 = (KS / d[:, None]).astype(np.float32).astype(np.float64)
            np.fill_diagonal(E32, 1.0)
            KS32 = E32 * d[:, None]  # E_ij = KS32_ij / KS32_ii = the stored FP32 E (to FP64 rounding)
            KN32 = KN.astype(np.float32).astype(np.float64)
            status, ref = orc.sim_dense(prob, KS32, KN32, total_step=STEPS, record_step=0)
            assert status == 0
            pp = dict(prob)
            pp["theta_i"] = prob["theta_i"] * (1 + 1e-15 * np.random.default_rng(2).standard_normal(prob["n"]))
            status, per = orc.sim_dense(pp, KS32, KN32, total_step=STEPS, record_step=0)
            assert status == 0
            dev = np.abs(per["trace_told"] - ref["trace_told"]) > 1e-10 * np.maximum(ref["trace_told"], 1.0)
            horizon = int(np.argmax(dev)) if dev.any() else STEPS
            on = np.nonzero(ref["trace_vmax"] >= 1e-3)[0]
            k = min(int(on[0]) if on.size else STEPS, horizon)
            assert k > 50, (seg.shape[0], horizon)
            np.testing.assert_allclose(tr["told"][:k], ref["trace_told"][:k], rtol=1e-9)
            np.testing.assert_allclose(tr["dt"][:k], ref["trace_dt"][:k], rtol=1e-7)
            np.testing.assert_allclose(tr["vmax"][:k], ref["trace_vmax"][:k], rtol=1e-6)
            print(f"[ensemble] member of {seg.shape[0]} elements: {k} steps compared (horizon {horizon}), "
                  f"t = {tr['told'][k - 1]:.6e} s")
    finally:
        e.close()
