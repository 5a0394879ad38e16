"""GPU: the 8-fault, 8,192-element dense network (config 1), main() of
scripts/Functions_RSFDFN3DMain_D.jl:2-324 at its full size.

* FP64 storage: the device loop tracks the oracle's main() for the first 150
  steps (told 1e-9, Dt 1e-7, max V 1e-6).
* The FP64 run's first event is well conditioned: perturbing the initial
  state by a relative 1e-15 moves it by less than 1e-8 of its time.
* Stated tolerance of this configuration's mixed form (FP32 storage, FP64
  state): its first event within 5e-5 of the FP64 run's (measured on MI355X:
  2.85e-5).  FP32 entries err by 2^-24, which the loop turns into this shift;
  a 1e-6 bar would need more than FP32 storage.
* The same bar for a mixed-storage H operator of the network (FP32 + FP16
  tails, ACA 1e-6, DistDiamRatioCrit 1.0).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(qd):
    from qdfn import configs
    cfg = configs.fault_network()
    seg = cfg["seg"]
    KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, cfg["mu"], cfg["nu"])
    ops = {p: qd.DenseOperator.from_matrices(KS, KN, p) for p in (qd.PREC_FP64, qd.PREC_FP32)}
    ks, _ = ops[qd.PREC_FP64].kself()
    prob = configs.problem_from_segments(cfg, ks, KS.sum(1), KN.sum(1))
    yield dict(KS=KS, KN=KN, ops=ops, prob=prob)
    for o in ops.values():
        o.close()


def _onset(qd, op, prob, steps=6000):
    from qdfn import configs
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=steps)
    sim.run_until(2 * configs.SEC_PER_YEAR, steps)
    tr = sim.trace()
    sim.close()
    return configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])


def test_network_fp64_tracks_oracle(net, qd, orc):
    steps = 150
    status, ref = orc.sim_dense(net["prob"], net["KS"], net["KN"], total_step=steps, record_step=0)
    assert status == 0
    sim = qd.Simulation(net["ops"][qd.PREC_FP64], net["prob"], record_step=0, trace_capacity=steps)
    sim.run(steps)
    tr = sim.trace()
    sim.close()
    np.testing.assert_allclose(tr["told"], ref["trace_told"], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"], ref["trace_dt"], rtol=1e-7)
    np.testing.assert_allclose(tr["vmax"], ref["trace_vmax"], rtol=1e-6)


def test_network_event_time_tolerance(net, qd):
    t64 = _onset(qd, net["ops"][qd.PREC_FP64], net["prob"])
    t32 = _onset(qd, net["ops"][qd.PREC_FP32], net["prob"])
    p = dict(net["prob"])
    p["theta_i"] = p["theta_i"] * (1 + 1e-15 * np.random.default_rng(1).standard_normal(p["theta_i"].size))
    tp = _onset(qd, net["ops"][qd.PREC_FP64], p)
    assert t64 is not None and t32 is not None and tp is not None
    print(f"[network] first event FP64 {t64:.9e} s, perturbed {tp:.9e} s, FP32 {t32:.9e} s: "
          f"FP32 rel. {abs(t32 - t64) / t64:.2e}, perturbation rel. {abs(tp - t64) / t64:.2e}")
    assert abs(tp - t64) <= 1e-8 * t64
    assert abs(t32 - t64) <= 5e-5 * t64


def test_network_mixed_h_event_time(net, qd):
    """The network on a mixed-storage H operator (DistDiamRatioCrit 1.0, the
    reference default; leaves of 64; ACA 1e-6): its first event within the
    same 5e-5 of the FP64 dense run's (measured on MI355X: 2.4e-5), at about a
    tenth of the FP32 dense operator's bytes per product."""
    from qdfn import configs
    cfg = configs.fault_network()
    seg = cfg["seg"]
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    segs = np.ascontiguousarray(seg[p])
    H = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], st["range_SR"], st["admissible"], aca_rtol=1e-6,
                              prec=qd.PREC_MIXED)
    try:
        assert H.info()["bytes_per_product"] < 0.15 * 2 * 4 * seg.shape[0] ** 2
        ks, _ = net["ops"][qd.PREC_FP64].kself()
        # the same problem in the cluster tree's element order
        prob = configs.problem_from_segments(dict(cfg, seg=segs), ks[p], net["KS"].sum(1)[p], net["KN"].sum(1)[p])
        th = _onset(qd, H, prob)
    finally:
        H.close()
    t64 = _onset(qd, net["ops"][qd.PREC_FP64], net["prob"])
    assert th is not None and t64 is not None
    print(f"[network] first event FP64 dense {t64:.9e} s, mixed H {th:.9e} s: rel. {abs(th - t64) / t64:.2e}")
    assert abs(th - t64) <= 5e-5 * t64
