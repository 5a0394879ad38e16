"""GPU parity of the loop's external forcing (InterpolateFromStressChange,
Functions_Solvers.jl:465-540, read at TOld + Dt by main() and at TOld by
main_H) and of loading faults (the last rows of SolveOneTimeStep,
Functions_Solvers.jl:370-377: V = V_Const, Disp += V_Const Dt, theta and
friction unchanged, EffNormalStress = 1e10).

* Multi-sample table: the 512-element single fault loaded by a six-sample
  table (uneven shear ramp plus a normal-stress history) whose run covers
  times before the first sample, between samples and after the last; the
  device loop tracks the oracle's step by step up to the first event, for
  both timings.
* Loading faults: the reference's own example geometry
  (tests/golden/Input_BulkFaultGeometry.txt: 4,005 elements, the last 5 of
  them loading faults at V_Const = 1e-9 m/s, no external table), stiffness
  assembled on the device, 2,000 steps through the initial rupture (one
  element starts at 0.03 m/s), its end and the first interseismic steps; the
  device loop tracks the oracle's and the loading rows hold the reference's
  values exactly.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _table(KS, KN, rate=1e-9):
    t = np.array([1e5, 1e6, 3e6, 8e6, 2e7, 3e7])
    w = np.array([1.0, 1.3, 0.9, 1.1, 1.05, 1.2])
    dp = -1e5 * np.array([0.0, 1.0, 3.0, 2.0, 5.0, 4.0])
    return {"t": t, "s": -np.outer(t * w * rate, KS.sum(1)), "n": -np.outer(t * rate, KN.sum(1)) + dp[:, None]}


@pytest.mark.parametrize("ext_at_told", [0, 1])
def test_multisample_table_loop(orc, qd, ext_at_told):
    from qdfn import configs
    cfg = configs.single_fault()
    seg = cfg["seg"]
    KS, KN = orc.stiffness_rect(seg, seg, cfg["mu"], cfg["nu"])
    _, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(cfg, ks, KS.sum(1), KN.sum(1))
    prob.update(ext=_table(KS, KN), ext_at_told=ext_at_told)
    steps = 1500
    st, ref = orc.sim_dense(prob, KS, KN, total_step=steps, record_step=0)
    assert st == 0
    op = qd.DenseOperator.from_matrices(KS, KN)
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=steps)
    sim.run(steps)
    tr = sim.trace()
    sim.close()
    op.close()
    k = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
    told = ref["trace_told"][:k]
    assert (told < 1e5).sum() > 20 and ((told > 1e5) & (told < 3e7)).sum() > 20 and (told > 3e7).sum() > 20
    np.testing.assert_allclose(tr["told"][:k], told, rtol=1e-9)
    np.testing.assert_allclose(tr["dt"][:k], ref["trace_dt"][:k], rtol=1e-7)
    np.testing.assert_allclose(tr["vmax"][:k], ref["trace_vmax"][:k], rtol=1e-6)
    e_ref = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
    e_gpu = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
    assert abs(e_gpu - e_ref) <= 1e-9 * e_ref


def test_unsorted_table_is_refused(qd):
    from qdfn import configs
    cfg = configs.single_fault(nx=8, nz=4)
    op = qd.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"])
    ks, _ = op.kself()
    prob = configs.problem_from_segments(cfg, ks, np.zeros(op.n), np.zeros(op.n))
    prob["ext"] = {"t": np.array([0.0, 2.0, 1.0]), "n": np.zeros((3, op.n)), "s": np.zeros((3, op.n))}
    with pytest.raises(qd.QdfnError) as e:
        qd.Simulation(op, prob)
    assert e.value.code == -1 and "non-decreasing" in str(e.value)
    op.close()


def test_loading_faults_loop(orc, qd):
    from qdfn import configs, inputs
    seg, L, mu, nu, rho, _, _, _, minns, _, n = inputs.ReadBulkInput(os.path.join(GOLD, "Input_BulkFaultGeometry.txt"))
    assert L == 5
    KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, mu, nu)
    _, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_bulk(seg, L, mu, rho, minns, ks)
    steps = 2000
    st, ref = orc.sim_dense(prob, KS, KN, total_step=steps, record_step=10)
    assert st == 0
    op = qd.DenseOperator.from_matrices(KS, KN)
    sim = qd.Simulation(op, prob, record_step=10, record_capacity=steps // 10, trace_capacity=steps)
    sim.run(steps)
    tr, stats, fin = sim.trace(), sim.stats(), sim.state()
    hist = sim.history()
    sim.close()
    op.close()
    assert stats["retry_passes"] == ref["retry_passes"] and stats["stalls"] == ref["stalls"]
    slow = int(np.nonzero(ref["trace_vmax"] < 1e-2)[0][0])
    assert 100 < slow < steps - 100  # the run leaves the rupture inside the window
    dev = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    print(f"[loading] told {dev(tr['told'][1:], ref['trace_told'][1:]):.2e} dt {dev(tr['dt'], ref['trace_dt']):.2e} "
          f"vmax {dev(tr['vmax'], ref['trace_vmax']):.2e}")
    # measured on MI355X: 3.0e-11 / 3.1e-10 / 2.9e-10
    np.testing.assert_allclose(tr["told"], ref["trace_told"], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"], ref["trace_dt"], rtol=1e-8)
    np.testing.assert_allclose(tr["vmax"], ref["trace_vmax"], rtol=1e-8)
    # loading rows: V = V_Const, Disp = V_Const * t (sum of accepted Dt), theta and friction unchanged, sigma 1e10
    t_end = stats["t"]
    vc = prob["loading_rate"][-L:]
    np.testing.assert_array_equal(fin["V"][-L:], vc)
    np.testing.assert_allclose(fin["Disp"][-L:], vc * t_end, rtol=1e-12)
    np.testing.assert_array_equal(fin["Theta"][-L:], prob["theta_i"][-L:])
    np.testing.assert_array_equal(fin["Friction"][-L:], prob["friction_i"][-L:])
    np.testing.assert_array_equal(fin["Sigma"][-L:], 1e10)
    np.testing.assert_allclose(hist["History_Disp"][:, -L:], ref["hist_disp"][:, -L:], rtol=1e-9)
