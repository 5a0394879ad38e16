"""GPU parity of the device-resident cycle loop (main(),
scripts/Functions_RSFDFN3DMain_D.jl:2-324) against the CPU oracle's main().

Single-fault workload (512 elements, back-slip loading).  FP64 storage must
reproduce the oracle's step sequence: the same accepted-step count, step
sizes and times to 1e-9 relative up to and through the first event, and the
first-event onset time (max V >= 1e-3 m/s) within 1e-9 of the oracle's.  The
FP32-storage run is held to the stated event-time tolerance instead.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 1200  # the first event starts near step 520


@pytest.fixture(scope="module")
def fault(orc):
    from qdfn import configs
    c = configs.single_fault()
    seg = c["seg"]
    KS, KN = orc.stiffness_rect(seg, seg, c["mu"], c["nu"])
    E, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(c, ks, KS.sum(1), KN.sum(1))
    st, ref = orc.sim_dense(prob, KS, KN, total_step=STEPS, record_step=10)
    assert st == 0
    return dict(c=c, KS=KS, KN=KN, prob=prob, ref=ref)


def _gpu_run(qd, fault, prec, timed=False, steps=STEPS, chunks=1):
    op = qd.DenseOperator.from_matrices(fault["KS"], fault["KN"], prec)
    sim = qd.Simulation(op, fault["prob"], record_step=10, record_capacity=steps // 10, trace_capacity=steps)
    per = steps // chunks
    for _ in range(chunks):
        sim.run(per, timed=timed)
    out = dict(stats=sim.stats(), trace=sim.trace(), hist=sim.history(), state=sim.state())
    sim.close()
    op.close()
    return out


def test_fp64_run_tracks_oracle(qd, fault):
    from qdfn import configs
    g = _gpu_run(qd, fault, qd.PREC_FP64)
    ref = fault["ref"]
    assert g["stats"]["status"] == 0
    assert g["stats"]["steps"] == STEPS
    tr = g["trace"]
    k = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
    kg = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
    assert k is not None and kg is not None
    assert abs(kg - k) <= 1e-9 * k
    # step-by-step agreement up to the first event
    first = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
    np.testing.assert_allclose(tr["told"][:first], ref["trace_told"][:first], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"][:first], ref["trace_dt"][:first], rtol=1e-7)
    np.testing.assert_allclose(tr["vmax"][:first], ref["trace_vmax"][:first], rtol=1e-6)
    assert g["stats"]["rhs_evals"] == ref["rhs_evals"] or first < STEPS
    # histories: the recorded rows before the event
    h = g["hist"]
    nrec = first // 10
    np.testing.assert_allclose(h["History_Time"][:nrec], ref["hist_time"][:nrec], rtol=1e-9)
    np.testing.assert_allclose(h["History_V"][:nrec], ref["hist_v"][:nrec], rtol=1e-6)
    np.testing.assert_allclose(h["History_Disp"][:nrec], ref["hist_disp"][:nrec], rtol=1e-6, atol=1e-12)


def test_graph_and_timed_modes_agree_bitwise(qd, fault):
    a = _gpu_run(qd, fault, qd.PREC_FP64, timed=False, steps=600, chunks=3)
    b = _gpu_run(qd, fault, qd.PREC_FP64, timed=True, steps=600, chunks=1)
    for k in ("told", "dt", "vmax"):
        np.testing.assert_array_equal(a["trace"][k], b["trace"][k])
    for k in ("V", "Disp", "Theta"):
        np.testing.assert_array_equal(a["state"][k], b["state"][k])
    # every accepted step ran exactly one coupling product
    assert b["stats"]["coupling_timed"] == b["stats"]["rhs_evals"]
    assert b["stats"]["coupling_ms_avg"] > 0


def test_fp32_storage_first_event(qd, fault):
    from qdfn import configs
    g = _gpu_run(qd, fault, qd.PREC_FP32)
    ref = fault["ref"]
    k = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
    kg = configs.first_event_time(g["trace"]["told"], g["trace"]["dt"], g["trace"]["vmax"])
    assert kg is not None
    # event-time tolerance: 1e-6 of the time to the first event
    assert abs(kg - k) <= 1e-6 * k, (kg, k)
