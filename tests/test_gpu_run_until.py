"""GPU: qdfn_sim_run_until — main()'s loop run to a simulated time with the
reference's periodic save of History_* (Functions_RSFDFN3DMain_D.jl:265-297):
records streamed to a host callback through a small device staging buffer
equal, bitwise, the histories of one run that keeps every record on the
device; the run ends with the step that crosses t_end; the callback can stop
the run; the record times follow the oracle's main()."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fault(qd, orc):
    from qdfn import configs
    cfg = configs.single_fault()
    seg = cfg["seg"]
    KS, KN = orc.stiffness_rect(seg, seg, cfg["mu"], cfg["nu"])
    _, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(cfg, ks, KS.sum(1), KN.sum(1))
    op = qd.DenseOperator.from_matrices(KS, KN)
    yield dict(op=op, prob=prob, KS=KS, KN=KN)
    op.close()


def test_streamed_records_equal_one_run(fault, qd, orc):
    from qdfn import configs
    t_end = 1.5 * configs.SEC_PER_YEAR
    got = []
    sim = qd.Simulation(fault["op"], fault["prob"], record_step=10, record_capacity=6, trace_capacity=0)
    st = sim.run_until(t_end, 10 ** 6, on_records=lambda first, h: got.append((first, h)) and False, graph_attempts=32)
    sim.close()
    assert st["status"] == 0 and st["t"] >= t_end and st["t"] - st["dt"] < t_end
    steps = st["steps"]
    firsts = [f for f, _ in got]
    counts = [h["History_Time"].size for _, h in got]
    assert firsts == list(np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert sum(counts) == steps // 10 and len(got) > 10
    cat = {k: np.concatenate([h[k] for _, h in got]) for k in got[0][1]}
    # the same run with every record on the device
    ref = qd.Simulation(fault["op"], fault["prob"], record_step=10, record_capacity=steps // 10 + 1)
    ref.run(steps)
    h = ref.history()
    ref.close()
    for k, v in cat.items():
        np.testing.assert_array_equal(v, h[k], err_msg=k)
    # against the oracle's main() over the same steps (record times = TOld of the recorded step)
    status, o = orc.sim_dense(fault["prob"], fault["KS"], fault["KN"], total_step=steps, record_step=10)
    assert status == 0
    first_event = int(np.nonzero(o["trace_vmax"] >= 1e-3)[0][0]) // 10 if (o["trace_vmax"] >= 1e-3).any() else None
    m = first_event if first_event else cat["History_Time"].size
    np.testing.assert_allclose(cat["History_Time"][:m], o["hist_time"][:m], rtol=1e-9)


def test_callback_stops_the_run(fault, qd):
    from qdfn import configs
    seen = []

    def stop_after_two(first, h):
        seen.append(first)
        return len(seen) >= 2

    sim = qd.Simulation(fault["op"], fault["prob"], record_step=5, record_capacity=8)
    st = sim.run_until(100 * configs.SEC_PER_YEAR, 10 ** 6, on_records=stop_after_two, graph_attempts=32)
    sim.close()
    assert len(seen) == 2 and st["status"] == 0
    assert st["t"] < 100 * configs.SEC_PER_YEAR and st["steps"] < 200


def test_staging_buffer_too_small_is_refused(fault, qd):
    sim = qd.Simulation(fault["op"], fault["prob"], record_step=1, record_capacity=8)
    with pytest.raises(qd.QdfnError) as e:
        sim.run_until(1e7, 1000, on_records=lambda f, h: False, graph_attempts=32)
    assert e.value.code == -1 and "record_capacity" in str(e.value)
    sim.close()


def test_periodic_result_file(fault, qd, tmp_path):
    """The reference's save(SaveResultFileName, "History_V", ...) every SaveStep:
    the streamed records appended to host histories and written as Result.jld2
    at each batch; the last file holds every record of the run, bitwise."""
    from qdfn import jld2
    path = tmp_path / "Result.jld2"
    acc = {}

    def on_records(first, h):
        for k, v in h.items():
            acc[k] = v if k not in acc else np.concatenate([acc[k], v])
        jld2.save_result(path, acc)
        return False

    sim = qd.Simulation(fault["op"], fault["prob"], record_step=10, record_capacity=8, trace_capacity=0)
    st = sim.run_until(0.5 * 3.15576e7, 10 ** 6, on_records=on_records, graph_attempts=32)
    sim.close()
    V, D, T = jld2.load(path, "History_V", "History_Disp", "History_Time")
    assert T.shape == (st["steps"] // 10, 1) and V.shape == (st["steps"] // 10, fault["op"].n)
    np.testing.assert_array_equal(V, acc["History_V"])
    np.testing.assert_array_equal(D, acc["History_Disp"])
    np.testing.assert_array_equal(T[:, 0], acc["History_Time"])
