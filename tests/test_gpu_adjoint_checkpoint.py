"""GPU: checkpointed recomputation of the adjoint's tape
(qdfn_sim_run_checkpointed + qdfn_sim_adjoint) on the 4,096-element adjoint
workload (64 x 64 elements of 50 m, per-element a, b, Dc): the forward keeps
the run state every `every` steps, the backward sweep replays each segment
from its checkpoint just before sweeping it.  The replay retraces the
forward's steps bit for bit, so loss, gradients and the growth indicator
equal those of the full recorded tape exactly: over the workload's 2,000-step
window through the first event with the event-time term (checkpoints every
50 steps), and for segment lengths that do not divide the window."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(qd):
    from qdfn import configs
    m = configs.adjoint_fault()
    op = qd.DenseOperator.from_rect_segments(m["seg"], m["mu"], m["nu"], qd.PREC_FP64)
    ks, _ = op.kself()
    ksum, knsum = op.matvec(np.ones(op.n), loading_form=False)
    prob = configs.problem_from_segments(m, ks, ksum, knsum)
    obs = np.sort(np.random.default_rng(41).choice(op.n, 32, replace=False))
    yield dict(op=op, prob=prob, obs=obs)
    op.close()


def _target(qd, S, K):
    sim = qd.Simulation(S["op"], S["prob"], record_step=1, record_capacity=K)
    sim.run(K)
    dK = sim.history()["History_Disp"][-1]
    sim.close()
    return dK[S["obs"]] * 1.1 + 1e-3


def _same(a, b):
    La, ga, ia = a
    Lb, gb, ib = b
    assert La == Lb
    for k in ("a", "b", "dc"):
        assert np.array_equal(ga[k], gb[k]), k
    for k in ("steps", "onset_step", "retried_steps", "output_mismatches", "select_mismatches", "growth_step"):
        assert ia[k] == ib[k], k
    assert ia["lambda_growth"] == ib["lambda_growth"]
    assert (np.isnan(ia["t_first"]) and np.isnan(ib["t_first"])) or ia["t_first"] == ib["t_first"]


def test_checkpointed_adjoint_equals_full_tape_2000_steps(qd, case):
    from qdfn import adjoint
    K = 2000
    tgt = _target(qd, case, K)
    args = (case["op"], case["prob"], K, case["obs"], tgt, 1.0, 1e7)
    full = adjoint.loss_and_grad(*args)
    ck = adjoint.loss_and_grad(*args, checkpoint_every=50)
    assert full[2]["onset_step"] > 0  # the window holds the first event and its time term
    _same(full, ck)


@pytest.mark.parametrize("every", [1, 50, 64, 137])
def test_checkpointed_adjoint_segment_lengths(qd, case, every):
    from qdfn import adjoint
    K = 137
    tgt = _target(qd, case, K)
    args = (case["op"], case["prob"], K, case["obs"], tgt)
    _same(adjoint.loss_and_grad(*args), adjoint.loss_and_grad(*args, checkpoint_every=every))


def test_checkpointed_run_misuse(qd, case):
    from qdfn import _lib
    sim = qd.Simulation(case["op"], case["prob"], record_step=0, trace_capacity=100)
    with pytest.raises(_lib.QdfnError):
        sim.run_checkpointed(200, 50)  # the trace must hold every step
    sim.run_checkpointed(100, 50)
    sim.run(5)  # advanced past the checkpointed window: the tape no longer matches
    with pytest.raises(_lib.QdfnError):
        sim.adjoint(case["prob"]["v_i"], case["prob"]["theta_i"], case["obs"], np.zeros(case["obs"].size))
    sim.close()
    started = qd.Simulation(case["op"], case["prob"], record_step=0, trace_capacity=100)
    started.run(3)
    with pytest.raises(_lib.QdfnError):
        started.run_checkpointed(50, 10)  # checkpoints start at the run's start
    started.close()
