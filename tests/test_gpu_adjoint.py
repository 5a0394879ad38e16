"""GPU: the discrete adjoint of the step loop (qdfn_sim_adjoint) against
central finite differences (relative step 1e-3) of the same loss, on a
256-element fault (16 x 16 elements of 50 m) with seeded per-element a, b, Dc:
  - 16 random parameters (a, b and Dc mixed), differences through the GPU
    forward: agreement to 1e-4 relative (floor: 1e-3 of the largest gradient
    entry of that parameter kind);
  - 4 parameters, differences through the oracle's FP64 main() (the CPU
    restatement of the reference loop) -- the stand-in for a ForwardDiff pass
    through the reference, which cannot run here.
The loss here is the squared slip misfit at 16 seeded observation elements
over the interseismic steps before the fault's first event; the event-time
term and the steps through an event (HighV solves) are checked on the
4,096-element workload by bench.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# The forward's secant solves stop at |V_trial / V - 1| < 1e-9, so the loss
# is smooth only down to relative parameter changes of about 1e-5: central
# differences with a 1e-6 step measure that tolerance (errors of 1e-3 and
# more), with 1e-3 they agree with the adjoint to ~1e-5.
REL_H = 1e-3
TOL = 1e-4


@pytest.fixture(scope="module")
def case(orc):
    from qdfn import configs
    m = configs.adjoint_fault(nx=16, nz=16, seed=11)
    seg = m["seg"]
    KS, KN = orc.stiffness_rect(seg, seg, m["mu"], m["nu"])
    E, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(m, ks, KS.sum(1), KN.sum(1))
    return dict(m=m, KS=KS, KN=KN, prob=prob)


@pytest.fixture(scope="module")
def setup(qd, case):
    from qdfn import adjoint
    op = qd.DenseOperator.from_matrices(case["KS"], case["KN"], qd.PREC_FP64)
    prob = case["prob"]
    # window: the interseismic steps before the first event (this small fault
    # nucleates in one step, a jump no difference quotient can follow)
    probe = qd.Simulation(op, prob, trace_capacity=4000)
    probe.run(4000)
    tr = probe.trace()
    probe.close()
    on = np.nonzero(tr["vmax"] >= 1e-3)[0]
    assert on.size and on[0] > 10, "the probe run must reach the first event"
    K = int(on[0]) - 2
    rng = np.random.default_rng(3)
    obs = np.sort(rng.choice(prob["n"], 16, replace=False))
    sim = qd.Simulation(op, prob, record_step=1, record_capacity=K)
    sim.run(K)
    dK = sim.history()["History_Disp"][-1]
    sim.close()
    target = dK[obs] * 1.1 + 1e-3
    t_target, w_time = 0.0, 0.0  # no onset inside this window
    yield dict(op=op, prob=prob, K=K, obs=obs, target=target, w_time=w_time, t_target=t_target)
    op.close()


def _fd(loss_fn, prob, kind, i):
    from qdfn import adjoint
    p = prob[kind].copy()
    h = REL_H * abs(p[i])
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    Lp = loss_fn(adjoint.with_params(prob, **{kind: pp}))
    Lm = loss_fn(adjoint.with_params(prob, **{kind: pm}))
    return (Lp - Lm) / (2 * h)


def test_adjoint_matches_finite_differences_gpu(qd, setup):
    from qdfn import adjoint
    S = setup
    L, g, info = adjoint.loss_and_grad(S["op"], S["prob"], S["K"], S["obs"], S["target"], S["w_time"],
                                       S["t_target"])
    assert info["steps"] == S["K"] and info["output_mismatches"] == 0 and info["select_mismatches"] == 0
    print("adjoint info", info)
    L_fwd, _ = adjoint.forward_loss(S["op"], S["prob"], S["K"], S["obs"], S["target"], S["w_time"], S["t_target"])
    assert L == pytest.approx(L_fwd, rel=1e-12)
    loss_fn = lambda q: adjoint.forward_loss(S["op"], q, S["K"], S["obs"], S["target"], S["w_time"],
                                             S["t_target"])[0]
    rng = np.random.default_rng(17)
    kinds = ["a", "b", "dc"]
    worst = 0.0
    for q in range(16):
        kind = kinds[q % 3]
        i = int(rng.integers(S["prob"]["n"]))
        fd = _fd(loss_fn, S["prob"], "dc" if kind == "dc" else kind, i)
        ad = g[kind][i]
        floor = 1e-3 * np.abs(g[kind]).max()
        err = abs(ad - fd) / max(abs(fd), floor)
        worst = max(worst, err)
        assert err <= TOL, (kind, i, ad, fd, err)


def test_adjoint_matches_finite_differences_oracle(qd, orc, case, setup):
    from qdfn import adjoint
    S = setup
    L, g, info = adjoint.loss_and_grad(S["op"], S["prob"], S["K"], S["obs"], S["target"], S["w_time"],
                                       S["t_target"])

    def loss_fn(q):
        st, r = orc.sim_dense(q, case["KS"], case["KN"], total_step=S["K"], record_step=1)
        assert st == 0
        hist = {"History_Time": r["hist_time"], "History_Dt": r["hist_dt"], "History_V": r["hist_v"],
                "History_Disp": r["hist_disp"]}
        return adjoint.misfit(hist, 0, S["obs"], S["target"], S["w_time"], S["t_target"])[0]

    rng = np.random.default_rng(29)
    for kind in ("a", "b", "dc", "a"):
        i = int(rng.integers(S["prob"]["n"]))
        fd = _fd(loss_fn, S["prob"], kind, i)
        ad = g[kind][i]
        floor = 1e-3 * np.abs(g[kind]).max()
        assert abs(ad - fd) / max(abs(fd), floor) <= TOL, (kind, i, ad, fd)


def test_adjoint_same_on_tiled_layout(qd, case, setup, monkeypatch):
    """The adjoint's transposed coupling reads the split-tiled dense layout
    through the same index as the forward: gradients bitwise equal to the
    column-major operator's."""
    grads = []
    for tiled in ("0", "1"):
        monkeypatch.setenv("QDFN_DENSE_TILED", tiled)
        op = qd.DenseOperator.from_matrices(case["KS"], case["KN"], qd.PREC_FP64)
        sim = qd.Simulation(op, setup["prob"], record_step=1, record_capacity=setup["K"])
        sim.run(setup["K"])
        _, g, _ = sim.adjoint(setup["prob"]["v_i"], setup["prob"]["theta_i"], setup["obs"], setup["target"])
        sim.close()
        op.close()
        grads.append(g)
    for k in ("a", "b", "dc"):
        np.testing.assert_array_equal(grads[0][k], grads[1][k])
