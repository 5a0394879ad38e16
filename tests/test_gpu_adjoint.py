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


def _dual_directional(od, prob, KS, KN, K, obs, target, delta, w_time=0.0, t_target=0.0, v_onset=1e-3):
    """dL/ds along delta by the forward-mode (dual-number) oracle: the stand-in
    for a ForwardDiff pass through the reference's CPU code."""
    st, r = od.sim_dense(prob, KS, KN, total_step=K, record_step=1, seed=delta)
    assert st == 0
    D, dD = r["hist_disp"][-1], r["d_hist_disp"][-1]
    res = D[obs] - target
    L = float(np.sum(res ** 2))
    dL = float(np.sum(2 * res * dD[obs]))
    if w_time > 0:
        v, dv = r["trace_vmax"], r["d_trace_vmax"]
        te, dte = r["trace_told"] + r["trace_dt"], r["d_trace_told"] + r["d_trace_dt"]
        k = int(np.nonzero(v >= v_onset)[0][0])
        assert k >= 1
        l0, l1, lt = np.log10(v[k - 1]), np.log10(v[k]), np.log10(v_onset)
        dl0, dl1 = dv[k - 1] / (v[k - 1] * np.log(10)), dv[k] / (v[k] * np.log(10))
        f = (lt - l0) / (l1 - l0)
        df = (-dl0 * (l1 - l0) - (lt - l0) * (dl1 - dl0)) / (l1 - l0) ** 2
        t = te[k - 1] + f * (te[k] - te[k - 1])
        dt = dte[k - 1] + df * (te[k] - te[k - 1]) + f * (dte[k] - dte[k - 1])
        L += w_time * (t - t_target) ** 2
        dL += 2 * w_time * (t - t_target) * dt
    return L, dL


def test_adjoint_dot_product_with_dual_forward(qd, case, setup):
    """<grad L, delta> from the device adjoint equals the forward-mode
    derivative of the oracle's main() along delta (random direction in a, b and
    Dc, each entry scaled by its parameter) to 1e-6 relative, on the
    interseismic window (two directions)."""
    from oracle import oracle_dual as od
    prob, K = setup["prob"], setup["K"]
    sim = qd.Simulation(setup["op"], prob, record_step=1, record_capacity=K)
    sim.run(K)
    L, g, info = sim.adjoint(prob["v_i"], prob["theta_i"], setup["obs"], setup["target"])
    sim.close()
    assert info["lambda_growth"] < 1e8 and info["growth_step"] == -1
    for seed in (1, 2):
        rng = np.random.default_rng(seed)
        delta = {k: prob[k] * rng.standard_normal(prob["n"]) for k in ("a", "b", "dc")}
        ad = sum(float(np.dot(g[k], delta[k])) for k in delta)
        Ld, dd = _dual_directional(od, prob, case["KS"], case["KN"], K, setup["obs"], setup["target"], delta)
        print(f"[adjoint] interseismic window ({K} steps): adjoint {ad:.12e} dual {dd:.12e} "
              f"rel {abs(ad - dd) / abs(dd):.2e}")
        assert abs(L - Ld) <= 1e-9 * abs(Ld)
        assert abs(ad - dd) <= 1e-6 * abs(dd)


@pytest.mark.parametrize("after", [3])
def test_adjoint_dot_product_through_event(qd, case, after):
    """The same identity on a window through the fault's first event (onset plus
    `after` HighV steps at RuptureDt), with the event-time term.  Deeper into
    the rupture the device and the CPU oracle stop following one discrete path:
    the secant solves stop at |V_trial / V - 1| < 1e-9, so ulp-level differences
    between the two libm implementations change iteration counts, which the
    rupture amplifies (losses apart by 3.7e-8 three steps after the onset,
    2.7e-5 after eight, 1.3e-3 after thirty) -- each derivative is then exact
    for its own path."""
    from oracle import oracle_dual as od
    from qdfn import adjoint
    prob = case["prob"]
    op = qd.DenseOperator.from_matrices(case["KS"], case["KN"], qd.PREC_FP64)
    probe = qd.Simulation(op, prob, trace_capacity=4000)
    probe.run(4000)
    tr = probe.trace()
    probe.close()
    on = int(np.nonzero(tr["vmax"] >= 1e-3)[0][0])
    K = on + after
    obs = np.sort(np.random.default_rng(5).choice(prob["n"], 16, replace=False))
    sim = qd.Simulation(op, prob, record_step=1, record_capacity=K)
    sim.run(K)
    h = sim.history()
    dK = h["History_Disp"][-1]
    target = dK[obs] * 1.1 + 1e-3
    _, t_first = adjoint.misfit(h, 0, obs, target)
    t_target = t_first * 1.05
    w_time = float(np.sum((dK[obs] - target) ** 2)) / (t_first - t_target) ** 2
    L, g, info = sim.adjoint(prob["v_i"], prob["theta_i"], obs, target, w_time, t_target)
    sim.close()
    op.close()
    rng = np.random.default_rng(3)
    delta = {k: prob[k] * rng.standard_normal(prob["n"]) for k in ("a", "b", "dc")}
    ad = sum(float(np.dot(g[k], delta[k])) for k in delta)
    Ld, dd = _dual_directional(od, prob, case["KS"], case["KN"], K, obs, target, delta, w_time, t_target)
    print(f"[adjoint] window through the event ({K} steps, onset {on}): adjoint {ad:.12e} dual {dd:.12e} "
          f"rel {abs(ad - dd) / abs(dd):.2e}, lambda growth {info['lambda_growth']:.3e}")
    # the two forwards agree to ~4e-8 (onset + 3 steps); the derivatives of their paths to ~1e-7
    assert abs(L - Ld) <= 1e-7 * abs(Ld)
    assert abs(ad - dd) <= 1e-5 * abs(dd)
