"""GPU: the batched ensemble (qdfn_ensemble_*) advances each member exactly as
the member advances alone: same step times, step sizes, max V and final state,
bit for bit (the batched kernels keep each run's split structure, column
order and fixed-order sums; the standalone run is itself pinned to the
oracle's main() in test_gpu_sim.py).  Members that share a geometry share one
operator and are read as one multi-right-hand-side product."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 400


@pytest.fixture(scope="module")
def members():
    from qdfn import configs
    # sizes 280-320 map onto a few grids, so several members share an operator
    return configs.ensemble_members(m=7, seed=5, n_min=280, n_max=320)


def _standalone(qd, op, prob, steps, t_end=None):
    sim = qd.Simulation(op, prob, record_step=10, trace_capacity=steps)
    if t_end is None:
        sim.run(steps)
    else:
        # advance step by step until TOld reaches t_end, as the ensemble's time limit does
        while sim.stats()["steps"] < steps and sim.stats()["t"] < t_end:
            sim.run(1, graph_attempts=1)
    out = dict(stats=sim.stats(), trace=sim.trace(), state=sim.state())
    sim.close()
    return out


def _ensemble(qd, members, steps, max_rhs, t_end=0.0):
    from qdfn.ensemble import problem_for_member
    e = qd.Ensemble.from_members(members, problem_for_member, max_rhs=max_rhs, trace_capacity=steps)
    probs = [problem_for_member(m, s.op) for m, s in zip(members, e.sims)]
    e.run(steps, t_end=t_end)
    info = e.info()
    res = [dict(stats=s.stats(), trace=s.trace(), state=s.state()) for s in e.sims]
    return e, probs, info, res


def _same(a, b):
    assert a["stats"]["steps"] == b["stats"]["steps"]
    assert a["stats"]["rhs_evals"] == b["stats"]["rhs_evals"]
    assert a["stats"]["retry_passes"] == b["stats"]["retry_passes"]
    for k in ("told", "dt", "vmax"):
        np.testing.assert_array_equal(a["trace"][k], b["trace"][k])
    for k in ("V", "Disp", "Theta", "Friction", "Sigma"):
        np.testing.assert_array_equal(a["state"][k], b["state"][k])


def test_members_share_geometry(members):
    grids = [m["grid"] for m in members]
    assert len(set(grids)) < len(grids)


@pytest.mark.parametrize("max_rhs", [0, 1, 4])
def test_ensemble_matches_standalone(qd, members, max_rhs):
    e, probs, info, res = _ensemble(qd, members, STEPS, max_rhs)
    try:
        assert info["groups"] <= len(members)
        if max_rhs:
            assert info["max_rhs"] == max_rhs
        for k, (s, prob) in enumerate(zip(e.sims, probs)):
            alone = _standalone(qd, s.op, prob, STEPS)
            assert res[k]["stats"]["status"] == 0 and res[k]["stats"]["steps"] == STEPS
            _same(res[k], alone)
    finally:
        e.close()


def test_ensemble_time_limit(qd, members):
    t_end = 3.0e8  # about 9.5 years
    e, probs, info, res = _ensemble(qd, members[:3], 5000, 0, t_end=t_end)
    try:
        for r in res:
            st = r["stats"]
            assert st["status"] == 0
            assert st["t"] >= t_end or st["steps"] == 5000
            # the step before the last started short of t_end
            if st["steps"] >= 2 and st["t"] >= t_end:
                assert r["trace"]["told"][-1] < t_end
        # the runs reach t_end after different step counts, so the work lists
        # are rebuilt from the runs still going before the last ones finish;
        # every run is still bitwise its standalone run
        assert len({r["stats"]["steps"] for r in res}) > 1
        for k, (s, prob) in enumerate(zip(e.sims, probs)):
            _same(res[k], _standalone(qd, s.op, prob, 5000, t_end=t_end))
    finally:
        e.close()
