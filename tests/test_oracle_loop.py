"""CPU: the oracle's step loops (orc_sim_run / orc_sim_hmat, restating main()
of scripts/Functions_RSFDFN3DMain_D.jl:2-324 and main_H of
Functions_RSFDFN3DMain_H.jl:2-383) and its external-stress interpolation
(orc_interp_external, InterpolateFromStressChange of
Functions_Solvers.jl:465-540), pinned by what the mathematics fixes:

* one element loaded by back-slip, with K coupling only, is the spring-slider
  with radiation damping and inertia
      M dV/dt = tau0 + K (V_pl t - D) - eta V - sigma f(V, theta),
      dD/dt = V,  dtheta/dt = 1 - V theta / Dc  (aging law),
  f = f0 + a ln(V/V0) + b ln(V0 theta/Dc).  scipy's Radau (rtol 1e-12) solves
  it through the first event.  The loop's first-event onset converges to
  Radau's at first order as DtCut doubles, with the external stress read at
  TOld + Dt (main) and at TOld (main_H); the rupture's peak V and the event's
  slip converge at first order as RuptureDt halves (the HighV solver and the
  rupture time step);
* the Dt controller's first steps are hand-computed: Accel = 0 makes
  |V/Accel| infinite, so DtRef = MaximumDt and Dt grows from 1 by x1.2 per
  step (1.2, 1.44, 1.728, ...) until it reaches MaximumDt, which it then
  keeps; every step after one whose max V exceeds SwitchV takes RuptureDt,
  and below SwitchV no step grows past 1.2 x the last unless it is at most
  RuptureDt;
* the interpolation equals numpy.interp on the table with (0, 0) prepended,
  for times before the first sample, on and between samples, and after the
  last (held);
* main_H's loop on a block partition of a dense matrix (all blocks dense)
  follows main()'s loop on the same matrix up to the first event (same
  external-stress timing): the two differ only in how the coupling is formed.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

MU, VS = 32e9, 3464.0
RHO = MU / VS ** 2
ETA = MU / 2 / np.sqrt(MU / RHO)
K, SIG, A, B, DC, VPL = 1e8, 50e6, 0.010, 0.015, 1e-3, 1e-9
M = 100.0 * RHO / 2
V0 = 1e-9
VI, THI = 1e-10, 1e-3 / 1e-10
FI = 0.6 + (A - B) * np.log(VI / 1e-6)
F0 = FI - A * np.log(VI / V0) - B * np.log(THI * V0 / DC)
TAU0 = SIG * FI


def _rhs(t, y):
    D, lv, th = y
    V = np.exp(lv)
    f = F0 + A * np.log(V / V0) + B * np.log(V0 * th / DC)
    return [V, (TAU0 + K * (VPL * t - D) - ETA * V - SIG * f) / M / V, 1 - V * th / DC]


@pytest.fixture(scope="module")
def radau():
    def onset(t, y):
        return y[1] - np.log(1e-3)
    onset.terminal = True
    tol = dict(method="Radau", rtol=1e-12, atol=[1e-14, 1e-12, 1e-6])
    s1 = solve_ivp(_rhs, (0, 3e10), [0.0, np.log(VI), THI], events=onset, **tol)
    t_on = s1.t_events[0][0]

    def end(t, y):
        return y[1] - np.log(1e-3)
    end.terminal, end.direction = True, -1
    s2 = solve_ivp(_rhs, (t_on, t_on + 1000), s1.y_events[0][0], events=end, max_step=0.01, **tol)
    return dict(onset=t_on, peak_v=float(np.exp(s2.y[1].max())), slip=float(s2.y_events[0][0][0]))


def _prob(dt_cut=5.0, rupture_dt=3e-3, maximum_dt=1e7, ext_at_told=0):
    return dict(n=1, n_loading=0, mass=np.array([M]), a=np.array([A]), b=np.array([B]), dc=np.array([DC]),
                theta_i=np.array([THI]), v_i=np.array([VI]), friction_i=np.array([FI]), sigma0=np.array([SIG]),
                loading_rate=np.zeros(1), mu=MU, rho=RHO, switch_v=1e-2, dt_cut=dt_cut, rupture_dt=rupture_dt,
                maximum_dt=maximum_dt, min_normal_stress=2e6, alpha_evo=0.0, evolution_dr=1, planar=1,
                ext_at_told=ext_at_told,
                ext={"t": np.array([0.0, 1e11]), "n": np.zeros((2, 1)), "s": np.array([[0.0], [K * VPL * 1e11]])})


def _run(orc, prob, steps, record_step=0):
    st, r = orc.sim_dense(prob, np.array([[-K]]), np.zeros((1, 1)), total_step=steps, record_step=record_step)
    assert st == 0
    return r


def _onset(r):
    from qdfn import configs
    return configs.first_event_time(r["trace_told"], r["trace_dt"], r["trace_vmax"])


@pytest.mark.parametrize("ext_at_told", [0, 1])
def test_loop_onset_converges_to_radau(orc, radau, ext_at_told):
    errs = []
    for dt_cut in (10.0, 20.0, 40.0, 80.0):
        r = _run(orc, _prob(dt_cut=dt_cut, ext_at_told=ext_at_told), 2000)
        t = _onset(r)
        assert t is not None
        errs.append((t - radau["onset"]) / radau["onset"])
    errs = np.array(errs)
    ratios = errs[:-1] / errs[1:]
    assert np.all((ratios > 1.7) & (ratios < 2.3)), errs  # first order in the time step
    assert abs(errs[-1]) < 5e-3
    # Richardson extrapolation removes the first-order term (two finest runs),
    # then the second-order term (three finest runs): what is left shrinks
    r1, r0 = 2 * errs[-1] - errs[-2], 2 * errs[-2] - errs[-3]
    r2 = (4 * r1 - r0) / 3
    print(f"[loop] onset errors {errs}, extrapolated {r1:.2e}, {r2:.2e}")
    assert abs(r1) < 2e-4 and abs(r2) < 3e-5 and abs(r2) < abs(r1), (errs, r1, r2)


def test_loop_event_converges_to_radau(orc, radau):
    pk, sl = [], []
    for rdt in (3e-3, 1.5e-3, 7.5e-4):
        r = _run(orc, _prob(dt_cut=40.0, rupture_dt=rdt), 1500 + int(4 / rdt), record_step=1)
        v, d = r["hist_v"].ravel(), r["hist_disp"].ravel()
        k0 = int(np.nonzero(v >= 1e-3)[0][0])
        k1 = k0 + int(np.nonzero(v[k0:] < 1e-3)[0][0])
        pk.append(v[k0:k1].max() / radau["peak_v"] - 1)
        sl.append(d[k1] / radau["slip"] - 1)
    pk, sl = np.array(pk), np.array(sl)
    assert np.all(pk > 0) and np.all(pk[:-1] / pk[1:] > 1.7) and np.all(pk[:-1] / pk[1:] < 2.3), pk
    assert pk[-1] < 1e-2
    assert np.all(np.abs(sl[1:]) < np.abs(sl[:-1])) and abs(sl[-1]) < 2e-3, sl


def test_dt_controller_first_steps_hand_computed(orc):
    r = _run(orc, _prob(maximum_dt=10.0), 20)
    # Dt = 1 at start; Accel = 0 -> |V/Accel| = Inf -> DtRef = MaximumDt = 10;
    # Dt < DtRef / 1.2 -> Dt * 1.2, else DtRef
    hand = [1.2, 1.44, 1.728, 2.0736, 2.48832, 2.985984, 3.5831808, 4.29981696, 5.159780352, 6.1917364224,
            7.43008370688, 8.916100448256] + [10.0] * 8
    np.testing.assert_allclose(r["trace_dt"], hand, rtol=1e-15)
    np.testing.assert_allclose(r["trace_told"], np.concatenate([[0.0], np.cumsum(hand)[:-1]]), rtol=1e-15)
    r = _run(orc, _prob(maximum_dt=1e7), 12)
    np.testing.assert_allclose(r["trace_dt"], hand[:12], rtol=1e-15)


def test_rupture_steps_take_rupture_dt(orc):
    rdt = 2.5e-3
    r = _run(orc, _prob(dt_cut=20.0, rupture_dt=rdt), 3000)
    assert r["retry_passes"] == 0
    fast = np.nonzero(r["trace_vmax"][:-1] > 1e-2)[0] + 1
    assert fast.size > 100
    np.testing.assert_array_equal(r["trace_dt"][fast], rdt)
    # below SwitchV, Dt = DtRef when DtRef <= RuptureDt, else at most 1.2 x the last Dt
    slow = np.nonzero(r["trace_vmax"][:-1] <= 1e-2)[0] + 1
    dt = r["trace_dt"]
    assert np.all(dt[slow] <= np.maximum(1.2 * dt[slow - 1], rdt) * (1 + 1e-15))
    assert np.any(dt[slow] < rdt) and np.any(dt[slow] > 1e6)


def test_interp_external_matches_numpy_interp(orc):
    rng = np.random.default_rng(3)
    n = 7
    t = np.array([2.0, 5.0, 5.5, 9.0, 20.0])
    en, es, ep = (rng.standard_normal((t.size, n)) for _ in range(3))
    times = [0.0, 0.7, 1.999, 2.0, 3.3, 5.0, 5.2, 5.5, 8.9, 9.0, 15.0, 20.0, 20.5, 1e6]
    for time in times:
        got = orc.interp_external(time, t, en, es, ep)
        for g, tab in zip(got, (en, es, ep)):
            want = np.array([np.interp(time, np.concatenate([[0.0], t]), np.concatenate([[0.0], tab[:, i]]))
                             for i in range(n)])
            np.testing.assert_allclose(g, want, rtol=4e-16 * 8, atol=1e-15)
    # one-sample table: a ramp from zero, then held
    one = orc.interp_external(1.0, t[:1], en[:1], es[:1], ep[:1])[1]
    np.testing.assert_allclose(one, es[0] * 0.5, rtol=1e-15)
    np.testing.assert_array_equal(orc.interp_external(7.0, t[:1], en[:1], es[:1], ep[:1])[1], es[0])


def test_hmat_loop_follows_dense_loop(orc):
    from qdfn import configs
    cfg = configs.single_fault(nx=16, nz=8)
    seg = cfg["seg"]
    KS, KN = orc.stiffness_rect(seg, seg, cfg["mu"], cfg["nu"])
    _, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(cfg, ks, KS.sum(1), KN.sum(1))
    prob["ext_at_told"] = 0
    n = seg.shape[0]
    h = n // 2
    rng = [[1, h, 1, h], [h + 1, n, 1, h], [1, h, h + 1, n], [h + 1, n, h + 1, n]]

    def blocks(Kx):
        return dict(n=n, range=np.array(rng), kind=np.zeros(4), rank=np.zeros(4),
                    a=[np.asfortranarray(Kx[r[2] - 1:r[3], r[0] - 1:r[1]]) for r in rng], b=[None] * 4)
    steps = 600
    s1, d = orc.sim_dense(prob, KS, KN, total_step=steps, record_step=0)
    s2, hm = orc.sim_hmat(prob, blocks(KS), blocks(KN), total_step=steps, record_step=0)
    assert s1 == 0 and s2 == 0
    k = int(np.nonzero(d["trace_vmax"] >= 1e-3)[0][0])
    assert k > 20
    np.testing.assert_allclose(hm["trace_told"][:k], d["trace_told"][:k], rtol=1e-10)
    np.testing.assert_allclose(hm["trace_vmax"][:k], d["trace_vmax"][:k], rtol=1e-7)
