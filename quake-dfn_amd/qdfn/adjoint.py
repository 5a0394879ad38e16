"""Slip / event-time misfit and its gradient with respect to per-element a,
b and Dc (qdfn_sim_adjoint).

    forward_loss(op, prob, steps, obs, d_target, ...)  the loss of a fresh run
    loss_and_grad(op, prob, steps, obs, d_target, ...)  loss, gradients, info
    misfit(hist, obs, d_target, ...)                    the loss of recorded histories

The loss is sum_o (Disp_K[obs_o] - d_target_o)^2 + w_time (t_first - t_target)^2,
t_first being the onset of the first event (first step whose max V over the
non-loading elements reaches v_onset, interpolated in log10 V between step
end times).  The gradient is the discrete adjoint of the device step loop.
"""
import numpy as np

from .sim import Simulation


def _onset(told, dt, vmax, v_onset):
    te = np.asarray(told) + np.asarray(dt)
    idx = np.nonzero(np.asarray(vmax) >= v_onset)[0]
    if idx.size == 0:
        return None
    k = int(idx[0])
    if k == 0:
        return float(te[0])
    l0, l1, lt = np.log10(vmax[k - 1]), np.log10(vmax[k]), np.log10(v_onset)
    f = (lt - l0) / (l1 - l0) if l1 != l0 else 1.0
    return float(te[k - 1] + f * (te[k] - te[k - 1]))


def misfit(hist, n_loading, obs, d_target, w_time=0.0, t_target=0.0, v_onset=1e-3):
    """The loss from a run's histories (History_Time, History_Dt, History_V,
    History_Disp with one row per step)."""
    d = hist["History_Disp"][-1]
    r = d[np.asarray(obs)] - np.asarray(d_target)
    L = float(np.sum(r * r))
    V = hist["History_V"]
    nn = V.shape[1] - n_loading
    vmax = V[:, :nn].max(axis=1)
    t = _onset(hist["History_Time"], hist["History_Dt"], vmax, v_onset)
    if w_time > 0.0 and t is not None:
        L += w_time * (t - t_target) ** 2
    return L, t


def _run(op, prob, steps):
    sim = Simulation(op, prob, record_step=1, record_capacity=steps)
    sim.run(steps)
    return sim


def forward_loss(op, prob, steps, obs, d_target, w_time=0.0, t_target=0.0, v_onset=1e-3):
    sim = _run(op, prob, steps)
    try:
        st = sim.stats()
        if st["status"] != 0:
            raise RuntimeError(f"run failed with status {st['status']}")
        return misfit(sim.history(), int(prob.get("n_loading", 0)), obs, d_target, w_time, t_target, v_onset)
    finally:
        sim.close()


def loss_and_grad(op, prob, steps, obs, d_target, w_time=0.0, t_target=0.0, v_onset=1e-3, checkpoint_every=None):
    """Loss, gradients and info of a fresh run of `steps` steps; with
    checkpoint_every the forward keeps a checkpoint every that many steps and
    the backward sweep recomputes each segment from its checkpoint."""
    if checkpoint_every:
        sim = Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=steps)
        sim.run_checkpointed(steps, checkpoint_every)
    else:
        sim = _run(op, prob, steps)
    try:
        return sim.adjoint(prob["v_i"], prob["theta_i"], obs, d_target, w_time, t_target, v_onset)
    finally:
        sim.close()


def with_params(prob, a=None, b=None, dc=None):
    """A copy of the problem with a, b, Dc replaced (Friction_i, Theta_i and
    V_i unchanged: f0 follows from them, as the run set-up computes it)."""
    q = dict(prob)
    for k, v in (("a", a), ("b", b), ("dc", dc)):
        if v is not None:
            q[k] = np.asarray(v, dtype=np.float64).copy()
    return q


def fd_check(op, prob, steps, obs, d_target, grads, params, w_time=0.0, t_target=0.0, v_onset=1e-3,
             rel_steps=(1e-3, 1e-4, 1e-5, 1e-6)):
    """Central differences of the loss for each (kind, index) in params at
    several relative steps.  The step loop's discrete choices (the Dt
    controller's argmin and growth rule, solver switches) make the loss
    piecewise smooth, and its secant solves make it noisy below ~1e-6
    relative, so for each parameter the step whose forward and backward
    one-sided differences agree best is taken (the scale on which the loss
    is smooth).  Returns one dict per parameter: adjoint, difference, step,
    smoothness (|forward - backward| / |central|) and relative error."""
    L0 = forward_loss(op, prob, steps, obs, d_target, w_time, t_target, v_onset)[0]
    out = []
    for kind, i in params:
        p = prob[kind]
        best = None
        for hr in rel_steps:
            h = hr * abs(p[i])
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            Lp = forward_loss(op, with_params(prob, **{kind: pp}), steps, obs, d_target, w_time, t_target,
                              v_onset)[0]
            Lm = forward_loss(op, with_params(prob, **{kind: pm}), steps, obs, d_target, w_time, t_target,
                              v_onset)[0]
            c = (Lp - Lm) / (2 * h)
            fw, bw = (Lp - L0) / h, (L0 - Lm) / h
            smooth = abs(fw - bw) / max(abs(c), 1e-300)
            if best is None or smooth < best["smoothness"]:
                best = {"kind": kind, "index": int(i), "adjoint": float(grads[kind][i]), "fd": float(c),
                        "rel_step": hr, "smoothness": float(smooth)}
        best["rel_err"] = abs(best["adjoint"] - best["fd"]) / max(abs(best["fd"]), 1e-300)
        out.append(best)
    return out
