// Rate-and-state local solve for one element, FP64, as device code.
// Semantics of scripts/Functions_Solvers.jl (EvolutionLaw :2-15,
// Solver_HighV_D :26-182, Solver_LowV_D :185-327): the same secant updates,
// the same failure branches and outcome codes (InstabilityThistime 0/2/3/4,
// FLAG_GoodToGo), the same 100/200-iteration relaxation rules, and the same
// operation order (the library is built without FMA contraction so that the
// device arithmetic rounds as the reference's does).
#pragma once
#include <hip/hip_runtime.h>

namespace qdfn {

struct ElemIn {
    double conv, d_old, f_old, th_old, v_old, f0, a, b, dc, v0, k, omega, mass, mu, sig, tl, rho, dtau,
        switch_v, alpha_evo, sig_old, tau0;
    int dr;
};

struct ElemOut {
    double v, f, d, th, sig, dt;
    int instab, ok, iters, stall;
};

__device__ __host__ inline double evolve_state(double th_old, double dt, double v, double dc, double alpha,
                                               double sig, double sig_old, double b, bool fast, int dr) {
    if (dr == 1) {
        if (fast) {
            const double ct = v / dc + alpha * (sig - sig_old) / dt / b / sig;
            return (th_old - 1.0 / ct) * exp(-v / dc * dt) + 1.0 / ct;
        }
        return (th_old + dt) / (1.0 + dt * v / dc + alpha * dt * (sig - sig_old) / dt / b / sig);
    }
    return th_old - v * th_old / dc * log(v * th_old / dc) * dt;
}

// Lumped-mass oscillator solve (Solver_HighV_D).
__device__ inline ElemOut solve_high_v(const ElemIn& e, double dt) {
    ElemOut o;
    int ok = 1, it = 0, code = 0, stall = 0;
    double crit = e.conv, diff = 10.0;
    const double dv = e.v_old / 100.0;
    double v = e.v_old / 1e10, d = e.d_old, f = e.f_old, th = e.th_old;
    if (v < 0.0) v = 1e-100;
    const double eta_over_sig = e.mu / 2.0 / sqrt(e.mu / e.rho) / e.sig;
    auto restore = [&]() { v = e.v_old; th = e.th_old; f = e.f_old; d = e.d_old; };
    // one oscillator evaluation: velocity implied by trial velocity vt
    auto implied_v = [&](double vt, double tht) {
        const double frd = e.f0 + e.a * log(vt / e.v0) + e.b * log(tht * e.v0 / e.dc) + eta_over_sig * vt -
                           e.dtau / e.sig;
        const double F = e.tl + (e.tau0 - frd * e.sig) / e.k;
        d = (e.d_old - F) * cos(e.omega * dt) + (e.v_old / e.omega) * sin(e.omega * dt) + F;
        return (d - e.d_old) / dt * 2.0 - e.v_old;
    };
    while (diff > crit) {
        it += 1;
        const double vt = v;
        th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, true, e.dr);
        if (th < 0.0) { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; break; }
        v = implied_v(v, th);
        const double r0 = vt - v;
        v = vt + dv;
        const double vt2 = v;
        th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, true, e.dr);
        if (th < 0.0) { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; break; }
        v = implied_v(v, th);
        const double r1 = vt2 - v;
        v = vt - r0 * dv / (r1 - r0);
        if (v < 0.0) {
            if (!(dt > 1e-5)) code = 4;
            dt = dt / 2.0; restore(); diff = 0.0; ok = 0;
        } else if (isnan(v)) {
            if (dt > 1e-5) { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; }
            else {
                dt = dt / 2.0;
                if (dt < 1e-10) stall += 1;
                code = 4; restore(); diff = 0.0; ok = 0;
            }
        } else {
            d = e.d_old + (e.v_old + v) / 2.0 * dt;
            th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, true, e.dr);
            if (th < 0.0) { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; break; }
            f = e.f0 + e.a * log(v / e.v0) + e.b * log(th * e.v0 / e.dc);
            diff = fabs(vt / v - 1.0);
            code = 0;
        }
        if (it % 100 == 0) {
            if (dt < 1e-5) { crit = crit * 10.0; code = 2; }
            else { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; }
        }
    }
    o.v = v; o.f = f; o.d = d; o.th = th; o.sig = e.sig; o.dt = dt;
    o.instab = code; o.ok = ok; o.iters = it; o.stall = stall;
    return o;
}

// Quasi-dynamic solve with radiation damping and inertia term (Solver_LowV_D).
__device__ inline ElemOut solve_low_v(const ElemIn& e, double dt) {
    ElemOut o;
    int ok = 1, it = 0, code = 0;
    double crit = e.conv, diff = 10.0;
    double v = e.v_old, d = e.d_old, f = e.f_old, th = e.th_old;
    const double eta = e.mu / 2.0 / sqrt(e.mu / e.rho);
    auto restore = [&]() { v = e.v_old; th = e.th_old; f = e.f_old; d = e.d_old; };
    // the three-way outcome of a negative or NaN state (:210-222 and repeats)
    auto bad_state = [&]() {
        if (dt < e.switch_v) { code = 3; restore(); ok = 0; }
        else if (e.v_old == 0.0) { restore(); ok = 1; }
        else { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; }
    };
    auto friction_of = [&](double vv, double acc) {
        return 1.0 / e.sig * (e.k * (e.tl - d) + e.tau0 - eta * vv + e.dtau - e.mass * acc);
    };
    while (diff > crit) {
        it += 1;
        const double vt = v;
        const double dv = vt / 1e10;
        double acc = (v - e.v_old) / dt;
        th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, false, e.dr);
        if (th < 0.0 || isnan(th)) { bad_state(); break; }
        d = e.d_old + (e.v_old + v) / 2.0 * dt;
        f = friction_of(v, acc);
        double vel = e.v0 * exp((f - e.f0 - e.b * log(e.v0 * th / e.dc)) / e.a);
        const double r0 = vt - vel;

        v = v + dv;
        const double vt2 = v;
        acc = (v - e.v_old) / dt;
        th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, false, e.dr);
        if (th < 0.0 || isnan(th)) { bad_state(); break; }
        d = e.d_old + (e.v_old + v) / 2.0 * dt;
        f = friction_of(v, acc);
        vel = e.v0 * exp((f - e.f0 - e.b * log(e.v0 * th / e.dc)) / e.a);
        const double r1 = vt2 - vel;

        v = vt - r0 * dv / (r1 - r0);
        if (e.dr != 1 && v < 0.0) v = exp(log(vt) - r0 * log(vt2 / vt) / (r1 - r0));
        acc = (v - e.v_old) / dt;
        d = e.d_old + (e.v_old + v) / 2.0 * dt;
        th = evolve_state(e.th_old, dt, v, e.dc, e.alpha_evo, e.sig, e.sig_old, e.b, false, e.dr);
        if (th < 0.0) { bad_state(); break; }
        if (isnan(th)) {
            if (e.v_old == 0.0) { restore(); ok = 1; }
            else if (dt < 1e-3) {
                if (e.v_old < 1e-7) {
                    v = 0.0; th = e.th_old + dt; f = e.f_old; d = e.d_old; diff = 0.0; ok = 1;
                } else { code = 3; restore(); ok = 0; }
            } else { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; }
            break;
        }
        f = friction_of(v, acc);
        diff = fabs(vt / v - 1.0);
        if (it % 200 == 0) {
            if (dt < 1e-4) { code = 3; restore(); ok = 0; }
            else if (dt < 1e1) { crit = crit * 2.0; }
            else { dt = dt / 2.0; restore(); diff = 0.0; ok = 0; }
            break;
        }
    }
    o.v = v; o.f = f; o.d = d; o.th = th; o.sig = e.sig; o.dt = dt;
    o.instab = code; o.ok = ok; o.iters = it; o.stall = 0;
    return o;
}

}  // namespace qdfn
