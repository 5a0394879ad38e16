"""Seeded element states for the local-solve tests (shared by CPU and GPU tests).

The states are drawn around the single-fault workload: 100 m elements,
mu = 32 GPa, rho = mu / 3464^2, K_Self ~ 3.4e8 Pa/m, sigma ~ 50 MPa, with
slip rates spanning 1e-12..1 m/s so both solvers (HighV above SwitchV = 1e-2)
and their failure branches are exercised.
"""
import numpy as np

MU = 32e9
RHO = MU / 3464.0 ** 2
PARAMS = dict(conv_crit=1e-9, v0=1e-9, mu=MU, rho=RHO, switch_v=1e-2, alpha_evo=0.0, evolution_dr=1)


def random_states(n, seed=0, stress=False):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.008, 0.012, n)
    b = a + rng.uniform(-0.004, 0.006, n)
    dc = rng.uniform(5e-4, 5e-3, n)
    v_old = 10.0 ** rng.uniform(-12, 0, n)
    th_old = dc / v_old * 10.0 ** rng.uniform(-1, 1, n)
    k = rng.uniform(2.5e8, 4e8, n)
    mass = np.full(n, 100.0 * RHO / 2)
    omega = np.sqrt(k / mass)
    sig = rng.uniform(30e6, 80e6, n)
    f_i = 0.6 + (a - b) * np.log(1e-9 / 1e-6)
    f0 = f_i - a * np.log(1e-9 / PARAMS["v0"]) - b * np.log(dc / 1e-9 * PARAMS["v0"] / dc)
    tau0 = sig * f_i
    disp_old = rng.uniform(0, 2.0, n)
    f_old = f0 + a * np.log(v_old / PARAMS["v0"]) + b * np.log(th_old * PARAMS["v0"] / dc)
    # loading displacement chosen near force balance so most solves converge
    tl = disp_old + (sig * f_old - tau0) / k + rng.normal(0, 1e-4, n)
    dtau = rng.normal(0, 1e4, n)
    sig_old = sig.copy()
    sw = (v_old > PARAMS["switch_v"]).astype(np.int32)
    lrate = np.full(n, 1e-9)
    if stress:
        # push a share of elements far from balance / into degenerate states
        m = rng.random(n)
        tl[m < 0.15] += rng.choice([-1.0, 1.0], (m < 0.15).sum()) * 10.0 ** rng.uniform(-2, 1, (m < 0.15).sum())
        v_old[(m >= 0.15) & (m < 0.2)] = 0.0
        th_old[(m >= 0.2) & (m < 0.25)] *= 1e8
        sw[(m >= 0.25) & (m < 0.35)] = 1
    return dict(a=a, b=b, dc=dc, v_old=v_old, th_old=th_old, k=k, mass=mass, omega=omega, sig=sig, f0=f0,
                tau0=tau0, disp_old=disp_old, f_old=f_old, tl=tl, dtau=dtau, sig_old=sig_old, sw=sw, lrate=lrate)
