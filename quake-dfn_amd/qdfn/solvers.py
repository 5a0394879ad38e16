"""Rate-and-state local solve with the reference's interface.

SolveOneTimeStep  scripts/Functions_Solvers.jl:334-380 (Solver_HighV_D :26,
Solver_LowV_D :185, EvolutionLaw :2).  The per-element secant solves run in
the HIP kernel of libqdfn_hip.so, one lane per element, in FP64.
"""
import numpy as np

from ._lib import StepParams, check, lib, ptr, ptr_i32


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def SolveOneTimeStep(ConvergenceCrit, DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc, V0, K_Self, Dt,
                     Omega, Mass, ShearModulus, EffNormalStress_i, Total_Loading_Disp, SolverSwitch, V, Friction,
                     Disp, Theta, EffNormalStress, Dt_All, InstabilityThistime, LoadingRate, LoadingFaultCount,
                     FaultCount, RockDensity, D_EffStress_Shear, SwitchV, Alpha_Evo, EffNormalStress_Old,
                     EvolutionDR, InitialShearStress):
    """Same argument list and return tuple as the reference:
    (V, Friction, Disp, Theta, EffNormalStress, Dt_All, InstabilityThistime, FLAG_GoodToGo).
    The output arrays passed in (V, Friction, ...) are filled in place when
    they are float64 numpy vectors of length FaultCount, as the reference
    mutates them; fresh arrays are returned either way."""
    n = int(FaultCount)
    p = StepParams(conv_crit=float(ConvergenceCrit), v0=float(V0), mu=float(ShearModulus), rho=float(RockDensity),
                   switch_v=float(SwitchV), alpha_evo=float(Alpha_Evo), evolution_dr=int(EvolutionDR))
    ins = [_f64(x) for x in (DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc, K_Self)]
    omega, mass, sig_i, tl = (_f64(x) for x in (Omega, Mass, EffNormalStress_i, Total_Loading_Disp))
    sw = np.ascontiguousarray(np.asarray(SolverSwitch).ravel(), dtype=np.int32)
    lr, dtau, sig_old, tau0 = (_f64(x) for x in (LoadingRate, D_EffStress_Shear, EffNormalStress_Old,
                                                 InitialShearStress))
    for arr in ins + [omega, mass, sig_i, tl, sw, lr, dtau, sig_old, tau0]:
        if arr.size != n:
            raise ValueError("all element arrays must have FaultCount entries")
    outs = [np.zeros(n) for _ in range(6)]
    ints = [np.zeros(n, dtype=np.int32) for _ in range(3)]
    stalls = check(lib().qdfn_solve_one_timestep(
        p, n, int(LoadingFaultCount), *[ptr(x) for x in ins], float(Dt), ptr(omega), ptr(mass), ptr(sig_i),
        ptr(tl), ptr_i32(sw), ptr(lr), ptr(dtau), ptr(sig_old), ptr(tau0), *[ptr(o) for o in outs],
        *[ptr_i32(i) for i in ints]))
    for dst, src in zip((V, Friction, Disp, Theta, EffNormalStress, Dt_All), outs):
        if isinstance(dst, np.ndarray) and dst.dtype == np.float64 and dst.size == n:
            dst.reshape(-1)[:] = src
    if isinstance(InstabilityThistime, np.ndarray) and InstabilityThistime.size == n:
        InstabilityThistime.reshape(-1)[:] = ints[0]
    SolveOneTimeStep.last_iterations = ints[2]
    SolveOneTimeStep.last_stalls = stalls
    return (*outs, ints[0], ints[1])
