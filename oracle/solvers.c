/* ORACLE — test infrastructure only (see okada.c header).
 *
 * CPU FP64 restatement of the rate-and-state local solve of
 * scripts/Functions_Solvers.jl: EvolutionLaw (:2-15), Solver_HighV_D (:26-182),
 * Solver_LowV_D (:185-327), SolveOneTimeStep (:334-380) and
 * InterpolateFromStressChange (:465-540).  Every branch, counter, cap and
 * threshold of the reference is kept, including its quirks:
 *   - LowV compares Dt (not V) with SwitchV (:211,:218,:238,:244,:270);
 *   - the 100/200-iteration checks run after a failure branch that already
 *     set maxDiff=0 (HighV :151-162), so Dt can be halved twice in one call;
 *   - HighV's readline() when Dt<1e-10 (:123-125) cannot block here; it is
 *     counted in *stall so the caller can report it, and the solve continues
 *     as the reference does after the key press.
 * Build with -ffp-contract=off. */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

/* Functions_Solvers.jl:2-15 */
double orc_evolution_law(double ThetaOld, double Dt, double V, double Dc, double Alpha_Evo,
                         double Sig, double SigOld, double b, int Fast1_Slow0, int EvolutionDR) {
    double Theta;
    if (EvolutionDR == 1) {
        if (Fast1_Slow0 == 1) {
            double Const_T = V / Dc + Alpha_Evo * (Sig - SigOld) / Dt / b / Sig;
            Theta = (ThetaOld - 1.0 / Const_T) * exp(-V / Dc * Dt) + 1.0 / Const_T;
        } else {
            Theta = (ThetaOld + Dt) / (1.0 + Dt * V / Dc + Alpha_Evo * Dt * (Sig - SigOld) / Dt / b / Sig);
        }
    } else {
        Theta = ThetaOld - V * ThetaOld / Dc * log(V * ThetaOld / Dc) * Dt;
    }
    return Theta;
}

#define RESET() do { V = VOld; Theta = ThetaOld; Friction = FrictionOld; Disp = DispOld; } while (0)

/* Functions_Solvers.jl:26-182 */
void orc_solver_highv(const orc_elem_in_t *in, double Dt, orc_elem_out_t *out) {
    const double ConvergenceCrit = in->conv_crit, DispOld = in->disp_old, FrictionOld = in->friction_old;
    const double ThetaOld = in->theta_old, VOld = in->v_old, Friction0 = in->friction0;
    const double a = in->a, b = in->b, Dc = in->dc, V0 = in->v0, K_Self = in->k_self;
    const double Omega = in->omega, ShearModulus = in->mu, EffNormalStress = in->sigma;
    const double Total_Loading_Disp = in->tl, RockDensity = in->rho, D_EffStress_Shear = in->dtau;
    const double Alpha_Evo = in->alpha_evo, EffNormalStress_Old = in->sigma_old;
    const double InitialShearStress = in->tau0;
    const int EvolutionDR = in->evolution_dr;

    int FLAG_GoodToGo = 1, Iteration = 0, Instab = 0, stall = 0;
    double maxDiff = 10.0, CCi = ConvergenceCrit;
    double DV = VOld / 100.0;
    double V = VOld / 1e10, Disp = DispOld, Friction = FrictionOld, Theta = ThetaOld;
    if (V < 0.0) V = 1e-100;

    while (maxDiff > CCi) {
        Iteration += 1;
        double VTest = V;
        Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 1, EvolutionDR);
        if (Theta < 0.0) { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; break; }
        double Friction_RadDamp = Friction0 + a * log(V / V0) + b * log(Theta * V0 / Dc) +
                                  ShearModulus / 2.0 / sqrt(ShearModulus / RockDensity) / EffNormalStress * V -
                                  D_EffStress_Shear / EffNormalStress;
        double F = Total_Loading_Disp + (InitialShearStress - Friction_RadDamp * EffNormalStress) / K_Self;
        Disp = (DispOld - F) * cos(Omega * Dt) + (VOld / Omega) * sin(Omega * Dt) + F;
        V = (Disp - DispOld) / Dt * 2.0 - VOld;
        double FOriginal = VTest - V;

        V = VTest + DV;
        double VTest2 = V;
        Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 1, EvolutionDR);
        if (Theta < 0.0) { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; break; }
        Friction_RadDamp = Friction0 + a * log(V / V0) + b * log(Theta * V0 / Dc) +
                           ShearModulus / 2.0 / sqrt(ShearModulus / RockDensity) / EffNormalStress * V -
                           D_EffStress_Shear / EffNormalStress;
        F = Total_Loading_Disp + (InitialShearStress - Friction_RadDamp * EffNormalStress) / K_Self;
        Disp = (DispOld - F) * cos(Omega * Dt) + (VOld / Omega) * sin(Omega * Dt) + F;
        V = (Disp - DispOld) / Dt * 2.0 - VOld;
        double FUpdated = VTest2 - V;
        V = VTest - FOriginal * DV / (FUpdated - FOriginal);

        if (V < 0.0) {
            if (Dt > 1e-5) { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
            else { Dt = Dt / 2.0; Instab = 4; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
        } else if (isnan(V)) {
            if (Dt > 1e-5) { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
            else {
                Dt = Dt / 2.0;
                if (Dt < 1e-10) stall += 1;
                Instab = 4; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0;
            }
        } else {
            Disp = DispOld + (VOld + V) / 2.0 * Dt;
            Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 1, EvolutionDR);
            if (Theta < 0.0) { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; break; }
            Friction = Friction0 + a * log(V / V0) + b * log(Theta * V0 / Dc);
            maxDiff = fabs(VTest / V - 1.0);
            Instab = 0;
        }
        if (Iteration % 100 == 0) {
            if (Dt < 1e-5) { CCi = CCi * 10.0; Instab = 2; }
            else { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
        }
    }
    out->v = V; out->friction = Friction; out->disp = Disp; out->theta = Theta;
    out->sigma = EffNormalStress; out->dt = Dt; out->instab = Instab; out->flag = FLAG_GoodToGo;
    out->iters = Iteration; out->stall = stall;
}

/* the three-way outcome shared by LowV's Theta<0 / NaN checks (:210-222 etc.) */
#define LOWV_FAIL3()                                                              \
    do {                                                                          \
        if (Dt < SwitchV) { Instab = 3; RESET(); FLAG_GoodToGo = 0; }             \
        else if (VOld == 0.0) { RESET(); FLAG_GoodToGo = 1; }                     \
        else { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }        \
    } while (0)

/* Functions_Solvers.jl:185-327 */
void orc_solver_lowv(const orc_elem_in_t *in, double Dt, orc_elem_out_t *out) {
    const double ConvergenceCrit = in->conv_crit, DispOld = in->disp_old, FrictionOld = in->friction_old;
    const double ThetaOld = in->theta_old, VOld = in->v_old, Friction0 = in->friction0;
    const double a = in->a, b = in->b, Dc = in->dc, V0 = in->v0, K_Self = in->k_self;
    const double Mass = in->mass, ShearModulus = in->mu, EffNormalStress = in->sigma;
    const double Total_Loading_Disp = in->tl, RockDensity = in->rho, D_EffStress_Shear = in->dtau;
    const double SwitchV = in->switch_v, Alpha_Evo = in->alpha_evo, EffNormalStress_Old = in->sigma_old;
    const double InitialShearStress = in->tau0;
    const int EvolutionDR = in->evolution_dr;

    int FLAG_GoodToGo = 1, Iteration = 0, Instab = 0;
    double maxDiff = 10.0, CCi = ConvergenceCrit;
    double V = VOld, Disp = DispOld, Friction = FrictionOld, Theta = ThetaOld;

    while (maxDiff > CCi) {
        Iteration += 1;
        double VTest = V;
        double DV = VTest / 1e10;
        double Accel = (V - VOld) / Dt;
        Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 0, EvolutionDR);
        if (Theta < 0.0) { LOWV_FAIL3(); break; }
        else if (isnan(Theta)) { LOWV_FAIL3(); break; }
        Disp = DispOld + (VOld + V) / 2.0 * Dt;
        Friction = 1.0 / EffNormalStress * (K_Self * (Total_Loading_Disp - Disp) + InitialShearStress -
                   ShearModulus / 2.0 / sqrt(ShearModulus / RockDensity) * V + D_EffStress_Shear - Mass * Accel);
        double Vel = V0 * exp((Friction - Friction0 - b * log(V0 * Theta / Dc)) / a);
        double FV = VTest - Vel;

        V = V + DV;
        double VTest2 = V;
        Accel = (V - VOld) / Dt;
        Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 0, EvolutionDR);
        if (Theta < 0.0) { LOWV_FAIL3(); break; }
        else if (isnan(Theta)) { LOWV_FAIL3(); break; }
        Disp = DispOld + (VOld + V) / 2.0 * Dt;
        Friction = 1.0 / EffNormalStress * (K_Self * (Total_Loading_Disp - Disp) + InitialShearStress -
                   ShearModulus / 2.0 / sqrt(ShearModulus / RockDensity) * V + D_EffStress_Shear - Mass * Accel);
        Vel = V0 * exp((Friction - Friction0 - b * log(V0 * Theta / Dc)) / a);
        double FV_DeV = VTest2 - Vel;

        V = VTest - FV * DV / (FV_DeV - FV);
        if (EvolutionDR != 1 && V < 0.0)
            V = exp(log(VTest) - FV * log(VTest2 / VTest) / (FV_DeV - FV));

        Accel = (V - VOld) / Dt;
        Disp = DispOld + (VOld + V) / 2.0 * Dt;
        Theta = orc_evolution_law(ThetaOld, Dt, V, Dc, Alpha_Evo, EffNormalStress, EffNormalStress_Old, b, 0, EvolutionDR);
        if (Theta < 0.0) { LOWV_FAIL3(); break; }
        else if (isnan(Theta)) {
            if (VOld == 0.0) { RESET(); FLAG_GoodToGo = 1; }
            else if (Dt < 1e-3) {
                if (VOld < 1e-7) {
                    V = 0.0; Theta = ThetaOld + Dt; Friction = FrictionOld; Disp = DispOld;
                    maxDiff = 0.0; FLAG_GoodToGo = 1;
                } else { Instab = 3; RESET(); FLAG_GoodToGo = 0; }
            } else { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
            break;
        }
        Friction = 1.0 / EffNormalStress * (K_Self * (Total_Loading_Disp - Disp) + InitialShearStress -
                   ShearModulus / 2.0 / sqrt(ShearModulus / RockDensity) * V + D_EffStress_Shear - Mass * Accel);
        maxDiff = fabs(VTest / V - 1.0);

        if (Iteration % 200 == 0) {
            if (Dt < 1e-4) { Instab = 3; RESET(); FLAG_GoodToGo = 0; break; }
            else if (Dt < 1e1) { CCi = CCi * 2.0; }
            else { Dt = Dt / 2.0; RESET(); maxDiff = 0.0; FLAG_GoodToGo = 0; }
            break;
        }
    }
    out->v = V; out->friction = Friction; out->disp = Disp; out->theta = Theta;
    out->sigma = EffNormalStress; out->dt = Dt; out->instab = Instab; out->flag = FLAG_GoodToGo;
    out->iters = Iteration; out->stall = 0;
}

/* SolveOneTimeStep, Functions_Solvers.jl:334-380.  Element arrays are length n;
 * the last n_loading rows are loading faults.  Writes V, Friction, Disp, Theta,
 * EffNormalStress, Dt_All, Instab, Flag; returns the number of readline()
 * stalls met (0 normally). */
int64_t orc_solve_one_timestep(const orc_step_params_t *p, int64_t n, int64_t n_loading,
                               const double *DispOld, const double *FrictionOld, const double *ThetaOld,
                               const double *VOld, const double *Friction0, const double *a,
                               const double *b, const double *Dc, const double *K_Self, double Dt,
                               const double *Omega, const double *Mass, const double *Sig_i,
                               const double *TL, const int32_t *SolverSwitch, const double *LoadingRate,
                               const double *D_EffStress_Shear, const double *Sig_old, const double *Tau0,
                               double *V, double *Friction, double *Disp, double *Theta, double *Sig,
                               double *Dt_All, int32_t *Instab, int32_t *Flag, int32_t *Iters) {
    int64_t stalls = 0;
    double DtOld = Dt;
    /* elements are independent: any thread count gives the same outputs */
#pragma omp parallel for schedule(static) reduction(+ : stalls)
    for (int64_t i = 0; i < n - n_loading; ++i) {
        orc_elem_in_t in;
        in.conv_crit = p->conv_crit; in.disp_old = DispOld[i]; in.friction_old = FrictionOld[i];
        in.theta_old = ThetaOld[i]; in.v_old = VOld[i]; in.friction0 = Friction0[i]; in.a = a[i];
        in.b = b[i]; in.dc = Dc[i]; in.v0 = p->v0; in.k_self = K_Self[i]; in.omega = Omega[i];
        in.mass = Mass[i]; in.mu = p->mu; in.sigma = Sig_i[i]; in.tl = TL[i]; in.rho = p->rho;
        in.dtau = D_EffStress_Shear[i]; in.switch_v = p->switch_v; in.alpha_evo = p->alpha_evo;
        in.sigma_old = Sig_old[i]; in.evolution_dr = p->evolution_dr; in.tau0 = Tau0[i];
        orc_elem_out_t o;
        if (SolverSwitch[i] == 1) orc_solver_highv(&in, DtOld, &o);
        else orc_solver_lowv(&in, DtOld, &o);
        V[i] = o.v; Friction[i] = o.friction; Disp[i] = o.disp; Theta[i] = o.theta; Sig[i] = o.sigma;
        Dt_All[i] = o.dt; Instab[i] = o.instab; Flag[i] = o.flag;
        if (Iters) Iters[i] = o.iters;
        stalls += o.stall;
    }
    for (int64_t i = n - n_loading; i < n; ++i) {
        V[i] = LoadingRate[i];
        Friction[i] = FrictionOld[i];
        Disp[i] = DispOld[i] + V[i] * Dt;
        Theta[i] = ThetaOld[i];
        Sig[i] = 1e10;
        Dt_All[i] = Dt;
        Instab[i] = 0;
        Flag[i] = 1;
        if (Iters) Iters[i] = 0;
    }
    return stalls;
}

/* InterpolateFromStressChange, Functions_Solvers.jl:465-540.  Tables are
 * nt x n row-per-time, column-major (time k, element i at tab[k + i*nt]),
 * i.e. the reference's (T x N) Julia matrices. */
void orc_interp_external(double Time, int64_t n, int64_t nt, const double *tarr,
                         const double *ExtN, const double *ExtS, const double *ExtP,
                         double *dSigN, double *dTau, double *dP) {
    int64_t timeIdx = 0;
    for (int64_t i = 1; i <= nt; ++i)
        if (Time >= tarr[i - 1]) timeIdx = i;
    int64_t lo = timeIdx, up = timeIdx + 1;
    double timediff = (timeIdx == 0) ? Time : Time - tarr[timeIdx - 1];
    for (int64_t e = 0; e < n; ++e) {
        if (timeIdx == 0) {
            double uN = ExtN[(up - 1) + e * nt], uS = ExtS[(up - 1) + e * nt], uP = ExtP[(up - 1) + e * nt];
            dSigN[e] = timediff / tarr[up - 1] * (uN - 0.0) + 0.0;
            dTau[e] = timediff / tarr[up - 1] * (uS - 0.0) + 0.0;
            dP[e] = timediff / tarr[up - 1] * (uP - 0.0) + 0.0;
        } else if (timeIdx == nt) {
            dSigN[e] = ExtN[(lo - 1) + e * nt];
            dTau[e] = ExtS[(lo - 1) + e * nt];
            dP[e] = ExtP[(lo - 1) + e * nt];
        } else {
            double span = tarr[up - 1] - tarr[lo - 1];
            double lN = ExtN[(lo - 1) + e * nt], lS = ExtS[(lo - 1) + e * nt], lP = ExtP[(lo - 1) + e * nt];
            double uN = ExtN[(up - 1) + e * nt], uS = ExtS[(up - 1) + e * nt], uP = ExtP[(up - 1) + e * nt];
            dSigN[e] = timediff / span * (uN - lN) + lN;
            dTau[e] = timediff / span * (uS - lS) + lS;
            dP[e] = timediff / span * (uP - lP) + lP;
        }
    }
}
