/* ORACLE — test infrastructure only (see okada.c header).
 *
 * CPU FP64 restatement of the earthquake-cycle step loop:
 *   main   (dense)    scripts/Functions_RSFDFN3DMain_D.jl:2-324
 *   main_H (H-matrix) scripts/Functions_RSFDFN3DMain_H.jl:2-383
 * The two loops differ only in how the coupling is evaluated
 * (dense E=K_ij/K_ii vs. H blocks divided by -K_Self afterwards) and in the
 * time at which external stress is sampled (TOld+Dt in D, TOld in H); both are
 * parameters here.  Histories are recorded every RecordStep steps exactly as
 * the reference does (History_Time holds TOld of the recorded step).
 *
 * The reference's readline() on more than 100 retries of one step
 * (Functions_RSFDFN3DMain_D.jl:255-258) becomes status ORC_ERR_RETRIES, and an
 * empty Dtmin candidate set (Julia's minimum() of an empty collection throws,
 * :165-168) becomes ORC_ERR_EMPTY_DTMIN. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* Dense coupling: TL = E*(Far - DispOld), SigProd = KN*DispOld with E the
 * zero-diagonal normalised shear matrix (Functions_RSFDFN3DMain_D.jl:54-63,
 * :199-202).  Column-major n x n.  Rows are split over OpenMP threads; each
 * row sums over columns in ascending order. */
void orc_dense_coupling(const orc_dense_t *op, const double *x, double *TL, double *SigProd,
                        int planar) {
    const int64_t n = op->n;
    double *negx = (double *)malloc(sizeof(double) * n);
    for (int64_t j = 0; j < n; ++j) negx[j] = 0.0 - x[j];
#pragma omp parallel for schedule(static)
    for (int64_t i0 = 0; i0 < n; i0 += 64) {
        int64_t i1 = i0 + 64 < n ? i0 + 64 : n;
        double accS[64], accN[64];
        for (int64_t i = i0; i < i1; ++i) { accS[i - i0] = 0.0; accN[i - i0] = 0.0; }
        for (int64_t j = 0; j < n; ++j) {
            const double *e = op->E + j * n, *k = op->KN + j * n;
            double xs = negx[j], xn = x[j];
            for (int64_t i = i0; i < i1; ++i) accS[i - i0] += e[i] * xs;
            if (!planar)
                for (int64_t i = i0; i < i1; ++i) accN[i - i0] += k[i] * xn;
        }
        for (int64_t i = i0; i < i1; ++i) {
            TL[i] = accS[i - i0];
            if (!planar) SigProd[i] = accN[i - i0];
        }
    }
    free(negx);
}

/* Build E and K_Self from K_S as main() does (:54-63). E is n x n col-major. */
void orc_dense_prepare(int64_t n, const double *KS, double *E, double *KSelf) {
    for (int64_t i = 0; i < n; ++i) KSelf[i] = fabs(KS[i + i * n]);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i)
            E[i + j * n] = (i == j) ? 0.0 : KS[i + j * n] / KS[i + i * n];
}

static int dtmin_candidates(const orc_sim_in_t *in, const double *V, const double *Accel,
                            const int64_t *idx, int64_t cnt, double *Vmax, double *Dtmin) {
    double vm = -INFINITY, dm = INFINITY;
    int found = 0;
    for (int64_t k = 0; k < cnt; ++k) {
        int64_t i = idx[k];
        if (k == 0 || V[i] > vm) vm = V[i];
        double r = fabs(V[i] / Accel[i]);
        if (isnan(r) || r == 0.0) continue;
        if (!found || r < dm) dm = r;
        found = 1;
    }
    *Vmax = vm;
    *Dtmin = dm / in->dt_cut;
    return found;
}

int orc_sim_run(const orc_sim_in_t *in, orc_coupling_fn coupling, const void *ctx, orc_sim_out_t *out) {
    const int64_t n = in->n, L = in->n_loading;
    const double V0 = 1e-9, ConvergenceCrit = 1e-9;
    int status = ORC_OK;
    double Dt = 1.0;
    int64_t nbytes = sizeof(double) * n;
#define ALLOC(name) double *name = (double *)calloc((size_t)n, sizeof(double))
    ALLOC(Tau0); ALLOC(Omega); ALLOC(Friction0); ALLOC(Friction); ALLOC(Accel); ALLOC(Disp);
    ALLOC(DispOld); ALLOC(Sig); ALLOC(SigOld); ALLOC(Sig_i); ALLOC(SigProd); ALLOC(dSigN);
    ALLOC(dTau); ALLOC(dP); ALLOC(Dt_All); ALLOC(VOld); ALLOC(V); ALLOC(Theta); ALLOC(ThetaOld);
    ALLOC(FrictionOld); ALLOC(TL); ALLOC(TLd); ALLOC(Far);
#undef ALLOC
    int32_t *Switch = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int32_t *Instab = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int32_t *Flag = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int64_t *unstable = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t n_unstable = 0;

    for (int64_t i = 0; i < n; ++i) {
        Tau0[i] = in->sigma0[i] * in->friction_i[i];
        Omega[i] = sqrt(in->k_self[i] / in->mass[i]);
        if (in->a[i] - in->b[i] < 0.0) unstable[n_unstable++] = i;
    }
    int only_unstable = n_unstable > 0;
    if (!only_unstable) {
        n_unstable = n - L;
        for (int64_t i = 0; i < n_unstable; ++i) unstable[i] = i;
    }
    for (int64_t i = 0; i < n; ++i) {
        Friction0[i] = in->friction_i[i] - in->a[i] * log(in->v_i[i] / V0) -
                       in->b[i] * log(in->theta_i[i] * V0 / in->dc[i]);
        Friction[i] = in->friction_i[i];
        SigOld[i] = in->sigma0[i];
        VOld[i] = in->v_i[i];
        V[i] = VOld[i];
        Theta[i] = in->theta_i[i];
        ThetaOld[i] = Theta[i];
        FrictionOld[i] = Friction[i];
    }
    orc_step_params_t sp;
    sp.conv_crit = ConvergenceCrit; sp.v0 = V0; sp.mu = in->mu; sp.rho = in->rho;
    sp.switch_v = in->switch_v; sp.alpha_evo = in->alpha_evo; sp.evolution_dr = in->evolution_dr;

    double TOld = 0.0, T = 0.0, DtRef = 1.0;
    int64_t rec = 0;
    out->rhs_evals = 0; out->retry_passes = 0; out->stalls = 0; out->steps_done = 0;
    for (int64_t step = 1; step <= in->total_step; ++step) {
        memcpy(V, VOld, nbytes);
        for (int64_t i = 0; i < n; ++i) Switch[i] = VOld[i] > in->switch_v ? 1 : 0;
        double Vmax, Dtmin;
        if (!dtmin_candidates(in, V, Accel, unstable, n_unstable, &Vmax, &Dtmin)) {
            status = ORC_ERR_EMPTY_DTMIN;
            break;
        }
        if (Vmax > in->switch_v) DtRef = in->rupture_dt;
        else if (Dtmin > in->maximum_dt) DtRef = in->maximum_dt;
        else DtRef = Dtmin;
        if (DtRef <= in->rupture_dt) Dt = DtRef;
        else if (Dt < DtRef / 1.2) Dt = Dt * 1.2;
        else Dt = DtRef;

        for (int64_t i = 0; i < n; ++i) Far[i] = 0.0;
        coupling(ctx, DispOld, TLd, SigProd, in->planar);
        out->rhs_evals += 1;
        int Iteration = 0;
        for (;;) {
            Iteration += 1;
            for (int64_t i = 0; i < n; ++i) Dt_All[i] = Dt;
            if (in->ext_nt > 0) {
                double t = in->ext_at_told ? TOld : TOld + Dt;
                orc_interp_external(t, n, in->ext_nt, in->ext_t, in->ext_n, in->ext_s, in->ext_p, dSigN, dTau, dP);
            }
            for (int64_t i = 0; i < n; ++i) {
                Sig_i[i] = (in->planar ? 0.0 : SigProd[i]) + in->sigma0[i] + dSigN[i];
                if (Sig_i[i] < in->min_normal_stress) Sig_i[i] = in->min_normal_stress;
                TL[i] = Far[i] + TLd[i];
            }
            out->stalls += orc_solve_one_timestep(&sp, n, L, DispOld, FrictionOld, ThetaOld, VOld, Friction0,
                                                  in->a, in->b, in->dc, in->k_self, Dt, Omega, in->mass, Sig_i,
                                                  TL, Switch, in->loading_rate, dTau, SigOld, Tau0, V, Friction,
                                                  Disp, Theta, Sig, Dt_All, Instab, Flag, NULL);
            int ok = 1;
            for (int64_t i = 0; i < n; ++i) if (Flag[i] == 0) { ok = 0; break; }
            if (!ok) {
                double m = Dt_All[0];
                for (int64_t i = 1; i < n; ++i) if (Dt_All[i] < m) m = Dt_All[i];
                Dt = m;
                for (int64_t i = 0; i < n; ++i) {
                    if (Instab[i] == 3) Switch[i] = 1;
                    else if (Instab[i] == 4) { Switch[i] = 0; Dt = 1e-3; }
                }
                out->retry_passes += 1;
            } else break;
            if (Iteration > 100) { status = ORC_ERR_RETRIES; break; }
        }
        if (status != ORC_OK) break;
        {
            double m = Dt_All[0];
            for (int64_t i = 1; i < n; ++i) if (Dt_All[i] < m) m = Dt_All[i];
            Dt = m;
        }
        if (in->trace_told) {
            double vm = -INFINITY;
            for (int64_t i = 0; i < n - L; ++i) if (V[i] > vm) vm = V[i];
            in->trace_told[step - 1] = TOld;
            in->trace_dt[step - 1] = Dt;
            in->trace_vmax[step - 1] = vm;
        }
        if (in->record_step > 0 && step % in->record_step == 0 && rec < in->n_record) {
            if (in->hist_time) in->hist_time[rec] = TOld;
            if (in->hist_dt) in->hist_dt[rec] = Dt;
            for (int64_t i = 0; i < n; ++i) {
                if (in->hist_v) in->hist_v[rec + i * in->n_record] = V[i];
                if (in->hist_disp) in->hist_disp[rec + i * in->n_record] = Disp[i];
                if (in->hist_theta) in->hist_theta[rec + i * in->n_record] = Theta[i];
                if (in->hist_sigma) in->hist_sigma[rec + i * in->n_record] = Sig[i];
            }
            rec += 1;
        }
        T = TOld + Dt;
        for (int64_t i = 0; i < n; ++i) {
            Accel[i] = (V[i] - VOld[i]) / Dt;
            DispOld[i] = Disp[i];
            ThetaOld[i] = Theta[i];
            VOld[i] = V[i];
            FrictionOld[i] = Friction[i];
            SigOld[i] = Sig[i];
        }
        TOld = T;
        out->steps_done = step;
    }
    out->t_end = TOld;
    out->dt_last = Dt;
    out->n_recorded = rec;
    if (out->final_state) {
        double *f = out->final_state;
        memcpy(f + 0 * n, VOld, nbytes);
        memcpy(f + 1 * n, DispOld, nbytes);
        memcpy(f + 2 * n, ThetaOld, nbytes);
        memcpy(f + 3 * n, FrictionOld, nbytes);
        memcpy(f + 4 * n, SigOld, nbytes);
        memcpy(f + 5 * n, Accel, nbytes);
    }
    free(Tau0); free(Omega); free(Friction0); free(Friction); free(Accel); free(Disp); free(DispOld);
    free(Sig); free(SigOld); free(Sig_i); free(SigProd); free(dSigN); free(dTau); free(dP); free(Dt_All);
    free(VOld); free(V); free(Theta); free(ThetaOld); free(FrictionOld); free(TL); free(TLd); free(Far);
    free(Switch); free(Instab); free(Flag); free(unstable);
    return status;
}

static void dense_cb(const void *ctx, const double *x, double *TL, double *SigProd, int planar) {
    orc_dense_coupling((const orc_dense_t *)ctx, x, TL, SigProd, planar);
}

/* main() of Functions_RSFDFN3DMain_D.jl with a dense operator built from K_S. */
int orc_sim_dense(const orc_dense_t *op, const orc_sim_in_t *in, orc_sim_out_t *out) {
    return orc_sim_run(in, dense_cb, op, out);
}
