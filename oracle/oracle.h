/* ORACLE — test infrastructure only.  CPU FP64 restatement of the reference
 * (limkjae/Quake-DFN, Julia) used by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; the product library never
 * links or loads it.  Each function cites the reference lines it restates. */
#ifndef QDFN_ORACLE_H
#define QDFN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_ERR_RETRIES = 1, ORC_ERR_EMPTY_DTMIN = 2 };

/* okada.c */
void orc_dc3d(double alpha, int64_t n, const double *X, const double *Y, const double *Z,
              double depth, double dip, double al1, double al2, double aw1, double aw2,
              double disl1, double disl2, double disl3, double *U, double *iret);
void orc_stiffness_rect(const double *src, int64_t ns, int64_t lds, const double *rcv, int64_t nr,
                        int64_t ldr, double mu, double nu, double *KS, double *KN, int64_t ldk);

/* tde.c */
void orc_tdstress_hs(int64_t n, const double *X, const double *Y, const double *Z, const double *P1,
                     const double *P2, const double *P3, double Ss, double Ds, double Ts, double mu,
                     double lambda, double *Stress, double *Strain);
void orc_rotverts_unitvectors(int64_t n, double *P1, double *P2, double *P3, const double *rake,
                              double *nrm, double *strike, double *dip, double *slip);
void orc_stiffness_tri(int64_t ns, const double *P1, const double *P2, const double *P3,
                       const double *rake, int64_t nr, const double *ctr, const double *nrm,
                       const double *slip, double mu, double lambda, double *KS, double *KN, int64_t ldk);

/* solvers.c */
typedef struct {
    double conv_crit, disp_old, friction_old, theta_old, v_old, friction0, a, b, dc, v0, k_self;
    double omega, mass, mu, sigma, tl, rho, dtau, switch_v, alpha_evo, sigma_old, tau0;
    int evolution_dr;
} orc_elem_in_t;
typedef struct {
    double v, friction, disp, theta, sigma, dt;
    int instab, flag, iters, stall;
} orc_elem_out_t;
typedef struct {
    double conv_crit, v0, mu, rho, switch_v, alpha_evo;
    int evolution_dr;
} orc_step_params_t;

double orc_evolution_law(double ThetaOld, double Dt, double V, double Dc, double Alpha_Evo, double Sig,
                         double SigOld, double b, int Fast1_Slow0, int EvolutionDR);
void orc_solver_highv(const orc_elem_in_t *in, double Dt, orc_elem_out_t *out);
void orc_solver_lowv(const orc_elem_in_t *in, double Dt, orc_elem_out_t *out);
int64_t orc_solve_one_timestep(const orc_step_params_t *p, int64_t n, int64_t n_loading,
                               const double *DispOld, const double *FrictionOld, const double *ThetaOld,
                               const double *VOld, const double *Friction0, const double *a,
                               const double *b, const double *Dc, const double *K_Self, double Dt,
                               const double *Omega, const double *Mass, const double *Sig_i,
                               const double *TL, const int32_t *SolverSwitch, const double *LoadingRate,
                               const double *D_EffStress_Shear, const double *Sig_old, const double *Tau0,
                               double *V, double *Friction, double *Disp, double *Theta, double *Sig,
                               double *Dt_All, int32_t *Instab, int32_t *Flag, int32_t *Iters);
void orc_interp_external(double Time, int64_t n, int64_t nt, const double *tarr, const double *ExtN,
                         const double *ExtS, const double *ExtP, double *dSigN, double *dTau, double *dP);

/* sim.c */
typedef struct {
    int64_t n;
    const double *E;  /* n x n col-major, zero diagonal, K_ij/K_ii */
    const double *KN; /* n x n col-major */
} orc_dense_t;

typedef struct {
    int64_t n, n_loading, total_step, record_step, n_record;
    const double *mass, *a, *b, *dc, *theta_i, *v_i, *friction_i, *sigma0, *loading_rate, *k_self;
    double mu, rho, switch_v, dt_cut, rupture_dt, maximum_dt, min_normal_stress, alpha_evo;
    int evolution_dr, planar, ext_at_told;
    int64_t ext_nt;                      /* 0: no external stress */
    const double *ext_t, *ext_n, *ext_s, *ext_p;
    double *hist_time, *hist_dt, *hist_v, *hist_disp, *hist_theta, *hist_sigma; /* n_record x n */
    double *trace_told, *trace_dt, *trace_vmax;                                 /* total_step */
} orc_sim_in_t;

typedef struct {
    int64_t rhs_evals, retry_passes, stalls, steps_done, n_recorded;
    double t_end, dt_last;
    double *final_state; /* optional 6n: V, Disp, Theta, Friction, Sigma, Accel */
} orc_sim_out_t;

typedef void (*orc_coupling_fn)(const void *ctx, const double *x, double *TL, double *SigProd, int planar);
void orc_dense_prepare(int64_t n, const double *KS, double *E, double *KSelf);
void orc_dense_coupling(const orc_dense_t *op, const double *x, double *TL, double *SigProd, int planar);
int orc_sim_run(const orc_sim_in_t *in, orc_coupling_fn coupling, const void *ctx, orc_sim_out_t *out);
int orc_sim_dense(const orc_dense_t *op, const orc_sim_in_t *in, orc_sim_out_t *out);

/* hmat.c */
typedef struct {
    int64_t n, nblk;
    const int64_t *range; /* nblk x 4, row-major per block: src_i src_f rcv_i rcv_f (1-based, inclusive) */
    const int64_t *kind;  /* 0 dense (m x k col-major in a), 1 low-rank a (m x r) * b (r x k) */
    const int64_t *rank;
    const double *const *a;
    const double *const *b;
} orc_hmat_t;
void orc_hmat_matvec(const orc_hmat_t *h, const double *x, double *y);
/* OpenMP threads of the parallel loops (the products and the local solves);
 * results do not depend on it.  Returns the previous count. */
int orc_set_threads(int n);
int orc_sim_hmat(const orc_hmat_t *shear_loading, const orc_hmat_t *normal, const orc_sim_in_t *in,
                 orc_sim_out_t *out);
/* node arrays: begin/end 1-based inclusive, level, centre (n_nodes x 3 row-major),
 * diameter; loading_begin = 0 when there is no loading fault.  Returns the block
 * count (or -1 when cap is exceeded). */
int64_t orc_build_hierarchy(int64_t n_nodes, const int64_t *node_begin, const int64_t *node_end,
                            const int64_t *node_level, const double *node_ctr, const double *node_diam,
                            int64_t total_levels, double crit, int64_t loading_begin, int64_t loading_end,
                            int64_t cap, int64_t *range_SR, int64_t *admissible);

/* catalog.c */
int64_t orc_event_fragments(int64_t n_rec, int64_t n, const double *time, const double *disp, int64_t ld,
                            const double *center, const double *area, double mu, double crit, int64_t cap,
                            double *ftime, double *floc, double *fmom, int32_t *felem);
int64_t orc_merge_fragments(int64_t F, const double *ftime, const double *floc, const double *fmom, double mt,
                            double md, double *mtime, double *mloc, double *mmom, double *mw, int32_t *fid);

#ifdef __cplusplus
}
#endif
#endif
