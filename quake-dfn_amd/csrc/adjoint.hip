// Hand-written discrete adjoint of the device-resident step loop (main(),
// scripts/Functions_RSFDFN3DMain_D.jl:146-319, with the local solves of
// scripts/Functions_Solvers.jl:26-327): the gradient of
//     L = sum_o (Disp_K[obs_o] - target_o)^2 + w_t (t_first - t_target)^2
// with respect to the per-element a, b and Dc, for the K accepted steps a run
// has taken.  t_first is the onset of the first event (first accepted step
// whose max V over the non-loading elements reaches v_onset, interpolated in
// log10 V against the step end times).
//
// The tape is the run's own per-step device history (TOld, Dt, V, Disp,
// Theta of every accepted step).  The reverse sweep takes, for each step from
// the last to the first:
//   k_adj_gather     the step's input state from the history
//   coupling         TL = E (0 - Disp_old), SigProd = K_N Disp_old (the
//                    forward kernels, recomputed rather than stored)
//   k_adj_select     adjoint of the next step's Dt control (:162-187): the
//                    argmin element of |V / Accel| and the 1.2x growth rule
//   k_adj_elem       one lane per element: the local solve's converged V is
//                    the root of R(V; inputs) = V - G(V; inputs) (the secant
//                    iteration's fixed point, LowV or HighV as the forward
//                    chose); one forward-mode dual evaluation with 12
//                    directions gives R_V, R_p and the explicit output
//                    partials, dV/dp = -R_p / R_V (implicit function
//                    theorem), and the adjoints of the step inputs follow
//   k_adj_transpose  Disp_old adjoint += -E^T lamTL + K_N^T lamSigProd
//   k_adj_scalar     time and Dt adjoints (fixed-order block sums)
// Discrete decisions (solver branch of the accepted attempt, Dt rule, argmin,
// effective normal stress floor, number of retries) are those the forward
// took: a retried step's Dt is the selected Dt times 2^-j, and the run's
// history records both and each element's solver.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "dense.h"
#include "dual.h"
#include "qdfn_internal.h"
#include "sim.h"
#include "sim_device.h"

namespace qdfn {
namespace {

constexpr int kAdjBlock = 256;
constexpr int kNd = 12;  // dual directions: V, V_old, D_old, th_old, TL, sig, dtau, dt, a, b, Dc, f0
typedef ad::Dual<kNd> D12;

struct AdjScal {
    double lam_dt_next;  // adjoint of the Dt selected for the step after the current one
    double lam_t_new;    // adjoint of TOld + Dt of the current step
    double lam_dt_acc;   // adjoint of the current step's Dt (being accumulated)
    int32_t select_mismatch;
    int32_t pad;
};

struct AdjArgs {
    int64_t n, n_loading, k, rec_cap;
    // history (element-major: [rec + i * rec_cap])
    const double *h_time, *h_dt, *h_v, *h_d, *h_th;
    const int8_t* h_sw;
    // gathered step inputs / outputs (contiguous)
    double *v_new, *v_old, *d_old, *th_old, *tl, *sigprod;
    // adjoints: of step outputs (cur) and of step inputs (prev)
    double *lv_cur, *ld_cur, *lth_cur, *lv_prev, *ld_prev, *lth_prev, *pend, *ltl, *lsig;
    double *ga, *gb, *gdc, *gf0;
    double *part_dt, *part_t;  // per-block sums
    AdjScal* sc;
    int32_t* mismatch;
    // run constants
    const double *a, *b, *dc, *f0, *kself, *omega, *mass, *tau0, *sigma0, *lrate;
    const int8_t* cand;
    const double *v_init, *th_init;
    // state before the first step of the tape held in h_* (seg_d null: zero)
    const double *seg_v, *seg_d, *seg_th;
    int64_t nt;
    const double *ext_t, *ext_n, *ext_s;
    double mu, rho, switch_v, dt_cut, rupture_dt, maximum_dt, min_sig, v0;
    int32_t ext_at_told, planar;
};

// the inputs and the accepted V of the step at tape record k (record -1: the
// state before the tape's first step)
__global__ __launch_bounds__(kAdjBlock) void k_adj_gather(AdjArgs A) {
    const int64_t i = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= A.n) return;
    const int64_t k = A.k;
    A.v_new[i] = A.h_v[k + i * A.rec_cap];
    if (k == 0) {
        A.v_old[i] = A.seg_v[i];
        A.d_old[i] = A.seg_d ? A.seg_d[i] : 0.0;
        A.th_old[i] = A.seg_th[i];
    } else {
        A.v_old[i] = A.h_v[(k - 1) + i * A.rec_cap];
        A.d_old[i] = A.h_d[(k - 1) + i * A.rec_cap];
        A.th_old[i] = A.h_th[(k - 1) + i * A.rec_cap];
    }
    A.pend[i] = 0.0;
}

// external stress and its time slope at time t for element i (the segment
// InterpolateFromStressChange, Functions_Solvers.jl:465-540, interpolates on)
__device__ inline void ext_and_slope(const AdjArgs& A, double t, int64_t i, double& dsn, double& dtau, double& sn,
                                     double& ss) {
    dsn = dtau = sn = ss = 0.0;
    if (A.nt == 0) return;
    const int64_t nt = A.nt, n = A.n, idx = ext_index(A.ext_t, nt, t);
    ext_interp_at(A.ext_t, A.ext_n, A.ext_s, nt, n, idx, t, i, dsn, dtau);
    if (idx == 0) {
        sn = A.ext_n[i] / A.ext_t[0];
        ss = A.ext_s[i] / A.ext_t[0];
    } else if (idx < nt) {
        const double span = A.ext_t[idx] - A.ext_t[idx - 1];
        sn = (A.ext_n[idx * n + i] - A.ext_n[(idx - 1) * n + i]) / span;
        ss = (A.ext_s[idx * n + i] - A.ext_s[(idx - 1) * n + i]) / span;
    }
}

// adjoint of the Dt selection that produced the next step's Dt, from this
// step's V (v_new), its V_old and accepted Dt (:162-187)
__global__ __launch_bounds__(1024) void k_adj_select(AdjArgs A, double dt_acc, double dt_next_rec) {
    __shared__ double sv[16], sd[16];
    __shared__ int64_t si[16];
    double vmax = -INFINITY, dmin = INFINITY;
    int64_t arg = -1;
    for (int64_t i = threadIdx.x; i < A.n; i += 1024) {
        if (!A.cand[i]) continue;
        const double vn = A.v_new[i];
        vmax = vn > vmax ? vn : vmax;
        const double acc = (vn - A.v_old[i]) / dt_acc;
        const double r = fabs(vn / acc);
        if (!(isnan(r) || r == 0.0) && r < dmin) {
            dmin = r;
            arg = i;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(vmax, o, 64);
        const double od = __shfl_xor(dmin, o, 64);
        const int64_t oi = __shfl_xor(arg, o, 64);
        vmax = ov > vmax ? ov : vmax;
        if (od < dmin || (od == dmin && oi >= 0 && (arg < 0 || oi < arg))) {
            dmin = od;
            arg = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sv[w] = vmax;
        sd[w] = dmin;
        si[w] = arg;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 1; q < 16; ++q) {
        vmax = sv[q] > vmax ? sv[q] : vmax;
        if (sd[q] < dmin || (sd[q] == dmin && si[q] >= 0 && (arg < 0 || si[q] < arg))) {
            dmin = sd[q];
            arg = si[q];
        }
    }
    // the scalar block is produced by the previous kernel: accessed through
    // volatile so that it is read with vector loads, never from the scalar cache
    volatile AdjScal& S = *A.sc;
    const double lam = S.lam_dt_next;
    // select_dt (sim_device.h) with its branch recorded
    const double dtmin = dmin / A.dt_cut;
    double dtref;
    int via_min = 0;
    if (vmax > A.switch_v) dtref = A.rupture_dt;
    else if (dtmin > A.maximum_dt) dtref = A.maximum_dt;
    else {
        dtref = dtmin;
        via_min = 1;
    }
    double dt_next, l_min = 0.0, l_acc = 0.0;
    if (dtref <= A.rupture_dt) {
        dt_next = dtref;
        if (via_min) l_min = lam;
    } else if (dt_acc < dtref / 1.2) {
        dt_next = dt_acc * 1.2;
        l_acc = 1.2 * lam;
    } else {
        dt_next = dtref;
        if (via_min) l_min = lam;
    }
    if (dt_next != dt_next_rec) S.select_mismatch += 1;
    S.lam_dt_acc += l_acc;
    if (l_min != 0.0 && arg >= 0) {
        // r = |u| dt / |u - w| with u = V_new, w = V_old of the argmin element
        const double u = A.v_new[arg], wv = A.v_old[arg], du = u - wv;
        const double lr = l_min / A.dt_cut;
        const double su = u > 0.0 ? 1.0 : (u < 0.0 ? -1.0 : 0.0);
        const double sd2 = du > 0.0 ? 1.0 : (du < 0.0 ? -1.0 : 0.0);
        const double adu = fabs(du);
        A.lv_cur[arg] += lr * (su * dt_acc / adu - fabs(u) * dt_acc * sd2 / (du * du));
        A.pend[arg] += lr * (fabs(u) * dt_acc * sd2 / (du * du));
        S.lam_dt_acc += lr * (fabs(u) / adu);
    }
}

struct ElemC {
    double k, mass, tau0, eta, v0, omega, mu, rho;
};

// R = V - G(V; p), Disp and Theta after the step, LowV (Solver_LowV_D :185-327)
template <class S>
__device__ inline void lowv_eval(const S& V, const S& Vo, const S& Do, const S& tho, const S& TL, const S& sig,
                                 const S& dtau, const S& dt, const S& a, const S& b, const S& dc, const S& f0,
                                 const ElemC& c, S& R, S& D, S& th) {
    const S acc = (V - Vo) / dt;
    th = (tho + dt) / (1.0 + dt * V / dc);
    D = Do + (Vo + V) / 2.0 * dt;
    const S f = 1.0 / sig * (c.k * (TL - D) + c.tau0 - c.eta * V + dtau - c.mass * acc);
    const S vel = c.v0 * exp((f - f0 - b * log(c.v0 * th / dc)) / a);
    R = V - vel;
}

// the same for the oscillator solve (Solver_HighV_D :26-182)
template <class S>
__device__ inline void highv_eval(const S& V, const S& Vo, const S& Do, const S& tho, const S& TL, const S& sig,
                                  const S& dtau, const S& dt, const S& a, const S& b, const S& dc, const S& f0,
                                  const ElemC& c, S& R, S& D, S& th) {
    const S ct = V / dc;
    th = (tho - 1.0 / ct) * exp(-V / dc * dt) + 1.0 / ct;
    const S eos = c.mu / 2.0 / ::sqrt(c.mu / c.rho) / sig;
    const S frd = f0 + a * log(V / c.v0) + b * log(th * c.v0 / dc) + eos * V - dtau / sig;
    const S F = TL + (c.tau0 - frd * sig) / c.k;
    const S Dosc = (Do - F) * cos(c.omega * dt) + (Vo / c.omega) * sin(c.omega * dt) + F;
    R = V - ((Dosc - Do) / dt * 2.0 - Vo);
    D = Do + (Vo + V) / 2.0 * dt;
}

__global__ __launch_bounds__(kAdjBlock) void k_adj_elem(AdjArgs A, double t_old, double dt) {
    __shared__ double sdt[kAdjBlock / 64], st_[kAdjBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    double l_dt = 0.0, l_t = 0.0;
    if (i < A.n) {
        const double lV = A.lv_cur[i], lD = A.ld_cur[i], lTh = A.lth_cur[i];
        const double t_ext = A.ext_at_told ? t_old : t_old + dt;
        double dsn, dtau, sn, ss;
        ext_and_slope(A, t_ext, i, dsn, dtau, sn, ss);
        const double sig_in = (A.planar ? 0.0 : A.sigprod[i]) + A.sigma0[i] + dsn;
        const bool floored = sig_in < A.min_sig;
        const double sig = floored ? A.min_sig : sig_in;
        double lp[kNd];
        for (int q = 0; q < kNd; ++q) lp[q] = 0.0;
        const double vrec = A.v_new[i], drec = A.h_d[A.k + i * A.rec_cap], threc = A.h_th[A.k + i * A.rec_cap];
        const bool high = A.h_sw[A.k + i * A.rec_cap] != 0;  // the solver of the accepted attempt
        if (i < A.n - A.n_loading && !high && vrec == 0.0 && drec == A.d_old[i] &&
            (threc == A.th_old[i] || threc == A.th_old[i] + dt)) {
            // Solver_LowV_D's stalled element (Functions_Solvers.jl:210-222, 276-290):
            // V = 0 with Disp kept, Theta kept (V_old = 0) or aged by Dt
            lp[1] = A.v_old[i] == 0.0 ? lV : 0.0;
            lp[2] = lD;
            lp[3] = lTh;
            if (threc != A.th_old[i]) lp[7] = lTh;
        } else if (i < A.n - A.n_loading) {
            ElemC c;
            c.k = A.kself[i];
            c.mass = A.mass[i];
            c.tau0 = A.tau0[i];
            c.eta = A.mu / 2.0 / ::sqrt(A.mu / A.rho);
            c.v0 = A.v0;
            c.omega = A.omega[i];
            c.mu = A.mu;
            c.rho = A.rho;
            const double vo = A.v_old[i];
            const D12 V = D12::var(vrec, 0), Vo = D12::var(vo, 1), Do = D12::var(A.d_old[i], 2);
            const D12 tho = D12::var(A.th_old[i], 3), TL = D12::var(0.0 + A.tl[i], 4), sg = D12::var(sig, 5);
            const D12 dta = D12::var(dtau, 6), dtv = D12::var(dt, 7), a = D12::var(A.a[i], 8);
            const D12 b = D12::var(A.b[i], 9), dc = D12::var(A.dc[i], 10), f0 = D12::var(A.f0[i], 11);
            D12 R, Dn, th;
            if (high) highv_eval(V, Vo, Do, tho, TL, sg, dta, dtv, a, b, dc, f0, c, R, Dn, th);
            else lowv_eval(V, Vo, Do, tho, TL, sg, dta, dtv, a, b, dc, f0, c, R, Dn, th);
            // consistency with the recorded step outputs
            if (!(fabs(Dn.v - drec) <= 1e-9 * (fabs(drec) + 1e-12) &&
                  fabs(th.v - threc) <= 1e-9 * (fabs(threc) + 1e-300)))
                atomicAdd(A.mismatch, 1);
            // implicit function theorem: dV/dp = -R_p / R_V
            const double rv = R.d[0];
            for (int q = 1; q < kNd; ++q) {
                const double dvdp = -R.d[q] / rv;
                lp[q] = lV * dvdp + lD * (Dn.d[q] + Dn.d[0] * dvdp) + lTh * (th.d[q] + th.d[0] * dvdp);
            }
        } else {
            // loading element: V = V_const, Disp += V_const Dt, Theta unchanged
            lp[2] = lD;
            lp[3] = lTh;
            lp[7] = A.lrate[i] * lD;
        }
        A.lv_prev[i] = lp[1] + A.pend[i];
        A.ld_prev[i] = lp[2];
        A.lth_prev[i] = lp[3];
        A.ltl[i] = lp[4];
        const double lsig_in = floored ? 0.0 : lp[5];
        A.lsig[i] = A.planar ? 0.0 : lsig_in;
        l_dt = lp[7];
        // time dependence of the external tables
        l_t = lsig_in * sn + lp[6] * ss;
        A.ga[i] += lp[8];
        A.gb[i] += lp[9];
        A.gdc[i] += lp[10];
        A.gf0[i] += lp[11];
    }
    for (int o = 32; o > 0; o >>= 1) {
        l_dt += __shfl_xor(l_dt, o, 64);
        l_t += __shfl_xor(l_t, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sdt[threadIdx.x >> 6] = l_dt;
        st_[threadIdx.x >> 6] = l_t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s1 = 0.0, s2 = 0.0;
        for (int q = 0; q < kAdjBlock / 64; ++q) {
            s1 += sdt[q];
            s2 += st_[q];
        }
        A.part_dt[blockIdx.x] = s1;
        A.part_t[blockIdx.x] = s2;
    }
}

// ld_prev[j] += sum_i (-E[i, j] ltl[i] + K_N[i, j] lsig[i]): one workgroup per column
template <class T>
__global__ __launch_bounds__(256) void k_adj_transpose(const T* __restrict__ E, const T* __restrict__ KN, int64_t ld,
                                                       int64_t n, int32_t planar, int32_t tiled, DenseTiles lay,
                                                       const double* __restrict__ ltl,
                                                       const double* __restrict__ lsig, double* __restrict__ ld_prev) {
    __shared__ double sh[4];
    const int64_t j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const int64_t o = tiled ? lay.index(i, j) : i + j * ld;
        s += -(double)E[o] * ltl[i];
        if (!planar) s += (double)KN[o] * lsig[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) ld_prev[j] += ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// time and Dt adjoints of the step; hands lam_dt / lam_t to the step before
__global__ void k_adj_scalar(AdjArgs A, int64_t nblk, double retry_factor) {
    if (threadIdx.x != 0) return;
    // block sums and the scalar block come from earlier kernels: volatile
    // (vector) loads, never the scalar cache
    const volatile double* pdt = A.part_dt;
    const volatile double* pt = A.part_t;
    double sdt = 0.0, st = 0.0;
    for (int64_t q = 0; q < nblk; ++q) {
        sdt += pdt[q];
        st += pt[q];
    }
    volatile AdjScal& S = *A.sc;
    const double lam_t_old = S.lam_t_new + st;
    double lam_dt = S.lam_dt_acc + S.lam_t_new + sdt + (A.ext_at_told ? 0.0 : st);
    // the accepted Dt is the selected one times 2^-j (j halvings over the
    // retries); a retry at the fixed 1e-3 (retry_factor 0) cuts the chain
    S.lam_dt_next = retry_factor * lam_dt;
    S.lam_t_new = lam_t_old;
    S.lam_dt_acc = 0.0;
}

__global__ void k_adj_add(double* p, int64_t i, double v, AdjScal* sc, double lam_t) {
    if (threadIdx.x != 0) return;
    if (p) *(volatile double*)(p + i) += v;
    if (sc) *(volatile double*)&sc->lam_t_new += lam_t;
}

// out[0] = max_i |lam[i]| (one 1,024-thread workgroup; the adjoint sweep's growth record)
__global__ __launch_bounds__(1024) void k_adj_maxabs(const double* __restrict__ lam, int64_t n, double* __restrict__ out) {
    __shared__ double sh[16];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(lam[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int k = 1; k < 16; ++k) r = fmax(r, sh[k]);
        out[0] = r;
    }
}

// gradient of f0 = Friction_i - a ln(V_i / V0) - b ln(Theta_i V0 / Dc) into a, b, Dc
__global__ __launch_bounds__(kAdjBlock) void k_adj_f0(AdjArgs A) {
    const int64_t i = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= A.n) return;
    const double l = A.gf0[i];
    A.ga[i] += l * -::log(A.v_init[i] / A.v0);
    A.gb[i] += l * -::log(A.th_init[i] * A.v0 / A.dc[i]);
    A.gdc[i] += l * (A.b[i] / A.dc[i]);
}

}  // namespace
}  // namespace qdfn

using namespace qdfn;

namespace {
// the replay run of checkpointed recomputation, destroyed on every return
struct ReplayRun {
    qdfn_sim* r = nullptr;
    ~ReplayRun() { qdfn_sim_destroy(r); }
};
}  // namespace

extern "C" int qdfn_sim_adjoint(qdfn_sim_t* s, const double* v_i, const double* theta_i, int64_t n_obs,
                                const int64_t* obs, const double* d_target, double w_time, double t_target,
                                double v_onset, double* loss, double* grad_a, double* grad_b, double* grad_dc,
                                qdfn_adjoint_info* info) {
    if (!s || !s->op || !v_i || !theta_i || n_obs < 0 || (n_obs > 0 && (!obs || !d_target)) || !loss ||
        !grad_a || !grad_b || !grad_dc)
        return fail(QDFN_ERR_ARG, "qdfn_sim_adjoint: bad arguments (dense run required)");
    const qdfn_sim_params& P = s->P;
    const int64_t n = P.n;
    for (int64_t q = 0; q < n_obs; ++q)
        if (obs[q] < 0 || obs[q] >= n) return fail(QDFN_ERR_ARG, "qdfn_sim_adjoint: observation index out of range");
    if (P.alpha_evo != 0.0 || P.evolution_dr != 1)
        return fail(QDFN_ERR_ARG, "qdfn_sim_adjoint: aging law with alpha = 0 only");
    QDFN_HIP(hipStreamSynchronize(s->stream));
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    const SimState st = s->h_st;
    if (st.status != QDFN_OK) return fail(QDFN_ERR_STATE, "qdfn_sim_adjoint: the run is in an error state");
    // tape: the run's own history of every step (record_step 1), or the
    // checkpoints of qdfn_sim_run_checkpointed, each segment recomputed from
    // its checkpoint into a replay run just before its backward sweep
    const bool ckm = s->ck_every > 0;
    const int64_t K = st.step, C = s->ck_every;
    if (ckm) {
        if (K < 1 || (int64_t)s->ck_state.size() != (K + C - 1) / C || st.ntrace != K)
            return fail(QDFN_ERR_STATE, "qdfn_sim_adjoint: the run was advanced after qdfn_sim_run_checkpointed");
    } else if (P.record_step != 1 || st.rec != st.step || st.step < 1) {
        return fail(QDFN_ERR_STATE, "qdfn_sim_adjoint: the run must record every step from its start "
                                    "(or be made by qdfn_sim_run_checkpointed)");
    }
    const int64_t cap = P.record_capacity;
    hipStream_t strm = s->stream;
    const SimArrays& SA = s->A;
    // per-step scalars and the V / Disp rows needed for the loss
    std::vector<double> h_time(K), h_dt(K), h_dtsel(K), dK(n), retry_factor(K);
    const int64_t nn = n - P.n_loading;
    std::vector<double> vmax(K, -INFINITY);
    std::vector<int64_t> vmax_arg(K, -1);
    if (ckm) {
        QDFN_HIP(hipMemcpy(h_time.data(), SA.t_told, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(h_dt.data(), SA.t_dt, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(h_dtsel.data(), SA.t_dtsel, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(vmax.data(), SA.t_vmax, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(dK.data(), SA.D + (int64_t)st.parity * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        QDFN_HIP(hipMemcpy(h_time.data(), SA.h_time, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(h_dt.data(), SA.h_dt, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(h_dtsel.data(), SA.h_dtsel, sizeof(double) * K, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy2D(dK.data(), sizeof(double), SA.h_d + (K - 1), sizeof(double) * cap, sizeof(double), n,
                             hipMemcpyDeviceToHost));
        // vmax per step over the non-loading elements and its argmax (onset time)
        std::vector<double> hv((size_t)cap * nn);
        QDFN_HIP(hipMemcpy(hv.data(), SA.h_v, sizeof(double) * cap * nn, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nn; ++i)
            for (int64_t k = 0; k < K; ++k) {
                const double v = hv[k + i * cap];
                if (v > vmax[k]) {
                    vmax[k] = v;
                    vmax_arg[k] = i;
                }
            }
    }
    int64_t retried = 0;
    for (int64_t k = 0; k < K; ++k) {
        const double q = h_dt[k] / h_dtsel[k];
        int ex = 0;
        const double m = frexp(q, &ex);
        retry_factor[k] = (m == 0.5 && ex <= 1) ? q : 0.0;  // exact power of two (halvings), else 1e-3 reset
        retried += h_dt[k] != h_dtsel[k];
    }
    double L = 0.0;
    std::vector<double> lam_dK(n, 0.0);
    for (int64_t q = 0; q < n_obs; ++q) {
        const double r = dK[obs[q]] - d_target[q];
        L += r * r;
        lam_dK[obs[q]] += 2.0 * r;
    }
    // onset: first k with vmax >= v_onset; t = te[k-1] + f (te[k] - te[k-1])
    int64_t kon = -1;
    for (int64_t k = 0; k < K; ++k)
        if (vmax[k] >= v_onset) {
            kon = k;
            break;
        }
    double t_first = NAN;
    // a seed's element (the step's argmax of V) is found from the tape when
    // the backward sweep reaches its step (i = -2 until then; -1: time only)
    struct Seed { int64_t k; int64_t i; double lv; double lt; };
    std::vector<Seed> seeds;
    if (w_time > 0.0 && kon >= 1) {
        const double te0 = h_time[kon - 1] + h_dt[kon - 1], te1 = h_time[kon] + h_dt[kon];
        const double l0 = log10(vmax[kon - 1]), l1 = log10(vmax[kon]), lt = log10(v_onset);
        const double f = l1 != l0 ? (lt - l0) / (l1 - l0) : 1.0;
        t_first = te0 + f * (te1 - te0);
        const double r = t_first - t_target;
        L += w_time * r * r;
        const double g = 2.0 * w_time * r;
        double df_dl0 = 0.0, df_dl1 = 0.0;
        if (l1 != l0) {
            df_dl0 = (lt - l1) / ((l1 - l0) * (l1 - l0));
            df_dl1 = -(lt - l0) / ((l1 - l0) * (l1 - l0));
        }
        const double ln10 = log(10.0);
        seeds.push_back({kon - 1, ckm ? -2 : vmax_arg[kon - 1], g * (te1 - te0) * df_dl0 / (vmax[kon - 1] * ln10),
                         g * (1.0 - f)});
        seeds.push_back({kon, ckm ? -2 : vmax_arg[kon], g * (te1 - te0) * df_dl1 / (vmax[kon] * ln10), g * f});
    } else if (w_time > 0.0 && kon == 0) {
        t_first = h_time[0] + h_dt[0];
        const double r = t_first - t_target;
        L += w_time * r * r;
        seeds.push_back({0, -1, 0.0, 2.0 * w_time * r});
    }
    // device work space
    DevBuf vec, grads, parts, scal, mm, init, lnorm;
    QDFN_HIP(lnorm.alloc(sizeof(double) * (K + 1)));
    const int64_t nblk = (n + kAdjBlock - 1) / kAdjBlock;
    if (vec.alloc(sizeof(double) * 15 * n) || grads.alloc(sizeof(double) * 4 * n) ||
        parts.alloc(sizeof(double) * 2 * nblk) || scal.alloc(sizeof(AdjScal)) || mm.alloc(sizeof(int32_t)) ||
        init.alloc(sizeof(double) * 2 * n))
        return fail(QDFN_ERR_NOMEM, "qdfn_sim_adjoint: work space");
    QDFN_HIP(hipMemset(vec.p, 0, sizeof(double) * 15 * n));
    QDFN_HIP(hipMemset(grads.p, 0, sizeof(double) * 4 * n));
    QDFN_HIP(hipMemset(scal.p, 0, sizeof(AdjScal)));
    QDFN_HIP(hipMemset(mm.p, 0, sizeof(int32_t)));
    QDFN_HIP(hipMemcpy(init.p, v_i, sizeof(double) * n, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(init.as<double>() + n, theta_i, sizeof(double) * n, hipMemcpyHostToDevice));
    double* w = vec.as<double>();
    AdjArgs A{};
    A.n = n;
    A.n_loading = P.n_loading;
    A.rec_cap = cap;
    A.h_time = SA.h_time; A.h_dt = SA.h_dt; A.h_v = SA.h_v; A.h_d = SA.h_d; A.h_th = SA.h_th; A.h_sw = SA.h_sw;
    A.v_new = w; A.v_old = w + n; A.d_old = w + 2 * n; A.th_old = w + 3 * n; A.tl = w + 4 * n;
    A.sigprod = w + 5 * n;
    double* lbuf[2][3] = {{w + 6 * n, w + 7 * n, w + 8 * n}, {w + 9 * n, w + 10 * n, w + 11 * n}};
    A.pend = w + 12 * n; A.ltl = w + 13 * n; A.lsig = w + 14 * n;
    double* g = grads.as<double>();
    A.ga = g; A.gb = g + n; A.gdc = g + 2 * n; A.gf0 = g + 3 * n;
    A.part_dt = parts.as<double>(); A.part_t = parts.as<double>() + nblk;
    A.sc = scal.as<AdjScal>();
    A.mismatch = mm.as<int32_t>();
    A.a = SA.a; A.b = SA.b; A.dc = SA.dc; A.f0 = SA.f0; A.kself = SA.kself; A.omega = SA.omega; A.mass = SA.mass;
    A.tau0 = SA.tau0; A.sigma0 = SA.sigma0; A.lrate = SA.lrate; A.cand = SA.cand;
    A.v_init = init.as<double>(); A.th_init = init.as<double>() + n;
    A.seg_v = A.v_init; A.seg_d = nullptr; A.seg_th = A.th_init;
    A.nt = SA.nt; A.ext_t = SA.ext_t; A.ext_n = SA.ext_n; A.ext_s = SA.ext_s;
    A.mu = SA.mu; A.rho = SA.rho; A.switch_v = SA.switch_v; A.dt_cut = SA.dt_cut; A.rupture_dt = SA.rupture_dt;
    A.maximum_dt = SA.maximum_dt; A.min_sig = SA.min_sig; A.v0 = SA.v0;
    A.ext_at_told = SA.ext_at_told; A.planar = SA.planar;
    // seed: adjoint of Disp after the last step
    QDFN_HIP(hipMemcpy(lbuf[0][1], lam_dK.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    double seed_norm = 0.0;
    for (double v : lam_dK) seed_norm = std::max(seed_norm, std::fabs(v));
    hipEvent_t e0, e1;
    QDFN_HIP(hipEventCreate(&e0));
    QDFN_HIP(hipEventCreate(&e1));
    QDFN_HIP(hipEventRecord(e0, strm));
    const qdfn_dense* op = s->op;
    const unsigned gb = (unsigned)nblk;
    int cur = 0;
    ReplayRun rp;
    int64_t seg_lo = K;  // first step of the recomputed segment (checkpoint mode)
    if (ckm) {
        rp.r = sim_clone(s, C);
        if (!rp.r) return QDFN_ERR_NOMEM;
    }
    for (int64_t k = K - 1; k >= 0; --k) {
        if (ckm && k < seg_lo) {
            // recompute segment j from its checkpoint: every step recorded
            const int64_t j = k / C, cnt = std::min(C, K - j * C);
            seg_lo = j * C;
            QDFN_HIP(hipStreamSynchronize(strm));
            int rc = sim_restore(s, j, rp.r);
            if (rc == QDFN_OK) rc = qdfn_sim_run(rp.r, cnt, 0, 32);
            if (rc) return rc;
            // the replay must retrace the forward's steps exactly
            std::vector<double> rdt(cnt), rt(cnt);
            QDFN_HIP(hipMemcpy(rdt.data(), rp.r->A.h_dt, sizeof(double) * cnt, hipMemcpyDeviceToHost));
            QDFN_HIP(hipMemcpy(rt.data(), rp.r->A.h_time, sizeof(double) * cnt, hipMemcpyDeviceToHost));
            for (int64_t q = 0; q < cnt; ++q)
                if (rdt[q] != h_dt[seg_lo + q] || rt[q] != h_time[seg_lo + q])
                    return fail(QDFN_ERR_STATE, "qdfn_sim_adjoint: the recomputed segment departs from the forward run");
            for (Seed& sd : seeds)
                if (sd.i == -2 && sd.k >= seg_lo && sd.k < seg_lo + cnt) {
                    std::vector<double> vrow(nn);
                    QDFN_HIP(hipMemcpy2D(vrow.data(), sizeof(double), rp.r->A.h_v + (sd.k - seg_lo), sizeof(double) * C,
                                         sizeof(double), nn, hipMemcpyDeviceToHost));
                    double vm = -INFINITY;
                    for (int64_t i = 0; i < nn; ++i)
                        if (vrow[i] > vm) {
                            vm = vrow[i];
                            sd.i = i;
                        }
                }
            const SimArrays& RA = rp.r->A;
            A.rec_cap = C;
            A.h_time = RA.h_time; A.h_dt = RA.h_dt; A.h_v = RA.h_v; A.h_d = RA.h_d; A.h_th = RA.h_th;
            A.h_sw = RA.h_sw;
            const double* ck = s->ckpt.as<double>() + j * 11 * n;
            const int64_t po = (int64_t)s->ck_state[j].parity * n;
            A.seg_v = ck + po;
            A.seg_d = ck + 2 * n + po;
            A.seg_th = ck + 4 * n + po;
        }
        A.k = ckm ? k - seg_lo : k;
        A.lv_cur = lbuf[cur][0]; A.ld_cur = lbuf[cur][1]; A.lth_cur = lbuf[cur][2];
        A.lv_prev = lbuf[cur ^ 1][0]; A.ld_prev = lbuf[cur ^ 1][1]; A.lth_prev = lbuf[cur ^ 1][2];
        hipLaunchKernelGGL(k_adj_gather, dim3(gb), dim3(kAdjBlock), 0, strm, A);
        for (const Seed& sd : seeds)
            if (sd.k == k)
                hipLaunchKernelGGL(k_adj_add, dim3(1), dim3(64), 0, strm, sd.i >= 0 ? A.lv_cur : nullptr, sd.i, sd.lv,
                                   A.sc, sd.lt);
        // TL = E (0 - Disp_old), SigProd = K_N Disp_old, as the forward
        int rc = launch_coupling(op, A.d_old, CouplingGate{}, strm);
        if (rc == QDFN_OK) rc = launch_coupling_finish(op, A.tl, A.sigprod, nullptr, strm);
        if (rc) return rc;
        if (k < K - 1)
            hipLaunchKernelGGL(k_adj_select, dim3(1), dim3(1024), 0, strm, A, h_dt[k], h_dtsel[k + 1]);
        hipLaunchKernelGGL(k_adj_elem, dim3(gb), dim3(kAdjBlock), 0, strm, A, h_time[k], h_dt[k]);
        if (op->prec == QDFN_PREC_FP64)
            hipLaunchKernelGGL(k_adj_transpose<double>, dim3((unsigned)n), dim3(256), 0, strm, op->E.as<double>(),
                               op->KN.as<double>(), op->ld, n, op->normal_zero, op->tiled, dense_tiles(op), A.ltl,
                               A.lsig, A.ld_prev);
        else
            hipLaunchKernelGGL(k_adj_transpose<float>, dim3((unsigned)n), dim3(256), 0, strm, op->E.as<float>(),
                               op->KN.as<float>(), op->ld, n, op->normal_zero, op->tiled, dense_tiles(op), A.ltl,
                               A.lsig, A.ld_prev);
        hipLaunchKernelGGL(k_adj_scalar, dim3(1), dim3(64), 0, strm, A, nblk, retry_factor[k]);
        hipLaunchKernelGGL(k_adj_maxabs, dim3(1), dim3(1024), 0, strm, A.ld_prev, n, lnorm.as<double>() + k);
        QDFN_HIP(hipGetLastError());
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_adj_f0, dim3(gb), dim3(kAdjBlock), 0, strm, A);
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipEventRecord(e1, strm));
    QDFN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    QDFN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    QDFN_HIP(hipMemcpy(grad_a, A.ga, sizeof(double) * n, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(grad_b, A.gb, sizeof(double) * n, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(grad_dc, A.gdc, sizeof(double) * n, hipMemcpyDeviceToHost));
    AdjScal hs;
    int32_t mism = 0;
    QDFN_HIP(hipMemcpy(&hs, A.sc, sizeof(AdjScal), hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(&mism, A.mismatch, sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> hn(K);
    QDFN_HIP(hipMemcpy(hn.data(), lnorm.p, sizeof(double) * K, hipMemcpyDeviceToHost));
    // growth of the Disp adjoint over the backward sweep (k = K - 1 down to 0)
    double growth = 0.0;
    int64_t growth_step = -1;
    const double ref = seed_norm > 0.0 ? seed_norm : 1.0;
    for (int64_t k = K - 1; k >= 0; --k) {
        const double gk = hn[k] / ref;
        if (gk > growth || std::isnan(gk)) growth = std::isnan(gk) ? INFINITY : gk;
        if (growth_step < 0 && !(gk <= 1e8)) growth_step = k;
    }
    *loss = L;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->steps = K;
        info->output_mismatches = mism;
        info->select_mismatches = hs.select_mismatch;
        info->onset_step = kon;
        info->retried_steps = retried;
        info->t_first = t_first;
        info->ms_backward = ms;
        info->lambda_growth = growth;
        info->growth_step = growth_step;
    }
    return QDFN_OK;
}
