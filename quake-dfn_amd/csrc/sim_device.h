// Device side of the earthquake-cycle loop (main() of
// scripts/Functions_RSFDFN3DMain_D.jl:2-324, main_H's identical control logic
// in scripts/Functions_RSFDFN3DMain_H.jl:180-383): run state, per-element
// attempt and the retry-or-accept decision, shared by the single-run kernels
// (sim.hip) and the batched ensemble kernels (ensemble.hip).
#pragma once
#include <math.h>

#include "friction_device.h"
#include "qdfn_internal.h"

namespace qdfn {

struct SimState {
    double dt, told;
    double t_end;           // > 0: the run also stops once TOld reaches it
    double dt_first;        // Dt of the current step's first attempt (the selected Dt)
    int64_t step, target, rhs_evals, retry_passes, stalls, rec, ntrace;
    int32_t need_coupling;  // gate for the coupling kernels
    int32_t iteration;      // attempt number within the current step (1-based)
    int32_t status;         // QDFN_OK or error code
    int32_t done;           // 1: requested steps reached or error
    int32_t parity;         // buffer holding the "Old" state
    int32_t pad;
};

// The fields of the run state that kernels read, loaded coherently (the
// previous kernel on the stream wrote them; see ctrl_load).
__device__ __forceinline__ SimState load_state(const SimState* st) {
    SimState S;
    S.dt = ctrl_load(&st->dt);
    S.told = ctrl_load(&st->told);
    S.t_end = ctrl_load(&st->t_end);
    S.dt_first = ctrl_load(&st->dt_first);
    S.step = ctrl_load(&st->step);
    S.target = ctrl_load(&st->target);
    S.rhs_evals = ctrl_load(&st->rhs_evals);
    S.retry_passes = ctrl_load(&st->retry_passes);
    S.stalls = 0;  // only ever updated by atomics
    S.rec = ctrl_load(&st->rec);
    S.ntrace = ctrl_load(&st->ntrace);
    S.need_coupling = 0;  // read by the coupling kernels only (through their gate)
    S.iteration = ctrl_load(&st->iteration);
    S.status = 0;  // read by the host only
    S.done = ctrl_load(&st->done);
    S.parity = ctrl_load(&st->parity);
    S.pad = 0;
    return S;
}

// Per-workgroup partial results of one attempt.
struct BlockPart {
    int32_t ok, any4, found, pad;
    double mindt, vmax, vmax_next, dtmin_next;
};

// Device arrays of a run.  Double-buffered state arrays are [2][n].
struct SimArrays {
    int64_t n, n_loading;
    // constants
    const double *a, *b, *dc, *f0, *kself, *omega, *mass, *tau0, *sigma0, *lrate;
    const int8_t* cand;  // Dt-control candidate set (unstable patch or all non-loading)
    // double-buffered state
    double *V, *D, *TH, *F, *SIG;
    // single
    double *accel, *tl, *sigprod;
    int32_t* sw;  // SolverSwitch to use if the next attempt retries this step
    int32_t* ran_log;  // per-attempt "coupling ran" marks of a timed batch
    BlockPart* parts;
    unsigned int* arrive;
    // external stress (nt == 0: none); tables sample-major: ext_n[k * n + i]
    int64_t nt;
    const double *ext_t, *ext_n, *ext_s;
    // histories
    int64_t rec_cap, trace_cap, record_step;
    double *h_time, *h_dt, *h_v, *h_d, *h_th, *h_sig;
    double* h_dtsel;  // selected Dt of each recorded step (before any retry)
    int8_t* h_sw;     // solver of each element in the recorded step's accepted attempt (1: HighV)
    double *t_told, *t_dt, *t_vmax, *t_dtsel;  // per-step trace (t_dtsel: the step's selected Dt)
    // H operator (null otherwise): its product's unit slots, hslots x [2][n]
    // row sums that this kernel adds in slot order (TL = (-sum_S) / K_S,ii,
    // SigmaProd = sum_N, the loading form of the product)
    const double* hpart;
    const double* hkdiag;
    int64_t hslots;
    // dense operator (null otherwise): its product's split partial sums,
    // dsplits x dld per matrix, that this kernel adds in the order of the
    // coupling finish (splits k = q, q + 4, ... summed per quarter q, then the
    // quarters in order): TL = -sum_S, SigmaProd = sum_N (0 when the
    // operator's K_N is zero)
    const double *dpartS, *dpartN;
    int64_t dld;
    int32_t dsplits, dnormal_zero;
    // parameters
    double mu, rho, switch_v, dt_cut, rupture_dt, maximum_dt, min_sig, alpha_evo, conv, v0;
    int32_t dr, planar, ext_at_told, pad;
};

// The reference's scan for InterpolateFromStressChange (Functions_Solvers.jl:
// 465-540): the 1-based index of the last sample time not after `time` (0 if
// none).  The times are non-decreasing (qdfn_sim_set_external checks), so a
// binary search gives the same index: the number of samples <= time.  A NaN
// time compares false everywhere and gives 0, as the scan does.
__device__ inline int64_t ext_index(const double* __restrict__ ext_t, int64_t nt, double time) {
    int64_t lo = 0, hi = nt;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (time >= ext_t[mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// InterpolateFromStressChange for element i of n at sample index idx
// (ext_index).  Tables are sample-major on the device (ext[k * n + i]), so
// the lanes of a wavefront read consecutive addresses.
__device__ inline void ext_interp_at(const double* __restrict__ ext_t, const double* __restrict__ ext_n,
                                     const double* __restrict__ ext_s, int64_t nt, int64_t n, int64_t idx,
                                     double time, int64_t i, double& dsn, double& dtau) {
    if (idx == 0) {
        const double td = time / ext_t[0];
        dsn = td * (ext_n[i] - 0.0) + 0.0;
        dtau = td * (ext_s[i] - 0.0) + 0.0;
    } else if (idx == nt) {
        dsn = ext_n[(idx - 1) * n + i];
        dtau = ext_s[(idx - 1) * n + i];
    } else {
        const double td = time - ext_t[idx - 1];
        const double span = ext_t[idx] - ext_t[idx - 1];
        const double ln = ext_n[(idx - 1) * n + i], un = ext_n[idx * n + i];
        const double ls = ext_s[(idx - 1) * n + i], us = ext_s[idx * n + i];
        dsn = td / span * (un - ln) + ln;
        dtau = td / span * (us - ls) + ls;
    }
}

// threads per step workgroup: a single run's k_step (its decision in
// k_decide; smaller workgroups spread the latency-bound solves over more CUs)
// and the ensemble's k_ens_step (decision by the last workgroup of each run);
// k_decide's one workgroup
constexpr int kRunStepBlock = 64;
constexpr int kEnsStepBlock = 256;
constexpr int kDecideBlock = 256;
constexpr int kSetupBlock = 1024;
constexpr int kTimedBatch = 64;

template <int NT, typename T, typename Op>
__device__ inline T block_reduce(T v, Op op, T* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = sh[0];
    for (int k = 1; k < NT / 64; ++k) r = op(r, sh[k]);
    __syncthreads();
    return r;
}

struct MinOp { __device__ double operator()(double a, double b) const { return b < a ? b : a; } };
struct MaxOp { __device__ double operator()(double a, double b) const { return b > a ? b : a; } };
struct AndOp { __device__ int operator()(int a, int b) const { return a & b; } };
struct OrOp { __device__ int operator()(int a, int b) const { return a | b; } };

// b folded into a (the seven quantities of a step's decision)
__device__ inline void fold_part(BlockPart& a, const BlockPart& b) {
    a.ok = AndOp()(a.ok, b.ok);
    a.any4 = OrOp()(a.any4, b.any4);
    a.found = OrOp()(a.found, b.found);
    a.mindt = MinOp()(a.mindt, b.mindt);
    a.vmax = MaxOp()(a.vmax, b.vmax);
    a.vmax_next = MaxOp()(a.vmax_next, b.vmax_next);
    a.dtmin_next = MinOp()(a.dtmin_next, b.dtmin_next);
}

// Workgroup reduction of all seven quantities at once (one barrier pair); per
// quantity the order of block_reduce: butterfly within the wavefront, then
// wavefronts in order.
template <int NT>
__device__ inline BlockPart block_reduce_part(BlockPart v, BlockPart* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        BlockPart o;
        o.ok = __shfl_xor(v.ok, off, 64);
        o.any4 = __shfl_xor(v.any4, off, 64);
        o.found = __shfl_xor(v.found, off, 64);
        o.mindt = __shfl_xor(v.mindt, off, 64);
        o.vmax = __shfl_xor(v.vmax, off, 64);
        o.vmax_next = __shfl_xor(v.vmax_next, off, 64);
        o.dtmin_next = __shfl_xor(v.dtmin_next, off, 64);
        fold_part(v, o);
    }
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    BlockPart r = sh[0];
    for (int k = 1; k < NT / 64; ++k) fold_part(r, sh[k]);
    __syncthreads();
    return r;
}

// DtRef / Dt selection of main() (:171-187); returns the new Dt.
__device__ inline double select_dt(const SimArrays& A, double dt_prev, double vmax, double dtmin) {
    double dtref;
    if (vmax > A.switch_v) dtref = A.rupture_dt;
    else if (dtmin > A.maximum_dt) dtref = A.maximum_dt;
    else dtref = dtmin;
    if (dtref <= A.rupture_dt) return dtref;
    if (dt_prev < dtref / 1.2) return dt_prev * 1.2;
    return dtref;
}

// Run-start step setup (:148-187) from the stored Old state and Accel; one
// workgroup of kSetupBlock threads.
__device__ __forceinline__ void setup_body(const SimArrays& A, SimState* __restrict__ st) {
    __shared__ double shd[kSetupBlock / 64];
    __shared__ int shi[kSetupBlock / 64];
    const SimState S = load_state(st);
    if (S.done) return;
    const int64_t n = A.n;
    const int64_t o = (int64_t)S.parity * n;
    double vmax = -INFINITY, dmin = INFINITY;
    int found = 0;
    for (int64_t i = threadIdx.x; i < n; i += kSetupBlock) {
        const double v = A.V[o + i];
        if (A.cand[i]) {
            vmax = v > vmax ? v : vmax;
            const double r = fabs(v / A.accel[i]);
            if (!(isnan(r) || r == 0.0)) {
                dmin = r < dmin ? r : dmin;
                found = 1;
            }
        }
    }
    vmax = block_reduce<kSetupBlock>(vmax, MaxOp(), shd);
    dmin = block_reduce<kSetupBlock>(dmin, MinOp(), shd);
    found = block_reduce<kSetupBlock>(found, OrOp(), shi);
    if (threadIdx.x == 0) {
        if (!found) {
            st->status = QDFN_ERR_EMPTY_DTMIN;
            st->done = 1;
            st->need_coupling = 0;
        } else {
            const double dt = select_dt(A, S.dt, vmax, dmin / A.dt_cut);
            st->dt = dt;
            st->dt_first = dt;
            st->iteration = 1;
            st->need_coupling = 1;
            st->rhs_evals = S.rhs_evals + 1;
        }
    }
}

// One attempt of SolveOneTimeStep (Functions_Solvers.jl:334-380) with the
// effective normal stress of main() (:212-226), plus the step decision, for
// workgroup blk of the nblk (kBlock threads each) that cover the run.
// The retry-or-accept decision of main() (:240-319) from the reduction `all`
// of every workgroup's part, taken by one thread: the run state S as the
// attempt read it.
__device__ inline void decide(const SimArrays& A, SimState* __restrict__ st, const SimState& S, const BlockPart& all) {
    const int p = S.parity;
    const double told = S.told;
    const int64_t step = S.step;
    const bool record = A.record_step > 0 && (step + 1) % A.record_step == 0 && S.rec < A.rec_cap;
    const int64_t rec = S.rec;
    const int OK = all.ok, ANY4 = all.any4, FOUND = all.found;
    const double MINDT = all.mindt, VMAX = all.vmax, VNEXT = all.vmax_next, DNEXT = all.dtmin_next;
    if (!OK) {  // retry the step with the reduced Dt (:240-258)
        st->dt = ANY4 ? 1e-3 : MINDT;
        st->need_coupling = 0;
        st->retry_passes = S.retry_passes + 1;
        if (S.iteration > 100) {
            st->status = QDFN_ERR_RETRIES;
            st->done = 1;
        }
        st->iteration = S.iteration + 1;
        return;
    }
    // accept (:262-319); Dt = min(Dt_All) == dt when every element converged
    const double dta = MINDT;
    if (S.ntrace < A.trace_cap) {
        A.t_told[S.ntrace] = told;
        A.t_dt[S.ntrace] = dta;
        A.t_vmax[S.ntrace] = VMAX;
        A.t_dtsel[S.ntrace] = S.dt_first;
        st->ntrace = S.ntrace + 1;
    }
    if (record) {
        A.h_time[rec] = told;
        A.h_dt[rec] = dta;
        A.h_dtsel[rec] = S.dt_first;
        st->rec = rec + 1;
    }
    st->told = told + dta;
    st->parity = 1 - p;
    st->step = step + 1;
    if (step + 1 >= S.target || (S.t_end > 0.0 && told + dta >= S.t_end)) {
        st->dt = dta;
        st->done = 1;
        st->need_coupling = 0;
        return;
    }
    if (!FOUND) {
        st->dt = dta;
        st->status = QDFN_ERR_EMPTY_DTMIN;
        st->done = 1;
        st->need_coupling = 0;
        return;
    }
    const double dt_next = select_dt(A, dta, VNEXT, DNEXT / A.dt_cut);
    st->dt = dt_next;
    st->dt_first = dt_next;
    st->iteration = 1;
    st->need_coupling = 1;
    st->rhs_evals = S.rhs_evals + 1;
}

// kLastDecides: the last workgroup to arrive takes the decision (after an
// agent-scope fence); otherwise the workgroups only write their parts and
// decide_body, in the next kernel on the stream, takes it.
template <bool kLastDecides, int kBlock>
__device__ __forceinline__ void step_body(const SimArrays& A, SimState* __restrict__ st, unsigned blk, unsigned nblk) {
    __shared__ BlockPart shp[kBlock / 64];
    __shared__ int s_last;
    const int64_t n = A.n;
    const int64_t i = (int64_t)blk * kBlock + threadIdx.x;
    // element constants, issued before the run state is read (their latency
    // overlaps the state's)
    const int64_t ic = i < n ? i : n - 1;
    const double c_f0 = A.f0[ic], c_a = A.a[ic], c_b = A.b[ic], c_dc = A.dc[ic], c_k = A.kself[ic];
    const double c_omega = A.omega[ic], c_mass = A.mass[ic], c_sigma0 = A.sigma0[ic], c_tau0 = A.tau0[ic];
    const SimState S = load_state(st);
    if (S.done) return;
    const int p = S.parity;
    const int64_t o = (int64_t)p * n, w = (int64_t)(1 - p) * n;
    const double dt = S.dt, told = S.told;
    const int fresh = S.iteration == 1;
    const int64_t step = S.step;
    const bool record = A.record_step > 0 && (step + 1) % A.record_step == 0 && S.rec < A.rec_cap;
    const int64_t rec = S.rec;

    int ok = 1, any4 = 0, found = 0;
    double mindt = INFINITY, vmax = -INFINITY, vmax_next = -INFINITY, dtmin_next = INFINITY;
    if (i < n) {
        const double v_old = A.V[o + i];
        double dsn = 0.0, dtau = 0.0;
        if (A.nt > 0) {
            const double te = A.ext_at_told ? told : told + dt;
            ext_interp_at(A.ext_t, A.ext_n, A.ext_s, A.nt, n, ext_index(A.ext_t, A.nt, te), te, i, dsn, dtau);
        }
        double tl_i, sp_i;
        if (A.hpart) {
            double sS = 0.0, sN = 0.0;
            for (int64_t u = 0; u < A.hslots; ++u) {
                sS += A.hpart[u * 2 * n + i];
                sN += A.hpart[u * 2 * n + n + i];
            }
            tl_i = (-sS) / A.hkdiag[i];
            A.tl[i] = tl_i;
            A.sigprod[i] = sN;
            sp_i = A.planar ? 0.0 : sN;
        } else if (A.dpartS) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, n0 = 0.0, n1 = 0.0, n2 = 0.0, n3 = 0.0;
            const double* pS = A.dpartS + i;
            const double* pN = A.dpartN + i;
            const int64_t ld = A.dld;
            int32_t k = 0;
            for (; k + 4 <= A.dsplits; k += 4) {
                s0 += pS[k * ld];
                s1 += pS[(k + 1) * ld];
                s2 += pS[(k + 2) * ld];
                s3 += pS[(k + 3) * ld];
                n0 += pN[k * ld];
                n1 += pN[(k + 1) * ld];
                n2 += pN[(k + 2) * ld];
                n3 += pN[(k + 3) * ld];
            }
            if (k < A.dsplits) { s0 += pS[k * ld]; n0 += pN[k * ld]; }
            if (k + 1 < A.dsplits) { s1 += pS[(k + 1) * ld]; n1 += pN[(k + 1) * ld]; }
            if (k + 2 < A.dsplits) { s2 += pS[(k + 2) * ld]; n2 += pN[(k + 2) * ld]; }
            tl_i = -(((s0 + s1) + s2) + s3);
            const double sn = A.dnormal_zero ? 0.0 : ((n0 + n1) + n2) + n3;
            A.tl[i] = tl_i;
            A.sigprod[i] = sn;
            sp_i = A.planar ? 0.0 : sn;
        } else {
            tl_i = A.tl[i];
            sp_i = A.planar ? 0.0 : A.sigprod[i];
        }
        double sig = sp_i + c_sigma0 + dsn;
        if (sig < A.min_sig) sig = A.min_sig;
        double vn;
        if (i < n - A.n_loading) {
            const int sw = fresh ? (v_old > A.switch_v ? 1 : 0) : A.sw[i];
            ElemIn e;
            e.conv = A.conv; e.d_old = A.D[o + i]; e.f_old = A.F[o + i]; e.th_old = A.TH[o + i];
            e.v_old = v_old; e.f0 = c_f0; e.a = c_a; e.b = c_b; e.dc = c_dc; e.v0 = A.v0;
            e.k = c_k; e.omega = c_omega; e.mass = c_mass; e.mu = A.mu; e.sig = sig;
            e.tl = 0.0 + tl_i; e.rho = A.rho; e.dtau = dtau; e.switch_v = A.switch_v; e.alpha_evo = A.alpha_evo;
            e.sig_old = A.SIG[o + i]; e.tau0 = c_tau0; e.dr = A.dr;
            const ElemOut r = (sw == 1) ? solve_high_v(e, dt) : solve_low_v(e, dt);
            A.V[w + i] = r.v; A.F[w + i] = r.f; A.D[w + i] = r.d; A.TH[w + i] = r.th; A.SIG[w + i] = r.sig;
            ok = r.ok;
            any4 = r.instab == 4;
            mindt = r.dt;
            // SolverSwitch if the step is retried (:242-249)
            A.sw[i] = r.instab == 3 ? 1 : (r.instab == 4 ? 0 : sw);
            if (r.stall) atomicAdd((unsigned long long*)&st->stalls, (unsigned long long)r.stall);
            vn = r.v;
            vmax = vn;
            if (record) {
                A.h_v[rec + i * A.rec_cap] = vn;
                A.h_d[rec + i * A.rec_cap] = r.d;
                A.h_th[rec + i * A.rec_cap] = r.th;
                A.h_sig[rec + i * A.rec_cap] = r.sig;
                A.h_sw[rec + i * A.rec_cap] = (int8_t)sw;
            }
        } else {
            vn = A.lrate[i];
            const double d = A.D[o + i] + vn * dt;
            A.V[w + i] = vn; A.F[w + i] = A.F[o + i]; A.D[w + i] = d; A.TH[w + i] = A.TH[o + i];
            A.SIG[w + i] = 1e10;
            mindt = dt;
            if (record) {
                A.h_v[rec + i * A.rec_cap] = vn;
                A.h_d[rec + i * A.rec_cap] = d;
                A.h_th[rec + i * A.rec_cap] = A.TH[o + i];
                A.h_sig[rec + i * A.rec_cap] = 1e10;
                A.h_sw[rec + i * A.rec_cap] = 0;
            }
        }
        // quantities of the next step's setup, valid if this attempt is accepted
        const double acc = (vn - v_old) / dt;
        A.accel[i] = acc;
        if (A.cand[i]) {
            vmax_next = vn;
            const double r = fabs(vn / acc);
            if (!(isnan(r) || r == 0.0)) {
                dtmin_next = r;
                found = 1;
            }
        }
    }
    BlockPart mine;
    mine.ok = ok; mine.any4 = any4; mine.found = found; mine.pad = 0;
    mine.mindt = mindt; mine.vmax = vmax; mine.vmax_next = vmax_next; mine.dtmin_next = dtmin_next;
    const BlockPart bpart = block_reduce_part<kBlock>(mine, shp);
    if (!kLastDecides) {
        if (threadIdx.x == 0) A.parts[blk] = bpart;
        return;
    }
    if (threadIdx.x == 0) {
        A.parts[blk] = bpart;
        __threadfence();
        const unsigned prev = atomicAdd(A.arrive, 1u);
        s_last = prev == nblk - 1;
    }
    __syncthreads();
    if (!s_last) return;
    // ---- last workgroup: the decision of main() ----
    // every thread folds a strided share of the workgroups' parts (written by
    // other workgroups: read through vector loads), then one reduction
    __threadfence();
    BlockPart all;
    all.ok = 1; all.any4 = 0; all.found = 0; all.pad = 0;
    all.mindt = INFINITY; all.vmax = -INFINITY; all.vmax_next = -INFINITY; all.dtmin_next = INFINITY;
    for (unsigned k = threadIdx.x; k < nblk; k += kBlock) {
        const volatile BlockPart* q = A.parts + k;
        BlockPart b;
        b.ok = q->ok; b.any4 = q->any4; b.found = q->found; b.pad = 0;
        b.mindt = q->mindt; b.vmax = q->vmax; b.vmax_next = q->vmax_next; b.dtmin_next = q->dtmin_next;
        fold_part(all, b);
    }
    all = block_reduce_part<kBlock>(all, shp);
    if (threadIdx.x != 0) return;
    *A.arrive = 0u;
    decide(A, st, S, all);
}

// The decision of step_body<false>'s attempt, one workgroup: the parts of the
// nblk workgroups folded in the same order as the last workgroup folds them.
// The kernel boundary orders the parts' writes before these reads, so no
// agent-scope fence (an L2 write-back on every XCD) is needed.
__device__ __forceinline__ void decide_body(const SimArrays& A, SimState* __restrict__ st, unsigned nblk) {
    __shared__ BlockPart shp[kDecideBlock / 64];
    const SimState S = load_state(st);
    if (S.done) return;
    BlockPart all;
    all.ok = 1; all.any4 = 0; all.found = 0; all.pad = 0;
    all.mindt = INFINITY; all.vmax = -INFINITY; all.vmax_next = -INFINITY; all.dtmin_next = INFINITY;
    for (unsigned k = threadIdx.x; k < nblk; k += kDecideBlock) fold_part(all, A.parts[k]);
    all = block_reduce_part<kDecideBlock>(all, shp);
    if (threadIdx.x != 0) return;
    decide(A, st, S, all);
}

}  // namespace qdfn
