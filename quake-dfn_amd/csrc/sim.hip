// Device-resident earthquake-cycle loop: main() of
// scripts/Functions_RSFDFN3DMain_D.jl:2-324 (and main_H's identical control
// logic, scripts/Functions_RSFDFN3DMain_H.jl:180-383) without host round trips.
//
// One "attempt" = three launches on the run's stream:
//   k_coupling        split coupling product of DispOld (gated: only on the
//                     first attempt of a step; retries reuse it, as the
//                     reference does), reading the parity-selected buffer
//   k_coupling_finish fixed-order sum of the splits -> TL, SigProd (gated)
//   k_step            one lane per element: HighV / LowV secant solve (FP64),
//                     then per-workgroup partial reductions; the last
//                     workgroup to arrive takes the retry-or-accept decision
//                     of main() (:240-258), records history / trace, swaps
//                     Old<-New by flipping the buffer parity, and selects the
//                     next step's Dt (:162-187) from quantities every lane
//                     computed on the assumption that its attempt is accepted
//                     (Accel, V/Accel, SolverSwitch inputs).
// All decisions live in a device SimState block, so attempts can be captured
// once into a HIP graph and replayed; the host reads the state once per
// replay.  Kernels return at once when the state says the requested number of
// steps is done, so over-issued attempts are harmless.
#include <string.h>

#include <algorithm>
#include <vector>

#include "dense.h"
#include "hmat.h"
#include "friction_device.h"
#include "qdfn_internal.h"

namespace qdfn {

struct SimState {
    double dt, told;
    int64_t step, target, rhs_evals, retry_passes, stalls, rec, ntrace;
    int32_t need_coupling;  // gate for the coupling kernels
    int32_t iteration;      // attempt number within the current step (1-based)
    int32_t status;         // QDFN_OK or error code
    int32_t done;           // 1: requested steps reached or error
    int32_t parity;         // buffer holding the "Old" state
    int32_t pad;
};

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
    // external stress (nt == 0: none)
    int64_t nt;
    const double *ext_t, *ext_n, *ext_s;
    // histories
    int64_t rec_cap, trace_cap, record_step;
    double *h_time, *h_dt, *h_v, *h_d, *h_th, *h_sig;
    double *t_told, *t_dt, *t_vmax;
    // parameters
    double mu, rho, switch_v, dt_cut, rupture_dt, maximum_dt, min_sig, alpha_evo, conv, v0;
    int32_t dr, planar, ext_at_told, pad;
};

// InterpolateFromStressChange (Functions_Solvers.jl:465-540) for one element.
__device__ inline void ext_interp(const SimArrays& A, double time, int64_t i, double& dsn, double& dtau) {
    int64_t idx = 0;
    for (int64_t k = 1; k <= A.nt; ++k)
        if (time >= A.ext_t[k - 1]) idx = k;
    const int64_t nt = A.nt;
    if (idx == 0) {
        const double td = time / A.ext_t[0];
        dsn = td * (A.ext_n[0 + i * nt] - 0.0) + 0.0;
        dtau = td * (A.ext_s[0 + i * nt] - 0.0) + 0.0;
    } else if (idx == nt) {
        dsn = A.ext_n[(idx - 1) + i * nt];
        dtau = A.ext_s[(idx - 1) + i * nt];
    } else {
        const double td = time - A.ext_t[idx - 1];
        const double span = A.ext_t[idx] - A.ext_t[idx - 1];
        const double ln = A.ext_n[(idx - 1) + i * nt], un = A.ext_n[idx + i * nt];
        const double ls = A.ext_s[(idx - 1) + i * nt], us = A.ext_s[idx + i * nt];
        dsn = td / span * (un - ln) + ln;
        dtau = td / span * (us - ls) + ls;
    }
}

constexpr int kStepBlock = 256;
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

// Run-start step setup (:148-187) from the stored Old state and Accel.
__global__ __launch_bounds__(kSetupBlock) void k_setup(SimArrays A, SimState* __restrict__ st) {
    __shared__ double shd[kSetupBlock / 64];
    __shared__ int shi[kSetupBlock / 64];
    if (st->done) return;
    const int64_t n = A.n;
    const int64_t o = (int64_t)st->parity * n;
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
            st->dt = select_dt(A, st->dt, vmax, dmin / A.dt_cut);
            st->iteration = 1;
            st->need_coupling = 1;
            st->rhs_evals += 1;
        }
    }
}

// One attempt of SolveOneTimeStep (Functions_Solvers.jl:334-380) with the
// effective normal stress of main() (:212-226), plus the step decision.
__global__ __launch_bounds__(kStepBlock) void k_step(SimArrays A, SimState* __restrict__ st) {
    __shared__ double shd[kStepBlock / 64];
    __shared__ int shi[kStepBlock / 64];
    __shared__ int s_last;
    if (st->done) return;
    const int64_t n = A.n;
    const int64_t i = (int64_t)blockIdx.x * kStepBlock + threadIdx.x;
    const int p = st->parity;
    const int64_t o = (int64_t)p * n, w = (int64_t)(1 - p) * n;
    const double dt = st->dt, told = st->told;
    const int fresh = st->iteration == 1;
    const int64_t step = st->step;
    const bool record = A.record_step > 0 && (step + 1) % A.record_step == 0 && st->rec < A.rec_cap;
    const int64_t rec = st->rec;

    int ok = 1, any4 = 0, found = 0;
    double mindt = INFINITY, vmax = -INFINITY, vmax_next = -INFINITY, dtmin_next = INFINITY;
    if (i < n) {
        const double v_old = A.V[o + i];
        double dsn = 0.0, dtau = 0.0;
        if (A.nt > 0) ext_interp(A, A.ext_at_told ? told : told + dt, i, dsn, dtau);
        double sig = (A.planar ? 0.0 : A.sigprod[i]) + A.sigma0[i] + dsn;
        if (sig < A.min_sig) sig = A.min_sig;
        double vn;
        if (i < n - A.n_loading) {
            const int sw = fresh ? (v_old > A.switch_v ? 1 : 0) : A.sw[i];
            ElemIn e;
            e.conv = A.conv; e.d_old = A.D[o + i]; e.f_old = A.F[o + i]; e.th_old = A.TH[o + i];
            e.v_old = v_old; e.f0 = A.f0[i]; e.a = A.a[i]; e.b = A.b[i]; e.dc = A.dc[i]; e.v0 = A.v0;
            e.k = A.kself[i]; e.omega = A.omega[i]; e.mass = A.mass[i]; e.mu = A.mu; e.sig = sig;
            e.tl = 0.0 + A.tl[i]; e.rho = A.rho; e.dtau = dtau; e.switch_v = A.switch_v; e.alpha_evo = A.alpha_evo;
            e.sig_old = A.SIG[o + i]; e.tau0 = A.tau0[i]; e.dr = A.dr;
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
    ok = block_reduce<kStepBlock>(ok, AndOp(), shi);
    any4 = block_reduce<kStepBlock>(any4, OrOp(), shi);
    found = block_reduce<kStepBlock>(found, OrOp(), shi);
    mindt = block_reduce<kStepBlock>(mindt, MinOp(), shd);
    vmax = block_reduce<kStepBlock>(vmax, MaxOp(), shd);
    vmax_next = block_reduce<kStepBlock>(vmax_next, MaxOp(), shd);
    dtmin_next = block_reduce<kStepBlock>(dtmin_next, MinOp(), shd);
    if (threadIdx.x == 0) {
        BlockPart& bp = A.parts[blockIdx.x];
        bp.ok = ok; bp.any4 = any4; bp.found = found;
        bp.mindt = mindt; bp.vmax = vmax; bp.vmax_next = vmax_next; bp.dtmin_next = dtmin_next;
        __threadfence();
        const unsigned prev = atomicAdd(A.arrive, 1u);
        s_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    // ---- last workgroup: the decision of main() ----
    __threadfence();
    *A.arrive = 0u;
    int OK = 1, ANY4 = 0, FOUND = 0;
    double MINDT = INFINITY, VMAX = -INFINITY, VNEXT = -INFINITY, DNEXT = INFINITY;
    for (unsigned k = 0; k < gridDim.x; ++k) {
        const BlockPart bp = A.parts[k];
        OK &= bp.ok; ANY4 |= bp.any4; FOUND |= bp.found;
        MINDT = bp.mindt < MINDT ? bp.mindt : MINDT;
        VMAX = bp.vmax > VMAX ? bp.vmax : VMAX;
        VNEXT = bp.vmax_next > VNEXT ? bp.vmax_next : VNEXT;
        DNEXT = bp.dtmin_next < DNEXT ? bp.dtmin_next : DNEXT;
    }
    if (!OK) {  // retry the step with the reduced Dt (:240-258)
        st->dt = ANY4 ? 1e-3 : MINDT;
        st->need_coupling = 0;
        st->retry_passes += 1;
        if (st->iteration > 100) {
            st->status = QDFN_ERR_RETRIES;
            st->done = 1;
        }
        st->iteration += 1;
        return;
    }
    // accept (:262-319); Dt = min(Dt_All) == dt when every element converged
    const double dta = MINDT;
    if (st->ntrace < A.trace_cap) {
        A.t_told[st->ntrace] = told;
        A.t_dt[st->ntrace] = dta;
        A.t_vmax[st->ntrace] = VMAX;
        st->ntrace += 1;
    }
    if (record) {
        A.h_time[rec] = told;
        A.h_dt[rec] = dta;
        st->rec = rec + 1;
    }
    st->told = told + dta;
    st->parity = 1 - p;
    st->step = step + 1;
    if (step + 1 >= st->target) {
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
    st->dt = select_dt(A, dta, VNEXT, DNEXT / A.dt_cut);
    st->iteration = 1;
    st->need_coupling = 1;
    st->rhs_evals += 1;
}

}  // namespace qdfn

using namespace qdfn;

struct GraphSlot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int attempts = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        attempts = 0;
    }
};

struct qdfn_sim {
    qdfn_dense* op = nullptr;   // dense coupling (main) ...
    qdfn_hmat* hop = nullptr;   // ... or H-matrix coupling (main_H)
    qdfn_sim_params P{};
    hipStream_t stream = nullptr;
    DevBuf consts, state2, single, ints, cand, ext, hist, trace, dstate, ranlog, parts;
    SimArrays A{};
    SimState* d_st = nullptr;
    SimState h_st{};
    bool started = false;
    GraphSlot graph;  // captured attempt sequence
    // timing of coupling kernels (mode 1)
    double coup_ms_sum = 0.0;
    int64_t coup_count = 0;
};

static int enqueue_attempt(qdfn_sim* s, int slot, hipEvent_t ev0, hipEvent_t ev1) {
    const int64_t n = s->P.n;
    CouplingGate g;
    g.gate = &s->d_st->need_coupling;
    g.parity = &s->d_st->parity;
    g.parity_stride = n;
    g.ran_log = slot >= 0 ? s->A.ran_log : nullptr;
    g.slot = slot;
    int rc;
    if (s->hop) {
        // timed mode: events around the phase-2 kernel (the dominant one)
        rc = launch_hmat(s->hop, s->A.D, g, 1, s->A.tl, s->A.sigprod, s->stream, nullptr, ev0, ev1);
        if (rc) return rc;
    } else {
        if (ev0) QDFN_HIP(hipEventRecord(ev0, s->stream));
        rc = launch_coupling(s->op, s->A.D, g, s->stream);
        if (rc) return rc;
        if (ev1) QDFN_HIP(hipEventRecord(ev1, s->stream));
        rc = launch_coupling_finish(s->op, s->A.tl, s->A.sigprod, g.gate, s->stream);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_step, dim3((unsigned)((n + kStepBlock - 1) / kStepBlock)), dim3(kStepBlock), 0, s->stream,
                       s->A, s->d_st);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

extern "C" void qdfn_sim_destroy(qdfn_sim_t* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->graph.reset();
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

// Shared set-up of main() / main_H(): kself = K_Self per element (host).
static qdfn_sim* sim_create(const char* who, int64_t n_op, const double* kself_h, int normal_zero,
                            const qdfn_sim_params* p, const double* mass, const double* a, const double* b,
                            const double* dc, const double* theta_i, const double* v_i, const double* friction_i,
                            const double* sigma0, const double* loading_rate) {
    if (!p || !mass || !a || !b || !dc || !theta_i || !v_i || !friction_i || !sigma0 || !loading_rate) {
        fail(QDFN_ERR_ARG, std::string(who) + ": null argument");
        return nullptr;
    }
    if (p->n != n_op || p->n_loading < 0 || p->n_loading > p->n || p->record_capacity < 0 ||
        p->trace_capacity < 0) {
        fail(QDFN_ERR_ARG, std::string(who) + ": size mismatch");
        return nullptr;
    }
    const int64_t n = p->n, L = p->n_loading;
    qdfn_sim* s = new qdfn_sim();
    s->P = *p;
    auto bail = [&](const char* msg) -> qdfn_sim_t* {
        fail(QDFN_ERR_HIP, msg);
        qdfn_sim_destroy(s);
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail("stream");
    // host-side constants (QUAKEDFN_RunARG.jl / main():42-87)
    const double V0 = 1e-9;
    std::vector<double> kself(kself_h, kself_h + n), hc(10 * n);
    std::vector<int8_t> cand(n, 0);
    int64_t n_unstable = 0;
    for (int64_t i = 0; i < n; ++i) n_unstable += (a[i] - b[i] < 0.0);
    for (int64_t i = 0; i < n; ++i) cand[i] = n_unstable > 0 ? (a[i] - b[i] < 0.0) : (i < n - L);
    double* C = hc.data();
    for (int64_t i = 0; i < n; ++i) {
        C[0 * n + i] = a[i];
        C[1 * n + i] = b[i];
        C[2 * n + i] = dc[i];
        C[3 * n + i] = friction_i[i] - a[i] * log(v_i[i] / V0) - b[i] * log(theta_i[i] * V0 / dc[i]);
        C[4 * n + i] = kself[i];
        C[5 * n + i] = sqrt(kself[i] / mass[i]);
        C[6 * n + i] = mass[i];
        C[7 * n + i] = sigma0[i] * friction_i[i];
        C[8 * n + i] = sigma0[i];
        C[9 * n + i] = loading_rate[i];
    }
    if (s->consts.alloc(sizeof(double) * 10 * n) ||
        hipMemcpy(s->consts.p, C, sizeof(double) * 10 * n, hipMemcpyHostToDevice))
        return bail("constants upload");
    // state [V, D, TH, F, SIG] x 2 buffers
    std::vector<double> h0(10 * n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        h0[0 * n + i] = v_i[i];        // V (old buffer 0)
        h0[2 * n + i] = 0.0;           // D
        h0[4 * n + i] = theta_i[i];    // TH
        h0[6 * n + i] = friction_i[i]; // F
        h0[8 * n + i] = sigma0[i];     // SIG (EffNormalStress_Old = InitialNormalStress)
    }
    if (s->state2.alloc(sizeof(double) * 10 * n) ||
        hipMemcpy(s->state2.p, h0.data(), sizeof(double) * 10 * n, hipMemcpyHostToDevice))
        return bail("state upload");
    // single: accel, tl, sigprod
    if (s->single.alloc(sizeof(double) * 3 * n) || hipMemset(s->single.p, 0, sizeof(double) * 3 * n))
        return bail("single alloc");
    // sw[n] + arrival counter
    if (s->ints.alloc(sizeof(int32_t) * (n + 1)) || hipMemset(s->ints.p, 0, sizeof(int32_t) * (n + 1)))
        return bail("int alloc");
    const int64_t nblk = (n + kStepBlock - 1) / kStepBlock;
    if (s->parts.alloc(sizeof(BlockPart) * nblk)) return bail("partials alloc");
    if (s->cand.alloc(n) || hipMemcpy(s->cand.p, cand.data(), n, hipMemcpyHostToDevice)) return bail("cand");
    const int64_t rc_cap = p->record_capacity, tc = p->trace_capacity;
    if (s->hist.alloc(sizeof(double) * (2 * rc_cap + 4 * rc_cap * n + 1))) return bail("history alloc");
    if (s->trace.alloc(sizeof(double) * (3 * tc + 1))) return bail("trace alloc");
    if (s->dstate.alloc(sizeof(SimState))) return bail("state alloc");
    s->d_st = s->dstate.as<SimState>();
    memset(&s->h_st, 0, sizeof(SimState));
    s->h_st.dt = 1.0;  // Dt = 1 (:45)
    if (hipMemcpy(s->d_st, &s->h_st, sizeof(SimState), hipMemcpyHostToDevice)) return bail("state init");

    SimArrays& A = s->A;
    A.n = n;
    A.n_loading = L;
    const double* cd = s->consts.as<double>();
    A.a = cd; A.b = cd + n; A.dc = cd + 2 * n; A.f0 = cd + 3 * n; A.kself = cd + 4 * n; A.omega = cd + 5 * n;
    A.mass = cd + 6 * n; A.tau0 = cd + 7 * n; A.sigma0 = cd + 8 * n; A.lrate = cd + 9 * n;
    A.cand = s->cand.as<int8_t>();
    double* sd = s->state2.as<double>();
    A.V = sd; A.D = sd + 2 * n; A.TH = sd + 4 * n; A.F = sd + 6 * n; A.SIG = sd + 8 * n;
    double* sg = s->single.as<double>();
    A.accel = sg; A.tl = sg + n; A.sigprod = sg + 2 * n;
    if (s->ranlog.alloc(sizeof(int32_t) * kTimedBatch) || hipMemset(s->ranlog.p, 0, sizeof(int32_t) * kTimedBatch))
        return bail("log alloc");
    A.ran_log = s->ranlog.as<int32_t>();
    int32_t* si = s->ints.as<int32_t>();
    A.sw = si;
    A.arrive = reinterpret_cast<unsigned int*>(si + n);
    A.parts = s->parts.as<BlockPart>();
    A.nt = 0; A.ext_t = A.ext_n = A.ext_s = nullptr;
    A.rec_cap = rc_cap; A.trace_cap = tc; A.record_step = p->record_step;
    double* hh = s->hist.as<double>();
    A.h_time = hh; A.h_dt = hh + rc_cap; A.h_v = hh + 2 * rc_cap; A.h_d = A.h_v + rc_cap * n;
    A.h_th = A.h_d + rc_cap * n; A.h_sig = A.h_th + rc_cap * n;
    double* tt = s->trace.as<double>();
    A.t_told = tt; A.t_dt = tt + tc; A.t_vmax = tt + 2 * tc;
    A.mu = p->mu; A.rho = p->rho; A.switch_v = p->switch_v; A.dt_cut = p->dt_cut; A.rupture_dt = p->rupture_dt;
    A.maximum_dt = p->maximum_dt; A.min_sig = p->min_normal_stress; A.alpha_evo = p->alpha_evo;
    A.conv = 1e-9; A.v0 = V0; A.dr = p->evolution_dr; A.planar = p->planar || normal_zero;
    A.ext_at_told = p->ext_at_told;
    return s;
}

extern "C" qdfn_sim_t* qdfn_sim_create_dense(qdfn_dense_t* op, const qdfn_sim_params* p, const double* mass,
                                             const double* a, const double* b, const double* dc,
                                             const double* theta_i, const double* v_i, const double* friction_i,
                                             const double* sigma0, const double* loading_rate) {
    if (!op) {
        fail(QDFN_ERR_ARG, "qdfn_sim_create_dense: null operator");
        return nullptr;
    }
    std::vector<double> kself(op->n);
    if (hipMemcpy(kself.data(), op->kself.p, sizeof(double) * op->n, hipMemcpyDeviceToHost) != hipSuccess) {
        fail(QDFN_ERR_HIP, "qdfn_sim_create_dense: kself download");
        return nullptr;
    }
    qdfn_sim* s = sim_create("qdfn_sim_create_dense", op->n, kself.data(), op->normal_zero, p, mass, a, b, dc,
                             theta_i, v_i, friction_i, sigma0, loading_rate);
    if (s) s->op = op;
    return s;
}

extern "C" qdfn_sim_t* qdfn_sim_create_hmat(qdfn_hmat_t* h, const qdfn_sim_params* p, const double* mass,
                                            const double* a, const double* b, const double* dc,
                                            const double* theta_i, const double* v_i, const double* friction_i,
                                            const double* sigma0, const double* loading_rate) {
    if (!h) {
        fail(QDFN_ERR_ARG, "qdfn_sim_create_hmat: null operator");
        return nullptr;
    }
    // K_Self = -K_S,ii from the diagonal blocks (StiffnessTransitionToLoading, Functions_Hmatrix.jl:459-477)
    std::vector<double> kself(h->n);
    for (int64_t i = 0; i < h->n; ++i) kself[i] = -h->kdiag_h[i];
    qdfn_sim* s = sim_create("qdfn_sim_create_hmat", h->n, kself.data(), h->normal_zero, p, mass, a, b, dc,
                             theta_i, v_i, friction_i, sigma0, loading_rate);
    if (s) s->hop = h;
    return s;
}

extern "C" int qdfn_sim_set_external(qdfn_sim_t* s, int64_t nt, const double* times, const double* ext_n,
                                     const double* ext_s) {
    if (!s || nt < 0 || (nt > 0 && (!times || !ext_n || !ext_s)))
        return fail(QDFN_ERR_ARG, "qdfn_sim_set_external: bad args");
    if (s->started) return fail(QDFN_ERR_STATE, "qdfn_sim_set_external: run already started");
    const int64_t n = s->P.n;
    if (nt == 0) {
        s->A.nt = 0;
        return QDFN_OK;
    }
    QDFN_HIP(s->ext.alloc(sizeof(double) * (nt + 2 * nt * n)));
    double* d = s->ext.as<double>();
    QDFN_HIP(hipMemcpy(d, times, sizeof(double) * nt, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d + nt, ext_n, sizeof(double) * nt * n, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d + nt + nt * n, ext_s, sizeof(double) * nt * n, hipMemcpyHostToDevice));
    s->A.nt = nt;
    s->A.ext_t = d;
    s->A.ext_n = d + nt;
    s->A.ext_s = d + nt + nt * n;
    return QDFN_OK;
}

extern "C" int qdfn_sim_run(qdfn_sim_t* s, int64_t n_steps, int32_t mode, int32_t graph_attempts) {
    if (!s || n_steps < 0) return fail(QDFN_ERR_ARG, "qdfn_sim_run: bad args");
    if (n_steps == 0) return QDFN_OK;
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    if (s->h_st.status != QDFN_OK) return fail(s->h_st.status, "qdfn_sim_run: run is in an error state");
    s->h_st.target = s->h_st.step + n_steps;
    s->h_st.done = 0;
    QDFN_HIP(hipMemcpy(s->d_st, &s->h_st, sizeof(SimState), hipMemcpyHostToDevice));
    s->started = true;
    // step setup for the first step of this call
    hipLaunchKernelGGL(k_setup, dim3(1), dim3(kSetupBlock), 0, s->stream, s->A, s->d_st);
    QDFN_HIP(hipGetLastError());
    auto take_times = [&](const std::vector<hipEvent_t>& ev, const std::vector<int32_t>& ran, int K) -> int {
        for (int k = 0; k < K; ++k) {
            if (!ran[k]) continue;  // gated off: the kernel returned at once
            float ms = 0.f;
            QDFN_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            s->coup_ms_sum += ms;
            s->coup_count += 1;
        }
        return QDFN_OK;
    };
    if (mode != 0 && mode != 1) return fail(QDFN_ERR_ARG, "qdfn_sim_run: mode must be 0 or 1");
    if (mode == 0) {
        const int K = std::max(1, graph_attempts > 0 ? graph_attempts : 32);
        GraphSlot& gs = s->graph;
        if (!gs.exec || gs.attempts != K) {
            gs.reset();
            QDFN_HIP(hipStreamSynchronize(s->stream));
            QDFN_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            int rc = QDFN_OK;
            for (int k = 0; k < K && rc == QDFN_OK; ++k) rc = enqueue_attempt(s, -1, nullptr, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (rc != QDFN_OK) return rc;
            if (e != hipSuccess) return fail(QDFN_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            gs.graph = g;
            QDFN_HIP(hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0));
            gs.attempts = K;
        }
        for (;;) {
            QDFN_HIP(hipGraphLaunch(gs.exec, s->stream));
            QDFN_HIP(hipMemcpyAsync(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost, s->stream));
            QDFN_HIP(hipStreamSynchronize(s->stream));
            if (s->h_st.done) break;
        }
    } else {
        const int K = kTimedBatch;
        std::vector<hipEvent_t> ev(2 * K);
        for (auto& e : ev) QDFN_HIP(hipEventCreate(&e));
        std::vector<int32_t> ran(K);
        int rc = QDFN_OK;
        for (;;) {
            // K attempts, each with HIP events around its coupling kernel
            for (int k = 0; k < K && rc == QDFN_OK; ++k) rc = enqueue_attempt(s, k, ev[2 * k], ev[2 * k + 1]);
            if (rc) break;
            QDFN_HIP(hipMemcpyAsync(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost, s->stream));
            QDFN_HIP(hipMemcpyAsync(ran.data(), s->A.ran_log, sizeof(int32_t) * K, hipMemcpyDeviceToHost, s->stream));
            QDFN_HIP(hipStreamSynchronize(s->stream));
            rc = take_times(ev, ran, K);
            if (rc || s->h_st.done) break;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (rc) return rc;
    }
    if (s->h_st.status != QDFN_OK) {
        const char* what = s->h_st.status == QDFN_ERR_RETRIES
                               ? "more than 100 retries in one step (reference blocks in readline)"
                               : "empty candidate set for min |V/Accel|";
        return fail(s->h_st.status, std::string("qdfn_sim_run: ") + what);
    }
    return QDFN_OK;
}

extern "C" int qdfn_sim_get_stats(qdfn_sim_t* s, qdfn_sim_stats* o) {
    if (!s || !o) return fail(QDFN_ERR_ARG, "qdfn_sim_get_stats: null argument");
    QDFN_HIP(hipStreamSynchronize(s->stream));
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    memset(o, 0, sizeof(*o));
    o->steps = s->h_st.step;
    o->rhs_evals = s->h_st.rhs_evals;
    o->retry_passes = s->h_st.retry_passes;
    o->stalls = s->h_st.stalls;
    o->n_recorded = s->h_st.rec;
    o->n_traced = s->h_st.ntrace;
    o->status = s->h_st.status;
    o->t = s->h_st.told;
    o->dt = s->h_st.dt;
    o->coupling_ms_avg = s->coup_count ? s->coup_ms_sum / s->coup_count : 0.0;
    o->coupling_timed = s->coup_count;
    return QDFN_OK;
}

extern "C" int qdfn_sim_get_state(qdfn_sim_t* s, double* V, double* Disp, double* Theta, double* Friction,
                                  double* Sigma, double* Accel) {
    if (!s) return fail(QDFN_ERR_ARG, "qdfn_sim_get_state: null run");
    QDFN_HIP(hipStreamSynchronize(s->stream));
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    const int64_t n = s->P.n, o = (int64_t)s->h_st.parity * n;
    const size_t b = sizeof(double) * n;
    if (V) QDFN_HIP(hipMemcpy(V, s->A.V + o, b, hipMemcpyDeviceToHost));
    if (Disp) QDFN_HIP(hipMemcpy(Disp, s->A.D + o, b, hipMemcpyDeviceToHost));
    if (Theta) QDFN_HIP(hipMemcpy(Theta, s->A.TH + o, b, hipMemcpyDeviceToHost));
    if (Friction) QDFN_HIP(hipMemcpy(Friction, s->A.F + o, b, hipMemcpyDeviceToHost));
    if (Sigma) QDFN_HIP(hipMemcpy(Sigma, s->A.SIG + o, b, hipMemcpyDeviceToHost));
    if (Accel) QDFN_HIP(hipMemcpy(Accel, s->A.accel, b, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_sim_get_history(qdfn_sim_t* s, int64_t count, double* time, double* dt, double* V, double* Disp,
                                    double* Theta, double* Sigma) {
    if (!s || count < 0 || count > s->P.record_capacity) return fail(QDFN_ERR_ARG, "qdfn_sim_get_history: bad count");
    if (count == 0) return QDFN_OK;
    QDFN_HIP(hipStreamSynchronize(s->stream));
    const int64_t n = s->P.n, cap = s->P.record_capacity;
    const SimArrays& A = s->A;
    if (time) QDFN_HIP(hipMemcpy(time, A.h_time, sizeof(double) * count, hipMemcpyDeviceToHost));
    if (dt) QDFN_HIP(hipMemcpy(dt, A.h_dt, sizeof(double) * count, hipMemcpyDeviceToHost));
    double* dst[4] = {V, Disp, Theta, Sigma};
    const double* src[4] = {A.h_v, A.h_d, A.h_th, A.h_sig};
    for (int k = 0; k < 4; ++k)
        if (dst[k])
            QDFN_HIP(hipMemcpy2D(dst[k], sizeof(double) * count, src[k], sizeof(double) * cap, sizeof(double) * count,
                                 n, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

int qdfn::sim_history_view(qdfn_sim_t* s, SimHistoryView* v) {
    if (!s || !v) return fail(QDFN_ERR_ARG, "sim_history_view: null argument");
    QDFN_HIP(hipStreamSynchronize(s->stream));
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    v->time = s->A.h_time;
    v->disp = s->A.h_d;
    v->n_rec = s->h_st.rec;
    v->ld = s->P.record_capacity;
    v->n = s->P.n;
    v->stream = s->stream;
    return QDFN_OK;
}

extern "C" int qdfn_sim_get_trace(qdfn_sim_t* s, int64_t count, double* told, double* dt, double* vmax) {
    if (!s || count < 0 || count > s->P.trace_capacity) return fail(QDFN_ERR_ARG, "qdfn_sim_get_trace: bad count");
    if (count == 0) return QDFN_OK;
    QDFN_HIP(hipStreamSynchronize(s->stream));
    if (told) QDFN_HIP(hipMemcpy(told, s->A.t_told, sizeof(double) * count, hipMemcpyDeviceToHost));
    if (dt) QDFN_HIP(hipMemcpy(dt, s->A.t_dt, sizeof(double) * count, hipMemcpyDeviceToHost));
    if (vmax) QDFN_HIP(hipMemcpy(vmax, s->A.t_vmax, sizeof(double) * count, hipMemcpyDeviceToHost));
    return QDFN_OK;
}
