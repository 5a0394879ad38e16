// Device-resident earthquake-cycle loop: main() of
// scripts/Functions_RSFDFN3DMain_D.jl:2-324 (and main_H's identical control
// logic, scripts/Functions_RSFDFN3DMain_H.jl:180-383) without host round trips.
//
// One "attempt" = launches on the run's stream:
//   coupling          the product of DispOld (gated: only on the first
//                     attempt of a step; retries reuse it, as the reference
//                     does), reading the parity-selected buffer: k_coupling
//                     (dense: split partial sums) or the two H phases (unit
//                     slots)
//   k_step            one lane per element: the split / slot sums -> TL,
//                     SigProd, then HighV / LowV secant solve (FP64), then
//                     per-workgroup partial reductions
//   k_decide          one workgroup: the retry-or-accept decision of main()
//                     (:240-258) from the parts, records history / trace,
//                     swaps Old<-New by flipping the buffer parity, and
//                     selects the next step's Dt (:162-187) from quantities
//                     every lane computed on the assumption that its attempt
//                     is accepted (Accel, V/Accel, SolverSwitch inputs).
// All decisions live in a device SimState block, so attempts can be captured
// once into a HIP graph and replayed; the host reads the state once per
// replay.  Kernels return at once when the state says the requested number of
// steps is done, so over-issued attempts are harmless.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "dense.h"
#include "hmat.h"
#include "friction_device.h"
#include "qdfn_internal.h"
#include "sim_device.h"
#include "sim.h"

namespace qdfn {

// Run-start step setup (:148-187).
__global__ __launch_bounds__(kSetupBlock) void k_setup(SimArrays A, SimState* __restrict__ st) { setup_body(A, st); }

// One attempt over the whole run (one workgroup per kRunStepBlock elements).
__global__ __launch_bounds__(kRunStepBlock) void k_step(SimArrays A, SimState* __restrict__ st) {
    step_body<false, kRunStepBlock>(A, st, blockIdx.x, gridDim.x);
}

// The attempt's decision from k_step's nblk workgroup parts.
__global__ __launch_bounds__(kDecideBlock) void k_decide(SimArrays A, SimState* __restrict__ st, unsigned nblk) {
    decide_body(A, st, nblk);
}

}  // namespace qdfn

using namespace qdfn;

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
        rc = launch_hmat(s->hop, s->A.D, g, 1, s->A.tl, s->A.sigprod, s->stream, nullptr, ev0, ev1, false);
        if (rc) return rc;
    } else {
        if (ev0) QDFN_HIP(hipEventRecord(ev0, s->stream));
        rc = launch_coupling(s->op, s->A.D, g, s->stream);
        if (rc) return rc;
        if (ev1) QDFN_HIP(hipEventRecord(ev1, s->stream));
    }
    const unsigned nblk = (unsigned)((n + kRunStepBlock - 1) / kRunStepBlock);
    hipLaunchKernelGGL(k_step, dim3(nblk), dim3(kRunStepBlock), 0, s->stream, s->A, s->d_st);
    QDFN_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(kDecideBlock), 0, s->stream, s->A, s->d_st, nblk);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

extern "C" void qdfn_sim_destroy(qdfn_sim_t* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->op) s->op->users -= 1;
    if (s->hop) s->hop->users -= 1;
    for (GraphSlot& g : s->graphs) g.reset();
    s->graphs.clear();
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

// Device buffers of a run of p->n elements and their wiring into SimArrays
// (values uploaded by the caller).  Returns nullptr after fail().
static qdfn_sim* sim_alloc(const char* who, const qdfn_sim_params* p) {
    const int64_t n = p->n, L = p->n_loading;
    qdfn_sim* s = new qdfn_sim();
    s->P = *p;
    auto bail = [&](const char* msg) -> qdfn_sim_t* {
        fail(QDFN_ERR_HIP, std::string(who) + ": " + msg);
        qdfn_sim_destroy(s);
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail("stream");
    // constants [a, b, Dc, f0, K_Self, Omega, Mass, tau0, sigma0, LoadingRate]
    if (s->consts.alloc(sizeof(double) * 10 * n)) return bail("constants alloc");
    // state [V, D, TH, F, SIG] x 2 buffers
    if (s->state2.alloc(sizeof(double) * 10 * n)) return bail("state alloc");
    // single: accel, tl, sigprod
    if (s->single.alloc(sizeof(double) * 3 * n) || hipMemset(s->single.p, 0, sizeof(double) * 3 * n))
        return bail("single alloc");
    // sw[n] + arrival counter
    if (s->ints.alloc(sizeof(int32_t) * (n + 1)) || hipMemset(s->ints.p, 0, sizeof(int32_t) * (n + 1)))
        return bail("int alloc");
    const int64_t nblk = (n + kRunStepBlock - 1) / kRunStepBlock;
    if (s->parts.alloc(sizeof(BlockPart) * nblk)) return bail("partials alloc");
    if (s->cand.alloc(n)) return bail("cand alloc");
    const int64_t rc_cap = p->record_capacity, tc = p->trace_capacity;
    if (s->hist.alloc(sizeof(double) * (3 * rc_cap + 4 * rc_cap * n + 1))) return bail("history alloc");
    if (s->histsw.alloc(rc_cap * n + 1)) return bail("history alloc");
    if (s->trace.alloc(sizeof(double) * (4 * tc + 1))) return bail("trace alloc");
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
    A.h_th = A.h_d + rc_cap * n; A.h_sig = A.h_th + rc_cap * n; A.h_dtsel = A.h_sig + rc_cap * n;
    A.h_sw = s->histsw.as<int8_t>();
    double* tt = s->trace.as<double>();
    A.t_told = tt; A.t_dt = tt + tc; A.t_vmax = tt + 2 * tc; A.t_dtsel = tt + 3 * tc;
    A.mu = p->mu; A.rho = p->rho; A.switch_v = p->switch_v; A.dt_cut = p->dt_cut; A.rupture_dt = p->rupture_dt;
    A.maximum_dt = p->maximum_dt; A.min_sig = p->min_normal_stress; A.alpha_evo = p->alpha_evo;
    A.conv = 1e-9; A.v0 = 1e-9; A.dr = p->evolution_dr; A.planar = p->planar;
    A.ext_at_told = p->ext_at_told;
    A.hpart = nullptr;
    A.hkdiag = nullptr;
    A.hslots = 0;
    A.dpartS = A.dpartN = nullptr;
    A.dld = 0;
    A.dsplits = 0;
    A.dnormal_zero = 0;
    return s;
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
    qdfn_sim* s = sim_alloc(who, p);
    if (!s) return nullptr;
    auto bail = [&](const char* msg) -> qdfn_sim_t* {
        fail(QDFN_ERR_HIP, msg);
        qdfn_sim_destroy(s);
        return nullptr;
    };
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
    if (hipMemcpy(s->consts.p, C, sizeof(double) * 10 * n, hipMemcpyHostToDevice)) return bail("constants upload");
    std::vector<double> h0(10 * n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        h0[0 * n + i] = v_i[i];        // V (old buffer 0)
        h0[2 * n + i] = 0.0;           // D
        h0[4 * n + i] = theta_i[i];    // TH
        h0[6 * n + i] = friction_i[i]; // F
        h0[8 * n + i] = sigma0[i];     // SIG (EffNormalStress_Old = InitialNormalStress)
    }
    if (hipMemcpy(s->state2.p, h0.data(), sizeof(double) * 10 * n, hipMemcpyHostToDevice)) return bail("state upload");
    if (hipMemcpy(s->cand.p, cand.data(), n, hipMemcpyHostToDevice)) return bail("cand");
    s->A.planar = p->planar || normal_zero;
    return s;
}

qdfn_sim* qdfn::sim_clone(const qdfn_sim* s, int64_t record_capacity) {
    qdfn_sim_params p = s->P;
    p.record_step = 1;
    p.record_capacity = record_capacity;
    p.trace_capacity = 0;
    qdfn_sim* r = sim_alloc("sim_clone", &p);
    if (!r) return nullptr;
    const int64_t n = p.n;
    auto bail = [&](const char* msg) -> qdfn_sim* {
        fail(QDFN_ERR_HIP, std::string("sim_clone: ") + msg);
        qdfn_sim_destroy(r);
        return nullptr;
    };
    if (hipMemcpy(r->consts.p, s->consts.p, sizeof(double) * 10 * n, hipMemcpyDeviceToDevice) ||
        hipMemcpy(r->cand.p, s->cand.p, n, hipMemcpyDeviceToDevice))
        return bail("constants copy");
    if (s->A.nt > 0) {
        const size_t bytes = sizeof(double) * (s->A.nt + 2 * s->A.nt * n);
        if (r->ext.alloc(bytes) || hipMemcpy(r->ext.p, s->ext.p, bytes, hipMemcpyDeviceToDevice)) return bail("ext");
        double* d = r->ext.as<double>();
        r->A.nt = s->A.nt;
        r->A.ext_t = d;
        r->A.ext_n = d + s->A.nt;
        r->A.ext_s = d + s->A.nt + s->A.nt * n;
    }
    r->A.conv = s->A.conv;
    r->A.planar = s->A.planar;
    r->op = s->op;
    r->hop = s->hop;
    r->A.hpart = s->A.hpart;
    r->A.hkdiag = s->A.hkdiag;
    r->A.hslots = s->A.hslots;
    r->A.dpartS = s->A.dpartS;
    r->A.dpartN = s->A.dpartN;
    r->A.dld = s->A.dld;
    r->A.dsplits = s->A.dsplits;
    r->A.dnormal_zero = s->A.dnormal_zero;
    if (r->op) r->op->users += 1;
    if (r->hop) r->hop->users += 1;
    r->started = true;
    return r;
}

int qdfn::sim_restore(const qdfn_sim* s, int64_t c, qdfn_sim* r) {
    const int64_t n = s->P.n;
    if (c < 0 || c >= (int64_t)s->ck_state.size()) return fail(QDFN_ERR_ARG, "sim_restore: no such checkpoint");
    const double* ck = s->ckpt.as<double>() + c * 11 * n;
    QDFN_HIP(hipMemcpy(r->state2.p, ck, sizeof(double) * 10 * n, hipMemcpyDeviceToDevice));
    QDFN_HIP(hipMemcpy(r->A.accel, ck + 10 * n, sizeof(double) * n, hipMemcpyDeviceToDevice));
    SimState st = s->ck_state[c];
    st.rec = 0;
    st.ntrace = 0;
    st.done = 1;
    r->h_st = st;
    QDFN_HIP(hipMemcpy(r->d_st, &st, sizeof(SimState), hipMemcpyHostToDevice));
    return QDFN_OK;
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
    if (s) {
        // the step kernel adds the coupling's split partials itself
        s->op = op;
        op->users += 1;
        s->A.dpartS = op->partS.as<double>();
        s->A.dpartN = op->partN.as<double>();
        s->A.dld = op->ld;
        s->A.dsplits = op->splits;
        s->A.dnormal_zero = op->normal_zero;
    }
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
    if (s) {
        s->hop = h;
        h->users += 1;
        s->A.hpart = h->part.as<double>();
        s->A.hkdiag = h->kdiag.as<double>();
        s->A.hslots = h->n_slots;
    }
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
    for (int64_t k = 1; k < nt; ++k)
        if (!(times[k] >= times[k - 1]))
            return fail(QDFN_ERR_ARG, "qdfn_sim_set_external: sample times must be non-decreasing (time " +
                                          std::to_string(k + 1) + " precedes time " + std::to_string(k) + ")");
    // the caller's nt x n column-major tables, stored sample-major on the device
    std::vector<double> h(nt + 2 * nt * n);
    std::copy(times, times + nt, h.begin());
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < nt; ++k) {
            h[nt + k * n + i] = ext_n[k + i * nt];
            h[nt + nt * n + k * n + i] = ext_s[k + i * nt];
        }
    QDFN_HIP(s->ext.alloc(sizeof(double) * h.size()));
    double* d = s->ext.as<double>();
    QDFN_HIP(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    s->A.nt = nt;
    s->A.ext_t = d;
    s->A.ext_n = d + nt;
    s->A.ext_s = d + nt + nt * n;
    return QDFN_OK;
}

// Hands the records held on the device (s->h_st.rec of them) to cb as host
// arrays and empties the device history; cb's return value is returned.
static int flush_records(qdfn_sim* s, qdfn_record_cb cb, void* user, int* cb_rc) {
    const int64_t cnt = s->h_st.rec, n = s->P.n;
    *cb_rc = 0;
    if (cnt == 0) return QDFN_OK;
    std::vector<double> t(cnt), dt(cnt), V(cnt * n), D(cnt * n), TH(cnt * n), SIG(cnt * n);
    int rc = qdfn_sim_get_history(s, cnt, t.data(), dt.data(), V.data(), D.data(), TH.data(), SIG.data());
    if (rc) return rc;
    *cb_rc = cb(user, s->rec_flushed, cnt, t.data(), dt.data(), V.data(), D.data(), TH.data(), SIG.data());
    s->rec_flushed += cnt;
    s->h_st.rec = 0;
    QDFN_HIP(hipMemcpy(s->d_st, &s->h_st, sizeof(SimState), hipMemcpyHostToDevice));
    return QDFN_OK;
}

// Advance up to n_steps accepted steps, or until TOld reaches t_end (> 0).
// With cb, the device history is a staging area handed to cb whenever the
// next batch of attempts could overflow it, and at the end.
static int run_impl(qdfn_sim* s, const char* who, int64_t n_steps, double t_end, int32_t mode,
                    int32_t graph_attempts, qdfn_record_cb cb, void* user) {
    if (n_steps == 0) return QDFN_OK;
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    if (s->h_st.status != QDFN_OK) return fail(s->h_st.status, std::string(who) + ": run is in an error state");
    if (t_end > 0.0 && s->h_st.told >= t_end) return QDFN_OK;
    s->h_st.target = s->h_st.step + n_steps;
    s->h_st.t_end = t_end;
    s->h_st.done = 0;
    QDFN_HIP(hipMemcpy(s->d_st, &s->h_st, sizeof(SimState), hipMemcpyHostToDevice));
    // records one batch of K attempts can add (each accepts at most one step)
    const int64_t rstep = s->P.record_step, cap = s->P.record_capacity;
    auto room_for = [&](int64_t attempts) { return rstep > 0 ? attempts / rstep + 1 : 0; };
    int cb_rc = 0;
    auto maybe_flush = [&](int64_t attempts, bool force) -> int {
        if (!cb || rstep <= 0 || cb_rc) return QDFN_OK;
        if (force || cap - s->h_st.rec < room_for(attempts)) return flush_records(s, cb, user, &cb_rc);
        return QDFN_OK;
    };
    if (cb && rstep > 0 && cap < room_for(std::max(1, graph_attempts > 0 ? graph_attempts : 32)))
        return fail(QDFN_ERR_ARG, std::string(who) + ": record_capacity must hold the records of one batch of attempts");
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
        if (s->graph_k != K) {
            for (GraphSlot& g : s->graphs) g.reset();
            s->graphs.clear();
            std::vector<int> sizes;
            for (int k = 1; k < K; k *= 2) sizes.push_back(k);
            sizes.push_back(K);
            QDFN_HIP(hipStreamSynchronize(s->stream));
            for (int size : sizes) {
                s->graphs.emplace_back();
                GraphSlot& gs = s->graphs.back();
                QDFN_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                int rc = QDFN_OK;
                for (int k = 0; k < size && rc == QDFN_OK; ++k) rc = enqueue_attempt(s, -1, nullptr, nullptr);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(s->stream, &g);
                if (rc != QDFN_OK) return rc;
                if (e != hipSuccess) return fail(QDFN_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                gs.graph = g;
                QDFN_HIP(hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0));
                gs.attempts = size;
            }
            s->graph_k = K;
        }
        for (;;) {
            int rc = maybe_flush(K, false);
            if (rc) return rc;
            if (cb_rc) break;
            // a batch of min(K, steps still wanted) attempts (each attempt
            // accepts at most one step; runs to t_end take K), composed of the
            // captured sizes largest first, then one read of the run state
            int64_t left = t_end > 0.0 ? K : std::min<int64_t>(K, std::max<int64_t>(1, s->h_st.target - s->h_st.step));
            for (size_t g = s->graphs.size(); g-- > 0 && left > 0;)
                while (s->graphs[g].attempts <= left) {
                    QDFN_HIP(hipGraphLaunch(s->graphs[g].exec, s->stream));
                    left -= s->graphs[g].attempts;
                }
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
            rc = maybe_flush(K, false);
            if (rc || cb_rc) break;
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
    {
        int rc = maybe_flush(0, true);
        if (rc) return rc;
    }
    if (s->h_st.status != QDFN_OK) {
        const char* what = s->h_st.status == QDFN_ERR_RETRIES
                               ? "more than 100 retries in one step (reference blocks in readline)"
                               : "empty candidate set for min |V/Accel|";
        return fail(s->h_st.status, std::string(who) + ": " + what);
    }
    return QDFN_OK;
}

extern "C" int qdfn_sim_run(qdfn_sim_t* s, int64_t n_steps, int32_t mode, int32_t graph_attempts) {
    if (!s || n_steps < 0) return fail(QDFN_ERR_ARG, "qdfn_sim_run: bad args");
    return run_impl(s, "qdfn_sim_run", n_steps, 0.0, mode, graph_attempts, nullptr, nullptr);
}

extern "C" int qdfn_sim_run_until(qdfn_sim_t* s, double t_end, int64_t max_steps, int32_t graph_attempts,
                                  qdfn_record_cb cb, void* user) {
    if (!s || max_steps < 0 || !(t_end > 0.0)) return fail(QDFN_ERR_ARG, "qdfn_sim_run_until: bad args");
    return run_impl(s, "qdfn_sim_run_until", max_steps, t_end, 0, graph_attempts, cb, user);
}

extern "C" int qdfn_sim_set_convergence(qdfn_sim_t* s, double conv_crit) {
    if (!s || !(conv_crit > 0.0) || !(conv_crit < 1.0))
        return fail(QDFN_ERR_ARG, "qdfn_sim_set_convergence: need a run and 0 < conv_crit < 1");
    if (s->started) return fail(QDFN_ERR_STATE, "qdfn_sim_set_convergence: run already started");
    s->A.conv = conv_crit;
    return QDFN_OK;
}

extern "C" int qdfn_sim_run_checkpointed(qdfn_sim_t* s, int64_t n_steps, int64_t every, int32_t graph_attempts) {
    if (!s || n_steps < 1 || every < 1) return fail(QDFN_ERR_ARG, "qdfn_sim_run_checkpointed: bad args");
    if (s->hop) return fail(QDFN_ERR_ARG, "qdfn_sim_run_checkpointed: dense runs only");
    QDFN_HIP(hipStreamSynchronize(s->stream));
    QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
    if (s->started || s->h_st.step != 0)
        return fail(QDFN_ERR_STATE, "qdfn_sim_run_checkpointed: the run must start here");
    if (s->P.trace_capacity < n_steps)
        return fail(QDFN_ERR_ARG, "qdfn_sim_run_checkpointed: trace_capacity must hold every step");
    const int64_t n = s->P.n, nck = (n_steps + every - 1) / every;
    s->ck_state.clear();
    s->ck_every = 0;
    QDFN_HIP(s->ckpt.alloc(sizeof(double) * 11 * n * nck));
    for (int64_t c = 0; c < nck; ++c) {
        double* ck = s->ckpt.as<double>() + c * 11 * n;
        QDFN_HIP(hipMemcpyAsync(ck, s->state2.p, sizeof(double) * 10 * n, hipMemcpyDeviceToDevice, s->stream));
        QDFN_HIP(hipMemcpyAsync(ck + 10 * n, s->A.accel, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
        QDFN_HIP(hipMemcpyAsync(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost, s->stream));
        QDFN_HIP(hipStreamSynchronize(s->stream));
        s->ck_state.push_back(s->h_st);
        const int64_t seg = std::min(every, n_steps - c * every);
        const int rc = run_impl(s, "qdfn_sim_run_checkpointed", seg, 0.0, 0, graph_attempts, nullptr, nullptr);
        if (rc) return rc;
    }
    s->ck_every = every;
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

extern "C" int qdfn_sim_get_solver_history(qdfn_sim_t* s, int64_t count, int8_t* solver) {
    if (!s || !solver || count < 0 || count > s->P.record_capacity)
        return fail(QDFN_ERR_ARG, "qdfn_sim_get_solver_history: bad arguments");
    if (count == 0) return QDFN_OK;
    QDFN_HIP(hipStreamSynchronize(s->stream));
    const int64_t cap = s->P.record_capacity;
    QDFN_HIP(hipMemcpy2D(solver, count, s->A.h_sw, cap, count, s->P.n, hipMemcpyDeviceToHost));
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
