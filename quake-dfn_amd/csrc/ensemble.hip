// Batched ensemble: many independent runs of main()
// (scripts/Functions_RSFDFN3DMain_D.jl:2-324), each with its own dense FP32
// operator, friction, adaptive Dt and retries, advanced together.  One
// attempt of every run is three batched launches:
//   k_ens_coupling      split coupling products of every run that needs one
//                       (all groups in one launch); runs sharing an operator
//                       (same geometry) are read as one multi-right-hand-side
//                       product of RG <= 4 runs, so the matrices stream once
//                       per RG runs
//   k_ens_finish        fixed-order split sums -> TL, SigProd of each run
//   k_ens_step          the per-element attempt and per-run decision of main()
// and k_ens_count counts the runs still going.  Attempts are captured into
// one HIP graph and replayed; runs that have finished return at once.
// Every run computes exactly what it computes alone: the split structure,
// the per-lane column order and the fixed-order sums are those of its own
// operator's k_coupling / k_coupling_finish, and the step is the same
// device code (sim_device.h), so an ensemble member's trajectory is bitwise
// the standalone run's.
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "dense.h"
#include "qdfn_internal.h"
#include "sim.h"
#include "sim_device.h"

namespace qdfn {
namespace {

constexpr int kEnsMaxR = 4;   // runs per shared-operator group (right-hand sides per pass)
constexpr int kEnsUnroll = 4; // columns per load batch

typedef float ens_f4 __attribute__((ext_vector_type(4)));

struct EnsInst {
    const int* gate;      // need_coupling of the run
    const int* parity;    // buffer parity of the run's DispOld
    const double* D;      // double-buffered Disp [2][n]
    double *partS, *partN, *tl, *sigprod;
    int64_t n, ld;
    int32_t splits, planar;
};

struct EnsGroup {
    const float *E, *KN;
    int64_t ld, n, cps;
    int32_t planar, ninst;
    int32_t inst[kEnsMaxR];
};

struct EnsTile {
    int32_t group, rb, split, pad;
};

struct EnsBlk {
    int32_t inst, blk;
};

// Split coupling of RG runs sharing (E, K_N): the loop of k_coupling<float>
// (4 rows per lane, columns in order, FP64 accumulation) with one
// accumulator set per run.  force = 1 ignores the gates (autotuning).
template <int RG>
__device__ __forceinline__ void ens_tile(const EnsTile& tl, const EnsGroup* __restrict__ groups,
                                         const EnsInst* __restrict__ insts, int32_t force) {
    const EnsGroup G = groups[tl.group];
    int run[RG];
    int any = 0;
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        run[r] = force ? 1 : ctrl_load(insts[G.inst[r]].gate);
        any |= run[r];
    }
    if (!any) return;
    const int64_t row = ((int64_t)tl.rb * 256 + threadIdx.x) * 4;
    if (row >= G.ld) return;
    const int64_t ld = G.ld, n = G.n;
    const int64_t j0 = (int64_t)tl.split * G.cps;
    const int64_t j1 = (j0 + G.cps < n) ? j0 + G.cps : n;
    const double* x[RG];
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        const EnsInst& I = insts[G.inst[r]];
        x[r] = I.D + (int64_t)ctrl_load(I.parity) * I.n;
    }
    double aS[RG][4], aN[RG][4];
#pragma unroll
    for (int r = 0; r < RG; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) aS[r][q] = aN[r][q] = 0.0;
    const float* pe = G.E + row + j0 * ld;
    const float* pk = G.KN + row + j0 * ld;
    int64_t j = j0;
    if (!G.planar) {
        for (; j + kEnsUnroll <= j1; j += kEnsUnroll) {
            ens_f4 ve[kEnsUnroll], vk[kEnsUnroll];
#pragma unroll
            for (int u = 0; u < kEnsUnroll; ++u) {
                ve[u] = __builtin_nontemporal_load((const ens_f4*)(pe + u * ld));
                vk[u] = __builtin_nontemporal_load((const ens_f4*)(pk + u * ld));
            }
#pragma unroll
            for (int u = 0; u < kEnsUnroll; ++u)
#pragma unroll
                for (int r = 0; r < RG; ++r) {
                    const double xj = x[r][j + u];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        aS[r][q] = fma((double)ve[u][q], xj, aS[r][q]);
                        aN[r][q] = fma((double)vk[u][q], xj, aN[r][q]);
                    }
                }
            pe += kEnsUnroll * ld;
            pk += kEnsUnroll * ld;
        }
        for (; j < j1; ++j) {
            const ens_f4 ve = __builtin_nontemporal_load((const ens_f4*)pe);
            const ens_f4 vk = __builtin_nontemporal_load((const ens_f4*)pk);
#pragma unroll
            for (int r = 0; r < RG; ++r) {
                const double xj = x[r][j];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    aS[r][q] = fma((double)ve[q], xj, aS[r][q]);
                    aN[r][q] = fma((double)vk[q], xj, aN[r][q]);
                }
            }
            pe += ld;
            pk += ld;
        }
    } else {
        for (; j + kEnsUnroll <= j1; j += kEnsUnroll) {
            ens_f4 ve[kEnsUnroll];
#pragma unroll
            for (int u = 0; u < kEnsUnroll; ++u) ve[u] = __builtin_nontemporal_load((const ens_f4*)(pe + u * ld));
#pragma unroll
            for (int u = 0; u < kEnsUnroll; ++u)
#pragma unroll
                for (int r = 0; r < RG; ++r) {
                    const double xj = x[r][j + u];
#pragma unroll
                    for (int q = 0; q < 4; ++q) aS[r][q] = fma((double)ve[u][q], xj, aS[r][q]);
                }
            pe += kEnsUnroll * ld;
        }
        for (; j < j1; ++j) {
            const ens_f4 ve = __builtin_nontemporal_load((const ens_f4*)pe);
#pragma unroll
            for (int r = 0; r < RG; ++r) {
                const double xj = x[r][j];
#pragma unroll
                for (int q = 0; q < 4; ++q) aS[r][q] = fma((double)ve[q], xj, aS[r][q]);
            }
            pe += ld;
        }
    }
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        if (!run[r]) continue;
        const EnsInst& I = insts[G.inst[r]];
        double* ps = I.partS + (int64_t)tl.split * ld + row;
        double* pn = I.partN + (int64_t)tl.split * ld + row;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ps[q] = aS[r][q];
            pn[q] = aN[r][q];
        }
    }
}

// All groups' tiles in one launch: tiles of groups of 4 runs first, then 3,
// 2 and 1 (t[r] holds the tiles of groups of r + 1 runs, n[r] of them).
struct EnsTiles {
    const EnsTile* t[kEnsMaxR];
    int32_t n[kEnsMaxR];
};

__global__ __launch_bounds__(256) void k_ens_coupling(EnsTiles T, const EnsGroup* __restrict__ groups,
                                                      const EnsInst* __restrict__ insts, int32_t force) {
    int32_t b = (int32_t)blockIdx.x;
    if (b < T.n[3]) return ens_tile<4>(T.t[3][b], groups, insts, force);
    b -= T.n[3];
    if (b < T.n[2]) return ens_tile<3>(T.t[2][b], groups, insts, force);
    b -= T.n[2];
    if (b < T.n[1]) return ens_tile<2>(T.t[1][b], groups, insts, force);
    b -= T.n[1];
    if (b < T.n[0]) ens_tile<1>(T.t[0][b], groups, insts, force);
}

// k_coupling_finish of each run: 64 rows per workgroup, its 4 wavefronts sum
// interleaved quarters of the splits, quarters added in order.
__global__ __launch_bounds__(256) void k_ens_finish(const EnsBlk* __restrict__ fins,
                                                    const EnsInst* __restrict__ insts, int32_t force) {
    __shared__ double sh[2][4][64];
    const EnsBlk f = fins[blockIdx.x];
    const EnsInst& I = insts[f.inst];
    if (!force && ctrl_load(I.gate) == 0) return;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t i = (int64_t)f.blk * 64 + lane;
    double s = 0.0, r = 0.0;
    if (i < I.n) {
        for (int k = q; k < I.splits; k += 4) {
            s += I.partS[(int64_t)k * I.ld + i];
            r += I.partN[(int64_t)k * I.ld + i];
        }
    }
    sh[0][q][lane] = s;
    sh[1][q][lane] = r;
    __syncthreads();
    if (q == 0 && i < I.n && !force) {
        const double ts = ((sh[0][0][lane] + sh[0][1][lane]) + sh[0][2][lane]) + sh[0][3][lane];
        const double tn = ((sh[1][0][lane] + sh[1][1][lane]) + sh[1][2][lane]) + sh[1][3][lane];
        I.tl[i] = -ts;
        I.sigprod[i] = I.planar ? 0.0 : tn;
    }
}

__global__ __launch_bounds__(kEnsStepBlock) void k_ens_step(const EnsBlk* __restrict__ blks,
                                                         const SimArrays* __restrict__ arrays,
                                                         SimState* const* __restrict__ states,
                                                         const uint32_t* __restrict__ nblk) {
    const EnsBlk b = blks[blockIdx.x];
    step_body<true, kEnsStepBlock>(arrays[b.inst], states[b.inst], (unsigned)b.blk, nblk[b.inst]);
}

__global__ __launch_bounds__(kSetupBlock) void k_ens_setup(const SimArrays* __restrict__ arrays,
                                                           SimState* const* __restrict__ states) {
    setup_body(arrays[blockIdx.x], states[blockIdx.x]);
}

// number of runs still going -> *out, and which ones -> mask[k]
__global__ __launch_bounds__(256) void k_ens_count(SimState* const* __restrict__ states, int32_t m,
                                                   int32_t* __restrict__ out, int32_t* __restrict__ mask) {
    __shared__ int sh[4];
    int c = 0;
    // the flags were written by earlier kernels: volatile (vector) loads, never
    // the scalar cache, which a one-workgroup kernel can find holding an old line
    for (int k = threadIdx.x; k < m; k += 256) {
        const int going = ctrl_load(&states[k]->done) ? 0 : 1;
        mask[k] = going;
        c += going;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) *out = sh[0] + sh[1] + sh[2] + sh[3];
}

}  // namespace
}  // namespace qdfn

using namespace qdfn;

struct qdfn_ensemble {
    std::vector<qdfn_sim*> sims;
    hipStream_t stream = nullptr;
    int32_t max_rhs = 0;
    DevBuf arrays, states, nblk, insts, groups, fins, steps, parts, live, livemask;
    DevBuf tiles[kEnsMaxR];
    int64_t ntiles[kEnsMaxR] = {0, 0, 0, 0};
    int64_t ngroups = 0, nfins = 0, nsteps = 0;
    int64_t bytes_per_sweep = 0;  // matrix bytes when every run needs its coupling
    int64_t listed = 0;           // runs in the current work lists (all, or those still going)
    GraphSlot graph;
    ~qdfn_ensemble() {
        graph.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Groups of at most rmax of the listed runs per operator and their tiles
// (runs: every run when null).
int build_groups(qdfn_ensemble* e, int rmax, const std::vector<int32_t>* runs = nullptr) {
    std::vector<EnsGroup> groups;
    std::vector<EnsTile> tiles[kEnsMaxR];
    std::map<const qdfn_dense*, std::vector<int32_t>> by_op;
    std::vector<const qdfn_dense*> order;
    const size_t nr = runs ? runs->size() : e->sims.size();
    for (size_t i = 0; i < nr; ++i) {
        const int32_t k = runs ? (*runs)[i] : (int32_t)i;
        const qdfn_dense* op = e->sims[k]->op;
        if (!by_op.count(op)) order.push_back(op);
        by_op[op].push_back(k);
    }
    int64_t bytes = 0;
    for (const qdfn_dense* op : order) {
        const std::vector<int32_t>& members = by_op[op];
        for (size_t c = 0; c < members.size(); c += rmax) {
            EnsGroup G{};
            G.E = op->E.as<float>();
            G.KN = op->KN.as<float>();
            G.ld = op->ld;
            G.n = op->n;
            G.cps = op->cols_per_split;
            G.planar = op->normal_zero;
            G.ninst = (int32_t)std::min<size_t>(rmax, members.size() - c);
            for (int r = 0; r < kEnsMaxR; ++r) G.inst[r] = r < G.ninst ? members[c + r] : members[c];
            const int64_t rb = (op->ld + 1023) / 1024;
            for (int64_t b = 0; b < rb; ++b)
                for (int32_t sp = 0; sp < op->splits; ++sp)
                    tiles[G.ninst - 1].push_back(EnsTile{(int32_t)groups.size(), (int32_t)b, sp, 0});
            bytes += (op->normal_zero ? 4 : 8) * op->ld * op->n;
            groups.push_back(G);
        }
    }
    QDFN_HIP(e->groups.alloc(sizeof(EnsGroup) * std::max<size_t>(groups.size(), 1)));
    QDFN_HIP(hipMemcpy(e->groups.p, groups.data(), sizeof(EnsGroup) * groups.size(), hipMemcpyHostToDevice));
    for (int r = 0; r < kEnsMaxR; ++r) {
        e->ntiles[r] = (int64_t)tiles[r].size();
        QDFN_HIP(e->tiles[r].alloc(sizeof(EnsTile) * std::max<size_t>(tiles[r].size(), 1)));
        if (!tiles[r].empty())
            QDFN_HIP(hipMemcpy(e->tiles[r].p, tiles[r].data(), sizeof(EnsTile) * tiles[r].size(),
                               hipMemcpyHostToDevice));
    }
    e->ngroups = (int64_t)groups.size();
    if (!runs) e->bytes_per_sweep = bytes;
    e->max_rhs = rmax;
    return QDFN_OK;
}

// The finish and step blocks of the listed runs (every run when null).
int build_blocks(qdfn_ensemble* e, const std::vector<int32_t>* runs = nullptr) {
    std::vector<EnsBlk> fins, steps;
    const size_t nr = runs ? runs->size() : e->sims.size();
    for (size_t i = 0; i < nr; ++i) {
        const int32_t k = runs ? (*runs)[i] : (int32_t)i;
        const int64_t n = e->sims[k]->P.n;
        for (int64_t b = 0; b < (n + 63) / 64; ++b) fins.push_back(EnsBlk{k, (int32_t)b});
        for (int64_t b = 0; b < (n + kEnsStepBlock - 1) / kEnsStepBlock; ++b) steps.push_back(EnsBlk{k, (int32_t)b});
    }
    e->nfins = (int64_t)fins.size();
    e->nsteps = (int64_t)steps.size();
    QDFN_HIP(e->fins.alloc(sizeof(EnsBlk) * std::max<size_t>(fins.size(), 1)));
    QDFN_HIP(e->steps.alloc(sizeof(EnsBlk) * std::max<size_t>(steps.size(), 1)));
    if (!fins.empty()) QDFN_HIP(hipMemcpy(e->fins.p, fins.data(), sizeof(EnsBlk) * fins.size(), hipMemcpyHostToDevice));
    if (!steps.empty())
        QDFN_HIP(hipMemcpy(e->steps.p, steps.data(), sizeof(EnsBlk) * steps.size(), hipMemcpyHostToDevice));
    e->listed = (int64_t)nr;
    return QDFN_OK;
}

int launch_coupling_all(qdfn_ensemble* e, int32_t force) {
    EnsTiles T;
    int64_t total = 0;
    for (int r = 0; r < kEnsMaxR; ++r) {
        T.t[r] = e->tiles[r].as<EnsTile>();
        T.n[r] = (int32_t)e->ntiles[r];
        total += e->ntiles[r];
    }
    if (total > 0)
        hipLaunchKernelGGL(k_ens_coupling, dim3((unsigned)total), dim3(256), 0, e->stream, T,
                           e->groups.as<EnsGroup>(), e->insts.as<EnsInst>(), force);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

int enqueue_attempt(qdfn_ensemble* e) {
    int rc = launch_coupling_all(e, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ens_finish, dim3((unsigned)e->nfins), dim3(256), 0, e->stream, e->fins.as<EnsBlk>(),
                       e->insts.as<EnsInst>(), 0);
    hipLaunchKernelGGL(k_ens_step, dim3((unsigned)e->nsteps), dim3(kEnsStepBlock), 0, e->stream,
                       e->steps.as<EnsBlk>(), e->arrays.as<SimArrays>(), e->states.as<SimState*>(),
                       e->nblk.as<uint32_t>());
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

// Autotuner: the group size (right-hand sides per pass) with the fastest
// forced coupling sweep of the whole ensemble: each candidate timed over
// three sweeps in each of three interleaved rounds, its fastest round kept
// (a single round picked a slower group size now and then).
int autotune(qdfn_ensemble* e, int32_t* best) {
    hipEvent_t a, b;
    QDFN_HIP(hipEventCreate(&a));
    QDFN_HIP(hipEventCreate(&b));
    float best_ms[kEnsMaxR + 1];
    for (float& v : best_ms) v = 1e30f;
    int rc = QDFN_OK;
    for (int round = 0; round < 3 && rc == QDFN_OK; ++round)
        for (int rmax = 1; rmax <= kEnsMaxR && rc == QDFN_OK; rmax *= 2) {
            if ((rc = build_groups(e, rmax))) break;
            if ((rc = launch_coupling_all(e, 1))) break;  // warm-up
            float ms = 0.f;
            if (hipEventRecord(a, e->stream) != hipSuccess) { rc = fail(QDFN_ERR_HIP, "autotune: event"); break; }
            for (int k = 0; k < 3 && rc == QDFN_OK; ++k) rc = launch_coupling_all(e, 1);
            if (rc) break;
            if (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                rc = fail(QDFN_ERR_HIP, "autotune: timing");
                break;
            }
            best_ms[rmax] = std::min(best_ms[rmax], ms);
        }
    *best = 1;
    for (int rmax = 2; rmax <= kEnsMaxR; rmax *= 2)
        if (best_ms[rmax] < best_ms[*best]) *best = rmax;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace

extern "C" qdfn_ensemble_t* qdfn_ensemble_create(int64_t m, qdfn_sim_t* const* sims, int32_t max_rhs) {
    if (m <= 0 || m > 0x7fffffff || !sims || max_rhs < 0 || max_rhs > kEnsMaxR) {
        fail(QDFN_ERR_ARG, "qdfn_ensemble_create: bad arguments");
        return nullptr;
    }
    for (int64_t k = 0; k < m; ++k) {
        const qdfn_sim* s = sims[k];
        if (!s || !s->op || s->op->prec != QDFN_PREC_FP32 || s->op->tiled) {
            fail(QDFN_ERR_ARG, "qdfn_ensemble_create: every run must be a dense run with a column-major FP32 "
                               "operator");
            return nullptr;
        }
        for (int64_t q = 0; q < k; ++q)
            if (sims[q] == s) {
                fail(QDFN_ERR_ARG, "qdfn_ensemble_create: a run appears twice");
                return nullptr;
            }
    }
    qdfn_ensemble* e = new qdfn_ensemble();
    e->sims.assign(sims, sims + m);
    auto bail = [&](int code, const char* msg) -> qdfn_ensemble_t* {
        if (code != QDFN_OK) fail(code, msg);
        delete e;
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(QDFN_ERR_HIP, "qdfn_ensemble_create: stream");
    // per-run tables
    std::vector<SimArrays> arrays((size_t)m);
    std::vector<SimState*> states((size_t)m);
    std::vector<uint32_t> nblk((size_t)m);
    std::vector<EnsInst> insts((size_t)m);
    int64_t part_total = 0;
    for (int64_t k = 0; k < m; ++k) part_total += 2 * (int64_t)sims[k]->op->splits * sims[k]->op->ld;
    if (e->parts.alloc(sizeof(double) * part_total))
        return bail(QDFN_ERR_NOMEM, "qdfn_ensemble_create: partial sums");
    double* pp = e->parts.as<double>();
    for (int64_t k = 0; k < m; ++k) {
        qdfn_sim* s = sims[k];
        const qdfn_dense* op = s->op;
        (void)hipStreamSynchronize(s->stream);
        arrays[k] = s->A;
        // TL and SigProd come from k_ens_finish here (with the whole ensemble's
        // elements the separate finish costs less than a longer step kernel)
        arrays[k].dpartS = arrays[k].dpartN = nullptr;
        states[k] = s->d_st;
        const int64_t n = s->P.n;
        nblk[k] = (uint32_t)((n + kEnsStepBlock - 1) / kEnsStepBlock);
        EnsInst& I = insts[k];
        I.gate = &s->d_st->need_coupling;
        I.parity = &s->d_st->parity;
        I.D = s->A.D;
        I.partS = pp;
        I.partN = pp + (int64_t)op->splits * op->ld;
        pp += 2 * (int64_t)op->splits * op->ld;
        I.tl = s->A.tl;
        I.sigprod = s->A.sigprod;
        I.n = op->n;
        I.ld = op->ld;
        I.splits = op->splits;
        I.planar = op->normal_zero;
    }
    if (e->arrays.alloc(sizeof(SimArrays) * m) || e->states.alloc(sizeof(SimState*) * m) ||
        e->nblk.alloc(sizeof(uint32_t) * m) || e->insts.alloc(sizeof(EnsInst) * m) ||
        e->live.alloc(sizeof(int32_t)) || e->livemask.alloc(sizeof(int32_t) * m))
        return bail(QDFN_ERR_NOMEM, "qdfn_ensemble_create: tables");
    if (hipMemcpy(e->arrays.p, arrays.data(), sizeof(SimArrays) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->states.p, states.data(), sizeof(SimState*) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->nblk.p, nblk.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->insts.p, insts.data(), sizeof(EnsInst) * m, hipMemcpyHostToDevice) != hipSuccess)
        return bail(QDFN_ERR_HIP, "qdfn_ensemble_create: table upload");
    if (build_blocks(e)) return bail(QDFN_OK, nullptr);
    int32_t rmax = max_rhs;
    if (rmax == 0 && autotune(e, &rmax)) return bail(QDFN_OK, nullptr);
    if (build_groups(e, rmax)) return bail(QDFN_OK, nullptr);
    return e;
}

extern "C" int qdfn_ensemble_info(qdfn_ensemble_t* e, int64_t* groups, int32_t* max_rhs, int64_t* bytes_per_sweep) {
    if (!e) return fail(QDFN_ERR_ARG, "qdfn_ensemble_info: null ensemble");
    if (groups) *groups = e->ngroups;
    if (max_rhs) *max_rhs = e->max_rhs;
    if (bytes_per_sweep) *bytes_per_sweep = e->bytes_per_sweep;
    return QDFN_OK;
}

extern "C" int qdfn_ensemble_run(qdfn_ensemble_t* e, int64_t n_steps, double t_end, int32_t graph_attempts) {
    if (!e || n_steps < 0 || !(t_end >= 0.0)) return fail(QDFN_ERR_ARG, "qdfn_ensemble_run: bad arguments");
    if (n_steps == 0) return QDFN_OK;
    const int64_t m = (int64_t)e->sims.size();
    for (qdfn_sim* s : e->sims) {
        QDFN_HIP(hipStreamSynchronize(s->stream));
        QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
        if (s->h_st.status != QDFN_OK) return fail(s->h_st.status, "qdfn_ensemble_run: a run is in an error state");
        s->h_st.target = s->h_st.step + n_steps;
        s->h_st.t_end = t_end;
        s->h_st.done = (t_end > 0.0 && s->h_st.told >= t_end) ? 1 : 0;
        QDFN_HIP(hipMemcpy(s->d_st, &s->h_st, sizeof(SimState), hipMemcpyHostToDevice));
        s->started = true;
    }
    hipLaunchKernelGGL(k_ens_setup, dim3((unsigned)m), dim3(kSetupBlock), 0, e->stream, e->arrays.as<SimArrays>(),
                       e->states.as<SimState*>());
    QDFN_HIP(hipGetLastError());
    const int K = std::max(1, graph_attempts > 0 ? graph_attempts : 16);
    GraphSlot& gs = e->graph;
    // K attempts of the listed runs, then the count of the runs still going
    auto capture = [&]() -> int {
        gs.reset();
        QDFN_HIP(hipStreamSynchronize(e->stream));
        QDFN_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = QDFN_OK;
        for (int k = 0; k < K && rc == QDFN_OK; ++k) rc = enqueue_attempt(e);
        if (rc == QDFN_OK)
            hipLaunchKernelGGL(k_ens_count, dim3(1), dim3(256), 0, e->stream, e->states.as<SimState*>(), (int32_t)m,
                               e->live.as<int32_t>(), e->livemask.as<int32_t>());
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &g);
        if (rc != QDFN_OK) return rc;
        if (ce != hipSuccess) return fail(QDFN_ERR_HIP, std::string("ensemble graph capture: ") + hipGetErrorString(ce));
        gs.graph = g;
        QDFN_HIP(hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0));
        gs.attempts = K;
        return QDFN_OK;
    };
    // every run is listed at the start of a call (an earlier call may have
    // left only the runs that were still going then)
    if (e->listed != m) {
        int rc = build_groups(e, e->max_rhs);
        if (!rc) rc = build_blocks(e);
        if (rc) return rc;
        gs.reset();
    }
    if (!gs.exec || gs.attempts != K) {
        const int rc = capture();
        if (rc) return rc;
    }
    // when a sixteenth of the listed runs have finished, the work lists are
    // rebuilt from those still going (a finished run's workgroups return at
    // once, but they are still dispatched); each run's own arithmetic is
    // unchanged.  Over 256 runs of 20 years: 283k RHS evals/s against 267k
    // rebuilding at half and 280k at an eighth; rebuilding at every finish
    // runs as fast as a sixteenth, with up to one rebuild per run.
    int32_t live = 1;
    std::vector<int32_t> mask((size_t)m);
    while (live > 0) {
        QDFN_HIP(hipGraphLaunch(gs.exec, e->stream));
        QDFN_HIP(hipMemcpyAsync(&live, e->live.p, sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
        QDFN_HIP(hipStreamSynchronize(e->stream));
        if (live > 0 && 16 * (int64_t)live <= 15 * e->listed) {
            QDFN_HIP(hipMemcpy(mask.data(), e->livemask.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
            std::vector<int32_t> going;
            for (int64_t k = 0; k < m; ++k)
                if (mask[k]) going.push_back((int32_t)k);
            int rc = build_groups(e, e->max_rhs, &going);
            if (!rc) rc = build_blocks(e, &going);
            if (!rc) rc = capture();
            if (rc) return rc;
        }
    }
    for (qdfn_sim* s : e->sims) {
        QDFN_HIP(hipMemcpy(&s->h_st, s->d_st, sizeof(SimState), hipMemcpyDeviceToHost));
        if (s->h_st.status != QDFN_OK) return fail(s->h_st.status, "qdfn_ensemble_run: a run stopped with an error");
    }
    return QDFN_OK;
}

extern "C" void qdfn_ensemble_destroy(qdfn_ensemble_t* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    delete e;
}
