// Dense coupling operator (scripts/Functions_RSFDFN3DMain_D.jl:54-63 build,
// :199-202 product).  Bound: HBM bandwidth — (P_E + P_N) * n^2 bytes per
// product for P = 8 (FP64) or 4 (FP32) bytes per stored entry, 2 flops per
// entry, i.e. 0.25 (FP64) / 0.5 (FP32) flop per byte against a machine
// balance near 10.
#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "assembly.h"
#include "dense.h"

namespace qdfn {

typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));

template <typename T> struct RowPack;
template <> struct RowPack<double> {
    static constexpr int R = 2;
    typedef vd2 V;
};
template <> struct RowPack<float> {
    static constexpr int R = 4;
    typedef vf4 V;
};

constexpr int kUnroll = 8;

// partS[split, row] = sum_{j in split} E[row, j] * x[j]; partN likewise with K_N.
template <typename T>
__global__ __launch_bounds__(256) void k_coupling(const T* __restrict__ E, const T* __restrict__ KN, int64_t ld,
                                                  int64_t n, int64_t cps, const double* __restrict__ xbase,
                                                  double* __restrict__ partS, double* __restrict__ partN,
                                                  int planar, int tiled, CouplingGate g) {
    const int run = g.gate ? ctrl_load(g.gate) : 1;
    if (g.ran_log && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) g.ran_log[g.slot] = run;
    if (!run) return;
    const double* __restrict__ x = xbase + (g.parity ? (int64_t)ctrl_load(g.parity) * g.parity_stride : 0);
    using P = RowPack<T>;
    using V = typename P::V;
    constexpr int R = P::R;
    const int64_t r0 = (int64_t)blockIdx.x * 256 * R;
    const int64_t row = r0 + (int64_t)threadIdx.x * R;
    const int64_t j0 = (int64_t)blockIdx.y * cps;
    const int64_t j1 = (j0 + cps < n) ? j0 + cps : n;
    if (row >= ld) return;
    double aS[R], aN[R];
#pragma unroll
    for (int r = 0; r < R; ++r) aS[r] = aN[r] = 0.0;
    // column stride: ld (column-major) or the tile's row count (DenseTiles)
    const int64_t rows = (ld - r0 < 256 * R) ? ld - r0 : 256 * R;
    const int64_t off = tiled ? j0 * ld + r0 * (j1 - j0) + (row - r0) : row + j0 * ld;
    const int64_t cs = tiled ? rows : ld;
    const T* pe = E + off;
    const T* pk = KN + off;
    int64_t j = j0;
    if (!planar) {
        for (; j + kUnroll <= j1; j += kUnroll) {
            V ve[kUnroll], vk[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                ve[u] = __builtin_nontemporal_load((const V*)(pe + u * cs));
                vk[u] = __builtin_nontemporal_load((const V*)(pk + u * cs));
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const double xj = x[j + u];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    aS[r] = fma((double)ve[u][r], xj, aS[r]);
                    aN[r] = fma((double)vk[u][r], xj, aN[r]);
                }
            }
            pe += kUnroll * cs;
            pk += kUnroll * cs;
        }
        for (; j < j1; ++j) {
            const V ve = __builtin_nontemporal_load((const V*)pe);
            const V vk = __builtin_nontemporal_load((const V*)pk);
            const double xj = x[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                aS[r] = fma((double)ve[r], xj, aS[r]);
                aN[r] = fma((double)vk[r], xj, aN[r]);
            }
            pe += cs;
            pk += cs;
        }
    } else {
        for (; j + kUnroll <= j1; j += kUnroll) {
            V ve[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) ve[u] = __builtin_nontemporal_load((const V*)(pe + u * cs));
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const double xj = x[j + u];
#pragma unroll
                for (int r = 0; r < R; ++r) aS[r] = fma((double)ve[u][r], xj, aS[r]);
            }
            pe += kUnroll * cs;
        }
        for (; j < j1; ++j) {
            const V ve = __builtin_nontemporal_load((const V*)pe);
            const double xj = x[j];
#pragma unroll
            for (int r = 0; r < R; ++r) aS[r] = fma((double)ve[r], xj, aS[r]);
            pe += cs;
        }
    }
    double* ps = partS + (int64_t)blockIdx.y * ld + row;
    double* pn = partN + (int64_t)blockIdx.y * ld + row;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ps[r] = aS[r];
        pn[r] = aN[r];
    }
}

// Sum of the split partials in a fixed order: each workgroup owns 64 rows;
// its 4 wavefronts sum interleaved quarters of the splits, then the four
// quarter sums are added in wavefront order.
__global__ __launch_bounds__(256) void k_coupling_finish(const double* __restrict__ partS,
                                                         const double* __restrict__ partN, int64_t ld, int64_t n,
                                                         int splits, int planar, double* __restrict__ tl,
                                                         double* __restrict__ sigprod, const int* __restrict__ gate) {
    if (gate && ctrl_load(gate) == 0) return;
    __shared__ double sh[2][4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    double s = 0.0, r = 0.0;
    if (i < n) {
        for (int k = q; k < splits; k += 4) {
            s += partS[(int64_t)k * ld + i];
            r += partN[(int64_t)k * ld + i];
        }
    }
    sh[0][q][lane] = s;
    sh[1][q][lane] = r;
    __syncthreads();
    if (q == 0 && i < n) {
        const double ts = ((sh[0][0][lane] + sh[0][1][lane]) + sh[0][2][lane]) + sh[0][3][lane];
        const double tn = ((sh[1][0][lane] + sh[1][1][lane]) + sh[1][2][lane]) + sh[1][3][lane];
        tl[i] = -ts;
        sigprod[i] = planar ? 0.0 : tn;
    }
}

int launch_coupling(const qdfn_dense* op, const double* d_x, const CouplingGate& g, hipStream_t s,
                    double* part_s, double* part_n) {
    const int R = op->prec == QDFN_PREC_FP64 ? 2 : 4;
    const int64_t rb = (op->ld + 256 * R - 1) / (256 * R);
    dim3 grid((unsigned)rb, (unsigned)op->splits);
    double* ps = part_s ? part_s : op->partS.as<double>();
    double* pn = part_n ? part_n : op->partN.as<double>();
    if (op->prec == QDFN_PREC_FP64)
        hipLaunchKernelGGL(k_coupling<double>, grid, dim3(256), 0, s, op->E.as<double>(), op->KN.as<double>(), op->ld,
                           op->n, op->cols_per_split, d_x, ps, pn, op->normal_zero, op->tiled, g);
    else
        hipLaunchKernelGGL(k_coupling<float>, grid, dim3(256), 0, s, op->E.as<float>(), op->KN.as<float>(), op->ld,
                           op->n, op->cols_per_split, d_x, ps, pn, op->normal_zero, op->tiled, g);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

int launch_coupling_finish(const qdfn_dense* op, double* d_tl, double* d_sigprod, const int* d_gate,
                           hipStream_t s, const double* part_s, const double* part_n) {
    hipLaunchKernelGGL(k_coupling_finish, dim3((unsigned)((op->n + 63) / 64)), dim3(256), 0, s,
                       part_s ? part_s : op->partS.as<double>(), part_n ? part_n : op->partN.as<double>(), op->ld,
                       op->n, op->splits, op->normal_zero, d_tl, d_sigprod, d_gate);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

// ---- operator build ------------------------------------------------------

__global__ void k_diag(const double* __restrict__ KS, int64_t ld, int64_t n, double* __restrict__ diag,
                       double* __restrict__ kself) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = KS[i + i * ld];
    diag[i] = d;
    kself[i] = fabs(d);
}

// E = K_S,ij / K_S,ii off the diagonal, 0 on it; K_N copied; both converted
// to T; padding rows (i >= n) zero.  Also flags any non-zero K_N entry (one
// atomic per workgroup at most).
template <typename T>
__global__ __launch_bounds__(256) void k_build(const double* __restrict__ KS, const double* __restrict__ KNin,
                                               int64_t ld, int64_t n, const double* __restrict__ diag,
                                               T* __restrict__ E, T* __restrict__ KN, int* __restrict__ nz,
                                               int tiled, DenseTiles lay) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
        int nonzero = 0;
        if (i < ld) {
            const int64_t idx = i + j * ld;
            double e = 0.0, k = 0.0;
            if (i < n) {
                e = (i == j) ? 0.0 : KS[idx] / diag[i];
                k = KNin[idx];
                nonzero = k != 0.0;
            }
            const int64_t o = tiled ? lay.index(i, j) : idx;
            E[o] = (T)e;
            KN[o] = (T)k;
        }
        if (__syncthreads_or(nonzero) && threadIdx.x == 0) atomicOr(nz, 1);
    }
}

// Column splits of the coupling kernel.  Each CU ends up with floor or ceil
// of (workgroups / 256) of them, so the kernel runs for the ceiling:
// workgroup counts just above a multiple of 256 leave most CUs idle at the end
// (20,000 elements in FP64: 280 workgroups of 11.7 MB took 1,270 us, 2,080
// workgroups 1,018 us; 8,192 elements: 256 workgroups 168 us, 320 took 190 us).
// Among the split counts giving 1.5-8 MB per workgroup (smaller ones where
// that leaves fewer than 256 workgroups), the one whose count wastes the
// least of the last round is taken (ties: nearest 3.5 MB per workgroup).
static int choose_splits(qdfn_dense* op) {
    const int R = op->prec == QDFN_PREC_FP64 ? 2 : 4;
    const int64_t rb = (op->ld + 256 * R - 1) / (256 * R);
    const int64_t esz = op->prec == QDFN_PREC_FP64 ? 8 : 4;
    const double total = 2.0 * esz * op->ld * op->n;
    const int64_t smax = std::max<int64_t>(1, op->n / 32);
    auto cps_of = [&](int64_t splits) { return round_up((op->n + splits - 1) / splits, kUnroll); };
    int64_t best_cps = cps_of(std::min<int64_t>(smax, std::max<int64_t>(1, (256 + rb - 1) / rb)));
    double best_waste = 1e30, best_dist = 1e30;
    for (int64_t sp = 1; sp <= smax; ++sp) {
        const int64_t cps = cps_of(sp), se = (op->n + cps - 1) / cps, W = rb * se;
        const double per = total / (double)W;
        if (per < std::min(1.5e6, 0.999 * total / 256.0) || per > 8e6) continue;
        const int64_t rounds = (W + 255) / 256;
        const double waste = (double)(rounds * 256 - W) / (double)W;
        const double dist = std::fabs(std::log(per / 3.5e6));
        if (waste < best_waste - 1e-12 || (waste <= best_waste + 1e-12 && dist < best_dist)) {
            best_waste = waste;
            best_dist = dist;
            best_cps = cps;
        }
    }
    if (const char* env = getenv("QDFN_COUPLING_WGS")) {
        const int64_t target = std::max<int64_t>(1, atoll(env));
        best_cps = cps_of(std::min<int64_t>(smax, std::max<int64_t>(1, (target + rb - 1) / rb)));
    }
    op->cols_per_split = best_cps;
    op->splits = (int32_t)std::max<int64_t>(1, (op->n + best_cps - 1) / best_cps);
    return QDFN_OK;
}

// Build the operator from device K_S / K_N (ld x n FP64, owned by the caller).
static int finish_build(qdfn_dense* op, double* d_KS, double* d_KN) {
    const int64_t n = op->n, ld = op->ld;
    hipStream_t s = op->stream;
    QDFN_HIP(op->kself.alloc(sizeof(double) * n));
    QDFN_HIP(op->diag.alloc(sizeof(double) * n));
    hipLaunchKernelGGL(k_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_KS, ld, n,
                       op->diag.as<double>(), op->kself.as<double>());
    QDFN_HIP(hipGetLastError());
    DevBuf nz;
    QDFN_HIP(nz.alloc(sizeof(int)));
    QDFN_HIP(hipMemsetAsync(nz.p, 0, sizeof(int), s));
    const size_t esz = op->prec == QDFN_PREC_FP64 ? sizeof(double) : sizeof(float);
    QDFN_HIP(op->E.alloc(esz * ld * n));
    QDFN_HIP(op->KN.alloc(esz * ld * n));
    choose_splits(op);
    op->tiled = (n >= kTiledMinN && op->prec == QDFN_PREC_FP64) ? 1 : 0;
    if (const char* e = getenv("QDFN_DENSE_TILED")) op->tiled = atoi(e) ? 1 : 0;
    const DenseTiles lay = dense_tiles(op);
    dim3 grid((unsigned)((ld + 255) / 256), (unsigned)std::min<int64_t>(n, 32768));
    if (op->prec == QDFN_PREC_FP64)
        hipLaunchKernelGGL(k_build<double>, grid, dim3(256), 0, s, d_KS, d_KN, ld, n, op->diag.as<double>(),
                           op->E.as<double>(), op->KN.as<double>(), nz.as<int>(), op->tiled, lay);
    else
        hipLaunchKernelGGL(k_build<float>, grid, dim3(256), 0, s, d_KS, d_KN, ld, n, op->diag.as<double>(),
                           op->E.as<float>(), op->KN.as<float>(), nz.as<int>(), op->tiled, lay);
    QDFN_HIP(hipGetLastError());
    int hnz = 0;
    QDFN_HIP(hipMemcpyAsync(&hnz, nz.p, sizeof(int), hipMemcpyDeviceToHost, s));
    QDFN_HIP(hipStreamSynchronize(s));
    op->normal_zero = hnz ? 0 : 1;
    QDFN_HIP(op->partS.alloc(sizeof(double) * op->splits * ld));
    QDFN_HIP(op->partN.alloc(sizeof(double) * op->splits * ld));
    QDFN_HIP(op->x.alloc(sizeof(double) * ld));
    QDFN_HIP(op->yS.alloc(sizeof(double) * ld));
    QDFN_HIP(op->yN.alloc(sizeof(double) * ld));
    return QDFN_OK;
}

static qdfn_dense* new_op(int64_t n, int32_t prec) {
    if (n <= 0 || (prec != QDFN_PREC_FP64 && prec != QDFN_PREC_FP32)) {
        fail(QDFN_ERR_ARG, "dense operator: n must be > 0 and prec FP64 or FP32");
        return nullptr;
    }
    qdfn_dense* op = new qdfn_dense();
    if (hipGetDevice(&op->device) != hipSuccess) {
        fail(QDFN_ERR_HIP, "dense operator: no current HIP device");
        delete op;
        return nullptr;
    }
    op->n = n;
    op->ld = round_up(n, 64);
    op->prec = prec;
    if (hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking) != hipSuccess) {
        fail(QDFN_ERR_HIP, "dense operator: stream creation failed");
        delete op;
        return nullptr;
    }
    return op;
}

}  // namespace qdfn

using namespace qdfn;

extern "C" void qdfn_dense_destroy(qdfn_dense_t* op) {
    if (!op) return;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    hipStream_t s = op->stream;
    delete op;
    if (s) (void)hipStreamDestroy(s);
}

extern "C" qdfn_dense_t* qdfn_dense_create(const double* KS, const double* KN, int64_t n, int64_t ld, int32_t prec) {
    if (!KS || !KN || ld < n) {
        fail(QDFN_ERR_ARG, "qdfn_dense_create: null matrix or ld < n");
        return nullptr;
    }
    qdfn_dense* op = new_op(n, prec);
    if (!op) return nullptr;
    DevBuf dks, dkn;
    const size_t b = sizeof(double) * op->ld * n;
    int rc = QDFN_OK;
    if (dks.alloc(b) != hipSuccess || dkn.alloc(b) != hipSuccess) rc = fail(QDFN_ERR_NOMEM, "qdfn_dense_create: out of device memory");
    if (rc == QDFN_OK &&
        (hipMemcpy2D(dks.p, sizeof(double) * op->ld, KS, sizeof(double) * ld, sizeof(double) * n, n,
                     hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy2D(dkn.p, sizeof(double) * op->ld, KN, sizeof(double) * ld, sizeof(double) * n, n,
                     hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(QDFN_ERR_HIP, "qdfn_dense_create: upload failed");
    if (rc == QDFN_OK) rc = finish_build(op, dks.as<double>(), dkn.as<double>());
    if (rc != QDFN_OK) {
        qdfn_dense_destroy(op);
        return nullptr;
    }
    return op;
}

extern "C" qdfn_dense_t* qdfn_dense_create_rect(const double* seg, int64_t n, int64_t ld_seg, double mu, double nu,
                                                int32_t prec) {
    if (!seg || ld_seg < n) {
        fail(QDFN_ERR_ARG, "qdfn_dense_create_rect: null table or ld_seg < n");
        return nullptr;
    }
    qdfn_dense* op = new_op(n, prec);
    if (!op) return nullptr;
    DevBuf drcv, dks, dkn;
    const size_t b = sizeof(double) * op->ld * n;
    int rc = QDFN_OK;
    if (drcv.alloc(sizeof(double) * ld_seg * 8) != hipSuccess || dks.alloc(b) != hipSuccess ||
        dkn.alloc(b) != hipSuccess)
        rc = fail(QDFN_ERR_NOMEM, "qdfn_dense_create_rect: out of device memory");
    if (rc == QDFN_OK && hipMemcpy(drcv.p, seg, sizeof(double) * ld_seg * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QDFN_ERR_HIP, "qdfn_dense_create_rect: upload failed");
    if (rc == QDFN_OK && hipMemsetAsync(dks.p, 0, b, op->stream) != hipSuccess) rc = fail(QDFN_ERR_HIP, "memset");
    if (rc == QDFN_OK && hipMemsetAsync(dkn.p, 0, b, op->stream) != hipSuccess) rc = fail(QDFN_ERR_HIP, "memset");
    if (rc == QDFN_OK)
        rc = assemble_rect_device(seg, n, ld_seg, drcv.as<double>(), n, ld_seg, mu, nu, dks.as<double>(),
                                  dkn.as<double>(), op->ld, op->stream);
    if (rc == QDFN_OK) rc = finish_build(op, dks.as<double>(), dkn.as<double>());
    if (rc != QDFN_OK) {
        qdfn_dense_destroy(op);
        return nullptr;
    }
    return op;
}

extern "C" qdfn_dense_t* qdfn_dense_create_tri(int64_t n, const double* P1, const double* P2, const double* P3,
                                               const double* rake, const double* ctr, const double* nrm,
                                               const double* slip, double mu, double lambda, int32_t prec) {
    if (!P1 || !P2 || !P3 || !rake || !ctr || !nrm || !slip) {
        fail(QDFN_ERR_ARG, "qdfn_dense_create_tri: null input");
        return nullptr;
    }
    qdfn_dense* op = new_op(n, prec);
    if (!op) return nullptr;
    DevBuf d1, d2, d3, dr, dc, dn, dsl, dks, dkn;
    const size_t v3 = sizeof(double) * 3 * n;
    const size_t b = sizeof(double) * op->ld * n;
    int rc = QDFN_OK;
    if (d1.alloc(v3) || d2.alloc(v3) || d3.alloc(v3) || dr.alloc(sizeof(double) * n) || dc.alloc(v3) ||
        dn.alloc(v3) || dsl.alloc(v3) || dks.alloc(b) || dkn.alloc(b))
        rc = fail(QDFN_ERR_NOMEM, "qdfn_dense_create_tri: out of device memory");
    const void* src[7] = {P1, P2, P3, rake, ctr, nrm, slip};
    DevBuf* dst[7] = {&d1, &d2, &d3, &dr, &dc, &dn, &dsl};
    for (int k = 0; rc == QDFN_OK && k < 7; ++k)
        if (hipMemcpy(dst[k]->p, src[k], k == 3 ? sizeof(double) * n : v3, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(QDFN_ERR_HIP, "qdfn_dense_create_tri: upload failed");
    if (rc == QDFN_OK && (hipMemsetAsync(dks.p, 0, b, op->stream) || hipMemsetAsync(dkn.p, 0, b, op->stream)))
        rc = fail(QDFN_ERR_HIP, "memset");
    if (rc == QDFN_OK)
        rc = assemble_tri_device(n, d1.as<double>(), d2.as<double>(), d3.as<double>(), dr.as<double>(), n,
                                 dc.as<double>(), dn.as<double>(), dsl.as<double>(), mu, lambda, dks.as<double>(),
                                 dkn.as<double>(), op->ld, op->stream);
    if (rc == QDFN_OK) rc = finish_build(op, dks.as<double>(), dkn.as<double>());
    if (rc != QDFN_OK) {
        qdfn_dense_destroy(op);
        return nullptr;
    }
    return op;
}

extern "C" int qdfn_dense_kself(qdfn_dense_t* op, double* kself, double* diag) {
    if (!op) return fail(QDFN_ERR_ARG, "qdfn_dense_kself: null operator");
    if (kself) QDFN_HIP(hipMemcpy(kself, op->kself.p, sizeof(double) * op->n, hipMemcpyDeviceToHost));
    if (diag) QDFN_HIP(hipMemcpy(diag, op->diag.p, sizeof(double) * op->n, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_dense_normal_zero(qdfn_dense_t* op) { return op ? op->normal_zero : QDFN_ERR_ARG; }

extern "C" int qdfn_dense_set_splits(qdfn_dense_t* op, int32_t splits) {
    if (!op || splits < 0) return fail(QDFN_ERR_ARG, "qdfn_dense_set_splits: bad arguments");
    if (op->tiled)
        return fail(QDFN_ERR_ARG, "qdfn_dense_set_splits: the split-tiled layout is fixed at the operator's build");
    if (op->users > 0)
        return fail(QDFN_ERR_STATE, "qdfn_dense_set_splits: runs exist on this operator (set the splits before "
                                    "creating them)");
    QDFN_HIP(hipStreamSynchronize(op->stream));
    if (splits == 0) {
        choose_splits(op);
    } else {
        const int64_t sp = std::min<int64_t>(splits, op->n);
        op->cols_per_split = round_up((op->n + sp - 1) / sp, kUnroll);
        op->splits = (int32_t)std::max<int64_t>(1, (op->n + op->cols_per_split - 1) / op->cols_per_split);
    }
    QDFN_HIP(op->partS.alloc(sizeof(double) * op->splits * op->ld));
    QDFN_HIP(op->partN.alloc(sizeof(double) * op->splits * op->ld));
    return op->splits;
}

__global__ void k_full_shear(const double* __restrict__ tl, const double* __restrict__ x,
                             const double* __restrict__ diag, int64_t n, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // K_S x = K_ii * (E x + x_i) = K_ii * (x_i - TL_i)
    y[i] = diag[i] * (x[i] - tl[i]);
}

extern "C" int qdfn_matvec(qdfn_dense_t* op, const double* x, double* yS, double* yN, int32_t loading_form) {
    if (!op || !x) return fail(QDFN_ERR_ARG, "qdfn_matvec: null argument");
    hipStream_t s = op->stream;
    QDFN_HIP(hipMemcpyAsync(op->x.p, x, sizeof(double) * op->n, hipMemcpyHostToDevice, s));
    int rc = launch_coupling(op, op->x.as<double>(), CouplingGate{}, s);
    if (rc) return rc;
    rc = launch_coupling_finish(op, op->yS.as<double>(), op->yN.as<double>(), nullptr, s);
    if (rc) return rc;
    if (!loading_form) {
        hipLaunchKernelGGL(k_full_shear, dim3((unsigned)((op->n + 255) / 256)), dim3(256), 0, s,
                           op->yS.as<double>(), op->x.as<double>(), op->diag.as<double>(), op->n,
                           op->yS.as<double>());
        QDFN_HIP(hipGetLastError());
    }
    if (yS) QDFN_HIP(hipMemcpyAsync(yS, op->yS.p, sizeof(double) * op->n, hipMemcpyDeviceToHost, s));
    if (yN) QDFN_HIP(hipMemcpyAsync(yN, op->yN.p, sizeof(double) * op->n, hipMemcpyDeviceToHost, s));
    QDFN_HIP(hipStreamSynchronize(s));
    return QDFN_OK;
}

extern "C" int qdfn_dense_time_coupling(qdfn_dense_t* op, int32_t reps, double* avg_ms) {
    if (!op || reps <= 0 || !avg_ms) return fail(QDFN_ERR_ARG, "qdfn_dense_time_coupling: bad args");
    hipStream_t s = op->stream;
    hipEvent_t a, b;
    QDFN_HIP(hipEventCreate(&a));
    QDFN_HIP(hipEventCreate(&b));
    int rc = launch_coupling(op, op->x.as<double>(), CouplingGate{}, s);  // warm
    QDFN_HIP(hipEventRecord(a, s));
    for (int r = 0; r < reps && rc == QDFN_OK; ++r) rc = launch_coupling(op, op->x.as<double>(), CouplingGate{}, s);
    QDFN_HIP(hipEventRecord(b, s));
    QDFN_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    QDFN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = (double)ms / reps;
    return rc;
}

// ---- coupling product that allocates nothing ------------------------------
//
// Split partial sums of qdfn_matvec_device_noalloc live in this file-scope
// device array (one copy per GPU, held until the process ends).  On the host,
// `scratch` names a host-side stand-in, not device memory: host code only
// ever uses the device address that hipGetSymbolAddress returns for the
// current device, kept per device below.

namespace qdfn {

constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices];  // device address of `scratch` per device (null: not yet looked up)
static std::mutex g_scratch_mu;

static int scratch_for_current_device(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(QDFN_ERR_ARG, "qdfn_matvec_device_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QDFN_ERR_HIP, std::string("qdfn_matvec_device_noalloc: hipGetSymbolAddress(scratch): ") +
                                          hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return QDFN_OK;
}

}  // namespace qdfn

extern "C" int qdfn_matvec_device_noalloc(qdfn_dense_t* op, const double* d_x, double* d_yS, double* d_yN,
                                          int32_t loading_form, void* stream) {
    if (!op || !d_x || !d_yS || !d_yN) return fail(QDFN_ERR_ARG, "qdfn_matvec_device_noalloc: null argument");
    int dev = -1;
    QDFN_HIP(hipGetDevice(&dev));
    if (dev != op->device)
        return fail(QDFN_ERR_ARG, "qdfn_matvec_device_noalloc: the operator lives on device " +
                                      std::to_string(op->device) + ", the current device is " + std::to_string(dev));
    const int64_t need = 2 * (int64_t)op->splits * op->ld;
    if (need > kNoAllocScratchDoubles)
        return fail(QDFN_ERR_ARG, "qdfn_matvec_device_noalloc: the operator's split partial sums need " +
                                      std::to_string(need * 8) + " bytes of scratch, above the " +
                                      std::to_string(kNoAllocScratchDoubles * 8) + "-byte maximum");
    double* base = nullptr;
    int rc = scratch_for_current_device(dev, &base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* ps = base;
    double* pn = base + (int64_t)op->splits * op->ld;
    // every split writes all ld rows of its partials before the finish reads
    // them, so nothing has to start at zero
    rc = launch_coupling(op, d_x, CouplingGate{}, s, ps, pn);
    if (rc) return rc;
    rc = launch_coupling_finish(op, d_yS, d_yN, nullptr, s, ps, pn);
    if (rc) return rc;
    if (!loading_form) {
        hipLaunchKernelGGL(k_full_shear, dim3((unsigned)((op->n + 255) / 256)), dim3(256), 0, s, d_yS, d_x,
                           op->diag.as<double>(), op->n, d_yS);
        QDFN_HIP(hipGetLastError());
    }
    return QDFN_OK;
}
