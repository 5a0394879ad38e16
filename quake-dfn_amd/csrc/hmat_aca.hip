// H-matrix construction from sampled kernel entries.
//
// HmatBuild_T / HmatBuild_R (scripts/Functions_BuildInputFile.jl:643-722,
// :725-800) assemble every block in full and compress the admissible ones with
// pqrfact.  Here no admissible block is assembled: a partially pivoted
// adaptive cross approximation asks only for the rows and columns it pivots
// on, and the entries of those lines come from the same per-pair code as the
// full assembly (rect_entry / tri_entry, evaluated by k_rect_lines /
// k_tde_lines).  Storage and entry evaluations then grow like the H-matrix,
// not like n^2.
//
// All parts of one phase (every admissible shear part, then every normal
// part) advance together, one step per launch pair:
//   lines kernel  - evaluates each active part's requested row or column;
//   k_aca_step    - one workgroup per part: residual of the line (entries
//                   minus the rank-r approximation), then
//                   row phase:    pivot column j* = argmax |residual row|,
//                                 v_r = residual row / pivot, request column j*;
//                   column phase: u_r = residual column, the Frobenius norm
//                                 estimate ||S_{r+1}||^2 = ||S_r||^2 +
//                                 2 sum_l (u_l.u_r)(v_l.v_r) + |u_r|^2 |v_r|^2,
//                                 stop test, next row i* = argmax |u_r| over
//                                 rows not yet used, request row i*;
//                   probe column: residual column of a probe (below).
// Stop test: a term with |u_r| |v_r| <= eps ||S||_F (eps = rtol / 2; ||S||_F
// the block's shear norm estimate, for the normal part as well; with
// atol > 0 also max|u_r| max|v_r| <= atol), or a negligible residual row or
// probe column (sqrt(m) |row| <= eps / 2 ||S||_F, sqrt(k) |column| likewise:
// the block's residual would be within half the tolerance if every row or
// column were as large).  Three such steps in a
// row stop the part.  After a step that meets the test the next request is a
// probe rather than the argmax-|u_r| row: alternately the unused row nearest
// to frac(q phi) m and the column frac(q phi) k (phi the golden ratio), so the
// test is confirmed away from the rows the pivots favour; a probe column that
// is not negligible restarts the cross at its largest residual row.  A part
// that reaches the rank cap falls back to full evaluation of its block and the
// full-pivot approximation the full build uses; inadmissible blocks are
// evaluated in full (both parts in one pass).  Approximations are recompressed
// on the host to the other half of the tolerance, as in the full build.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "assembly.h"
#include "hmat.h"

namespace qdfn {

namespace {

constexpr int kSThreads = 256;
constexpr int kSWaves = kSThreads / 64;
constexpr int kRankCap = 1024;  // sampled parts: at most this rank (LDS arrays of the step kernel)
constexpr int64_t kFallbackMaxEntries = int64_t(1) << 28;  // largest block evaluated in full at the cap
constexpr int kLineTile = 64;
constexpr int kStopHits = 3;  // the stop test on a term, then on a probe row and a probe column

// one part (shear or normal) of one admissible block
struct SJob {
    int64_t u_off, v_off, smp_off, used_off;  // U (m x rmax, column-major), V (rmax x k, row-major), samples, row flags
    int32_t m, k, r0, s0, rmax, which;
    double ref;  // threshold scale (normal parts: the block's shear norm); < 0: the running estimate
};

// device-written progress of one part
struct SState {
    int32_t rank, istar, jstar, phase, status, hits, zrows, probe, pcols, pad0;
    double norm2, vv, vmax;
};

enum : int32_t { kActive = 0, kConverged = 1, kFallback = 2 };

__device__ inline void better_max(double& v, int32_t& id, double& sv, double v2, int32_t id2, double sv2) {
    if (v2 > v || (v2 == v && id2 < id)) {
        v = v2;
        id = id2;
        sv = sv2;
    }
}

// Workgroup argmax (smallest index on ties, with the signed value) and sum,
// in a fixed order; every thread gets the result.
__device__ void wg_argmax_sum(double& v, int32_t& id, double& sv, double& ss) {
    __shared__ double sh_v[kSWaves], sh_s[kSWaves], sh_q[kSWaves];
    __shared__ int32_t sh_i[kSWaves];
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, 64);
        const int32_t i2 = __shfl_xor(id, o, 64);
        const double s2 = __shfl_xor(sv, o, 64);
        better_max(v, id, sv, v2, i2, s2);
        ss += __shfl_xor(ss, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sh_v[wave] = v;
        sh_i[wave] = id;
        sh_s[wave] = sv;
        sh_q[wave] = ss;
    }
    __syncthreads();
    v = sh_v[0];
    id = sh_i[0];
    sv = sh_s[0];
    ss = sh_q[0];
    for (int w = 1; w < kSWaves; ++w) {
        better_max(v, id, sv, sh_v[w], sh_i[w], sh_s[w]);
        ss += sh_q[w];
    }
    __syncthreads();
}

__device__ inline void request(SampleLine* lines, int32_t job, int32_t rcv, int32_t src, int32_t drcv, int32_t dsrc,
                               int32_t len, int32_t which, int64_t out) {
    volatile SampleLine* L = lines + job;
    L->rcv = rcv;
    L->src = src;
    L->drcv = drcv;
    L->dsrc = dsrc;
    L->which = which;
    L->out = out;
    L->len = len;
}

__global__ __launch_bounds__(64) void k_aca_init(const SJob* __restrict__ jobs, int32_t njobs, SState* st,
                                                 SampleLine* lines, uint8_t* __restrict__ used) {
    const int32_t j = blockIdx.x;
    if (j >= njobs) return;
    const SJob J = jobs[j];
    for (int32_t i = threadIdx.x; i < J.m; i += 64) used[J.used_off + i] = 0;
    if (threadIdx.x == 0) {
        volatile SState* s = st + j;
        s->rank = 0;
        s->istar = 0;
        s->jstar = 0;
        s->phase = 0;
        s->status = kActive;
        s->hits = 0;
        s->zrows = 0;
        s->probe = 0;
        s->pcols = 0;
        s->norm2 = 0.0;
        s->vv = 0.0;
        s->vmax = 0.0;
        request(lines, j, J.r0, J.s0, 0, 1, J.k, J.which, J.smp_off);
    }
}

// Next probe after a step that met the stop test: even probes are rows, the
// unused row nearest (cyclically) to frac((q + 1) phi) m; odd probes are
// column frac((q + 1) phi) k, q = probe / 2.  Called by every thread; thread 0
// writes the request.  Returns false when no unused row is left.
__device__ bool next_probe(const SJob& J, volatile SState* vs, SampleLine* lines, int32_t jb, int32_t probe,
                           const uint8_t* used) {
    const int32_t q = probe >> 1;
    const double f = (double)(q + 1) * 0.6180339887498949;
    const double frac = f - floor(f);
    if (probe & 1) {
        const int32_t jp = min((int32_t)(frac * J.k), J.k - 1);
        if (threadIdx.x == 0) {
            vs->probe = probe + 1;
            vs->jstar = jp;
            vs->phase = 2;
            request(lines, jb, J.r0, J.s0 + jp, 1, 0, J.m, J.which, J.smp_off);
        }
        return true;
    }
    const int32_t start = min((int32_t)(frac * J.m), J.m - 1);
    double bmax = -1.0, bsg = 0.0, bd = 0.0;
    int32_t bid = 0x7fffffff;
    for (int32_t i = threadIdx.x; i < J.m; i += kSThreads)
        if (!used[i]) better_max(bmax, bid, bsg, (double)(J.m - (i - start + J.m) % J.m), i, 0.0);
    wg_argmax_sum(bmax, bid, bsg, bd);
    if (bid == 0x7fffffff) return false;
    if (threadIdx.x == 0) {
        vs->probe = probe + 1;
        vs->istar = bid;
        vs->phase = 0;
        request(lines, jb, J.r0 + bid, J.s0, 0, 1, J.k, J.which, J.smp_off);
    }
    return true;
}

__device__ void finish(volatile SState* vs, SampleLine* lines, int32_t jb, int32_t status) {
    if (threadIdx.x == 0) {
        vs->status = status;
        request(lines, jb, 0, 0, 0, 0, 0, 0, 0);
    }
}

// One step of every active part (see the file comment).  State, lines and
// factor entries written by earlier launches are read through vector loads;
// every decision is computed identically by all threads from reduced values.
__global__ __launch_bounds__(kSThreads) void k_aca_step(const int32_t* act,
                                                        const SJob* __restrict__ jobs, SState* st, SampleLine* lines,
                                                        double* U, double* V, double* smp_pool, uint8_t* used_pool,
                                                        double eps, double atol) {
    __shared__ double sh_l[kRankCap];   // U row i* (row phase) or V column j* (column phases)
    __shared__ double sh_du[kRankCap];  // u_l . u_r
    __shared__ double sh_dv[kRankCap];  // v_l . v_r
    __shared__ int32_t sh_status;
    // the parts still active at the last count (rewritten between launches)
    const int32_t jb = *(const volatile int32_t*)(act + blockIdx.x);
    const SJob J = jobs[jb];
    volatile SState* vs = st + jb;
    if (threadIdx.x == 0) sh_status = vs->status;
    __syncthreads();
    if (sh_status != kActive) return;
    const int32_t rank = vs->rank, phase = vs->phase, istar = vs->istar, jstar = vs->jstar, hits = vs->hits;
    const int32_t probe = vs->probe;
    const double norm2 = vs->norm2;
    const int32_t m = J.m, k = J.k;
    double* Uj = U + J.u_off;
    double* Vj = V + J.v_off;
    double* smp = smp_pool + J.smp_off;
    uint8_t* used = used_pool + J.used_off;
    const double ref = J.ref >= 0.0 ? J.ref : sqrt(fmax(norm2, 0.0));
    // a residual row (column) is negligible when the block's residual would be
    // within half the tolerance if every row (column) were as large:
    // sqrt(m) |row| <= eps / 2 ||S||_F (sqrt(k) |column| for a column)
    auto negligible = [&](double amax, double line2, int32_t other) {
        return amax <= 0.0 || sqrt(line2 * (double)other) <= 0.5 * eps * ref || (atol > 0.0 && amax <= atol);
    };
    const int tid = threadIdx.x;
    if (phase == 0) {
        // ---- row i*: residual, pivot column
        for (int32_t l = tid; l < rank; l += kSThreads) sh_l[l] = Uj[istar + (int64_t)l * m];
        __syncthreads();
        double amax = -1.0, piv = 0.0, rr = 0.0;
        int32_t aid = 0x7fffffff;
        for (int32_t j = tid; j < k; j += kSThreads) {
            double r = smp[j];
            for (int32_t l = 0; l < rank; ++l) r -= sh_l[l] * Vj[(int64_t)l * k + j];
            smp[j] = r;
            better_max(amax, aid, piv, fabs(r), j, r);
            rr += r * r;
        }
        wg_argmax_sum(amax, aid, piv, rr);
        if (tid == 0) used[istar] = 1;
        if (!negligible(amax, rr, m)) {
            double vv = 0.0, vmax = -1.0, vsg = 0.0;
            int32_t vid = 0x7fffffff;
            const double inv = 1.0 / piv;
            for (int32_t j = tid; j < k; j += kSThreads) {
                const double v = smp[j] * inv;
                Vj[(int64_t)rank * k + j] = v;
                vv += v * v;
                better_max(vmax, vid, vsg, fabs(v), j, v);
            }
            wg_argmax_sum(vmax, vid, vsg, vv);
            if (tid == 0) {
                vs->vv = vv;
                vs->vmax = vmax;
                vs->jstar = aid;
                vs->phase = 1;
                request(lines, jb, J.r0, J.s0 + aid, 1, 0, m, J.which, J.smp_off);
            }
            return;
        }
        // negligible row: a step within tolerance without a term
        if (tid == 0) {
            vs->zrows = vs->zrows + 1;
            vs->hits = hits + 1;
        }
        __syncthreads();  // used[istar] visible to the probe search
        if (hits + 1 >= kStopHits || !next_probe(J, vs, lines, jb, probe, used)) finish(vs, lines, jb, kConverged);
        return;
    }
    // ---- a column: phase 1 after a pivot row (j*), phase 2 a probe column
    for (int32_t l = tid; l < rank; l += kSThreads) sh_l[l] = Vj[(int64_t)l * k + jstar];
    __syncthreads();
    if (phase == 2) {
        double amax = -1.0, asg = 0.0, cc = 0.0;
        int32_t aid = 0x7fffffff;
        for (int32_t i = tid; i < m; i += kSThreads) {
            if (used[i]) continue;  // rows already pivoted on are reproduced exactly
            double u = smp[i];
            for (int32_t l = 0; l < rank; ++l) u -= Uj[i + (int64_t)l * m] * sh_l[l];
            better_max(amax, aid, asg, fabs(u), i, u);
            cc += u * u;
        }
        wg_argmax_sum(amax, aid, asg, cc);
        if (tid == 0) vs->pcols = vs->pcols + 1;
        if (aid != 0x7fffffff && !negligible(amax, cc, k)) {
            // restart the cross at the probe column's largest residual row
            if (tid == 0) {
                vs->hits = 0;
                vs->istar = aid;
                vs->phase = 0;
                request(lines, jb, J.r0 + aid, J.s0, 0, 1, k, J.which, J.smp_off);
            }
            return;
        }
        if (tid == 0) vs->hits = hits + 1;
        if (hits + 1 >= kStopHits || !next_probe(J, vs, lines, jb, probe, used)) finish(vs, lines, jb, kConverged);
        return;
    }
    // phase 1: u_r = residual column (v_r was divided by the pivot)
    double umax = -1.0, usg = 0.0, uu = 0.0;
    int32_t uid = 0x7fffffff;
    double* ur = Uj + (int64_t)rank * m;
    for (int32_t i = tid; i < m; i += kSThreads) {
        double u = smp[i];
        for (int32_t l = 0; l < rank; ++l) u -= Uj[i + (int64_t)l * m] * sh_l[l];
        ur[i] = u;
        uu += u * u;
        better_max(umax, uid, usg, fabs(u), i, u);
    }
    wg_argmax_sum(umax, uid, usg, uu);  // also orders the stores to ur before the reads below
    // cross terms with the earlier terms: wavefront w takes l = w, w + 4, ...
    const int lane = tid & 63, wave = tid >> 6;
    const double* vr = Vj + (int64_t)rank * k;
    for (int32_t l = wave; l < rank; l += kSWaves) {
        double du = 0.0, dv = 0.0;
        const double* ul = Uj + (int64_t)l * m;
        const double* vl = Vj + (int64_t)l * k;
        for (int32_t i = lane; i < m; i += 64) du += ul[i] * ur[i];
        for (int32_t j = lane; j < k; j += 64) dv += vl[j] * vr[j];
        for (int o = 32; o > 0; o >>= 1) {
            du += __shfl_xor(du, o, 64);
            dv += __shfl_xor(dv, o, 64);
        }
        if (lane == 0) {
            sh_du[l] = du;
            sh_dv[l] = dv;
        }
    }
    __syncthreads();
    double cross = 0.0;
    for (int32_t l = 0; l < rank; ++l) cross += sh_du[l] * sh_dv[l];
    const double vv = vs->vv, vmax = vs->vmax;
    const double n2 = norm2 + 2.0 * cross + uu * vv;
    const double ref2 = J.ref >= 0.0 ? J.ref : sqrt(fmax(n2, 0.0));
    const bool hit = sqrt(uu * vv) <= eps * ref2 || (atol > 0.0 && umax * vmax <= atol);
    const int32_t nh = hit ? hits + 1 : 0;
    const int32_t r1 = rank + 1;
    // next pivot row: the largest |u_r| among unused rows
    double bmax = -1.0, bsg = 0.0, bd = 0.0;
    int32_t bid = 0x7fffffff;
    for (int32_t i = tid; i < m; i += kSThreads)
        if (!used[i]) better_max(bmax, bid, bsg, fabs(ur[i]), i, 0.0);
    wg_argmax_sum(bmax, bid, bsg, bd);
    if (tid == 0) {
        vs->norm2 = n2;
        vs->rank = r1;
        vs->hits = nh;
        vs->phase = 0;
    }
    if (nh >= kStopHits || bid == 0x7fffffff) {
        finish(vs, lines, jb, kConverged);
    } else if (r1 >= J.rmax) {
        finish(vs, lines, jb, kFallback);
    } else if (hit) {
        // confirm the stop test on a probe instead of the largest-|u| row
        if (!next_probe(J, vs, lines, jb, probe, used)) finish(vs, lines, jb, kConverged);
    } else if (tid == 0) {
        vs->istar = bid;
        request(lines, jb, J.r0 + bid, J.s0, 0, 1, k, J.which, J.smp_off);
    }
}

// status of every part (written by the step kernel: read through vector loads)
__global__ __launch_bounds__(256) void k_aca_status(const SState* st, int32_t njobs, int8_t* __restrict__ out) {
    const int32_t j = (int32_t)blockIdx.x * 256 + (int32_t)threadIdx.x;
    if (j >= njobs) return;
    const volatile SState* s = st + j;
    out[j] = (int8_t)s->status;
}

// progress of a build on stderr (QDFN_HMAT_VERBOSE set), wall time since the first line
void vlog(const char* fmt, ...) {
    static const bool on = getenv("QDFN_HMAT_VERBOSE") != nullptr;
    if (!on) return;
    static const auto t0 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[hmat build %9.1f ms] ", ms);
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    fflush(stderr);
    va_end(ap);
}

int launch_lines(const EntrySource& es, const SampleLine* d_lines, const SampleTile* d_tiles, int64_t ntiles,
                 double* d_outS, double* d_outN, hipStream_t stream) {
    return es.kind == 0 ? launch_rect_lines(es, d_lines, d_tiles, ntiles, d_outS, d_outN, stream)
                        : launch_tri_lines(es, d_lines, d_tiles, ntiles, d_outS, d_outN, stream);
}

int64_t workspace_budget() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return int64_t(256) << 20;
    return std::max<int64_t>(int64_t(256) << 20, std::min<int64_t>(int64_t(16) << 30, (int64_t)free_b / 3));
}

struct BuildStats {
    int64_t entries = 0;     // receiver-source pairs evaluated
    int64_t steps = 0;       // launch pairs of the sampled approximation
    int64_t sampled = 0, fallback = 0, dense = 0;
};

// Evaluate blocks in full: part w (0, 1, or 2 = both) of each listed block
// into pools (column-major, leading dimension m), offsets in doubles.
int eval_full_blocks(const EntrySource& es, const qdfn_hmat* h, const std::vector<int64_t>& blks, int which,
                     const std::vector<int64_t>& off, double* d_S, double* d_N, BuildStats& bs, hipStream_t stream) {
    std::vector<SampleLine> lines;
    std::vector<SampleTile> tiles;
    for (size_t i = 0; i < blks.size(); ++i) {
        const HBlock& b = h->blocks[blks[i]];
        const int32_t m = (int32_t)(b.r1 - b.r0 + 1), k = (int32_t)(b.s1 - b.s0 + 1);
        for (int32_t c = 0; c < k; ++c) {
            SampleLine L;
            L.rcv = (int32_t)b.r0;
            L.src = (int32_t)(b.s0 + c);
            L.drcv = 1;
            L.dsrc = 0;
            L.len = m;
            L.which = which;
            L.out = off[i] + (int64_t)c * m;
            for (int32_t f = 0; f < m; f += kLineTile) tiles.push_back({(int32_t)lines.size(), f});
            lines.push_back(L);
        }
        bs.entries += (int64_t)m * k;
    }
    if (tiles.empty()) return QDFN_OK;
    DevBuf dl, dt;
    QDFN_HIP(dl.alloc(sizeof(SampleLine) * lines.size()));
    QDFN_HIP(dt.alloc(sizeof(SampleTile) * tiles.size()));
    QDFN_HIP(hipMemcpy(dl.p, lines.data(), sizeof(SampleLine) * lines.size(), hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dt.p, tiles.data(), sizeof(SampleTile) * tiles.size(), hipMemcpyHostToDevice));
    int rc = launch_lines(es, dl.as<SampleLine>(), dt.as<SampleTile>(), (int64_t)tiles.size(), d_S, d_N, stream);
    if (rc) return rc;
    QDFN_HIP(hipStreamSynchronize(stream));
    return QDFN_OK;
}

// Inadmissible blocks: both parts evaluated in full, stored dense.
int dense_blocks(const EntrySource& es, qdfn_hmat* h, const std::vector<int64_t>& blks, BuildStats& bs) {
    const int64_t budget = workspace_budget();
    size_t pos = 0;
    while (pos < blks.size()) {
        std::vector<int64_t> sel, off;
        int64_t used = 0;
        while (pos < blks.size()) {
            const HBlock& b = h->blocks[blks[pos]];
            const int64_t mk = (b.r1 - b.r0 + 1) * (b.s1 - b.s0 + 1);
            if (!sel.empty() && (used + mk) * 16 > budget) break;
            sel.push_back(blks[pos]);
            off.push_back(used);
            used += mk;
            ++pos;
        }
        DevBuf dS, dN;
        QDFN_HIP(dS.alloc(sizeof(double) * used));
        QDFN_HIP(dN.alloc(sizeof(double) * used));
        int rc = eval_full_blocks(es, h, sel, 2, off, dS.as<double>(), dN.as<double>(), bs, h->stream);
        if (rc) return rc;
        std::vector<double> hS(used), hN(used);
        QDFN_HIP(hipMemcpy(hS.data(), dS.p, sizeof(double) * used, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(hN.data(), dN.p, sizeof(double) * used, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < sel.size(); ++i) {
            HBlock& b = h->blocks[sel[i]];
            const int64_t mk = (b.r1 - b.r0 + 1) * (b.s1 - b.s0 + 1);
            for (int w = 0; w < 2; ++w) {
                const double* a = (w ? hN : hS).data() + off[i];
                b.kind[w] = QDFN_HBLK_DENSE;
                b.rank[w] = 0;
                b.a[w].assign(a, a + mk);
                b.b[w].clear();
            }
        }
        bs.dense += (int64_t)sel.size();
    }
    return QDFN_OK;
}

// Parts of the listed blocks through the full-pivot approximation of the full
// build, after evaluation in full.  both: shear and normal parts, scaled by the
// exact shear norm (returned in shear_norm); otherwise the normal part only,
// scaled by shear_norm.
int fallback_blocks(const EntrySource& es, qdfn_hmat* h, const std::vector<int64_t>& blks, bool both,
                    double rtol, double atol, std::vector<double>& shear_norm, BuildStats& bs) {
    for (int64_t q : blks) {
        const HBlock& b = h->blocks[q];
        const int64_t mk = (b.r1 - b.r0 + 1) * (b.s1 - b.s0 + 1);
        if (mk > kFallbackMaxEntries)
            return fail(QDFN_ERR_ARG, "sampled H build: a block of " + std::to_string(mk) +
                                          " entries reached the rank cap of the cross approximation (a looser "
                                          "aca_rtol or a finer cluster tree)");
    }
    const int64_t budget = workspace_budget() / 2;
    size_t pos = 0;
    while (pos < blks.size()) {
        std::vector<int64_t> sel, off;
        int64_t used = 0;
        while (pos < blks.size()) {
            const HBlock& b = h->blocks[blks[pos]];
            const int64_t mk = (b.r1 - b.r0 + 1) * (b.s1 - b.s0 + 1);
            if (!sel.empty() && (used + mk) * 16 > budget) break;
            sel.push_back(blks[pos]);
            off.push_back(used);
            used += mk;
            ++pos;
        }
        DevBuf dS, dN;
        if (both) QDFN_HIP(dS.alloc(sizeof(double) * used));
        QDFN_HIP(dN.alloc(sizeof(double) * used));
        int rc;
        {
            ScopeMs timer(&h->build_phase_ms[0]);
            rc = eval_full_blocks(es, h, sel, both ? 2 : 1, off, dS.as<double>(), dN.as<double>(), bs, h->stream);
        }
        if (rc) return rc;
        std::vector<double> sc = shear_norm;
        if (both) {
            std::vector<const double*> src;
            std::vector<int64_t> ld;
            std::vector<int32_t> mm, kk;
            for (size_t i = 0; i < sel.size(); ++i) {
                const HBlock& b = h->blocks[sel[i]];
                src.push_back(dS.as<double>() + off[i]);
                ld.push_back(b.r1 - b.r0 + 1);
                mm.push_back((int32_t)(b.r1 - b.r0 + 1));
                kk.push_back((int32_t)(b.s1 - b.s0 + 1));
            }
            std::vector<double> nrm;
            rc = block_norms(src, ld, mm, kk, nrm, h->stream);
            if (rc) return rc;
            for (size_t i = 0; i < sel.size(); ++i) {
                sc[sel[i]] = nrm[i];
                shear_norm[sel[i]] = nrm[i];
            }
        }
        std::vector<DeviceBlock> lr;
        for (size_t i = 0; i < sel.size(); ++i) {
            const HBlock& b = h->blocks[sel[i]];
            const int64_t m = b.r1 - b.r0 + 1;
            if (both) lr.push_back({sel[i], 0, dS.as<double>() + off[i], m});
            lr.push_back({sel[i], 1, dN.as<double>() + off[i], m});
        }
        rc = compress_device_blocks(h, lr, sc, rtol, atol);
        if (rc) return rc;
        bs.fallback += (int64_t)lr.size();
        vlog("full evaluation and full-pivot approximation: %zu blocks (%lld entries each part)", sel.size(),
             (long long)used);
    }
    return QDFN_OK;
}

// Sampled approximation of part w of the listed admissible blocks.  ref[q]:
// threshold scale (< 0: the running estimate, for shear parts).  On return
// est[q] holds the final norm estimate and `fb` the blocks left to the
// fallback.
int sampled_parts(const EntrySource& es, qdfn_hmat* h, const std::vector<int64_t>& blks, int w,
                  const std::vector<double>& ref, double rtol, double atol, std::vector<double>& est,
                  std::vector<int64_t>& fb, BuildStats& bs) {
    ScopeMs timer(&h->build_phase_ms[1]);
    const int64_t budget = workspace_budget();
    const double eps = 0.5 * rtol;
    size_t pos = 0;
    while (pos < blks.size()) {
        // a batch of parts whose workspaces fit the budget
        std::vector<SJob> jobs;
        std::vector<int64_t> sel;
        std::vector<SampleTile> tiles;
        int64_t nu = 0, nv = 0, ns = 0, nb = 0;
        int32_t max_steps = 0;
        while (pos < blks.size()) {
            const HBlock& b = h->blocks[blks[pos]];
            const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
            const int64_t rmax = std::min<int64_t>(kRankCap, aca_rank_cap(m, k));
            const int64_t need = (m * rmax + rmax * k + std::max(m, k)) * 8 + m;
            if (!sel.empty() && (nu + nv + ns) * 8 + nb + need > budget) break;
            SJob J;
            J.u_off = nu;
            J.v_off = nv;
            J.smp_off = ns;
            J.used_off = nb;
            J.m = (int32_t)m;
            J.k = (int32_t)k;
            J.r0 = (int32_t)b.r0;
            J.s0 = (int32_t)b.s0;
            J.rmax = (int32_t)rmax;
            J.which = w;
            J.ref = ref[blks[pos]];
            nu += m * rmax;
            nv += rmax * k;
            ns += std::max(m, k);
            nb += m;
            const int32_t jl = (int32_t)jobs.size();
            for (int64_t f = 0; f < std::max(m, k); f += kLineTile) tiles.push_back({jl, (int32_t)f});
            // two steps per term and at most two probes after each
            max_steps = std::max<int32_t>(max_steps, (int32_t)(4 * rmax + 8));
            jobs.push_back(J);
            sel.push_back(blks[pos]);
            ++pos;
        }
        const int32_t nj = (int32_t)jobs.size();
        DevBuf dj, dst, dl, dt, dU, dV, dsm, dused, dact, dstat;
        QDFN_HIP(dj.alloc(sizeof(SJob) * nj));
        QDFN_HIP(dst.alloc(sizeof(SState) * nj));
        QDFN_HIP(dl.alloc(sizeof(SampleLine) * nj));
        QDFN_HIP(dt.alloc(sizeof(SampleTile) * tiles.size()));
        QDFN_HIP(dU.alloc(sizeof(double) * nu));
        QDFN_HIP(dV.alloc(sizeof(double) * nv));
        QDFN_HIP(dsm.alloc(sizeof(double) * ns));
        QDFN_HIP(dused.alloc(nb));
        QDFN_HIP(dact.alloc(sizeof(int32_t) * nj));
        QDFN_HIP(dstat.alloc(nj));
        std::vector<int32_t> act(nj);
        for (int32_t i = 0; i < nj; ++i) act[i] = i;
        QDFN_HIP(hipMemcpy(dj.p, jobs.data(), sizeof(SJob) * nj, hipMemcpyHostToDevice));
        QDFN_HIP(hipMemcpy(dt.p, tiles.data(), sizeof(SampleTile) * tiles.size(), hipMemcpyHostToDevice));
        QDFN_HIP(hipMemcpy(dact.p, act.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_aca_init, dim3(nj), dim3(64), 0, h->stream, dj.as<SJob>(), nj, dst.as<SState>(),
                           dl.as<SampleLine>(), dused.as<uint8_t>());
        QDFN_HIP(hipGetLastError());
        int32_t nact = nj;
        int64_t ntl = (int64_t)tiles.size();
        std::vector<int8_t> stat(nj);
        // each part needs at most max_steps line/step pairs; every 8 steps the
        // launches are narrowed to the parts still active
        for (int32_t it = 0; it < max_steps && nact > 0; ++it) {
            int rc = launch_lines(es, dl.as<SampleLine>(), dt.as<SampleTile>(), ntl, dsm.as<double>(),
                                  dsm.as<double>(), h->stream);
            if (rc) return rc;
            hipLaunchKernelGGL(k_aca_step, dim3(nact), dim3(kSThreads), 0, h->stream, dact.as<int32_t>(),
                               dj.as<SJob>(), dst.as<SState>(), dl.as<SampleLine>(), dU.as<double>(),
                               dV.as<double>(), dsm.as<double>(), dused.as<uint8_t>(), eps, atol);
            QDFN_HIP(hipGetLastError());
            ++bs.steps;
            if ((it & 7) == 7) {
                hipLaunchKernelGGL(k_aca_status, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, h->stream,
                                   dst.as<SState>(), nj, dstat.as<int8_t>());
                QDFN_HIP(hipGetLastError());
                QDFN_HIP(hipMemcpyAsync(stat.data(), dstat.p, nj, hipMemcpyDeviceToHost, h->stream));
                QDFN_HIP(hipStreamSynchronize(h->stream));
                act.clear();
                tiles.clear();
                for (int32_t i = 0; i < nj; ++i) {
                    if (stat[i] != kActive) continue;
                    const int32_t ln = std::max(jobs[i].m, jobs[i].k);
                    for (int32_t f = 0; f < ln; f += kLineTile) tiles.push_back({i, f});
                    act.push_back(i);
                }
                nact = (int32_t)act.size();
                ntl = (int64_t)tiles.size();
                if (nact > 0) {
                    QDFN_HIP(hipMemcpyAsync(dact.p, act.data(), sizeof(int32_t) * nact, hipMemcpyHostToDevice,
                                            h->stream));
                    QDFN_HIP(hipMemcpyAsync(dt.p, tiles.data(), sizeof(SampleTile) * ntl, hipMemcpyHostToDevice,
                                            h->stream));
                }
            }
        }
        QDFN_HIP(hipStreamSynchronize(h->stream));
        vlog("sampled %s parts: batch of %d, %d steps", w ? "normal" : "shear", nj, (int)bs.steps);
        std::vector<SState> s(nj);
        QDFN_HIP(hipMemcpy(s.data(), dst.p, sizeof(SState) * nj, hipMemcpyDeviceToHost));
        // the factors of the converged parts, in one packed copy
        std::vector<PackItem> items;
        for (int32_t i = 0; i < nj; ++i)
            if (s[i].status == kConverged && s[i].rank > 0) {
                const SJob& J = jobs[i];
                items.push_back({dU.as<double>() + J.u_off, J.m, J.m, s[i].rank, 0});
                items.push_back({dV.as<double>() + J.v_off, J.k, J.k, s[i].rank, 0});  // row-major r x k
            }
        std::vector<double> hp;
        int rc = pack_to_host(items, hp, h->stream);
        if (rc) return rc;
        size_t ii = 0;
        for (int32_t i = 0; i < nj; ++i) {
            const SJob& J = jobs[i];
            const int64_t q = sel[i];
            // entries sampled: a row and a column per term, the negligible rows, the probe columns
            bs.entries += (int64_t)s[i].rank * (J.m + J.k) + (int64_t)s[i].zrows * J.k + (int64_t)s[i].pcols * J.m;
            est[q] = std::sqrt(std::max(s[i].norm2, 0.0));
            if (s[i].status != kConverged) {  // rank cap (or step limit): full evaluation
                fb.push_back(q);
                continue;
            }
            HBlock& b = h->blocks[q];
            const int64_t m = J.m, k = J.k, r = s[i].rank;
            b.kind[w] = QDFN_HBLK_LOWRANK;
            b.rank[w] = r;
            b.b[w].resize(r * k);
            if (r > 0) {
                const double* u = hp.data() + items[ii++].dst;
                const double* v = hp.data() + items[ii++].dst;
                b.a[w].assign(u, u + m * r);
                for (int64_t l = 0; l < r; ++l)
                    for (int64_t c = 0; c < k; ++c) b.b[w][l + c * r] = v[l * k + c];
            } else {
                b.a[w].clear();
            }
            ++bs.sampled;
        }
    }
    return QDFN_OK;
}

int build_sampled(qdfn_hmat* h, const EntrySource& es, int64_t nblk, const int64_t* range, const int64_t* adm,
                  double rtol, double atol) {
    const auto t0 = std::chrono::steady_clock::now();
    init_blocks(h, nblk, range, adm);
    BuildStats bs;
    // with tail storage the inadmissible blocks off the diagonal (near field)
    // are evaluated in full and compressed like the full build's
    const bool near_lr = tail_mode(h->prec) > 0;
    std::vector<int64_t> dense, lowr, near;
    for (int64_t q = 0; q < nblk; ++q) {
        const HBlock& b = h->blocks[q];
        const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
        if (!b.adm || aca_rank_cap(m, k) < 1) {
            if (m * k > 0x7fffffffLL)
                return fail(QDFN_ERR_ARG, "sampled H build: an inadmissible block exceeds 2^31 - 1 entries");
            const bool diag = b.s0 == b.r0 && b.s1 == b.r1;
            if (near_lr && !b.adm && !diag && aca_rank_cap(m, k) >= 1 && m * k <= kFallbackMaxEntries)
                near.push_back(q);
            else
                dense.push_back(q);
        } else {
            lowr.push_back(q);  // sampled whatever its size
        }
    }
    int rc;
    {
        ScopeMs timer(&h->build_phase_ms[0]);
        rc = dense_blocks(es, h, dense, bs);
    }
    if (rc) return rc;
    std::vector<double> ref(nblk, -1.0), estS(nblk, 0.0), estN(nblk, 0.0), shear_norm(nblk, 0.0);
    if (!near.empty()) {
        const int64_t fb0 = bs.fallback;  // not a fallback of the sampling
        rc = fallback_blocks(es, h, near, true, rtol, atol, shear_norm, bs);
        if (rc) return rc;
        bs.fallback = fb0;
    }
    vlog("inadmissible blocks: %zu evaluated in full and stored dense, %zu near-field ones compressed; %zu "
         "admissible", dense.size(), near.size(), lowr.size());
    // shear parts: the threshold scale is each block's running norm estimate
    std::vector<int64_t> fbS, fbN;
    rc = sampled_parts(es, h, lowr, 0, ref, rtol, atol, estS, fbS, bs);
    if (rc) return rc;
    for (int64_t q : lowr) shear_norm[q] = estS[q];
    // blocks whose shear part reached the rank cap: one full evaluation gives
    // both parts, both through the full-pivot approximation
    rc = fallback_blocks(es, h, fbS, true, rtol, atol, shear_norm, bs);
    if (rc) return rc;
    vlog("shear parts: %lld sampled, %zu blocks at the rank cap evaluated in full", (long long)bs.sampled,
         fbS.size());
    // normal parts of the other blocks, scaled by the shear norm of the same block
    std::vector<bool> fellS(nblk, false);
    for (int64_t q : fbS) fellS[q] = true;
    std::vector<int64_t> lowrN;
    for (int64_t q : lowr)
        if (!fellS[q]) {
            lowrN.push_back(q);
            ref[q] = shear_norm[q];
        }
    rc = sampled_parts(es, h, lowrN, 1, ref, rtol, atol, estN, fbN, bs);
    if (rc) return rc;
    rc = fallback_blocks(es, h, fbN, false, rtol, atol, shear_norm, bs);
    if (rc) return rc;
    vlog("normal parts done: %zu at the rank cap evaluated in full", fbN.size());
    // recompression of the sampled parts to the other half of the tolerance
    if (rtol > 0) {
        std::vector<bool> fell[2] = {fellS, std::vector<bool>(nblk, false)};
        for (int64_t q : fbS) fell[1][q] = true;
        for (int64_t q : fbN) fell[1][q] = true;
        std::vector<DeviceBlock> lr;
        std::vector<size_t> items;
        std::vector<double> budgets;
        for (int64_t q : lowr)
            for (int w = 0; w < 2; ++w) {
                if (fell[w][q] || h->blocks[q].rank[w] < 2) continue;
                items.push_back(lr.size());
                lr.push_back({q, w, nullptr, 0});
                budgets.push_back(0.5 * rtol * shear_norm[q]);
            }
        recompress_parallel(h, lr, items, budgets);
        vlog("recompressed %zu sampled parts", items.size());
    }
    h->build_entries = bs.entries;
    h->build_steps = bs.steps;
    h->build_sampled = bs.sampled;
    h->build_fallback = bs.fallback;
    h->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return QDFN_OK;
}

}  // namespace

}  // namespace qdfn

using namespace qdfn;

extern "C" qdfn_hmat_t* qdfn_hmat_build_rect_sampled(const double* seg, int64_t n, int64_t ld, double mu, double nu,
                                                     int64_t nblk, const int64_t* range_SR,
                                                     const int64_t* admissible, double aca_rtol, double aca_atol,
                                                     int32_t prec) {
    if (!seg || ld < n) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_rect_sampled: null table or ld < n");
        return nullptr;
    }
    if (n > 0x7fffffffLL) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_rect_sampled: more than 2^31 - 1 elements");
        return nullptr;
    }
    if (check_blocks(n, nblk, range_SR, admissible)) return nullptr;
    qdfn_hmat* h = new_hmat(n, prec);
    if (!h) return nullptr;
    EntrySource es;
    int rc = make_rect_entries(es, seg, n, ld, mu, nu, h->stream);
    if (rc == QDFN_OK) rc = build_sampled(h, es, nblk, range_SR, admissible, aca_rtol, aca_atol);
    if (rc == QDFN_OK) rc = hmat_layout(h);
    if (rc != QDFN_OK) {
        qdfn_hmat_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" qdfn_hmat_t* qdfn_hmat_build_tri_sampled(int64_t n, const double* P1, const double* P2, const double* P3,
                                                    const double* rake, const double* ctr, const double* nrm,
                                                    const double* slip, double mu, double lambda, int64_t nblk,
                                                    const int64_t* range_SR, const int64_t* admissible,
                                                    double aca_rtol, double aca_atol, int32_t prec) {
    if (!P1 || !P2 || !P3 || !rake || !ctr || !nrm || !slip) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_tri_sampled: null input");
        return nullptr;
    }
    if (n > 0x7fffffffLL) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_tri_sampled: more than 2^31 - 1 elements");
        return nullptr;
    }
    if (check_blocks(n, nblk, range_SR, admissible)) return nullptr;
    for (int64_t s = 0; s < n; ++s)
        if (P1[s + 2 * n] > 0 || P2[s + 2 * n] > 0 || P3[s + 2 * n] > 0) {
            fail(QDFN_ERR_ARG, "Half-space solution: Z coordinates must be negative!");
            return nullptr;
        }
    qdfn_hmat* h = new_hmat(n, prec);
    if (!h) return nullptr;
    EntrySource es;
    int rc = make_tri_entries(es, n, P1, P2, P3, rake, ctr, nrm, slip, mu, lambda, h->stream);
    if (rc == QDFN_OK) rc = build_sampled(h, es, nblk, range_SR, admissible, aca_rtol, aca_atol);
    if (rc == QDFN_OK) rc = hmat_layout(h);
    if (rc != QDFN_OK) {
        qdfn_hmat_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" int qdfn_hmat_build_info(qdfn_hmat_t* h, qdfn_hmat_build_stats* out) {
    if (!h || !out) return fail(QDFN_ERR_ARG, "qdfn_hmat_build_info: null argument");
    out->entries_evaluated = h->build_entries;
    out->steps = h->build_steps;
    out->parts_sampled = h->build_sampled;
    out->blocks_fallback = h->build_fallback;
    out->ms_build = h->build_ms;
    for (int i = 0; i < 5; ++i) out->ms_phase[i] = h->build_phase_ms[i];
    return QDFN_OK;
}
