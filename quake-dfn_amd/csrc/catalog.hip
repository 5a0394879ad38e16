// Event catalog on the device: get_event_fragments and merge_fragments of
// scripts/Functions_Plot.jl:430-510 and :513-579, run on histories that stay in
// HBM (a run's own records, or host histories copied up once).
//
// Fragments (Functions_Plot.jl:467-494): for every element, the recorded slip
// rate dD/dt between consecutive records is compared with DispRateCrits; each
// rising edge opens a fragment at the later record's time, and the moment
// increments dD * FaultArea * ShearModulus are summed while the rate stays
// above the threshold.  One lane per element scans its own column
// sequentially, so every sum has the reference's order.  Fragments are
// produced in element-major order (two passes: count, exclusive scan, write)
// and then stably sorted by time (LSD radix sort, stable like Julia's
// sortperm), which reproduces the reference's fragment order exactly.
//
// Merge (:513-579): fragment i joins the event of the nearest earlier
// fragment j with t_i - t_j < MergeTimeCriteria and |x_i - x_j| <
// MergeDistanceCriteria, found by scanning back and stopping at the first j
// outside the time window; otherwise it opens a new event.  The scan is
// independent per fragment (one lane each); the resulting parent forest is
// resolved by pointer jumping, events are numbered by an exclusive scan over
// the roots (fragment order, as the reference numbers them), and each event's
// moment is summed over its fragments in fragment order after a stable sort
// by event id, so the sums are bitwise the reference's.
#include <math.h>
#include <string.h>

#include <hipcub/hipcub.hpp>
#include <vector>

#include "qdfn_internal.h"

using namespace qdfn;

namespace {

constexpr int kCatBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kCatBlock - 1) / kCatBlock); }

// pass 1: fragments per element
__global__ void __launch_bounds__(kCatBlock) k_frag_count(int64_t n_rec, int64_t n, const double* __restrict__ time,
                                                          const double* __restrict__ disp, int64_t ld, double crit,
                                                          int64_t* __restrict__ count) {
    const int64_t f = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (f >= n) return;
    const double* d = disp + f * ld;
    int on = 0;
    int64_t c = 0;
    double dp = d[0], tp = time[0];
    for (int64_t t = 1; t < n_rec; ++t) {
        const double dc = d[t], tc = time[t];
        const double rate = (dc - dp) / (tc - tp);
        if (rate > crit) {
            if (!on) {
                on = 1;
                ++c;
            }
        } else {
            on = 0;
        }
        dp = dc;
        tp = tc;
    }
    count[f] = c;
}

// pass 2: write each element's fragments at its offset (element-major order)
__global__ void __launch_bounds__(kCatBlock) k_frag_write(int64_t n_rec, int64_t n, const double* __restrict__ time,
                                                          const double* __restrict__ disp, int64_t ld, double crit,
                                                          const double* __restrict__ area, double mu,
                                                          const int64_t* __restrict__ off,
                                                          const int64_t* __restrict__ cnt, double* __restrict__ ftime,
                                                          double* __restrict__ fmom, int32_t* __restrict__ felem,
                                                          int32_t* __restrict__ iota) {
    const int64_t f = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (f >= n) return;
    const double* d = disp + f * ld;
    const double af = area[f];
    int on = 0;
    const int64_t k0 = off[f], k1 = k0 + cnt[f];  // this element's slots
    int64_t k = k0 - 1;
    double mom = 0.0;
    double dp = d[0], tp = time[0];
    for (int64_t t = 1; t < n_rec; ++t) {
        const double dc = d[t], tc = time[t];
        const double inc = dc - dp;
        const double rate = inc / (tc - tp);
        if (rate > crit) {
            const double dm = inc * af * mu;
            if (!on) {
                if (k >= k0) fmom[k] = mom;
                on = 1;
                ++k;
                if (k >= k1) return;  // the count pass saw fewer (cannot happen: same arithmetic)
                ftime[k] = tc;
                felem[k] = (int32_t)f;
                iota[k] = (int32_t)k;
                mom = dm;
            } else {
                mom = mom + dm;
            }
        } else {
            on = 0;
        }
        dp = dc;
        tp = tc;
    }
    if (k >= k0) fmom[k] = mom;
}

// sorted order -> fragment arrays (time, location = FaultCenter of the element, moment)
__global__ void __launch_bounds__(kCatBlock) k_frag_gather(int64_t F, int64_t n, const int32_t* __restrict__ perm,
                                                           const double* __restrict__ mom_in,
                                                           const int32_t* __restrict__ elem_in,
                                                           const double* __restrict__ center,
                                                           double* __restrict__ loc, double* __restrict__ mom,
                                                           int32_t* __restrict__ elem) {
    const int64_t i = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (i >= F) return;
    const int32_t j = perm[i];
    const int32_t e = elem_in[j];
    mom[i] = mom_in[j];
    elem[i] = e;
    for (int c = 0; c < 3; ++c) loc[i + F * c] = center[e + n * c];
}

// merge: nearest earlier fragment inside the time window and the distance
__global__ void __launch_bounds__(kCatBlock) k_merge_parent(int64_t F, const double* __restrict__ t,
                                                            const double* __restrict__ loc, double mt, double md,
                                                            int32_t* __restrict__ rep, int32_t* __restrict__ root) {
    const int64_t i = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (i >= F) return;
    const double ti = t[i], xi = loc[i], yi = loc[i + F], zi = loc[i + 2 * F];
    int64_t parent = -1;
    for (int64_t j = i - 1; j >= 0; --j) {
        if (!(ti - t[j] < mt)) break;
        const double dx = xi - loc[j], dy = yi - loc[j + F], dz = zi - loc[j + 2 * F];
        // 2-norm of a 3-vector: unscaled sum of squares in index order
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist < md) {
            parent = j;
            break;
        }
    }
    rep[i] = (int32_t)(parent < 0 ? i : parent);
    root[i] = parent < 0;
}

__global__ void __launch_bounds__(kCatBlock) k_jump(int64_t F, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (i >= F) return;
    b[i] = a[a[i]];
}

__global__ void __launch_bounds__(kCatBlock) k_merge_assign(int64_t F, int64_t E, const int32_t* __restrict__ rep,
                                                            const int32_t* __restrict__ root,
                                                            const int32_t* __restrict__ rank,
                                                            const double* __restrict__ t,
                                                            const double* __restrict__ loc, int32_t* __restrict__ fid,
                                                            double* __restrict__ mtime, double* __restrict__ mloc) {
    const int64_t i = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (i >= F) return;
    fid[i] = rank[rep[i]];
    if (root[i]) {
        const int32_t e = rank[i];
        mtime[e] = t[i];
        for (int c = 0; c < 3; ++c) mloc[e + E * c] = loc[i + F * c];
    }
}

__global__ void __launch_bounds__(kCatBlock) k_seg_start(int64_t F, const int32_t* __restrict__ sid,
                                                         int64_t* __restrict__ start) {
    const int64_t i = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (i >= F) return;
    if (i == 0 || sid[i] != sid[i - 1]) start[sid[i]] = i;
}

// EventMoment_Bulk in fragment order, then (log10 M - 9.05) / 1.5
__global__ void __launch_bounds__(kCatBlock) k_event_moment(int64_t F, int64_t E, const int64_t* __restrict__ start,
                                                            const double* __restrict__ smom,
                                                            double* __restrict__ mmom, double* __restrict__ mw) {
    const int64_t e = (int64_t)blockIdx.x * kCatBlock + threadIdx.x;
    if (e >= E) return;
    const int64_t end = e + 1 < E ? start[e + 1] : F;
    double s = 0.0;
    for (int64_t k = start[e]; k < end; ++k) s = s + smom[k];
    mmom[e] = s;
    mw[e] = (log10(s) - 9.05) / 1.5;
}

int bits_for(int64_t v) {
    int b = 1;
    while (b < 31 && (int64_t(1) << b) <= v) ++b;
    return b;
}

}  // namespace

struct qdfn_catalog {
    hipStream_t stream = nullptr;
    int64_t F = 0, E = -1;
    DevBuf ftime, floc, fmom, felem;         // fragments, time-sorted
    DevBuf fid, mtime, mloc, mmom, mw;       // merged events
    ~qdfn_catalog() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// fragments from device histories into c (stream-ordered after `after`)
int build_fragments(qdfn_catalog* c, int64_t n_rec, int64_t n, const double* d_time, const double* d_disp, int64_t ld,
                    const double* center, const double* area, double mu, double crit) {
    c->F = 0;
    if (n_rec < 2 || n == 0) return QDFN_OK;
    hipStream_t s = c->stream;
    DevBuf cnt, off, d_area, d_center, tmp;
    if (cnt.alloc(sizeof(int64_t) * n) || off.alloc(sizeof(int64_t) * n) || d_area.alloc(sizeof(double) * n) ||
        d_center.alloc(sizeof(double) * 3 * n))
        return fail(QDFN_ERR_NOMEM, "event catalog: allocation failed");
    QDFN_HIP(hipMemcpyAsync(d_area.p, area, sizeof(double) * n, hipMemcpyHostToDevice, s));
    QDFN_HIP(hipMemcpyAsync(d_center.p, center, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_frag_count, dim3(grid_for(n)), dim3(kCatBlock), 0, s, n_rec, n, d_time, d_disp, ld, crit,
                       cnt.as<int64_t>());
    QDFN_HIP(hipGetLastError());
    size_t tb = 0;
    QDFN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)n, s));
    if (tmp.alloc(tb)) return fail(QDFN_ERR_NOMEM, "event catalog: scan workspace");
    QDFN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)n, s));
    int64_t last_off = 0, last_cnt = 0;
    QDFN_HIP(hipMemcpyAsync(&last_off, off.as<int64_t>() + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    QDFN_HIP(hipMemcpyAsync(&last_cnt, cnt.as<int64_t>() + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    QDFN_HIP(hipStreamSynchronize(s));
    const int64_t F = last_off + last_cnt;
    if (F == 0) return QDFN_OK;
    if (F >= (int64_t(1) << 31) - 1) return fail(QDFN_ERR_ARG, "event catalog: more than 2^31 fragments");
    DevBuf t0, m0, e0, i0, perm;
    if (t0.alloc(sizeof(double) * F) || m0.alloc(sizeof(double) * F) || e0.alloc(sizeof(int32_t) * F) ||
        i0.alloc(sizeof(int32_t) * F) || perm.alloc(sizeof(int32_t) * F) || c->ftime.alloc(sizeof(double) * F) ||
        c->floc.alloc(sizeof(double) * 3 * F) || c->fmom.alloc(sizeof(double) * F) ||
        c->felem.alloc(sizeof(int32_t) * F))
        return fail(QDFN_ERR_NOMEM, "event catalog: fragment allocation failed");
    hipLaunchKernelGGL(k_frag_write, dim3(grid_for(n)), dim3(kCatBlock), 0, s, n_rec, n, d_time, d_disp, ld, crit,
                       d_area.as<double>(), mu, off.as<int64_t>(), cnt.as<int64_t>(), t0.as<double>(),
                       m0.as<double>(),
                       e0.as<int32_t>(), i0.as<int32_t>());
    QDFN_HIP(hipGetLastError());
    tb = 0;
    QDFN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, t0.as<double>(), c->ftime.as<double>(),
                                                i0.as<int32_t>(), perm.as<int32_t>(), (int)F, 0, 64, s));
    if (tmp.alloc(tb)) return fail(QDFN_ERR_NOMEM, "event catalog: sort workspace");
    QDFN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, t0.as<double>(), c->ftime.as<double>(), i0.as<int32_t>(),
                                                perm.as<int32_t>(), (int)F, 0, 64, s));
    hipLaunchKernelGGL(k_frag_gather, dim3(grid_for(F)), dim3(kCatBlock), 0, s, F, n, perm.as<int32_t>(),
                       m0.as<double>(), e0.as<int32_t>(), d_center.as<double>(), c->floc.as<double>(),
                       c->fmom.as<double>(), c->felem.as<int32_t>());
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipStreamSynchronize(s));
    c->F = F;
    return QDFN_OK;
}

qdfn_catalog* new_catalog() {
    qdfn_catalog* c = new (std::nothrow) qdfn_catalog();
    if (!c) {
        fail(QDFN_ERR_NOMEM, "event catalog: out of host memory");
        return nullptr;
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        fail(QDFN_ERR_HIP, "event catalog: hipStreamCreate failed");
        delete c;
        return nullptr;
    }
    return c;
}

}  // namespace

extern "C" qdfn_catalog_t* qdfn_catalog_from_sim(qdfn_sim_t* sim, const double* fault_center,
                                                 const double* fault_area, double shear_modulus,
                                                 double disp_rate_crit) {
    if (!sim || !fault_center || !fault_area) {
        fail(QDFN_ERR_ARG, "qdfn_catalog_from_sim: null argument");
        return nullptr;
    }
    SimHistoryView v;
    if (sim_history_view(sim, &v)) return nullptr;
    qdfn_catalog* c = new_catalog();
    if (!c) return nullptr;
    if (build_fragments(c, v.n_rec, v.n, v.time, v.disp, v.ld, fault_center, fault_area, shear_modulus,
                        disp_rate_crit)) {
        delete c;
        return nullptr;
    }
    return c;
}

extern "C" qdfn_catalog_t* qdfn_catalog_from_history(int64_t n_rec, int64_t n, const double* time,
                                                     const double* disp, int64_t ld, const double* fault_center,
                                                     const double* fault_area, double shear_modulus,
                                                     double disp_rate_crit) {
    if (n_rec < 0 || n < 0 || (n_rec > 0 && n > 0 && (!time || !disp || ld < n_rec || !fault_center || !fault_area))) {
        fail(QDFN_ERR_ARG, "qdfn_catalog_from_history: bad arguments");
        return nullptr;
    }
    qdfn_catalog* c = new_catalog();
    if (!c) return nullptr;
    DevBuf dt, dd;
    int rc = QDFN_OK;
    if (n_rec >= 2 && n > 0) {
        if (dt.alloc(sizeof(double) * n_rec) || dd.alloc(sizeof(double) * n_rec * n)) {
            rc = fail(QDFN_ERR_NOMEM, "qdfn_catalog_from_history: allocation failed");
        } else if (hipMemcpy(dt.p, time, sizeof(double) * n_rec, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy2D(dd.p, sizeof(double) * n_rec, disp, sizeof(double) * ld, sizeof(double) * n_rec, n,
                               hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(QDFN_ERR_HIP, "qdfn_catalog_from_history: copy failed");
        } else {
            rc = build_fragments(c, n_rec, n, dt.as<double>(), dd.as<double>(), n_rec, fault_center, fault_area,
                                 shear_modulus, disp_rate_crit);
        }
    }
    if (rc) {
        delete c;
        return nullptr;
    }
    return c;
}

extern "C" qdfn_catalog_t* qdfn_catalog_from_fragments(int64_t count, const double* time, const double* location,
                                                       const double* moment) {
    if (count < 0 || count >= (int64_t(1) << 31) - 1 || (count > 0 && (!time || !location || !moment))) {
        fail(QDFN_ERR_ARG, "qdfn_catalog_from_fragments: bad arguments");
        return nullptr;
    }
    qdfn_catalog* c = new_catalog();
    if (!c) return nullptr;
    if (count > 0) {
        std::vector<int32_t> minus1((size_t)count, -1);
        if (c->ftime.alloc(sizeof(double) * count) || c->floc.alloc(sizeof(double) * 3 * count) ||
            c->fmom.alloc(sizeof(double) * count) || c->felem.alloc(sizeof(int32_t) * count) ||
            hipMemcpy(c->ftime.p, time, sizeof(double) * count, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->floc.p, location, sizeof(double) * 3 * count, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->fmom.p, moment, sizeof(double) * count, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->felem.p, minus1.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess) {
            fail(QDFN_ERR_HIP, "qdfn_catalog_from_fragments: allocation or copy failed");
            delete c;
            return nullptr;
        }
    }
    c->F = count;
    return c;
}

extern "C" int qdfn_catalog_fragments(qdfn_catalog_t* c, int64_t* count, double* time, double* location,
                                      double* moment, int32_t* element) {
    if (!c || !count) return fail(QDFN_ERR_ARG, "qdfn_catalog_fragments: null argument");
    *count = c->F;
    const int64_t F = c->F;
    if (F == 0) return QDFN_OK;
    if (time) QDFN_HIP(hipMemcpy(time, c->ftime.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    if (location) QDFN_HIP(hipMemcpy(location, c->floc.p, sizeof(double) * 3 * F, hipMemcpyDeviceToHost));
    if (moment) QDFN_HIP(hipMemcpy(moment, c->fmom.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    if (element) QDFN_HIP(hipMemcpy(element, c->felem.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_catalog_merge(qdfn_catalog_t* c, double merge_time, double merge_distance, int64_t* count) {
    if (!c) return fail(QDFN_ERR_ARG, "qdfn_catalog_merge: null catalog");
    const int64_t F = c->F;
    c->E = -1;
    if (F == 0) c->E = 0;
    if (F > 0) {
        hipStream_t s = c->stream;
        DevBuf repA, repB, root, rank, tmp, sid, smom, start, iid;
        if (repA.alloc(sizeof(int32_t) * F) || repB.alloc(sizeof(int32_t) * F) || root.alloc(sizeof(int32_t) * F) ||
            rank.alloc(sizeof(int32_t) * F) || sid.alloc(sizeof(int32_t) * F) || smom.alloc(sizeof(double) * F) ||
            c->fid.alloc(sizeof(int32_t) * F))
            return fail(QDFN_ERR_NOMEM, "qdfn_catalog_merge: allocation failed");
        hipLaunchKernelGGL(k_merge_parent, dim3(grid_for(F)), dim3(kCatBlock), 0, s, F, c->ftime.as<double>(),
                           c->floc.as<double>(), merge_time, merge_distance, repA.as<int32_t>(),
                           root.as<int32_t>());
        QDFN_HIP(hipGetLastError());
        // parent links point to smaller indices, so ceil(log2 F) rounds reach every root
        int32_t* a = repA.as<int32_t>();
        int32_t* b = repB.as<int32_t>();
        for (int r = 0; r < bits_for(F); ++r) {
            hipLaunchKernelGGL(k_jump, dim3(grid_for(F)), dim3(kCatBlock), 0, s, F, a, b);
            std::swap(a, b);
        }
        QDFN_HIP(hipGetLastError());
        size_t tb = 0;
        QDFN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, root.as<int32_t>(), rank.as<int32_t>(), (int)F, s));
        if (tmp.alloc(tb)) return fail(QDFN_ERR_NOMEM, "qdfn_catalog_merge: scan workspace");
        QDFN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, root.as<int32_t>(), rank.as<int32_t>(), (int)F, s));
        int32_t lr = 0, lo = 0;
        QDFN_HIP(hipMemcpyAsync(&lr, rank.as<int32_t>() + F - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        QDFN_HIP(hipMemcpyAsync(&lo, root.as<int32_t>() + F - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        QDFN_HIP(hipStreamSynchronize(s));
        const int64_t E = (int64_t)lr + lo;
        if (c->mtime.alloc(sizeof(double) * E) || c->mloc.alloc(sizeof(double) * 3 * E) ||
            c->mmom.alloc(sizeof(double) * E) || c->mw.alloc(sizeof(double) * E) ||
            start.alloc(sizeof(int64_t) * E))
            return fail(QDFN_ERR_NOMEM, "qdfn_catalog_merge: allocation failed");
        hipLaunchKernelGGL(k_merge_assign, dim3(grid_for(F)), dim3(kCatBlock), 0, s, F, E, a, root.as<int32_t>(),
                           rank.as<int32_t>(), c->ftime.as<double>(), c->floc.as<double>(), c->fid.as<int32_t>(),
                           c->mtime.as<double>(), c->mloc.as<double>());
        QDFN_HIP(hipGetLastError());
        tb = 0;
        QDFN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, c->fid.as<int32_t>(), sid.as<int32_t>(),
                                                    c->fmom.as<double>(), smom.as<double>(), (int)F, 0,
                                                    bits_for(E), s));
        if (tmp.alloc(tb)) return fail(QDFN_ERR_NOMEM, "qdfn_catalog_merge: sort workspace");
        QDFN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, c->fid.as<int32_t>(), sid.as<int32_t>(),
                                                    c->fmom.as<double>(), smom.as<double>(), (int)F, 0,
                                                    bits_for(E), s));
        hipLaunchKernelGGL(k_seg_start, dim3(grid_for(F)), dim3(kCatBlock), 0, s, F, sid.as<int32_t>(),
                           start.as<int64_t>());
        hipLaunchKernelGGL(k_event_moment, dim3(grid_for(E)), dim3(kCatBlock), 0, s, F, E, start.as<int64_t>(),
                           smom.as<double>(), c->mmom.as<double>(), c->mw.as<double>());
        QDFN_HIP(hipGetLastError());
        QDFN_HIP(hipStreamSynchronize(s));
        c->E = E;
    }
    if (count) *count = c->E;
    return QDFN_OK;
}

extern "C" int qdfn_catalog_events(qdfn_catalog_t* c, int64_t* count, double* time, double* magnitude,
                                   double* location, double* moment, int32_t* fragment_event) {
    if (!c || !count) return fail(QDFN_ERR_ARG, "qdfn_catalog_events: null argument");
    if (c->E < 0) return fail(QDFN_ERR_ARG, "qdfn_catalog_events: qdfn_catalog_merge has not run");
    *count = c->E;
    const int64_t E = c->E, F = c->F;
    if (E == 0) return QDFN_OK;
    if (time) QDFN_HIP(hipMemcpy(time, c->mtime.p, sizeof(double) * E, hipMemcpyDeviceToHost));
    if (magnitude) QDFN_HIP(hipMemcpy(magnitude, c->mw.p, sizeof(double) * E, hipMemcpyDeviceToHost));
    if (location) QDFN_HIP(hipMemcpy(location, c->mloc.p, sizeof(double) * 3 * E, hipMemcpyDeviceToHost));
    if (moment) QDFN_HIP(hipMemcpy(moment, c->mmom.p, sizeof(double) * E, hipMemcpyDeviceToHost));
    if (fragment_event) QDFN_HIP(hipMemcpy(fragment_event, c->fid.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" void qdfn_catalog_destroy(qdfn_catalog_t* c) { delete c; }
