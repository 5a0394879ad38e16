// H-matrix operator: host block store and device layout (see hmat.hip).
#pragma once
#include <chrono>
#include <vector>

#include "dense.h"
#include "sim_device.h"
#include "qdfn_internal.h"

namespace qdfn {

// One block of ElementRange_SR (0-based inclusive ranges) with its shear (0)
// and normal (1) parts: dense a (m x k) or low-rank a = U (m x r), b = V (r x k),
// column-major FP64 as built or imported.
struct HBlock {
    int64_t s0 = 0, s1 = 0, r0 = 0, r1 = 0;
    int adm = 0;
    int kind[2] = {0, 0};
    int64_t rank[2] = {0, 0};
    std::vector<double> a[2], b[2];
    // recompressed low-rank parts are in SVD form (U = Q Sigma with columns of
    // decreasing norm sigma_j, V with orthonormal rows); slack = the part of
    // its Frobenius error budget the truncation left unused (0: not in SVD form)
    double slack[2] = {0.0, 0.0};
    // components [0, r_hi) of a low-rank part are stored in the FP64 / FP32
    // pool, [r_hi, r_8) in the 16-bit pool, [r_8, rank) in the 8-bit pool (set
    // by hmat_layout); with QDFN_PREC_MIXED / _MIXED8, tail component r_hi + j
    // is U's column times fu[j] and V's row times fv[j], its t row scaled by
    // fc[j] (choose_tails)
    int64_t r_hi[2] = {0, 0};
    int64_t r_8[2] = {0, 0};
    std::vector<double> fu[2], fv[2], fc[2];
};

// Phase-2 unit (one workgroup): columns [0, k) of one panel of a row group,
// the group being rows [out, out + m) of one matrix (kind 1 shear, 2
// normal).  The columns are column-major at off in the FP64 / FP32 pool
// (lo 0; column stride ld = m rounded up to four), in the 16-bit pool (lo 1)
// or in the 8-bit pool (lo 2; ld = m rounded up to eight in both); column c
// multiplies w[idx[z + c]], w = [x; t].  Unit u of its group writes its row sums to slot u.
struct HUnit {
    int64_t off, z;
    int32_t k, ld, lo, out, m, kind, u, pad;
};

// phase-1 wavefront task: rows [0, m) of a column-major V sub-panel slice at
// data_off (column stride ld, a multiple of the entries per lane load; K
// columns multiplying x[xoff + c]); row q writes t[out + q], or vpart[part + q]
// when the sub-panel is split into column segments (part >= 0); lpr = lanes per
// column; lo: the sub-panel is in the 16-bit (1) or 8-bit (2) pool
struct VTask {
    int64_t data_off;
    int32_t xoff, out, m, K, ld, lpr, part, lo;
};

// phase-1 finish of one row of a column-split panel:
// t[out] = sum over q < nseg of vpart[part0 + q * stride]
struct VFin {
    int64_t part0;
    int32_t stride, nseg, out, pad;
};

}  // namespace qdfn

struct qdfn_hmat {
    int64_t n = 0;
    int32_t prec = QDFN_PREC_FP64;
    int32_t normal_zero = 0;
    std::vector<qdfn::HBlock> blocks;
    std::vector<double> kdiag_h;          // K_S,ii (0 where no diagonal block)
    qdfn::DevBuf kdiag;
    qdfn::DevBuf pool, idx, t;            // panel entries (V sub-panels then row-group panels), gather indices, t
    qdfn::DevBuf pool_lo;                 // 16-bit V sub-panels, then row groups' U panels (BF16_FAR: BF16; MIXED: FP16)
    qdfn::DevBuf pool_8;                  // QDFN_PREC_MIXED8: FP8 (E4M3) V sub-panels, then row groups' U panels
    qdfn::DevBuf tscale;                  // QDFN_PREC_MIXED(8): scale of each t row (16- and 8-bit components, else 1)
    qdfn::DevBuf units, part;             // phase-2 units, their slots (n_slots x [y_S; y_N] row sums)
    qdfn::DevBuf vtasks, vfins, vpart;    // phase-1 wavefront tasks, split-panel finish, partial sums
    qdfn::DevBuf x, yS, yN;               // vectors for qdfn_hmat_matvec
    int64_t n_t = 0, n_units = 0, n_slots = 0, n_vtasks = 0, n_vfins = 0;
    int64_t stored_entries = 0;
    int64_t device_bytes = 0;
    int64_t phase_bytes = 0;              // HBM bytes one product reads and writes
    int64_t alg_bytes1 = 0, alg_bytes2 = 0; // algorithmic bytes of the phase-1 / phase-2 main kernels
    // their parts: stored entries (unpadded), x slices and t (phase 1); stored
    // entries, index list (4 B per column), gathered right-hand side (8 B per
    // column), outputs (phase 2)
    int64_t p1_entries = 0, p1_other = 0, p2_entries = 0, p2_index = 0, p2_gather = 0, p2_out = 0;
    hipStream_t stream = nullptr;
    int32_t users = 0;                    // runs created on it (their step kernels read part and kdiag)
    // construction statistics
    int64_t build_entries = 0, build_steps = 0, build_sampled = 0, build_fallback = 0;
    double build_ms = 0.0;
    double build_phase_ms[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // see qdfn_hmat_build_stats
};

namespace qdfn {
// Adds the wall time of its scope to *acc (milliseconds).
struct ScopeMs {
    double* acc;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit ScopeMs(double* a) : acc(a) {}
    ~ScopeMs() { *acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

// Both phases of one product on stream s.  form 1: yS = (H_load (0 - x)) ./ K_ii,
// form 0: yS = H_S x; yN = H_N x.  The gate skips the product as for the dense
// coupling.
// Events (optional): ev0 before phase 1, ev1 between the phases, ev2 after
// phase 2.  finish: k_hsum forms yS / yN from the unit slots; otherwise the
// caller's next kernel does (the step kernel: the slots' sums in slot order).
int launch_hmat(const qdfn_hmat* h, const double* d_x, const CouplingGate& g, int32_t form, double* d_yS,
                double* d_yN, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, hipEvent_t ev2, bool finish = true);

// ---- construction helpers shared by the full and the sampled builds ----
qdfn_hmat* new_hmat(int64_t n, int32_t prec);
int check_blocks(int64_t n, int64_t nblk, const int64_t* range, const int64_t* adm);
// blocks from ElementRange_SR (1-based) and Admissible
void init_blocks(qdfn_hmat* h, int64_t nblk, const int64_t* range, const int64_t* adm);
// device panels and product tasks from h->blocks
int hmat_layout(qdfn_hmat* h);
// rank cap of the cross approximation of an m x k block (fp16_tail: the
// break-even rank of FP32 + FP16-tail factors against the dense FP32 block)
int64_t aca_rank_cap(int64_t m, int64_t k, bool fp16_tail = false);
// share of a block's tolerance the full build's cross approximation may use
// (fp16_tail: QDFN_PREC_MIXED storage)
double aca_share(bool fp16_tail);
// U V (m x r, r x k) -> factors in SVD form within `budget` (Frobenius): the
// shortest, or with tail storage (tail_mode > 0) those whose tiered storage
// is smallest; returns the unused budget (budget - ||dropped tail||_F, >= 0), or
// 0 when the factors were left as they were (r < 2 or no budget)
double recompress(int64_t m, int64_t k, int64_t& r, std::vector<double>& U, std::vector<double>& V, double budget,
                  int tail);
// Rounding of the components of a low-rank part in SVD form (U m x r, column
// a of norm sig; V r x k with leading dimension ldv, orthonormal rows), each
// relative to its own scale: eu32 / ev32 of U's column / V's row copied to
// FP32, eu16 / ev16 of the FP16 unit vectors of QDFN_PREC_MIXED
struct CompRound {
    double sig, eu32, ev32, eu16, ev16, eu8, ev8;
};
std::vector<CompRound> component_rounding(const double* U, const double* V, int64_t m, int64_t k, int64_t r,
                                          int64_t ldv);
// Storage tiers of a part in SVD form whose stored factors stay within
// `slack` of the exact ones (see hmat.hip): components [0, r_hi) in FP32,
// [r_hi, r_8) in FP16, [r_8, r) in FP8 (fp8 false: r_8 = r)
void tail_tiers(const std::vector<CompRound>& c, int64_t r, double slack, bool fp8, int64_t& r_hi, int64_t& r_8);
// the bytes of an m x k part of r components stored in those tiers
inline int64_t tier_bytes(int64_t m, int64_t k, int64_t r, int64_t r_hi, int64_t r_8) {
    return (m + k) * (4 * r_hi + 2 * (r_8 - r_hi) + (r - r_8));
}
// tail storage of a precision: 0 none, 1 FP16 tails (QDFN_PREC_MIXED), 2 FP16
// and FP8 tails (QDFN_PREC_MIXED8)
inline int tail_mode(int32_t prec) { return prec == QDFN_PREC_MIXED ? 1 : prec == QDFN_PREC_MIXED8 ? 2 : 0; }
// FP8 E4M3 (OCP) rounding of the host layout, and its value
uint8_t to_fp8(double d);
double from_fp8(uint8_t b);

// part w of block q, held in device memory (column-major, leading dimension ld)
struct DeviceBlock {
    int64_t q;
    int w;
    const double* src;
    int64_t ld;
};
// a rows x cols column-major device matrix (leading dimension ld) to pack at dst
struct PackItem {
    const double* src;
    int64_t ld;
    int32_t rows, cols;
    int64_t dst;
};
// Copy the items to the host in one transfer: item i lands at host[items[i].dst]
// (offsets assigned here, in order).
int pack_to_host(std::vector<PackItem>& items, std::vector<double>& host, hipStream_t stream);

// Frobenius norms of device blocks (fixed-order sums)
int block_norms(const std::vector<const double*>& src, const std::vector<int64_t>& ld,
                const std::vector<int32_t>& m, const std::vector<int32_t>& k, std::vector<double>& out,
                hipStream_t stream);
// full-pivot cross approximation + recompression of the listed parts;
// bnorm[q] is block q's threshold scale ||K_S,b||_F.  Parts that reach the
// rank cap are stored dense.
int compress_device_blocks(qdfn_hmat* h, const std::vector<DeviceBlock>& lr, const std::vector<double>& bnorm,
                           double rtol, double atol);
// recompress(items[i]) with budgets[i] on the host threads
void recompress_parallel(qdfn_hmat* h, const std::vector<DeviceBlock>& lr, const std::vector<size_t>& items,
                         const std::vector<double>& budgets);
}  // namespace qdfn
