// H-matrix operator: host block store and device layout (see hmat.hip).
#pragma once
#include <vector>

#include "dense.h"
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
};

// phase-1 job: one row of a low-rank factor's V (row-major, zero-padded to a
// multiple of 4 entries at offset `off` of the V pool), applied to x[src, src + k)
struct HVJob {
    int64_t off, src;
    int32_t k, pad;
};

// phase-2 panel: rows [row0, row0 + m) of one matrix (which: 0 shear, 1 normal),
// row-major m x ld at offset `off` of the panel pool; column c multiplies
// w[idx[idx_off + c]] with w = [x; t]
struct HPanel {
    int64_t off, idx_off, row0;
    int32_t m, K, ld, which;
};

}  // namespace qdfn

struct qdfn_hmat {
    int64_t n = 0;
    int32_t prec = QDFN_PREC_FP64;
    int32_t normal_zero = 0;
    std::vector<qdfn::HBlock> blocks;
    std::vector<double> kdiag_h;          // K_S,ii (0 where no diagonal block)
    qdfn::DevBuf kdiag;
    qdfn::DevBuf vpool, jobs, t;          // phase 1
    qdfn::DevBuf ppool, panels, idx;      // phase 2
    qdfn::DevBuf x, yS, yN;               // vectors for qdfn_hmat_matvec
    int64_t n_jobs = 0, n_panels = 0;
    int64_t stored_entries = 0;
    int64_t device_bytes = 0;
    int64_t phase_bytes = 0;              // HBM bytes one product reads and writes
    hipStream_t stream = nullptr;
};

namespace qdfn {
// Both phases of one product on stream s.  form 1: yS = (H_load (0 - x)) ./ K_ii,
// form 0: yS = H_S x; yN = H_N x.  The gate skips the product as for the dense
// coupling.
int launch_hmat(const qdfn_hmat* h, const double* d_x, const CouplingGate& g, int32_t form, double* d_yS,
                double* d_yN, hipStream_t s);
}  // namespace qdfn
