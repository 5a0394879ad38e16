// Dense coupling operator: device layout and launch helpers.
//
// HBM layout: E (zero-diagonal K_S,ij/K_S,ii) and K_N, ld = round_up(n, 64)
// rows, in FP64 or FP32, column-major or (large operators) split-tiled
// (DenseTiles).  The coupling kernel streams both matrices once per RHS
// evaluation, each lane owning 2 (FP64) or 4 (FP32) consecutive rows and
// reading them with one 16-byte load per column, so a wavefront moves a
// contiguous 1 KiB per load instruction.  The column range is split across
// workgroups (split-K); each split writes its partial sums and a second,
// small kernel adds the splits in a fixed order, so the result is
// deterministic run to run.
#pragma once
#include "qdfn_internal.h"

struct qdfn_dense {
    int64_t n = 0, ld = 0;
    int32_t prec = QDFN_PREC_FP64;
    int32_t normal_zero = 0;
    int32_t splits = 1;
    int64_t cols_per_split = 0;
    int32_t tiled = 0;          // 1: split-tiled layout (qdfn::DenseTiles), 0: column-major ld x n
    qdfn::DevBuf E, KN;         // ld x n, element type by prec
    qdfn::DevBuf kself, diag;   // n doubles: |K_ii| and signed K_ii
    qdfn::DevBuf partS, partN;  // splits x ld doubles
    qdfn::DevBuf x, yS, yN;     // scratch vectors for qdfn_matvec
    hipStream_t stream = nullptr;
    int device = 0;             // HIP device the operator's memory lives on
    int32_t users = 0;          // runs created on it (their step kernels read partS / partN)
};

namespace qdfn {

// Split-tiled layout of an ld x n operator (FP64 operators of at least
// kTiledMinN elements; 8,192 elements: coupling 168 -> 158 us): the columns are cut into the coupling kernel's splits
// of cps columns and the rows into its row blocks of rb_rows rows; the tile of
// (row block b, split s) holds rows [b rb_rows, +rows) x columns [s cps, +w)
// column-major with leading dimension rows, and the tiles follow each other
// split by split, row block by row block.  A coupling workgroup then streams
// one contiguous region instead of w columns ld entries apart.
constexpr int64_t kTiledMinN = 6144;
struct DenseTiles {
    int64_t ld, n, cps, rb_rows;
    __host__ __device__ int64_t index(int64_t i, int64_t j) const {
        const int64_t j0 = (j / cps) * cps, w = (n - j0 < cps) ? n - j0 : cps;
        const int64_t r0 = (i / rb_rows) * rb_rows, rows = (ld - r0 < rb_rows) ? ld - r0 : rb_rows;
        return j0 * ld + r0 * w + (j - j0) * rows + (i - r0);
    }
};
inline DenseTiles dense_tiles(const qdfn_dense* op) {
    const int64_t R = op->prec == QDFN_PREC_FP64 ? 2 : 4;
    return DenseTiles{op->ld, op->n, op->cols_per_split, 256 * R};
}

// Optional device-side control of a coupling launch:
//   gate   - when non-null and *gate == 0 the kernel returns at once (the
//            step controller reuses the previous coupling on retries);
//   parity - when non-null the input vector is x + (*parity) * parity_stride
//            (the double-buffered DispOld of a run);
//   ran_log[slot] - set to 1/0 by the kernel: whether it streamed the matrices.
struct CouplingGate {
    const int* gate = nullptr;
    const int* parity = nullptr;
    int64_t parity_stride = 0;
    int* ran_log = nullptr;
    int slot = 0;
};

// Launch the split coupling kernel: partial sums of E*x and K_N*x, into the
// operator's partS / partN, or into part_s / part_n (splits x ld doubles
// each) when given.
int launch_coupling(const qdfn_dense* op, const double* d_x, const CouplingGate& g, hipStream_t s,
                    double* part_s = nullptr, double* part_n = nullptr);
// Finish: TL = -(sum of E*x partials), SigProd = sum of K_N*x partials
// (TL = E*(0-x) as Functions_RSFDFN3DMain_D.jl:199); partials as above.
int launch_coupling_finish(const qdfn_dense* op, double* d_tl, double* d_sigprod, const int* d_gate,
                           hipStream_t s, const double* part_s = nullptr, const double* part_n = nullptr);

}  // namespace qdfn
