// Internal entry points shared between translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qdfn {

// Rectangle stiffness into device buffers (column-major, leading dim ldk).
// src is a HOST segment table; d_rcv a DEVICE table with >= 8 columns.
int assemble_rect_device(const double* src_host, int64_t ns, int64_t lds, const double* d_rcv, int64_t nr,
                         int64_t ldr, double mu, double nu, double* d_KS, double* d_KN, int64_t ldk,
                         hipStream_t stream);

// Triangle stiffness into device buffers; all inputs are DEVICE arrays
// (n x 3 column-major with leading dimension = count).
int assemble_tri_device(int64_t ns, const double* d_P1, const double* d_P2, const double* d_P3,
                        const double* d_rake, int64_t nr, const double* d_ctr, const double* d_nrm,
                        const double* d_slip, double mu, double lambda, double* d_KS, double* d_KN, int64_t ldk,
                        hipStream_t stream);

}  // namespace qdfn
