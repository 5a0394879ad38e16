// Internal entry points shared between translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qdfn_internal.h"

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

// ---- sampled entries (H-matrix construction without full assembly) --------
//
// A sample line is a run of `len` stiffness entries: entry e pairs receiver
// rcv + e * drcv with source src + e * dsrc (element indices, 0-based).  which
// 0 writes the shear entry to outS[out + e], 1 the normal entry to
// outN[out + e], 2 both.  len 0: nothing to evaluate.  Lines may be rewritten
// on the device between launches (the cross approximation chooses its rows and
// columns there).
struct SampleLine {
    int32_t rcv, src, drcv, dsrc, len, which;
    int64_t out;
};
// 64 entries of one line, starting at entry `first`
struct SampleTile {
    int32_t line, first;
};

// Lines are written by earlier kernels: read them through vector loads so a
// uniform (scalar-cache) load never returns a line from before the rewrite.
__device__ __forceinline__ bool read_line(const SampleLine* lines, const SampleTile& t, SampleLine& L) {
    const volatile SampleLine* v = lines + t.line;
    L.len = v->len;
    if (t.first >= L.len) return false;
    L.rcv = v->rcv;
    L.src = v->src;
    L.drcv = v->drcv;
    L.dsrc = v->dsrc;
    L.which = v->which;
    L.out = v->out;
    return true;
}

__device__ __forceinline__ void write_entry(const SampleLine& L, int32_t e, double ks, double kn,
                                            double* __restrict__ outS, double* __restrict__ outN) {
    if (L.which != 1) outS[L.out + e] = ks;
    if (L.which != 0) outN[L.out + e] = kn;
}

// Per-element source constants on the device, for line evaluation.
struct EntrySource {
    int kind = 0;                 // 0 rectangles, 1 triangles
    DevBuf srcs;                  // RectSource[n] or TriSource[n]
    DevBuf tab;                   // rectangles: receiver table (n x 8); triangles: ctr, nrm, slip (n x 3 each)
    int64_t n = 0;
    double mu = 0.0, lam = 0.0;
};

// Rectangles: seg is a HOST segment table (n x >= 8, leading dimension ld).
int make_rect_entries(EntrySource& es, const double* seg, int64_t n, int64_t ld, double mu, double nu,
                      hipStream_t stream);
// Triangles: HOST arrays as qdfn_tde_stiffness_tri takes them.
int make_tri_entries(EntrySource& es, int64_t n, const double* P1, const double* P2, const double* P3,
                     const double* rake, const double* ctr, const double* nrm, const double* slip, double mu,
                     double lambda, hipStream_t stream);
// Evaluate the lines the tiles cover (no synchronisation).
int launch_rect_lines(const EntrySource& es, const SampleLine* d_lines, const SampleTile* d_tiles, int64_t ntiles,
                      double* d_outS, double* d_outN, hipStream_t stream);
int launch_tri_lines(const EntrySource& es, const SampleLine* d_lines, const SampleTile* d_tiles, int64_t ntiles,
                     double* d_outS, double* d_outN, hipStream_t stream);

}  // namespace qdfn
