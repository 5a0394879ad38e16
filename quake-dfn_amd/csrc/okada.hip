// Okada rectangle kernels:
//   k_dc3d            - Okada_DC3D_Vector (scripts/Functions_OKADA3D.jl:464-654),
//                       one thread per receiver point.
//   k_stiffness_rect  - StiffnessMatrix_ByParts_Calculation_Rec
//                       (scripts/Functions_BuildInputFile.jl:386-512) fused:
//                       source frame -> DC3D gradients -> Hooke -> receiver
//                       rotation -> shear/normal traction, one thread per
//                       (receiver, source) pair, nothing staged through HBM
//                       but the two output entries.
//   k_rect_lines      - the same entries along sample lines (rows / columns of
//                       H-matrix blocks chosen by the cross approximation).
// Bound: FP64 VALU (transcendental-heavy); output is 16 B per pair.
#include "okada_device.h"
#include "qdfn_internal.h"
#include "assembly.h"

namespace qdfn {

__global__ __launch_bounds__(256) void k_dc3d(DipConst k, int64_t n, const double* __restrict__ X,
                                              const double* __restrict__ Y, const double* __restrict__ Z,
                                              double depth, double al1, double al2, double aw1, double aw2,
                                              SlipW s, double* __restrict__ U, double* __restrict__ iret) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[12], ir;
    dc3d_point<true>(k, X[i], Y[i], Z[i], depth, al1, al2, aw1, aw2, s, u, &ir);
#pragma unroll
    for (int c = 0; c < 12; ++c) U[i + (int64_t)c * n] = u[c];
    iret[i] = ir;
}

// Per-source constants, computed once on the host side of the launch.
struct RectSource {
    double x, y, depth, al, aw, cs, sn, strike;
    DipConst k;
    SlipW s;
};

// One stiffness entry: shear / normal traction at receiver r (row r of the
// receiver table, leading dimension ldr) from unit slip on source `src`.
__device__ __forceinline__ void rect_entry(const RectSource& src, const double* __restrict__ rcv, int64_t r,
                                           int64_t ldr, double mu, double lam, double& ks, double& kn) {
    const double xd = rcv[r] - src.x, yd = rcv[r + ldr] - src.y;
    const double X = xd * src.cs - yd * src.sn;
    const double Y = xd * src.sn + yd * src.cs;
    const double Z = -rcv[r + 2 * ldr];
    double u[12];
    dc3d_point<false>(src.k, X, Y, Z, src.depth, src.al, src.al, src.aw, src.aw, src.s, u, nullptr);
    // u: 3 UXX, 4 UYX, 5 UZX, 6 UXY, 7 UYY, 8 UZY, 9 UXZ, 10 UYZ, 11 UZZ
    const double tr = u[3] + u[7] + u[11];
    double S[3][3];
    S[0][0] = lam * tr + 2.0 * mu * u[3];
    S[1][1] = lam * tr + 2.0 * mu * u[7];
    S[2][2] = lam * tr + 2.0 * mu * u[11];
    S[0][1] = S[1][0] = (u[6] + u[4]) * mu;
    S[0][2] = S[2][0] = (u[9] + u[5]) * mu;
    S[1][2] = S[2][1] = (u[10] + u[8]) * mu;
    // R = R_dip(-dip_r) * R_strike(-(strike_r - strike_s))   (:480-490)
    const double rel = rcv[r + 5 * ldr] - src.strike;
    const double rdip = rcv[r + 6 * ldr];
    const double c1 = cos_deg(-rel), s1 = sin_deg(-rel);
    const double c2 = cos_deg(-rdip), s2 = sin_deg(-rdip);
    const double Rs[3][3] = {{c1, -s1, 0.0}, {s1, c1, 0.0}, {0.0, 0.0, 1.0}};
    const double Rd[3][3] = {{1.0, 0.0, 0.0}, {0.0, c2, -s2}, {0.0, s2, c2}};
    double R[3][3], RS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) R[i][j] = Rd[i][0] * Rs[0][j] + Rd[i][1] * Rs[1][j] + Rd[i][2] * Rs[2][j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) RS[i][j] = R[i][0] * S[0][j] + R[i][1] * S[1][j] + R[i][2] * S[2][j];
    // only row 3 column 3 and column 3 of rows 1-2 of R S R' are needed
    const double s13 = RS[0][0] * R[2][0] + RS[0][1] * R[2][1] + RS[0][2] * R[2][2];
    const double s23 = RS[1][0] * R[2][0] + RS[1][1] * R[2][1] + RS[1][2] * R[2][2];
    const double s33 = RS[2][0] * R[2][0] + RS[2][1] * R[2][1] + RS[2][2] * R[2][2];
    const double rrake = rcv[r + 7 * ldr];
    kn = -s33;
    ks = cos_deg(rrake) * s13 + sin_deg(rrake) * s23;
}

__global__ __launch_bounds__(256) void k_stiffness_rect(const RectSource* __restrict__ srcs, int64_t ns,
                                                        const double* __restrict__ rcv, int64_t nr, int64_t ldr,
                                                        double mu, double lam, double* __restrict__ KS,
                                                        double* __restrict__ KN, int64_t ldk, int64_t rblocks) {
    const int64_t bid = blockIdx.x;
    const int64_t s = bid / rblocks;
    const int64_t r = (bid - s * rblocks) * 256 + threadIdx.x;
    if (s >= ns || r >= nr) return;
    const RectSource src = srcs[s];
    double ks, kn;
    rect_entry(src, rcv, r, ldr, mu, lam, ks, kn);
    KN[r + s * ldk] = kn;
    KS[r + s * ldk] = ks;
}

// Entries along sample lines (sampled H-matrix construction): wavefront w of
// workgroup b takes tile 4 b + w, 64 consecutive entries of one line.
__global__ __launch_bounds__(256) void k_rect_lines(const RectSource* __restrict__ srcs,
                                                    const double* __restrict__ rcv, int64_t ldr, double mu,
                                                    double lam, const SampleLine* lines,
                                                    const SampleTile* __restrict__ tiles, int64_t ntiles,
                                                    double* __restrict__ outS, double* __restrict__ outN) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    const SampleTile tl = tiles[t];
    SampleLine L;
    if (!read_line(lines, tl, L)) return;
    const int32_t e = tl.first + (int32_t)(threadIdx.x & 63);
    if (e >= L.len) return;
    const int64_t r = L.rcv + (int64_t)e * L.drcv, s = L.src + (int64_t)e * L.dsrc;
    double ks, kn;
    rect_entry(srcs[s], rcv, r, ldr, mu, lam, ks, kn);
    write_entry(L, e, ks, kn, outS, outN);
}

static SlipW make_slip(double d1, double d2, double d3) {
    SlipW s;
    const double d[3] = {d1, d2, d3};
    for (int k = 0; k < 3; ++k) {
        s.w[k] = d[k] / QDFN_TWO_PI;
        s.f[k] = d[k] != 0.0 ? 1.0 : 0.0;
    }
    return s;
}

// Per-source constants of a HOST segment table (n x >= 8, leading dimension lds).
static RectSource* rect_sources(const double* src_host, int64_t ns, int64_t lds, double alpha) {
    RectSource* h = new RectSource[ns];
    for (int64_t s = 0; s < ns; ++s) {
        RectSource& q = h[s];
        q.x = src_host[s];
        q.y = src_host[s + lds];
        q.depth = src_host[s + 2 * lds];
        q.al = src_host[s + 3 * lds] / 2.0;
        q.aw = src_host[s + 4 * lds] / 2.0;
        q.strike = src_host[s + 5 * lds];
        q.cs = cos_deg(-q.strike);
        q.sn = sin_deg(-q.strike);
        q.k = make_dip_const(alpha, src_host[s + 6 * lds]);
        const double rake = src_host[s + 7 * lds];
        q.s = make_slip(cos_deg(rake), sin_deg(rake), 0.0);
    }
    return h;
}

// Device-side assembly into caller-provided device buffers (used by the
// public entry points and by the dense operator's on-device build).
int assemble_rect_device(const double* src_host, int64_t ns, int64_t lds, const double* d_rcv, int64_t nr,
                         int64_t ldr, double mu, double nu, double* d_KS, double* d_KN, int64_t ldk,
                         hipStream_t stream) {
    if (ns <= 0 || nr <= 0) return QDFN_OK;
    const double lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
    const double alpha = (lam + mu) / (lam + 2.0 * mu);
    RectSource* h = rect_sources(src_host, ns, lds, alpha);
    DevBuf d_src;
    hipError_t e = d_src.alloc(sizeof(RectSource) * ns);
    if (e == hipSuccess) e = hipMemcpyAsync(d_src.p, h, sizeof(RectSource) * ns, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    delete[] h;
    if (e != hipSuccess) return fail(QDFN_ERR_HIP, std::string("stiffness_rect upload: ") + hipGetErrorString(e));
    const int64_t rblocks = (nr + 255) / 256;
    const int64_t blocks = rblocks * ns;
    if (blocks > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "stiffness_rect: problem too large for one launch");
    hipLaunchKernelGGL(k_stiffness_rect, dim3((unsigned)blocks), dim3(256), 0, stream, d_src.as<RectSource>(), ns,
                       d_rcv, nr, ldr, mu, lam, d_KS, d_KN, ldk, rblocks);
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipStreamSynchronize(stream));  // d_src is released on return
    return QDFN_OK;
}

int make_rect_entries(EntrySource& es, const double* seg, int64_t n, int64_t ld, double mu, double nu,
                      hipStream_t stream) {
    if (n <= 0 || ld < n || !seg) return fail(QDFN_ERR_ARG, "rectangle entries: bad table");
    es.kind = 0;
    es.n = n;
    es.mu = mu;
    es.lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
    const double alpha = (es.lam + mu) / (es.lam + 2.0 * mu);
    RectSource* h = rect_sources(seg, n, ld, alpha);
    hipError_t e = es.srcs.alloc(sizeof(RectSource) * n);
    if (e == hipSuccess) e = es.tab.alloc(sizeof(double) * n * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(es.srcs.p, h, sizeof(RectSource) * n, hipMemcpyHostToDevice, stream);
    // receiver table with leading dimension n
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(es.tab.p, sizeof(double) * n, seg, sizeof(double) * ld, sizeof(double) * n, 8,
                             hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    delete[] h;
    if (e != hipSuccess) return fail(QDFN_ERR_HIP, std::string("rectangle entries upload: ") + hipGetErrorString(e));
    return QDFN_OK;
}

int launch_rect_lines(const EntrySource& es, const SampleLine* d_lines, const SampleTile* d_tiles, int64_t ntiles,
                      double* d_outS, double* d_outN, hipStream_t stream) {
    if (ntiles <= 0) return QDFN_OK;
    const int64_t blocks = (ntiles + 3) / 4;
    if (blocks > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "rectangle lines: too many tiles for one launch");
    hipLaunchKernelGGL(k_rect_lines, dim3((unsigned)blocks), dim3(256), 0, stream, es.srcs.as<RectSource>(),
                       es.tab.as<double>(), es.n, es.mu, es.lam, d_lines, d_tiles, ntiles, d_outS, d_outN);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

}  // namespace qdfn

using namespace qdfn;

extern "C" int qdfn_okada_dc3d(double alpha, int64_t n, const double* X, const double* Y, const double* Z,
                               double depth, double dip, double al1, double al2, double aw1, double aw2,
                               double disl1, double disl2, double disl3, double* U, double* iret) {
    if (n < 0 || (n > 0 && (!X || !Y || !Z || !U || !iret))) return fail(QDFN_ERR_ARG, "qdfn_okada_dc3d: bad args");
    if (n == 0) return QDFN_OK;
    for (int64_t i = 0; i < n; ++i)
        if (Z[i] > 0.0) { set_error("qdfn_okada_dc3d: positive Z given (reference prints a warning)"); break; }
    DevBuf dx, dy, dz, du, di;
    const size_t b = sizeof(double) * n;
    QDFN_HIP(dx.alloc(b)); QDFN_HIP(dy.alloc(b)); QDFN_HIP(dz.alloc(b));
    QDFN_HIP(du.alloc(b * 12)); QDFN_HIP(di.alloc(b));
    QDFN_HIP(hipMemcpy(dx.p, X, b, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dy.p, Y, b, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dz.p, Z, b, hipMemcpyHostToDevice));
    const DipConst k = make_dip_const(alpha, dip);
    const SlipW s = make_slip(disl1, disl2, disl3);
    hipLaunchKernelGGL(k_dc3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, k, n, dx.as<double>(),
                       dy.as<double>(), dz.as<double>(), depth, al1, al2, aw1, aw2, s, du.as<double>(),
                       di.as<double>());
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipMemcpy(U, du.p, b * 12, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(iret, di.p, b, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_okada_stiffness_rect(const double* src_seg, int64_t ns, int64_t ld_src, const double* rcv_seg,
                                         int64_t nr, int64_t ld_rcv, double mu, double nu, double* KS, double* KN,
                                         int64_t ldk) {
    if (ns < 0 || nr < 0 || ld_src < ns || ld_rcv < nr || ldk < nr || !src_seg || !rcv_seg || !KS || !KN)
        return fail(QDFN_ERR_ARG, "qdfn_okada_stiffness_rect: bad sizes or null pointers");
    if (ns == 0 || nr == 0) return QDFN_OK;
    DevBuf drcv, dks, dkn;
    QDFN_HIP(drcv.alloc(sizeof(double) * ld_rcv * 8));
    QDFN_HIP(hipMemcpy(drcv.p, rcv_seg, sizeof(double) * ld_rcv * 8, hipMemcpyHostToDevice));
    const size_t ob = sizeof(double) * (size_t)ldk * (size_t)ns;
    QDFN_HIP(dks.alloc(ob));
    QDFN_HIP(dkn.alloc(ob));
    int rc = assemble_rect_device(src_seg, ns, ld_src, drcv.as<double>(), nr, ld_rcv, mu, nu, dks.as<double>(),
                                  dkn.as<double>(), ldk, 0);
    if (rc != QDFN_OK) return rc;
    // copy back only the nr rows of each column (the caller may hold ldk > nr)
    QDFN_HIP(hipMemcpy2D(KS, sizeof(double) * ldk, dks.p, sizeof(double) * ldk, sizeof(double) * nr, ns,
                         hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy2D(KN, sizeof(double) * ldk, dkn.p, sizeof(double) * ldk, sizeof(double) * nr, ns,
                         hipMemcpyDeviceToHost));
    return QDFN_OK;
}
