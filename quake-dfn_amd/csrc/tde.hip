// The angular-dislocation and free-surface-correction strain expressions
// below follow the published solution of Nikkhoo & Walter (2015), as carried
// in the reference's scripts/Functions_TDstressHS.jl, whose notice reads:
//
//   Copyright (c) 2014 Mehdi Nikkhoo
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files
//   (the "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
//
// Triangular-dislocation kernels (scripts/Functions_TDstressHS.jl,
// scripts/Functions_BuildInputFile.jl:338-377,551-638):
//   k_tri_prepare      per-source geometry (TriSource), one lane per triangle
//   k_tde_points       TDstressHS at n points for one triangle
//   k_tde_stiffness    fused TDstressHS -> traction -> normal / slip projection,
//                      one lane per (receiver, source) pair
//   k_tde_lines        the same entries along sample lines (rows / columns of
//                      H-matrix blocks chosen by the cross approximation)
//   k_rotverts         RotVerts_UnitVectors, one lane per triangle
// Bound: FP64 VALU (six angular-dislocation strains and six free-surface
// corrections per pair); output 16 B per pair.
#include <vector>

#include "assembly.h"
#include "qdfn_internal.h"
#include "tde_fsc.h"

namespace qdfn {

__device__ __host__ inline void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __host__ inline double norm3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Strike/dip/normal basis of a triangle (TDstressFS :149-166 /
// TDstress_HarFunc :307-318).  image_flip: horizontal image triangles above
// the surface reverse their strike vector.
__device__ __forceinline__ void tri_basis(const double P1[3], const double P2[3], const double P3[3], bool image_rule,
                                 double Vn[3], double Vs[3], double Vd[3]) {
    const double u[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    const double w[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    cross3(u, w, Vn);
    const double nn = norm3(Vn);
    for (int k = 0; k < 3; ++k) Vn[k] /= nn;
    const double ez[3] = {0, 0, 1};
    cross3(ez, Vn, Vs);
    if (norm3(Vs) == 0) {
        Vs[0] = 0.0;
        Vs[1] = Vn[2];
        Vs[2] = 0.0;
        if (image_rule && P1[2] > 0) Vs[1] = -Vs[1];
    }
    const double ns = norm3(Vs);
    for (int k = 0; k < 3; ++k) Vs[k] /= ns;
    cross3(Vn, Vs, Vd);
}

__device__ __forceinline__ void make_frame(const double P1[3], const double P2[3], const double P3[3], TdFrame& f) {
    double Vn[3], Vs[3], Vd[3];
    tri_basis(P1, P2, P3, true, Vn, Vs, Vd);
    for (int k = 0; k < 3; ++k) {
        f.A[k] = Vn[k];
        f.A[3 + k] = Vs[k];
        f.A[6 + k] = Vd[k];
        f.P2[k] = P2[k];
    }
    double d1[3] = {P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2]};
    double d3[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
    double p1[3], p3[3];
    for (int r = 0; r < 3; ++r) {
        p1[r] = f.A[3 * r] * d1[0] + f.A[3 * r + 1] * d1[1] + f.A[3 * r + 2] * d1[2];
        p3[r] = f.A[3 * r] * d3[0] + f.A[3 * r + 1] * d3[1] + f.A[3 * r + 2] * d3[2];
    }
    double s12[3], s13[3], s23[3], e12[3], e13[3], e23[3];
    for (int k = 0; k < 3; ++k) {
        s12[k] = 0.0 - p1[k];
        s13[k] = p3[k] - p1[k];
        s23[k] = p3[k] - 0.0;
    }
    const double n12 = norm3(s12), n13 = norm3(s13), n23 = norm3(s23);
    for (int k = 0; k < 3; ++k) {
        e12[k] = s12[k] / n12;
        e13[k] = s13[k] / n13;
        e23[k] = s23[k] / n23;
    }
    const double ang[3] = {acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]),
                           acos(-e12[0] * e23[0] - e12[1] * e23[1] - e12[2] * e23[2]),
                           acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2])};
    for (int k = 0; k < 3; ++k) {
        const double al = -QDFN_PI + ang[k];
        f.sinA[k] = sin(al);
        f.cosA[k] = cos(al);
    }
    f.tv[0][0] = p1[1];
    f.tv[0][1] = p1[2];
    f.tv[1][0] = 0.0;
    f.tv[1][1] = 0.0;
    f.tv[2][0] = p3[1];
    f.tv[2][1] = p3[2];
    f.sv[0][0] = e13[1];
    f.sv[0][1] = e13[2];
    f.sv[1][0] = -e12[1];
    f.sv[1][1] = -e12[2];
    f.sv[2][0] = -e23[1];
    f.sv[2][1] = -e23[2];
    f.flip_xz = 0;
    f.pad = 0;
}

__device__ __forceinline__ void make_edge(const double PA[3], const double PB[3], const double b[3], FscEdge& e) {
    const double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    const double beta = acos(-sv[2] / norm3(sv));
    e.beta = beta;
    e.degenerate = (fabs(beta) < 1e-16 || fabs(QDFN_PI - beta) < 1e-16) ? 1 : 0;
    const double h = sqrt(sv[0] * sv[0] + sv[1] * sv[1]);
    e.ey1x = sv[0] / h;
    e.ey1y = sv[1] / h;
    for (int k = 0; k < 3; ++k) e.PA[k] = PA[k];
    e.PB3 = PB[2];
    // A = [ey1 ey2 ey3] with ey2 = ey3 x ey1 = (ey1y, -ey1x, 0): A is symmetric
    // (rows (ey1x, ey1y, 0), (ey1y, -ey1x, 0), (0, 0, -1)).
    auto toA = [&](const double v[3], double o[3]) {
        o[0] = e.ey1x * v[0] + e.ey1y * v[1];
        o[1] = e.ey1y * v[0] - e.ey1x * v[1];
        o[2] = -v[2];
    };
    toA(sv, e.yAB);
    toA(b, e.bb);
    e.sinB = sin(beta);
    e.cosB = cos(beta);
    e.cotB = 1.0 / tan(beta);
    const double b2 = QDFN_PI - beta;
    e.sinB2 = sin(b2);
    e.cosB2 = cos(b2);
    e.cotB2 = 1.0 / tan(b2);
}

// P: ns x 3 column-major vertices; slip from rake as StiffnessMatrix_ByParts_Calculation_Tri
// (Ss = -cosd(rake), Ds = sind(rake), Ts = 0) unless slip3 is given (Ss, Ds, Ts per source).
__global__ void k_tri_prepare(int64_t ns, const double* __restrict__ P1, const double* __restrict__ P2,
                              const double* __restrict__ P3, const double* __restrict__ rake,
                              const double* __restrict__ slip3, double nu, TriSource* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    double p1[3], p2[3], p3[3];
    for (int k = 0; k < 3; ++k) {
        p1[k] = P1[s + k * ns];
        p2[k] = P2[s + k * ns];
        p3[k] = P3[s + k * ns];
    }
    TriSource t;
    if (slip3) {
        t.by = slip3[3 * s];
        t.bz = slip3[3 * s + 1];
        t.bx = slip3[3 * s + 2];
    } else {
        t.by = -cos_deg(rake[s]);
        t.bz = sin_deg(rake[s]);
        t.bx = 0.0;
    }
    t.nu = nu;
    make_frame(p1, p2, p3, t.main);
    const double q1[3] = {p1[0], p1[1], -p1[2]}, q2[3] = {p2[0], p2[1], -p2[2]}, q3[3] = {p3[0], p3[1], -p3[2]};
    make_frame(q1, q2, q3, t.image);
    t.image.flip_xz = (q1[2] == 0 && q2[2] == 0 && q3[2] == 0) ? 1 : 0;
    // slip in EFCS for the harmonic correction (TDstress_HarFunc :321-322)
    double Vn[3], Vs[3], Vd[3];
    tri_basis(p1, p2, p3, false, Vn, Vs, Vd);
    const double b[3] = {t.bx * Vn[0] + t.by * Vs[0] + t.bz * Vd[0], t.bx * Vn[1] + t.by * Vs[1] + t.bz * Vd[1],
                         t.bx * Vn[2] + t.by * Vs[2] + t.bz * Vd[2]};
    make_edge(p1, p2, b, t.edge[0]);
    make_edge(p2, p3, b, t.edge[1]);
    make_edge(p3, p1, b, t.edge[2]);
    out[s] = t;
}

// TensTrans (:338-362) with the matrix given by its three columns c0, c1, c2.
__device__ __forceinline__ void tens_rot(const double T[6], const double c0[3], const double c1[3], const double c2[3],
                                double O[6]) {
    const double A[9] = {c0[0], c0[1], c0[2], c1[0], c1[1], c1[2], c2[0], c2[1], c2[2]};
    const double xx = T[0], yy = T[1], zz = T[2], xy = T[3], xz = T[4], yz = T[5];
    O[0] = A[0] * A[0] * xx + 2 * A[0] * A[3] * xy + 2 * A[0] * A[6] * xz + 2 * A[3] * A[6] * yz + A[3] * A[3] * yy + A[6] * A[6] * zz;
    O[1] = A[1] * A[1] * xx + 2 * A[1] * A[4] * xy + 2 * A[1] * A[7] * xz + 2 * A[4] * A[7] * yz + A[4] * A[4] * yy + A[7] * A[7] * zz;
    O[2] = A[2] * A[2] * xx + 2 * A[2] * A[5] * xy + 2 * A[2] * A[8] * xz + 2 * A[5] * A[8] * yz + A[5] * A[5] * yy + A[8] * A[8] * zz;
    O[3] = A[0] * A[1] * xx + (A[0] * A[4] + A[1] * A[3]) * xy + (A[0] * A[7] + A[1] * A[6]) * xz +
           (A[7] * A[3] + A[6] * A[4]) * yz + A[4] * A[3] * yy + A[6] * A[7] * zz;
    O[4] = A[0] * A[2] * xx + (A[0] * A[5] + A[2] * A[3]) * xy + (A[0] * A[8] + A[2] * A[6]) * xz +
           (A[8] * A[3] + A[6] * A[5]) * yz + A[5] * A[3] * yy + A[6] * A[8] * zz;
    O[5] = A[1] * A[2] * xx + (A[2] * A[4] + A[1] * A[5]) * xy + (A[2] * A[7] + A[1] * A[8]) * xz +
           (A[7] * A[5] + A[8] * A[4]) * yz + A[4] * A[5] * yy + A[7] * A[8] * zz;
}

// one angular dislocation of TDSetupS (:433-464), strain accumulated into acc (TDCS);
// (s1, s2) is the side direction (y, z), tv the vertex (y, z).
__device__ __forceinline__ void setup_add(double x, double y, double z, double sinA, double cosA, double bx,
                                          double by, double bz, double nu, double tv0, double tv1, double s1,
                                          double s2, double acc[6]) {
    // A = [s2 -s1; s1 s2]
    const double y1 = s2 * (y - tv0) - s1 * (z - tv1);
    const double z1 = s1 * (y - tv0) + s2 * (z - tv1);
    const double by1 = s2 * by - s1 * bz;
    const double bz1 = s1 * by + s2 * bz;
    double e[6], o[6];
    ang_strain(x, y1, z1, sinA, cosA, bx, by1, bz1, nu, e);
    // B = [1 0 0; 0 s2 s1; 0 -s1 s2] by columns
    const double c0[3] = {1.0, 0.0, 0.0}, c1[3] = {0.0, s2, -s1}, c2[3] = {0.0, s1, s2};
    tens_rot(e, c0, c1, c2, o);
    for (int k = 0; k < 6; ++k) acc[k] += o[k];
}

// TDstressFS (:133-286) strain in EFCS at one point for one frame.
__device__ __forceinline__ void frame_strain(double X, double Y, double Z, const TdFrame& f, double bx, double by,
                                             double bz, double nu, double E[6]) {
    const double d[3] = {X - f.P2[0], Y - f.P2[1], Z - f.P2[2]};
    const double x = f.A[0] * d[0] + f.A[1] * d[1] + f.A[2] * d[2];
    const double y = f.A[3] * d[0] + f.A[4] * d[1] + f.A[5] * d[2];
    const double z = f.A[6] * d[0] + f.A[7] * d[1] + f.A[8] * d[2];
    // trimodefinder(y, z, x, p1[2:3], p2[2:3], p3[2:3]) with p2 = 0 (:393-430)
    const double a1 = f.tv[0][0], a2 = f.tv[0][1], c1 = f.tv[2][0], c2 = f.tv[2][1];
    const double den = (0.0 - c2) * (a1 - c1) + (c1 - 0.0) * (a2 - c2);
    const double la = ((0.0 - c2) * (y - c1) + (c1 - 0.0) * (z - c2)) / den;
    const double lb = ((c2 - a2) * (y - c1) + (a1 - c1) * (z - c2)) / den;
    const double lc = 1 - la - lb;
    int m = 1;
    if (la <= 0 && lb > lc && lc > la) m = -1;
    if (lb <= 0 && lc > la && la > lb) m = -1;
    if (lc <= 0 && la > lb && lb > lc) m = -1;
    if (la == 0 && lb >= 0 && lc >= 0) m = 0;
    if (la >= 0 && lb == 0 && lc >= 0) m = 0;
    if (la >= 0 && lb >= 0 && lc == 0) m = 0;
    if (m == 0 && x != 0) m = 1;
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (m == 0) {
        for (int k = 0; k < 6; ++k) e[k] = NAN;
    } else {
        // configuration I (m = 1) uses -e13, e12, e23; II uses e13, -e12, -e23
        const double sg = (m == 1) ? -1.0 : 1.0;
#pragma unroll 1
        for (int k = 0; k < 3; ++k)
            setup_add(x, y, z, f.sinA[k], f.cosA[k], bx, by, bz, nu, f.tv[k][0], f.tv[k][1], sg * f.sv[k][0],
                      sg * f.sv[k][1], e);
    }
    const double Vn[3] = {f.A[0], f.A[1], f.A[2]}, Vs[3] = {f.A[3], f.A[4], f.A[5]}, Vd[3] = {f.A[6], f.A[7], f.A[8]};
    tens_rot(e, Vn, Vs, Vd, E);
}

// AngSetupFSC_S (:467-574) strain at one point for one edge: v = v(PB) - v(PA),
// both evaluated in the configuration chosen by the sign of beta * y1 at PA.
__device__ __forceinline__ void edge_fsc_strain(double X, double Y, double Z, const FscEdge& g, double nu,
                                                double E[6]) {
    if (g.degenerate) {
        for (int k = 0; k < 6; ++k) E[k] = 0.0;
        return;
    }
    const double dx = X - g.PA[0], dy = Y - g.PA[1], dz = Z - g.PA[2];
    const double yA[3] = {g.ey1x * dx + g.ey1y * dy, g.ey1y * dx - g.ey1x * dy, -dz};
    const bool cfg1 = g.beta * yA[0] >= 0;
    // configuration I: mirrored angle pi - beta, (y1, y2, b1, b2) negated, xz and yz negated
    const double sgn = cfg1 ? -1.0 : 1.0;
    const double sinB = cfg1 ? g.sinB2 : g.sinB, cosB = cfg1 ? g.cosB2 : g.cosB, cotB = cfg1 ? g.cotB2 : g.cotB;
    double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
        const double y0 = p ? yA[0] - g.yAB[0] : yA[0];
        const double y1 = p ? yA[1] - g.yAB[1] : yA[1];
        const double y2 = p ? yA[2] - g.yAB[2] : yA[2];
        double w[6];
        ang_strain_fsc(sgn * y0, sgn * y1, y2, sinB, cosB, cotB, sgn * g.bb[0], sgn * g.bb[1], g.bb[2], nu,
                       p ? -g.PB3 : -g.PA[2], w);
        if (cfg1) {
            w[4] = -w[4];
            w[5] = -w[5];
        }
        for (int k = 0; k < 6; ++k) v[k] = p ? v[k] + w[k] : v[k] - w[k];
    }
    // TensTrans with A' (= A, symmetric): columns (ey1x, ey1y, 0), (ey1y, -ey1x, 0), (0, 0, -1)
    const double c0[3] = {g.ey1x, g.ey1y, 0.0}, c1[3] = {g.ey1y, -g.ey1x, 0.0}, c2[3] = {0.0, 0.0, -1.0};
    tens_rot(v, c0, c1, c2, E);
}

// Hooke's law (TDstressHS :122-129 and AngSetupFSC_S :566-572), accumulated.
__device__ __forceinline__ void add_hooke(const double e[6], double mu, double lam, double S[6], double E[6]) {
    const double tr = e[0] + e[1] + e[2];
    S[0] += 2 * mu * e[0] + lam * tr;
    S[1] += 2 * mu * e[1] + lam * tr;
    S[2] += 2 * mu * e[2] + lam * tr;
    S[3] += 2 * mu * e[3];
    S[4] += 2 * mu * e[4];
    S[5] += 2 * mu * e[5];
    for (int k = 0; k < 6; ++k) E[k] += e[k];
}

// TDstressHS (:2-130): strain and stress at one point (EFCS, xx yy zz xy xz yz)
// = main + image (full space, per frame) + the harmonic free-surface correction
// (the sum of the three edges' stresses and strains).
__device__ __forceinline__ void tde_point(double X, double Y, double Z, const TriSource& t, double mu, double lam,
                                          double S[6], double E[6]) {
    for (int k = 0; k < 6; ++k) {
        S[k] = 0.0;
        E[k] = 0.0;
    }
#pragma unroll 1
    for (int fi = 0; fi < 2; ++fi) {
        const TdFrame& f = fi ? t.image : t.main;
        double e[6];
        frame_strain(X, Y, Z, f, t.bx, t.by, t.bz, t.nu, e);
        if (f.flip_xz) {
            e[4] = -e[4];
            e[5] = -e[5];
        }
        add_hooke(e, mu, lam, S, E);
    }
    double Sf[6] = {0, 0, 0, 0, 0, 0}, Ef[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int j = 0; j < 3; ++j) {
        double Ek[6];
        edge_fsc_strain(X, Y, Z, t.edge[j], t.nu, Ek);
        add_hooke(Ek, mu, lam, Sf, Ef);
    }
    for (int k = 0; k < 6; ++k) {
        S[k] += Sf[k];
        E[k] += Ef[k];
    }
}

__global__ __launch_bounds__(256) void k_tde_points(int64_t n, const double* __restrict__ X,
                                                    const double* __restrict__ Y, const double* __restrict__ Z,
                                                    const TriSource* __restrict__ src, double mu, double lam,
                                                    double* __restrict__ S, double* __restrict__ E) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[6], e[6];
    tde_point(X[i], Y[i], Z[i], src[0], mu, lam, s, e);
    for (int k = 0; k < 6; ++k) {
        S[i + k * n] = s[k];
        E[i + k * n] = e[k];
    }
}

// StiffnessMatrix_ByParts_Calculation_Tri (:551-594): t = sigma n_r,
// K_N = t . n_r (tension positive), K_S = t . slip_r.  One entry: receiver r
// (ctr / nrm / slip n x 3, leading dimension nr) and source triangle src.
__device__ __forceinline__ void tri_entry(const TriSource& src, const double* __restrict__ ctr,
                                          const double* __restrict__ nrm, const double* __restrict__ slip,
                                          int64_t r, int64_t nr, double mu, double lam, double& ks, double& kn) {
    double S[6], E[6];
    tde_point(ctr[r], ctr[r + nr], ctr[r + 2 * nr], src, mu, lam, S, E);
    const double n0 = nrm[r], n1 = nrm[r + nr], n2 = nrm[r + 2 * nr];
    const double t0 = S[0] * n0 + S[3] * n1 + S[4] * n2;
    const double t1 = S[3] * n0 + S[1] * n1 + S[5] * n2;
    const double t2 = S[4] * n0 + S[5] * n1 + S[2] * n2;
    kn = t0 * n0 + t1 * n1 + t2 * n2;
    ks = t0 * slip[r] + t1 * slip[r + nr] + t2 * slip[r + 2 * nr];
}

__global__ __launch_bounds__(256) void k_tde_stiffness(const TriSource* __restrict__ src, int64_t ns,
                                                       const double* __restrict__ ctr, const double* __restrict__ nrm,
                                                       const double* __restrict__ slip, int64_t nr, double mu,
                                                       double lam, double* __restrict__ KS, double* __restrict__ KN,
                                                       int64_t ldk, int64_t rblocks) {
    const int64_t bid = blockIdx.x;
    const int64_t s = bid / rblocks;
    const int64_t r = (bid - s * rblocks) * 256 + threadIdx.x;
    if (s >= ns || r >= nr) return;
    double ks, kn;
    tri_entry(src[s], ctr, nrm, slip, r, nr, mu, lam, ks, kn);
    KN[r + s * ldk] = kn;
    KS[r + s * ldk] = ks;
}

// Entries along sample lines (sampled H-matrix construction): wavefront w of
// workgroup b takes tile 4 b + w, 64 consecutive entries of one line.
__global__ __launch_bounds__(256) void k_tde_lines(const TriSource* __restrict__ src, const double* __restrict__ ctr,
                                                   const double* __restrict__ nrm, const double* __restrict__ slip,
                                                   int64_t n, double mu, double lam, const SampleLine* lines,
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
    tri_entry(src[s], ctr, nrm, slip, r, n, mu, lam, ks, kn);
    write_entry(L, e, ks, kn, outS, outN);
}

// RotVerts_UnitVectors (:338-377), P1..P3 and outputs n x 3 column-major.
__global__ void k_rotverts(int64_t n, double* __restrict__ P1, double* __restrict__ P2, double* __restrict__ P3,
                           const double* __restrict__ rake, double* __restrict__ nrm, double* __restrict__ strike,
                           double* __restrict__ dip, double* __restrict__ slip, int* __restrict__ reversed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p1[3], p2[3], p3[3], u[3], w[3], c[3];
    for (int k = 0; k < 3; ++k) {
        p1[k] = P1[i + k * n];
        p2[k] = P2[i + k * n];
        p3[k] = P3[i + k * n];
    }
    for (int k = 0; k < 3; ++k) {
        u[k] = p2[k] - p1[k];
        w[k] = p3[k] - p1[k];
    }
    cross3(u, w, c);
    double cn = norm3(c);
    if (c[1] / cn <= 0) {
        for (int k = 0; k < 3; ++k) {
            const double tmp = p1[k];
            p1[k] = p2[k];
            p2[k] = tmp;
        }
        atomicAdd(reversed, 1);
    }
    for (int k = 0; k < 3; ++k) {
        P1[i + k * n] = p1[k];
        P2[i + k * n] = p2[k];
        u[k] = p2[k] - p1[k];
        w[k] = p3[k] - p1[k];
    }
    cross3(u, w, c);
    cn = norm3(c);
    const double N[3] = {c[0] / cn, c[1] / cn, c[2] / cn};
    const double ez[3] = {0, 0, 1};
    double s[3], d[3];
    cross3(N, ez, s);
    const double sn = norm3(s);
    for (int k = 0; k < 3; ++k) s[k] /= sn;
    cross3(s, N, d);
    const double dn = norm3(d);
    for (int k = 0; k < 3; ++k) d[k] /= dn;
    const double cr = cos_deg(rake[i]), sr = sin_deg(rake[i]);
    for (int k = 0; k < 3; ++k) {
        nrm[i + k * n] = N[k];
        strike[i + k * n] = s[k];
        dip[i + k * n] = d[k];
        slip[i + k * n] = s[k] * cr + d[k] * sr;
    }
}

int prepare_sources(int64_t ns, const double* d_P1, const double* d_P2, const double* d_P3, const double* d_rake,
                    const double* d_slip3, double nu, DevBuf& out, hipStream_t stream) {
    QDFN_HIP(out.alloc(sizeof(TriSource) * (ns > 0 ? ns : 1)));
    if (ns == 0) return QDFN_OK;
    hipLaunchKernelGGL(k_tri_prepare, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, stream, ns, d_P1, d_P2, d_P3,
                       d_rake, d_slip3, nu, out.as<TriSource>());
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

int assemble_tri_device(int64_t ns, const double* d_P1, const double* d_P2, const double* d_P3, const double* d_rake,
                        int64_t nr, const double* d_ctr, const double* d_nrm, const double* d_slip, double mu,
                        double lambda, double* d_KS, double* d_KN, int64_t ldk, hipStream_t stream) {
    if (ns <= 0 || nr <= 0) return QDFN_OK;
    const double nu = 1 / (1 + lambda / mu) / 2;
    DevBuf src;
    int rc = prepare_sources(ns, d_P1, d_P2, d_P3, d_rake, nullptr, nu, src, stream);
    if (rc) return rc;
    const int64_t rblocks = (nr + 255) / 256;
    const int64_t blocks = rblocks * ns;
    if (blocks > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "triangle assembly: problem too large for one launch");
    hipLaunchKernelGGL(k_tde_stiffness, dim3((unsigned)blocks), dim3(256), 0, stream, src.as<TriSource>(), ns, d_ctr,
                       d_nrm, d_slip, nr, mu, lambda, d_KS, d_KN, ldk, rblocks);
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipStreamSynchronize(stream));
    return QDFN_OK;
}

int make_tri_entries(EntrySource& es, int64_t n, const double* P1, const double* P2, const double* P3,
                     const double* rake, const double* ctr, const double* nrm, const double* slip, double mu,
                     double lambda, hipStream_t stream) {
    if (n <= 0 || !P1 || !P2 || !P3 || !rake || !ctr || !nrm || !slip)
        return fail(QDFN_ERR_ARG, "triangle entries: bad inputs");
    es.kind = 1;
    es.n = n;
    es.mu = mu;
    es.lam = lambda;
    const double nu = 1 / (1 + lambda / mu) / 2;
    DevBuf d1, d2, d3, dr;
    const size_t v3 = sizeof(double) * 3 * n;
    QDFN_HIP(d1.alloc(v3));
    QDFN_HIP(d2.alloc(v3));
    QDFN_HIP(d3.alloc(v3));
    QDFN_HIP(dr.alloc(sizeof(double) * n));
    QDFN_HIP(es.tab.alloc(3 * v3));
    QDFN_HIP(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(dr.p, rake, sizeof(double) * n, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(es.tab.as<double>(), ctr, v3, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(es.tab.as<double>() + 3 * n, nrm, v3, hipMemcpyHostToDevice, stream));
    QDFN_HIP(hipMemcpyAsync(es.tab.as<double>() + 6 * n, slip, v3, hipMemcpyHostToDevice, stream));
    int rc = prepare_sources(n, d1.as<double>(), d2.as<double>(), d3.as<double>(), dr.as<double>(), nullptr, nu,
                             es.srcs, stream);
    if (rc) return rc;
    QDFN_HIP(hipStreamSynchronize(stream));  // d1..dr are released on return
    return QDFN_OK;
}

int launch_tri_lines(const EntrySource& es, const SampleLine* d_lines, const SampleTile* d_tiles, int64_t ntiles,
                     double* d_outS, double* d_outN, hipStream_t stream) {
    if (ntiles <= 0) return QDFN_OK;
    const int64_t blocks = (ntiles + 3) / 4;
    if (blocks > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "triangle lines: too many tiles for one launch");
    const double* tab = es.tab.as<double>();
    hipLaunchKernelGGL(k_tde_lines, dim3((unsigned)blocks), dim3(256), 0, stream, es.srcs.as<TriSource>(), tab,
                       tab + 3 * es.n, tab + 6 * es.n, es.n, es.mu, es.lam, d_lines, d_tiles, ntiles, d_outS,
                       d_outN);
    QDFN_HIP(hipGetLastError());
    return QDFN_OK;
}

}  // namespace qdfn

using namespace qdfn;

extern "C" int qdfn_tde_stress_hs(int64_t n, const double* X, const double* Y, const double* Z, const double* P1,
                                  const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                                  double lambda, double* Stress, double* Strain) {
    if (n < 0 || !P1 || !P2 || !P3 || (n > 0 && (!X || !Y || !Z || !Stress || !Strain)))
        return fail(QDFN_ERR_ARG, "qdfn_tde_stress_hs: bad args");
    if (P1[2] > 0 || P2[2] > 0 || P3[2] > 0) return fail(QDFN_ERR_ARG, "Half-space solution: Z coordinates must be negative!");
    for (int64_t i = 0; i < n; ++i)
        if (Z[i] > 0) return fail(QDFN_ERR_ARG, "Half-space solution: Z coordinates must be negative!");
    if (n == 0) return QDFN_OK;
    const size_t b = sizeof(double) * n;
    DevBuf dx, dy, dz, ds, de, dp, dsl, src;
    QDFN_HIP(dx.alloc(b)); QDFN_HIP(dy.alloc(b)); QDFN_HIP(dz.alloc(b));
    QDFN_HIP(ds.alloc(6 * b)); QDFN_HIP(de.alloc(6 * b)); QDFN_HIP(dp.alloc(9 * sizeof(double)));
    QDFN_HIP(dsl.alloc(3 * sizeof(double)));
    QDFN_HIP(hipMemcpy(dx.p, X, b, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dy.p, Y, b, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dz.p, Z, b, hipMemcpyHostToDevice));
    const double hp[9] = {P1[0], P1[1], P1[2], P2[0], P2[1], P2[2], P3[0], P3[1], P3[2]};  // 1 x 3 each
    const double hs[3] = {Ss, Ds, Ts};
    QDFN_HIP(hipMemcpy(dp.p, hp, sizeof(hp), hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dsl.p, hs, sizeof(hs), hipMemcpyHostToDevice));
    const double nu = 1 / (1 + lambda / mu) / 2;
    int rc = prepare_sources(1, dp.as<double>(), dp.as<double>() + 3, dp.as<double>() + 6, nullptr, dsl.as<double>(),
                             nu, src, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tde_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, dx.as<double>(),
                       dy.as<double>(), dz.as<double>(), src.as<TriSource>(), mu, lambda, ds.as<double>(),
                       de.as<double>());
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipMemcpy(Stress, ds.p, 6 * b, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(Strain, de.p, 6 * b, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_tde_stiffness_tri(int64_t ns, const double* P1, const double* P2, const double* P3,
                                      const double* rake, int64_t nr, const double* ctr, const double* nrm,
                                      const double* slip, double mu, double lambda, double* KS, double* KN,
                                      int64_t ldk) {
    if (ns < 0 || nr < 0 || ldk < nr || !P1 || !P2 || !P3 || !rake || !ctr || !nrm || !slip || !KS || !KN)
        return fail(QDFN_ERR_ARG, "qdfn_tde_stiffness_tri: bad sizes or null pointers");
    if (ns == 0 || nr == 0) return QDFN_OK;
    for (int64_t s = 0; s < ns; ++s)
        if (P1[s + 2 * ns] > 0 || P2[s + 2 * ns] > 0 || P3[s + 2 * ns] > 0)
            return fail(QDFN_ERR_ARG, "Half-space solution: Z coordinates must be negative!");
    const size_t v3s = sizeof(double) * 3 * ns, v3r = sizeof(double) * 3 * nr;
    DevBuf d1, d2, d3, dr, dc, dn, dsl, dks, dkn;
    QDFN_HIP(d1.alloc(v3s)); QDFN_HIP(d2.alloc(v3s)); QDFN_HIP(d3.alloc(v3s)); QDFN_HIP(dr.alloc(sizeof(double) * ns));
    QDFN_HIP(dc.alloc(v3r)); QDFN_HIP(dn.alloc(v3r)); QDFN_HIP(dsl.alloc(v3r));
    const size_t ob = sizeof(double) * (size_t)ldk * (size_t)ns;
    QDFN_HIP(dks.alloc(ob)); QDFN_HIP(dkn.alloc(ob));
    QDFN_HIP(hipMemcpy(d1.p, P1, v3s, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d2.p, P2, v3s, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d3.p, P3, v3s, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dr.p, rake, sizeof(double) * ns, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dc.p, ctr, v3r, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dn.p, nrm, v3r, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dsl.p, slip, v3r, hipMemcpyHostToDevice));
    int rc = assemble_tri_device(ns, d1.as<double>(), d2.as<double>(), d3.as<double>(), dr.as<double>(), nr,
                                 dc.as<double>(), dn.as<double>(), dsl.as<double>(), mu, lambda, dks.as<double>(),
                                 dkn.as<double>(), ldk, 0);
    if (rc) return rc;
    QDFN_HIP(hipMemcpy2D(KS, sizeof(double) * ldk, dks.p, sizeof(double) * ldk, sizeof(double) * nr, ns,
                         hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy2D(KN, sizeof(double) * ldk, dkn.p, sizeof(double) * ldk, sizeof(double) * nr, ns,
                         hipMemcpyDeviceToHost));
    return QDFN_OK;
}

extern "C" int qdfn_rotverts_unitvectors(int64_t n, double* P1, double* P2, double* P3, const double* rake,
                                         double* nrm, double* strike, double* dip, double* slip) {
    if (n < 0 || (n > 0 && (!P1 || !P2 || !P3 || !rake || !nrm || !strike || !dip || !slip)))
        return fail(QDFN_ERR_ARG, "qdfn_rotverts_unitvectors: bad args");
    if (n == 0) return 0;
    const size_t v3 = sizeof(double) * 3 * n;
    DevBuf buf, cnt;
    QDFN_HIP(buf.alloc(v3 * 7 + sizeof(double) * n));
    QDFN_HIP(cnt.alloc(sizeof(int)));
    double* d = buf.as<double>();
    double *d1 = d, *d2 = d + 3 * n, *d3 = d + 6 * n, *dn = d + 9 * n, *ds = d + 12 * n, *dd = d + 15 * n,
           *dsl = d + 18 * n, *dr = d + 21 * n;
    QDFN_HIP(hipMemcpy(d1, P1, v3, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d2, P2, v3, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(d3, P3, v3, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dr, rake, sizeof(double) * n, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemset(cnt.p, 0, sizeof(int)));
    hipLaunchKernelGGL(k_rotverts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, d1, d2, d3, dr, dn, ds, dd,
                       dsl, cnt.as<int>());
    QDFN_HIP(hipGetLastError());
    double* outs[7] = {P1, P2, P3, nrm, strike, dip, slip};
    for (int k = 0; k < 7; ++k) QDFN_HIP(hipMemcpy(outs[k], d + 3 * n * k, v3, hipMemcpyDeviceToHost));
    int rev = 0;
    QDFN_HIP(hipMemcpy(&rev, cnt.p, sizeof(int), hipMemcpyDeviceToHost));
    return rev;
}
