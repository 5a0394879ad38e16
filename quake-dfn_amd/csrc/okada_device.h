// Okada (1992) DC3D displacement / displacement-gradient field of a finite
// rectangular dislocation in an elastic half-space, as a per-point device
// routine.  Semantics follow the reference's Julia DC3D
// (scripts/Functions_OKADA3D.jl:12-654), including its singular-point guards
// and its NaN propagation:
//   * the 1e-6 zeroing of XI/ET/Q applies only to the corner-geometry terms
//     (DCCON2_V :57-67); the A/B/C parts use the caller's XI/ET/Q;
//   * blended branches (UB's CD==0 substitution :283-307, zero-slip terms
//     :230/:339/:425) are evaluated as products with 0/1 flags so that NaN/Inf
//     in an unselected branch still propagates, as in the reference.
// Everything is FP64: near-field cancellation in these formulas needs it.
#pragma once
#include <hip/hip_runtime.h>

namespace qdfn {

#define QDFN_TWO_PI 6.283185307179586

struct DipConst {  // DCCON0_V, Functions_OKADA3D.jl:12-48
    double a1, a2, a3, a4, a5, sd, cd, ss, cc, sc;
};

__host__ __device__ inline DipConst make_dip_const(double alpha, double dip_deg) {
    DipConst k;
    k.a1 = (1.0 - alpha) / 2.0;
    k.a2 = alpha / 2.0;
    k.a3 = (1.0 - alpha) / alpha;
    k.a4 = 1.0 - alpha;
    k.a5 = alpha;
    const double r = dip_deg * (QDFN_TWO_PI / 360.0);
    double sd = sin(r), cd = cos(r);
    const double small = fabs(cd) < 1.0e-6 ? 1.0 : 0.0;
    const double big = fabs(cd) >= 1.0e-6 ? 1.0 : 0.0;
    const double sgn = (sd > 0.0 ? 1.0 : 0.0) + sd * (sd == 0.0 ? 1.0 : 0.0) - (sd < 0.0 ? 1.0 : 0.0);
    cd = 0.0 * small + cd * big;
    sd = small * sgn + big * sd;
    k.sd = sd;
    k.cd = cd;
    k.ss = sd * sd;
    k.cc = cd * cd;
    k.sc = sd * cd;
    return k;
}

// Corner geometry (DCCON2_V, :51-195) from the zeroed copies of XI, ET, Q.
struct Corner {
    double xi2, et2, q2, r2, r, r3, r5, y, d, tt, alx, x11, x32, ale, y11, y32;
    double ey, ez, fy, fz, gy, gz, hy, hz;
};

__device__ __forceinline__ double zero_small(double v) {
    return 0.0 * (fabs(v) < 1.0e-6 ? 1.0 : 0.0) + v * (fabs(v) >= 1.0e-6 ? 1.0 : 0.0);
}

__device__ __forceinline__ void corner_terms(double xi_in, double et_in, double q_in, double sd, double cd,
                                             Corner& g) {
    const double xi = zero_small(xi_in), et = zero_small(et_in), q = zero_small(q_in);
    g.xi2 = xi * xi;
    g.et2 = et * et;
    g.q2 = q * q;
    g.r2 = g.xi2 + g.et2 + g.q2;
    g.r = sqrt(g.r2);
    g.r3 = g.r * g.r2;
    g.r5 = g.r3 * g.r2;
    g.y = et * cd + q * sd;
    g.d = et * sd - q * cd;
    g.tt = (q != 0.0) ? atan(xi * et / (q * g.r)) : 0.0;
    // log singularity at the negative XI (resp. ET) edge extension (:107-182)
    if (xi < 0.0 && q == 0.0 && et == 0.0) {
        g.alx = -log(g.r - xi);
        g.x11 = 0.0;
        g.x32 = 0.0;
    } else {
        const double rxi = g.r + xi;
        g.alx = log(rxi);
        g.x11 = 1.0 / (g.r * rxi);
        g.x32 = (g.r + rxi) * g.x11 * g.x11 / g.r;
    }
    if (et < 0.0 && q == 0.0 && xi == 0.0) {
        g.ale = -log(g.r - et);
        g.y11 = 0.0;
        g.y32 = 0.0;
    } else {
        const double ret = g.r + et;
        g.ale = log(ret);
        g.y11 = 1.0 / (g.r * ret);
        g.y32 = (g.r + ret) * g.y11 * g.y11 / g.r;
    }
    g.ey = sd / g.r - g.y * q / g.r3;
    g.ez = cd / g.r + g.d * q / g.r3;
    g.fy = g.d / g.r3 + g.xi2 * g.y32 * sd;
    g.fz = g.y / g.r3 + g.xi2 * g.y32 * cd;
    g.gy = 2.0 * g.x11 * sd - g.y * q * g.x32;
    g.gz = 2.0 * g.x11 * cd + g.d * q * g.x32;
    g.hy = g.d * q * g.x32 + xi * q * g.y32 * sd;
    g.hz = g.y * q * g.x32 + xi * q * g.y32 * cd;
}

// Slip weights: w[k] = DISL_k / 2pi and flag f[k] = (DISL_k != 0); a part's
// column c accumulates  sum_k (w_k * du_k[c]) * f_k  in k order.
struct SlipW {
    double w[3], f[3];
};

#define QDFN_ADD3(dst, e1, e2, e3) \
    dst = ((dst + s.w[0] * (e1) * s.f[0]) + s.w[1] * (e2) * s.f[1]) + s.w[2] * (e3) * s.f[2]

// Part A (UA_V, :199-262).  u[0..2] displacements, u[3..11] gradients.
template <bool DISP>
__device__ __forceinline__ void part_a(double XI, double ET, double Q, const DipConst& k, const Corner& g,
                                       const SlipW& s, double u[12]) {
    const double A1 = k.a1, A2 = k.a2, SD = k.sd, CD = k.cd;
    const double R = g.r, R3 = g.r3, Y = g.y, D = g.d, X11 = g.x11, Y11 = g.y11, Y32 = g.y32;
    const double XY = XI * Y11, QX = Q * X11, QY = Q * Y11;
    for (int c = 0; c < 12; ++c) u[c] = 0.0;
    if (DISP) {
        QDFN_ADD3(u[0], g.tt / 2.0 + A2 * XI * QY, A2 * Q / R, -A1 * g.ale - A2 * Q * QY);
        QDFN_ADD3(u[1], A2 * Q / R, g.tt / 2.0 + A2 * ET * QX, -A1 * g.alx - A2 * Q * QX);
        QDFN_ADD3(u[2], A1 * g.ale - A2 * Q * QY, A1 * g.alx - A2 * Q * QX, g.tt / 2.0 - A2 * (ET * QX + XI * QY));
    }
    QDFN_ADD3(u[3], -A1 * QY - A2 * g.xi2 * Q * Y32, -A2 * XI * Q / R3, -A1 * XY + A2 * XI * g.q2 * Y32);
    QDFN_ADD3(u[4], -A2 * XI * Q / R3, -QY / 2.0 - A2 * ET * Q / R3, -A1 / R + A2 * g.q2 / R3);
    QDFN_ADD3(u[5], A1 * XY + A2 * XI * g.q2 * Y32, A1 / R + A2 * g.q2 / R3, -A1 * QY - A2 * Q * g.q2 * Y32);
    QDFN_ADD3(u[6], A1 * XY * SD + A2 * XI * g.fy + D / 2.0 * X11, A2 * g.ey,
              -A1 * (CD / R + QY * SD) - A2 * Q * g.fy);
    QDFN_ADD3(u[7], A2 * g.ey, A1 * D * X11 + XY / 2.0 * SD + A2 * ET * g.gy, -A1 * Y * X11 - A2 * Q * g.gy);
    QDFN_ADD3(u[8], A1 * (CD / R + QY * SD) - A2 * Q * g.fy, A1 * Y * X11 - A2 * Q * g.gy,
              A1 * (D * X11 + XY * SD) + A2 * Q * g.hy);
    QDFN_ADD3(u[9], A1 * XY * CD + A2 * XI * g.fz + Y / 2.0 * X11, A2 * g.ez,
              A1 * (SD / R - QY * CD) - A2 * Q * g.fz);
    QDFN_ADD3(u[10], A2 * g.ez, A1 * Y * X11 + XY / 2.0 * CD + A2 * ET * g.gz, A1 * D * X11 - A2 * Q * g.gz);
    QDFN_ADD3(u[11], -A1 * (SD / R - QY * CD) - A2 * Q * g.fz, -A1 * D * X11 - A2 * Q * g.gz,
              A1 * (Y * X11 + XY * CD) + A2 * Q * g.hz);
}

// Part B (UB_V, :266-375).
template <bool DISP>
__device__ __forceinline__ void part_b(double XI, double ET, double Q, const DipConst& k, const Corner& g,
                                       const SlipW& s, double u[12]) {
    const double A3 = k.a3, SD = k.sd, CD = k.cd, SS = k.ss, SC = k.sc;
    const double R = g.r, R3 = g.r3, Y = g.y, D = g.d, X11 = g.x11, Y11 = g.y11, Y32 = g.y32;
    const double RD = R + D;
    const double D11 = 1.0 / (R * RD);
    const double AJ2 = XI * Y / RD * D11;
    const double AJ5 = -(D + Y * Y / RD) * D11;
    const double f1 = (CD != 0.0) ? 1.0 : 0.0, f2 = (CD == 0.0) ? 1.0 : 0.0;
    const double g1 = (XI == 0.0) ? 1.0 : 0.0, g2 = (XI != 0.0) ? 1.0 : 0.0;
    const double cdx = f1 * CD + f2 * 1.0e-12;
    const double ccx = f1 * k.cc + f2 * 1.0e-12;
    const double X = sqrt(g.xi2 + g.q2);
    double at = atan((ET * (X + Q * cdx) + X * (R + X) * SD) / (XI * (R + X) * cdx));
    if (isnan(at)) at = 0.0;
    const double RD2 = RD * RD;
    const double AI4 = f1 * (g1 * 0.0 + g2 * (1.0 / ccx * (XI / RD * SC + 2.0 * at))) + f2 * (XI * Y / RD2 / 2.0);
    const double AI3 = f1 * ((Y * cdx / RD - g.ale + SD * log(RD)) / ccx) + f2 * ((ET / RD + Y * Q / RD2 - g.ale) / 2.0);
    const double AK1 = f1 * (XI * (D11 - Y11 * SD) / cdx) + f2 * (XI * Q / RD * D11);
    const double AK3 = f1 * ((Q * Y11 - Y * D11) / cdx) + f2 * (SD / RD * (g.xi2 * D11 - 1.0));
    const double AJ3 = f1 * ((AK1 - AJ2 * SD) / cdx) + f2 * (-XI / RD2 * (g.q2 * D11 - 0.5));
    const double AJ6 = f1 * ((AK3 - AJ5 * SD) / cdx) + f2 * (-Y / RD2 * (g.xi2 * D11 - 0.5));
    const double XY = XI * Y11;
    const double AI1 = -XI / RD * CD - AI4 * SD;
    const double AI2 = log(RD) + AI3 * SD;
    const double AK2 = 1.0 / R + AK3 * SD;
    const double AK4 = XY * CD - AK1 * SD;
    const double AJ1 = AJ5 * CD - AJ6 * SD;
    const double AJ4 = -XY - AJ2 * CD + AJ3 * SD;
    const double QX = Q * X11, QY = Q * Y11;
    for (int c = 0; c < 12; ++c) u[c] = 0.0;
    if (DISP) {
        QDFN_ADD3(u[0], -XI * QY - g.tt - A3 * AI1 * SD, -Q / R + A3 * AI3 * SC, Q * QY - A3 * AI3 * SS);
        QDFN_ADD3(u[1], -Q / R + A3 * Y / RD * SD, -ET * QX - g.tt - A3 * XI / RD * SC, Q * QX + A3 * XI / RD * SS);
        QDFN_ADD3(u[2], Q * QY - A3 * AI2 * SD, Q * QX + A3 * AI4 * SC, ET * QX + XI * QY - g.tt - A3 * AI4 * SS);
    }
    QDFN_ADD3(u[3], g.xi2 * Q * Y32 - A3 * AJ1 * SD, XI * Q / R3 + A3 * AJ4 * SC, -XI * g.q2 * Y32 - A3 * AJ4 * SS);
    QDFN_ADD3(u[4], XI * Q / R3 - A3 * AJ2 * SD, ET * Q / R3 + QY + A3 * AJ5 * SC, -g.q2 / R3 - A3 * AJ5 * SS);
    QDFN_ADD3(u[5], -XI * g.q2 * Y32 - A3 * AJ3 * SD, -g.q2 / R3 + A3 * AJ6 * SC, Q * g.q2 * Y32 - A3 * AJ6 * SS);
    QDFN_ADD3(u[6], -XI * g.fy - D * X11 + A3 * (XY + AJ4) * SD, -g.ey + A3 * AJ1 * SC, Q * g.fy - A3 * AJ1 * SS);
    QDFN_ADD3(u[7], -g.ey + A3 * (1.0 / R + AJ5) * SD, -ET * g.gy - XY * SD + A3 * AJ2 * SC, Q * g.gy - A3 * AJ2 * SS);
    QDFN_ADD3(u[8], Q * g.fy - A3 * (QY - AJ6) * SD, Q * g.gy + A3 * AJ3 * SC, -Q * g.hy - A3 * AJ3 * SS);
    QDFN_ADD3(u[9], -XI * g.fz - Y * X11 + A3 * AK1 * SD, -g.ez - A3 * AK3 * SC, Q * g.fz + A3 * AK3 * SS);
    QDFN_ADD3(u[10], -g.ez + A3 * Y * D11 * SD, -ET * g.gz - XY * CD - A3 * XI * D11 * SC,
              Q * g.gz + A3 * XI * D11 * SS);
    QDFN_ADD3(u[11], Q * g.fz + A3 * AK2 * SD, Q * g.gz - A3 * AK4 * SC, -Q * g.hz + A3 * AK4 * SS);
}

// Part C (UC_V, :378-460); Z is the receiver depth coordinate (<= 0).
__device__ __forceinline__ void part_c(double XI, double ET, double Q, double Z, const DipConst& k,
                                       const Corner& g, const SlipW& s, double u[12]) {
    const double A4 = k.a4, A5 = k.a5, SD = k.sd, CD = k.cd, SS = k.ss, CC = k.cc, SC = k.sc;
    const double R = g.r, R2 = g.r2, R3 = g.r3, R5 = g.r5, Y = g.y, D = g.d;
    const double X11 = g.x11, X32 = g.x32, Y11 = g.y11, Y32 = g.y32, XI2 = g.xi2, Q2 = g.q2;
    const double C = D + Z;
    const double X53 = (8.0 * R2 + 9.0 * R * XI + 3.0 * XI2) * X11 * X11 * X11 / R2;
    const double Y53 = (8.0 * R2 + 9.0 * R * ET + 3.0 * g.et2) * Y11 * Y11 * Y11 / R2;
    const double H = Q * CD - Z;
    const double Z32 = SD / R3 - H * Y32;
    const double Z53 = 3.0 * SD / R5 - H * Y53;
    const double Y0 = Y11 - XI2 * Y32;
    const double Z0 = Z32 - XI2 * Z53;
    const double PPY = CD / R3 + Q * Y32 * SD;
    const double PPZ = SD / R3 - Q * Y32 * CD;
    const double QQ = Z * Y32 + Z32 + Z0;
    const double QQY = 3.0 * C * D / R5 - QQ * SD;
    const double QQZ = 3.0 * C * Y / R5 - QQ * CD + Q * Y32;
    const double XY = XI * Y11, QY = Q * Y11;
    const double QR = 3.0 * Q / R5;
    const double CDR = (C + D) / R3;
    const double YY0 = Y / R3 - Y0 * CD;
    for (int c = 0; c < 12; ++c) u[c] = 0.0;
    {  // displacement entries are always formed: the image-source gradient
       // columns 10-12 add them (Functions_OKADA3D.jl:619-621)
        QDFN_ADD3(u[0], A4 * XY * CD - A5 * XI * Q * Z32, A4 * CD / R - QY * SD - A5 * C * Q / R3,
                  -A4 * (SD / R + QY * CD) - A5 * (Z * Y11 - Q2 * Z32));
        QDFN_ADD3(u[1], A4 * (CD / R + 2.0 * QY * SD) - A5 * C * Q / R3, A4 * Y * X11 - A5 * C * ET * Q * X32,
                  A4 * 2.0 * XY * SD + D * X11 - A5 * C * (X11 - Q2 * X32));
        QDFN_ADD3(u[2], A4 * QY * CD - A5 * (C * ET / R3 - Z * Y11 + XI2 * Z32),
                  -D * X11 - XY * SD - A5 * C * (X11 - Q2 * X32),
                  A4 * (Y * X11 + XY * CD) + A5 * Q * (C * ET * X32 + XI * Z32));
    }
    QDFN_ADD3(u[3], A4 * Y0 * CD - A5 * Q * Z0, -A4 * XI / R3 * CD + A5 * C * XI * QR + XI * Q * Y32 * SD,
              A4 * XI / R3 * SD + XI * Q * Y32 * CD + A5 * XI * (3.0 * C * ET / R5 - 2.0 * Z32 - Z0));
    QDFN_ADD3(u[4], -A4 * XI * (CD / R3 + 2.0 * Q * Y32 * SD) + A5 * C * XI * QR,
              -A4 * Y / R3 + A5 * C * ET * QR,
              A4 * 2.0 * Y0 * SD - D / R3 + A5 * C / R3 * (1.0 - 3.0 * Q2 / R2));
    QDFN_ADD3(u[5], -A4 * XI * Q * Y32 * CD + A5 * XI * (3.0 * C * ET / R5 - QQ),
              D / R3 - Y0 * SD + A5 * C / R3 * (1.0 - 3.0 * Q2 / R2),
              -A4 * YY0 - A5 * (C * ET * QR - Q * Z0));
    QDFN_ADD3(u[6], -A4 * XI * PPY * CD - A5 * XI * QQY,
              -A4 * ET / R3 + Y0 * SS - A5 * (CDR * SD - C * Y * QR),
              A4 * (Q / R3 + Y0 * SC) + A5 * (Z / R3 * CD + C * D * QR - Q * Z0 * SD));
    QDFN_ADD3(u[7], A4 * 2.0 * (D / R3 - Y0 * SD) * SD - Y / R3 * CD - A5 * (CDR * SD - ET / R3 - C * Y * QR),
              A4 * (X11 - Y * Y * X32) - A5 * C * ((D + 2.0 * Q * CD) * X32 - Y * ET * Q * X53),
              -A4 * 2.0 * XI * PPY * SD - Y * D * X32 + A5 * C * ((Y + 2.0 * Q * SD) * X32 - Y * Q2 * X53));
    QDFN_ADD3(u[8], -A4 * Q / R3 + YY0 * SD + A5 * (CDR * CD + C * D * QR - (Y0 * CD + Q * Z0) * SD),
              XI * PPY * SD + Y * D * X32 + A5 * C * ((Y + 2.0 * Q * SD) * X32 - Y * Q2 * X53),
              -A4 * (XI * PPY * CD - X11 + Y * Y * X32) +
                  A5 * (C * ((D + 2.0 * Q * CD) * X32 - Y * ET * Q * X53) + XI * QQY));
    QDFN_ADD3(u[9], A4 * XI * PPZ * CD - A5 * XI * QQZ, -Q / R3 + Y0 * SC - A5 * (CDR * CD + C * D * QR),
              -ET / R3 + Y0 * CC - A5 * (Z / R3 * SD - C * Y * QR - Y0 * SS + Q * Z0 * CD));
    QDFN_ADD3(u[10], A4 * 2.0 * (Y / R3 - Y0 * CD) * SD + D / R3 * CD - A5 * (CDR * CD + C * D * QR),
              A4 * Y * D * X32 - A5 * C * ((Y - 2.0 * Q * SD) * X32 + D * ET * Q * X53),
              A4 * 2.0 * XI * PPZ * SD - X11 + D * D * X32 - A5 * C * ((D - 2.0 * Q * CD) * X32 - D * Q2 * X53));
    QDFN_ADD3(u[11], YY0 * CD - A5 * (CDR * SD - C * Y * QR - Y0 * SS + Q * Z0 * CD),
              -XI * PPZ * SD + X11 - D * D * X32 - A5 * C * ((D - 2.0 * Q * CD) * X32 - D * Q2 * X53),
              A4 * (XI * PPZ * CD + Y * D * X32) +
                  A5 * (C * ((Y - 2.0 * Q * SD) * X32 + D * ET * Q * X53) + XI * QQZ));
}
#undef QDFN_ADD3

// Full DC3D at one receiver (real source + image source, 4 corners each).
// out[0..11] = UX,UY,UZ,UXX,UYX,UZX,UXY,UYY,UZY,UXZ,UYZ,UZZ; *iret as the
// reference's IRET (:532-540, :647-648).  With DISP=false only the nine
// gradient entries are formed.
template <bool DISP>
__device__ inline void dc3d_point(const DipConst& k, double X, double Y, double Z, double depth, double al1,
                                  double al2, double aw1, double aw2, const SlipW& s, double out[12],
                                  double* iret) {
    const double SD = k.sd, CD = k.cd;
    double U[12];
    for (int c = 0; c < 12; ++c) U[c] = 0.0;
    double ir = 0.0;
    Corner g;
    double pa[12], pb[12], pc[12];
    // real source
    {
        const double d = depth + Z;
        const double p = Y * CD + d * SD;
        const double q = Y * SD - d * CD;
        const bool jxi = (X + al1) * (X - al2) <= 0.0;
        const bool jet = (p + aw1) * (p - aw2) <= 0.0;
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            const int K = kc / 2 + 1, J = kc % 2 + 1;
            const double et = (K == 1) ? p + aw1 : p - aw2;
            const double xi = (J == 1) ? X + al1 : X - al2;
            corner_terms(xi, et, q, SD, CD, g);
            if (iret) {
                const bool q0 = fabs(q) <= 1.0e-12;
                if ((jxi && q0 && fabs(et) <= 1.0e-12) || (jet && q0 && fabs(xi) <= 1.0e-12)) ir += 1.0;
            }
            part_a<DISP>(xi, et, q, k, g, s, pa);
            const double sg = (J + K != 3) ? 1.0 : -1.0;
#pragma unroll
            for (int c = (DISP ? 0 : 3); c < 12; c += 3) {
                double t0 = -pa[c];
                double t1 = -pa[c + 1] * CD + pa[c + 2] * SD;
                double t2 = -pa[c + 1] * SD - pa[c + 2] * CD;
                if (c == 9) { t0 = -t0; t1 = -t1; t2 = -t2; }
                if (sg > 0.0) { U[c] = U[c] + t0; U[c + 1] = U[c + 1] + t1; U[c + 2] = U[c + 2] + t2; }
                else          { U[c] = U[c] - t0; U[c + 1] = U[c + 1] - t1; U[c + 2] = U[c + 2] - t2; }
            }
        }
    }
    // image source
    {
        const double d = depth - Z;
        const double p = Y * CD + d * SD;
        const double q = Y * SD - d * CD;
#pragma unroll 1
        for (int kc = 0; kc < 4; ++kc) {
            const int K = kc / 2 + 1, J = kc % 2 + 1;
            const double et = (K == 1) ? p + aw1 : p - aw2;
            const double xi = (J == 1) ? X + al1 : X - al2;
            corner_terms(xi, et, q, SD, CD, g);
            part_a<DISP>(xi, et, q, k, g, s, pa);
            part_b<DISP>(xi, et, q, k, g, s, pb);
            part_c(xi, et, q, Z, k, g, s, pc);
            const double sg = (J + K != 3) ? 1.0 : -1.0;
#pragma unroll
            for (int c = (DISP ? 0 : 3); c < 12; c += 3) {
                double t0 = pa[c] + pb[c] + Z * pc[c];
                double t1 = (pa[c + 1] + pb[c + 1] + Z * pc[c + 1]) * CD - (pa[c + 2] + pb[c + 2] + Z * pc[c + 2]) * SD;
                double t2 = (pa[c + 1] + pb[c + 1] - Z * pc[c + 1]) * SD + (pa[c + 2] + pb[c + 2] - Z * pc[c + 2]) * CD;
                if (c == 9) {
                    t0 = t0 + pc[0];
                    t1 = t1 + pc[1] * CD - pc[2] * SD;
                    t2 = t2 - pc[1] * SD - pc[2] * CD;
                }
                if (sg > 0.0) { U[c] = U[c] + t0; U[c + 1] = U[c + 1] + t1; U[c + 2] = U[c + 2] + t2; }
                else          { U[c] = U[c] - t0; U[c + 1] = U[c + 1] - t1; U[c + 2] = U[c + 2] - t2; }
            }
        }
    }
    for (int c = 0; c < 12; ++c) out[c] = U[c];
    if (iret) *iret = ir >= 1.0 ? 1.0 : 0.0;
}

}  // namespace qdfn
