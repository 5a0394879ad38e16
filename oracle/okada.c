/* ORACLE — test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py use it, and only as the checker.
 *
 * CPU FP64 restatement of Okada (1992) DC3D as the reference implements it
 * in scripts/Functions_OKADA3D.jl (which is itself a translation of Okada's
 * Fortran), plus the rectangle stiffness assembly of
 * scripts/Functions_BuildInputFile.jl:386-548.
 *
 * The restatement keeps the reference's IEEE semantics where they are
 * observable:
 *   - DCCON2 zeroes |XI|,|ET|,|Q| < 1e-6 in its *local* copies only; the UA/UB/UC
 *     kernels receive the caller's un-zeroed XI/ET/Q (Julia rebinds, it does not
 *     write back; Functions_OKADA3D.jl:57-67 vs :530-544).
 *   - branch selections are arithmetic blends (c*A + (1-c)*B style, e.g. UB's
 *     CD==0 substitution :283-307), so NaN/Inf in the unselected branch still
 *     propagates exactly as in the reference; the same holds for the slip
 *     components with DISL==0 (:230, :339, :425: U += DISL/2pi * DU * flag).
 *   - singular-point guards (ALX/ALE special cases :117-182, TT :94-101,
 *     UB ATAN NaN->0 :295-296) follow the reference, not Okada's Fortran.
 * Build with -ffp-contract=off so products/sums round like Julia's. */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "degtrig.h"
#include "oracle.h"

#define PI2 6.283185307179586

typedef struct {
    double alp1, alp2, alp3, alp4, alp5, sd, cd, sdsd, cdcd, sdcd;
} dccon0_t;

/* Functions_OKADA3D.jl:12-48 */
static dccon0_t dccon0(double alpha, double dip) {
    dccon0_t c;
    const double EPS = 1.0e-6;
    c.alp1 = (1.0 - alpha) / 2.0;
    c.alp2 = alpha / 2.0;
    c.alp3 = (1.0 - alpha) / alpha;
    c.alp4 = 1.0 - alpha;
    c.alp5 = alpha;
    double p18 = PI2 / 360.0;
    double sd = sin(dip * p18);
    double cd = cos(dip * p18);
    double c1 = fabs(cd) < EPS, c2 = fabs(cd) >= EPS;
    double s1 = sd > 0.0, s2 = sd == 0.0, s3 = sd < 0.0;
    cd = 0.0 * c1 + cd * c2;
    sd = c1 * (1.0 * s1 + sd * s2 + (-1.0) * 1.0 * s3) + c2 * sd;
    c.sd = sd;
    c.cd = cd;
    c.sdsd = sd * sd;
    c.cdcd = cd * cd;
    c.sdcd = sd * cd;
    return c;
}

typedef struct {
    double xi2, et2, q2, r2, r, r3, r5, y, d, tt, alx, x11, x32, ale, y11, y32;
    double ey, ez, fy, fz, gy, gz, hy, hz;
} dccon2_t;

/* Functions_OKADA3D.jl:51-195 (one receiver) */
static void dccon2(double XI, double ET, double Q, double SD, double CD, dccon2_t *o) {
    const double EPS = 0.000001;
    double c1, c2;
    c1 = fabs(XI) < EPS; c2 = fabs(XI) >= EPS; XI = 0.0 * c1 + XI * c2;
    c1 = fabs(ET) < EPS; c2 = fabs(ET) >= EPS; ET = 0.0 * c1 + ET * c2;
    c1 = fabs(Q) < EPS;  c2 = fabs(Q) >= EPS;  Q = 0.0 * c1 + Q * c2;

    double XI2 = XI * XI, ET2 = ET * ET, Q2 = Q * Q;
    double R2 = XI2 + ET2 + Q2;
    double R = sqrt(R2);
    double R3 = R * R2, R5 = R3 * R2;
    double Y = ET * CD + Q * SD;
    double D = ET * SD - Q * CD;

    double TT;
    if (!(Q != 0.0)) TT = 0.0;
    else TT = (double)(Q == 0.0) * 0.0 + 1.0 * atan(XI * ET / (Q * R));

    /* ALX, X11, X32 (:107-143) */
    double k4 = (double)((XI < 0.0) && (Q == 0.0) && (ET == 0.0));
    double RXI = R + XI;
    double ALX, X11, X32;
    if (k4 != 0.0) {
        ALX = (-log(R - XI)) * k4;
        X11 = 0.0 * k4;
        X32 = 0.0 * k4;
    } else {
        double k5 = 1.0 - k4;
        ALX = log(RXI) * k5;
        X11 = (1.0 / (R * RXI)) * k5;
        X32 = ((R + RXI) * X11 * X11 / R) * k5;
    }
    /* ALE, Y11, Y32 (:145-182) */
    double e4 = (double)((ET < 0.0) && (Q == 0.0) && (XI == 0.0));
    double RET = R + ET;
    double ALE, Y11, Y32;
    if (e4 != 0.0) {
        ALE = (-log(R - ET)) * e4;
        Y11 = 0.0 * e4;
        Y32 = 0.0 * e4;
    } else {
        double e5 = 1.0 - e4;
        ALE = log(RET) * e5;
        Y11 = (1.0 / (R * RET)) * e5;
        Y32 = ((R + RET) * Y11 * Y11 / R) * e5;
    }
    o->xi2 = XI2; o->et2 = ET2; o->q2 = Q2; o->r2 = R2; o->r = R; o->r3 = R3; o->r5 = R5;
    o->y = Y; o->d = D; o->tt = TT; o->alx = ALX; o->x11 = X11; o->x32 = X32;
    o->ale = ALE; o->y11 = Y11; o->y32 = Y32;
    o->ey = SD / R - Y * Q / R3;
    o->ez = CD / R + D * Q / R3;
    o->fy = D / R3 + XI2 * Y32 * SD;
    o->fz = Y / R3 + XI2 * Y32 * CD;
    o->gy = 2.0 * X11 * SD - Y * Q * X32;
    o->gz = 2.0 * X11 * CD + D * Q * X32;
    o->hy = D * Q * X32 + XI * Q * Y32 * SD;
    o->hz = Y * Q * X32 + XI * Q * Y32 * CD;
}

static void accum(double U[12], double disl, const double du[12]) {
    double flag = (double)(disl != 0.0);
    double s = disl / PI2;
    for (int k = 0; k < 12; ++k) U[k] = U[k] + s * du[k] * flag;
}

/* Functions_OKADA3D.jl:199-262 */
static void ua(double XI, double ET, double Q, double D1, double D2, double D3,
               const dccon0_t *c, const dccon2_t *g, double U[12]) {
    const double ALP1 = c->alp1, ALP2 = c->alp2, SD = c->sd, CD = c->cd;
    const double XI2 = g->xi2, Q2 = g->q2, R = g->r, R3 = g->r3, Y = g->y, D = g->d;
    const double TT = g->tt, ALX = g->alx, X11 = g->x11, ALE = g->ale, Y11 = g->y11, Y32 = g->y32;
    const double EY = g->ey, EZ = g->ez, FY = g->fy, FZ = g->fz, GY = g->gy, GZ = g->gz;
    const double HY = g->hy, HZ = g->hz;
    double du[12];
    for (int k = 0; k < 12; ++k) U[k] = 0.0;
    double XY = XI * Y11, QX = Q * X11, QY = Q * Y11;

    du[0] = TT / 2.0 + ALP2 * XI * QY;
    du[1] = ALP2 * Q / R;
    du[2] = ALP1 * ALE - ALP2 * Q * QY;
    du[3] = -ALP1 * QY - ALP2 * XI2 * Q * Y32;
    du[4] = -ALP2 * XI * Q / R3;
    du[5] = ALP1 * XY + ALP2 * XI * Q2 * Y32;
    du[6] = ALP1 * XY * SD + ALP2 * XI * FY + D / 2.0 * X11;
    du[7] = ALP2 * EY;
    du[8] = ALP1 * (CD / R + QY * SD) - ALP2 * Q * FY;
    du[9] = ALP1 * XY * CD + ALP2 * XI * FZ + Y / 2.0 * X11;
    du[10] = ALP2 * EZ;
    du[11] = -ALP1 * (SD / R - QY * CD) - ALP2 * Q * FZ;
    accum(U, D1, du);

    du[0] = ALP2 * Q / R;
    du[1] = TT / 2.0 + ALP2 * ET * QX;
    du[2] = ALP1 * ALX - ALP2 * Q * QX;
    du[3] = -ALP2 * XI * Q / R3;
    du[4] = -QY / 2.0 - ALP2 * ET * Q / R3;
    du[5] = ALP1 / R + ALP2 * Q2 / R3;
    du[6] = ALP2 * EY;
    du[7] = ALP1 * D * X11 + XY / 2.0 * SD + ALP2 * ET * GY;
    du[8] = ALP1 * Y * X11 - ALP2 * Q * GY;
    du[9] = ALP2 * EZ;
    du[10] = ALP1 * Y * X11 + XY / 2.0 * CD + ALP2 * ET * GZ;
    du[11] = -ALP1 * D * X11 - ALP2 * Q * GZ;
    accum(U, D2, du);

    du[0] = -ALP1 * ALE - ALP2 * Q * QY;
    du[1] = -ALP1 * ALX - ALP2 * Q * QX;
    du[2] = TT / 2.0 - ALP2 * (ET * QX + XI * QY);
    du[3] = -ALP1 * XY + ALP2 * XI * Q2 * Y32;
    du[4] = -ALP1 / R + ALP2 * Q2 / R3;
    du[5] = -ALP1 * QY - ALP2 * Q * Q2 * Y32;
    du[6] = -ALP1 * (CD / R + QY * SD) - ALP2 * Q * FY;
    du[7] = -ALP1 * Y * X11 - ALP2 * Q * GY;
    du[8] = ALP1 * (D * X11 + XY * SD) + ALP2 * Q * HY;
    du[9] = ALP1 * (SD / R - QY * CD) - ALP2 * Q * FZ;
    du[10] = ALP1 * D * X11 - ALP2 * Q * GZ;
    du[11] = ALP1 * (Y * X11 + XY * CD) + ALP2 * Q * HZ;
    accum(U, D3, du);
}

/* Functions_OKADA3D.jl:266-375 */
static void ub(double XI, double ET, double Q, double D1, double D2, double D3,
               const dccon0_t *c, const dccon2_t *g, double U[12]) {
    const double ALP3 = c->alp3, SD = c->sd, SDSD = c->sdsd, SDCD = c->sdcd;
    double CD = c->cd, CDCD = c->cdcd;
    const double XI2 = g->xi2, Q2 = g->q2, R = g->r, R3 = g->r3, Y = g->y, D = g->d;
    const double TT = g->tt, X11 = g->x11, ALE = g->ale, Y11 = g->y11, Y32 = g->y32;
    const double EY = g->ey, EZ = g->ez, FY = g->fy, FZ = g->fz, GY = g->gy, GZ = g->gz;
    const double HY = g->hy, HZ = g->hz;
    double du[12];
    for (int k = 0; k < 12; ++k) U[k] = 0.0;

    double RD = R + D;
    double D11 = 1.0 / (R * RD);
    double AJ2 = XI * Y / RD * D11;
    double AJ5 = -(D + Y * Y / RD) * D11;

    double c1 = CD != 0.0, c2 = CD == 0.0;
    double s1 = XI == 0.0, s2 = XI != 0.0;
    double tempCD = CD, tempCDCD = CDCD;
    CD = c1 * CD + c2 * 1.0e-12;
    CDCD = c1 * CDCD + c2 * 1.0e-12;

    double X = sqrt(XI2 + Q2);
    double ATAN = atan((ET * (X + Q * CD) + X * (R + X) * SD) / (XI * (R + X) * CD));
    if (isnan(ATAN)) ATAN = 0.0;
    double RD2 = RD * RD;
    double AI4 = c1 * (s1 * 0.0 + s2 * (1.0 / CDCD * (XI / RD * SDCD + 2.0 * ATAN))) +
                 c2 * (XI * Y / RD2 / 2.0);
    double AI3 = c1 * ((Y * CD / RD - ALE + SD * log(RD)) / CDCD) +
                 c2 * ((ET / RD + Y * Q / RD2 - ALE) / 2.0);
    double AK1 = c1 * (XI * (D11 - Y11 * SD) / CD) + c2 * (XI * Q / RD * D11);
    double AK3 = c1 * ((Q * Y11 - Y * D11) / CD) + c2 * (SD / RD * (XI2 * D11 - 1.0));
    double AJ3 = c1 * ((AK1 - AJ2 * SD) / CD) + c2 * (-XI / RD2 * (Q2 * D11 - 1.0 / 2.0));
    double AJ6 = c1 * ((AK3 - AJ5 * SD) / CD) + c2 * (-Y / RD2 * (XI2 * D11 - 1.0 / 2.0));
    CD = tempCD;
    CDCD = tempCDCD;

    double XY = XI * Y11;
    double AI1 = -XI / RD * CD - AI4 * SD;
    double AI2 = log(RD) + AI3 * SD;
    double AK2 = 1.0 / R + AK3 * SD;
    double AK4 = XY * CD - AK1 * SD;
    double AJ1 = AJ5 * CD - AJ6 * SD;
    double AJ4 = -XY - AJ2 * CD + AJ3 * SD;
    double QX = Q * X11, QY = Q * Y11;

    du[0] = -XI * QY - TT - ALP3 * AI1 * SD;
    du[1] = -Q / R + ALP3 * Y / RD * SD;
    du[2] = Q * QY - ALP3 * AI2 * SD;
    du[3] = XI2 * Q * Y32 - ALP3 * AJ1 * SD;
    du[4] = XI * Q / R3 - ALP3 * AJ2 * SD;
    du[5] = -XI * Q2 * Y32 - ALP3 * AJ3 * SD;
    du[6] = -XI * FY - D * X11 + ALP3 * (XY + AJ4) * SD;
    du[7] = -EY + ALP3 * (1.0 / R + AJ5) * SD;
    du[8] = Q * FY - ALP3 * (QY - AJ6) * SD;
    du[9] = -XI * FZ - Y * X11 + ALP3 * AK1 * SD;
    du[10] = -EZ + ALP3 * Y * D11 * SD;
    du[11] = Q * FZ + ALP3 * AK2 * SD;
    accum(U, D1, du);

    du[0] = -Q / R + ALP3 * AI3 * SDCD;
    du[1] = -ET * QX - TT - ALP3 * XI / RD * SDCD;
    du[2] = Q * QX + ALP3 * AI4 * SDCD;
    du[3] = XI * Q / R3 + ALP3 * AJ4 * SDCD;
    du[4] = ET * Q / R3 + QY + ALP3 * AJ5 * SDCD;
    du[5] = -Q2 / R3 + ALP3 * AJ6 * SDCD;
    du[6] = -EY + ALP3 * AJ1 * SDCD;
    du[7] = -ET * GY - XY * SD + ALP3 * AJ2 * SDCD;
    du[8] = Q * GY + ALP3 * AJ3 * SDCD;
    du[9] = -EZ - ALP3 * AK3 * SDCD;
    du[10] = -ET * GZ - XY * CD - ALP3 * XI * D11 * SDCD;
    du[11] = Q * GZ - ALP3 * AK4 * SDCD;
    accum(U, D2, du);

    du[0] = Q * QY - ALP3 * AI3 * SDSD;
    du[1] = Q * QX + ALP3 * XI / RD * SDSD;
    du[2] = ET * QX + XI * QY - TT - ALP3 * AI4 * SDSD;
    du[3] = -XI * Q2 * Y32 - ALP3 * AJ4 * SDSD;
    du[4] = -Q2 / R3 - ALP3 * AJ5 * SDSD;
    du[5] = Q * Q2 * Y32 - ALP3 * AJ6 * SDSD;
    du[6] = Q * FY - ALP3 * AJ1 * SDSD;
    du[7] = Q * GY - ALP3 * AJ2 * SDSD;
    du[8] = -Q * HY - ALP3 * AJ3 * SDSD;
    du[9] = Q * FZ + ALP3 * AK3 * SDSD;
    du[10] = Q * GZ + ALP3 * XI * D11 * SDSD;
    du[11] = -Q * HZ + ALP3 * AK4 * SDSD;
    accum(U, D3, du);
}

/* Functions_OKADA3D.jl:378-460 */
static void uc(double XI, double ET, double Q, double D1, double D2, double D3, double Z,
               const dccon0_t *c, const dccon2_t *g, double U[12]) {
    const double ALP4 = c->alp4, ALP5 = c->alp5, SD = c->sd, CD = c->cd;
    const double SDSD = c->sdsd, CDCD = c->cdcd, SDCD = c->sdcd;
    const double XI2 = g->xi2, ET2 = g->et2, Q2 = g->q2, R2 = g->r2, R = g->r, R3 = g->r3;
    const double R5 = g->r5, Y = g->y, D = g->d, X11 = g->x11, X32 = g->x32;
    const double Y11 = g->y11, Y32 = g->y32;
    double du[12];
    for (int k = 0; k < 12; ++k) U[k] = 0.0;

    double C = D + Z;
    double X53 = (8.0 * R2 + 9.0 * R * XI + 3.0 * XI2) * X11 * X11 * X11 / R2;
    double Y53 = (8.0 * R2 + 9.0 * R * ET + 3.0 * ET2) * Y11 * Y11 * Y11 / R2;
    double H = Q * CD - Z;
    double Z32 = SD / R3 - H * Y32;
    double Z53 = 3.0 * SD / R5 - H * Y53;
    double Y0 = Y11 - XI2 * Y32;
    double Z0 = Z32 - XI2 * Z53;
    double PPY = CD / R3 + Q * Y32 * SD;
    double PPZ = SD / R3 - Q * Y32 * CD;
    double QQ = Z * Y32 + Z32 + Z0;
    double QQY = 3.0 * C * D / R5 - QQ * SD;
    double QQZ = 3.0 * C * Y / R5 - QQ * CD + Q * Y32;
    double XY = XI * Y11, QY = Q * Y11;
    double QR = 3.0 * Q / R5;
    double CDR = (C + D) / R3;
    double YY0 = Y / R3 - Y0 * CD;

    du[0] = ALP4 * XY * CD - ALP5 * XI * Q * Z32;
    du[1] = ALP4 * (CD / R + 2.0 * QY * SD) - ALP5 * C * Q / R3;
    du[2] = ALP4 * QY * CD - ALP5 * (C * ET / R3 - Z * Y11 + XI2 * Z32);
    du[3] = ALP4 * Y0 * CD - ALP5 * Q * Z0;
    du[4] = -ALP4 * XI * (CD / R3 + 2.0 * Q * Y32 * SD) + ALP5 * C * XI * QR;
    du[5] = -ALP4 * XI * Q * Y32 * CD + ALP5 * XI * (3.0 * C * ET / R5 - QQ);
    du[6] = -ALP4 * XI * PPY * CD - ALP5 * XI * QQY;
    du[7] = ALP4 * 2.0 * (D / R3 - Y0 * SD) * SD - Y / R3 * CD - ALP5 * (CDR * SD - ET / R3 - C * Y * QR);
    du[8] = -ALP4 * Q / R3 + YY0 * SD + ALP5 * (CDR * CD + C * D * QR - (Y0 * CD + Q * Z0) * SD);
    du[9] = ALP4 * XI * PPZ * CD - ALP5 * XI * QQZ;
    du[10] = ALP4 * 2.0 * (Y / R3 - Y0 * CD) * SD + D / R3 * CD - ALP5 * (CDR * CD + C * D * QR);
    du[11] = YY0 * CD - ALP5 * (CDR * SD - C * Y * QR - Y0 * SDSD + Q * Z0 * CD);
    accum(U, D1, du);

    du[0] = ALP4 * CD / R - QY * SD - ALP5 * C * Q / R3;
    du[1] = ALP4 * Y * X11 - ALP5 * C * ET * Q * X32;
    du[2] = -D * X11 - XY * SD - ALP5 * C * (X11 - Q2 * X32);
    du[3] = -ALP4 * XI / R3 * CD + ALP5 * C * XI * QR + XI * Q * Y32 * SD;
    du[4] = -ALP4 * Y / R3 + ALP5 * C * ET * QR;
    du[5] = D / R3 - Y0 * SD + ALP5 * C / R3 * (1.0 - 3.0 * Q2 / R2);
    du[6] = -ALP4 * ET / R3 + Y0 * SDSD - ALP5 * (CDR * SD - C * Y * QR);
    du[7] = ALP4 * (X11 - Y * Y * X32) - ALP5 * C * ((D + 2.0 * Q * CD) * X32 - Y * ET * Q * X53);
    du[8] = XI * PPY * SD + Y * D * X32 + ALP5 * C * ((Y + 2.0 * Q * SD) * X32 - Y * Q2 * X53);
    du[9] = -Q / R3 + Y0 * SDCD - ALP5 * (CDR * CD + C * D * QR);
    du[10] = ALP4 * Y * D * X32 - ALP5 * C * ((Y - 2.0 * Q * SD) * X32 + D * ET * Q * X53);
    du[11] = -XI * PPZ * SD + X11 - D * D * X32 - ALP5 * C * ((D - 2.0 * Q * CD) * X32 - D * Q2 * X53);
    accum(U, D2, du);

    du[0] = -ALP4 * (SD / R + QY * CD) - ALP5 * (Z * Y11 - Q2 * Z32);
    du[1] = ALP4 * 2.0 * XY * SD + D * X11 - ALP5 * C * (X11 - Q2 * X32);
    du[2] = ALP4 * (Y * X11 + XY * CD) + ALP5 * Q * (C * ET * X32 + XI * Z32);
    du[3] = ALP4 * XI / R3 * SD + XI * Q * Y32 * CD + ALP5 * XI * (3.0 * C * ET / R5 - 2.0 * Z32 - Z0);
    du[4] = ALP4 * 2.0 * Y0 * SD - D / R3 + ALP5 * C / R3 * (1.0 - 3.0 * Q2 / R2);
    du[5] = -ALP4 * YY0 - ALP5 * (C * ET * QR - Q * Z0);
    du[6] = ALP4 * (Q / R3 + Y0 * SDCD) + ALP5 * (Z / R3 * CD + C * D * QR - Q * Z0 * SD);
    du[7] = -ALP4 * 2.0 * XI * PPY * SD - Y * D * X32 + ALP5 * C * ((Y + 2.0 * Q * SD) * X32 - Y * Q2 * X53);
    du[8] = -ALP4 * (XI * PPY * CD - X11 + Y * Y * X32) +
            ALP5 * (C * ((D + 2.0 * Q * CD) * X32 - Y * ET * Q * X53) + XI * QQY);
    du[9] = -ET / R3 + Y0 * CDCD - ALP5 * (Z / R3 * SD - C * Y * QR - Y0 * SDSD + Q * Z0 * CD);
    du[10] = ALP4 * 2.0 * XI * PPZ * SD - X11 + D * D * X32 - ALP5 * C * ((D - 2.0 * Q * CD) * X32 - D * Q2 * X53);
    du[11] = ALP4 * (XI * PPZ * CD + Y * D * X32) +
             ALP5 * (C * ((Y - 2.0 * Q * SD) * X32 + D * ET * Q * X53) + XI * QQZ);
    accum(U, D3, du);
}

/* Okada_DC3D_Vector, Functions_OKADA3D.jl:464-654, for one receiver point. */
static void dc3d_point(double ALPHA, double X, double Y, double Z, double DEPTH, double DIP,
                       double AL1, double AL2, double AW1, double AW2,
                       double D1, double D2, double D3, double U[12], double *iret) {
    dccon0_t c = dccon0(ALPHA, DIP);
    const double SD = c.sd, CD = c.cd;
    double DUA[12], DUB[12], DUC[12], DU[12];
    dccon2_t g;
    for (int k = 0; k < 12; ++k) U[k] = 0.0;
    double IRET = 0.0;

    /* real-source contribution (:495-567) */
    double D = DEPTH + Z;
    double P = Y * CD + D * SD;
    double Q = Y * SD - D * CD;
    double JXI = (double)((X + AL1) * (X - AL2) <= 0.0);
    double JET = (double)((P + AW1) * (P - AW2) <= 0.0);
    for (int K = 1; K <= 2; ++K) {
        double ET = (K == 1) ? P + AW1 : P - AW2;
        for (int J = 1; J <= 2; ++J) {
            double XI = (J == 1) ? X + AL1 : X - AL2;
            dccon2(XI, ET, Q, SD, CD, &g);
            double cq1 = fabs(Q) <= 1.0e-12, cet1 = fabs(ET) <= 1.0e-12, cxi1 = fabs(XI) <= 1.0e-12;
            double cc1 = (JXI == 1.0) * cq1 * cet1;
            double cc3 = (JET == 1.0) * cq1 * cxi1;
            IRET = IRET + (double)((cc1 + cc3) >= 1.0);
            ua(XI, ET, Q, D1, D2, D3, &c, &g, DUA);
            for (int I = 0; I < 12; I += 3) {
                DU[I] = -DUA[I];
                DU[I + 1] = -DUA[I + 1] * CD + DUA[I + 2] * SD;
                DU[I + 2] = -DUA[I + 1] * SD - DUA[I + 2] * CD;
                if (I < 9) continue;
                DU[I] = -DU[I];
                DU[I + 1] = -DU[I + 1];
                DU[I + 2] = -DU[I + 2];
            }
            if (J + K != 3) for (int k = 0; k < 12; ++k) U[k] = U[k] + DU[k];
            else            for (int k = 0; k < 12; ++k) U[k] = U[k] - DU[k];
        }
    }

    /* image-source contribution (:574-634) */
    D = DEPTH - Z;
    P = Y * CD + D * SD;
    Q = Y * SD - D * CD;
    for (int K = 1; K <= 2; ++K) {
        double ET = (K == 1) ? P + AW1 : P - AW2;
        for (int J = 1; J <= 2; ++J) {
            double XI = (J == 1) ? X + AL1 : X - AL2;
            dccon2(XI, ET, Q, SD, CD, &g);
            ua(XI, ET, Q, D1, D2, D3, &c, &g, DUA);
            ub(XI, ET, Q, D1, D2, D3, &c, &g, DUB);
            uc(XI, ET, Q, D1, D2, D3, Z, &c, &g, DUC);
            for (int I = 0; I < 12; I += 3) {
                DU[I] = DUA[I] + DUB[I] + Z * DUC[I];
                DU[I + 1] = (DUA[I + 1] + DUB[I + 1] + Z * DUC[I + 1]) * CD -
                            (DUA[I + 2] + DUB[I + 2] + Z * DUC[I + 2]) * SD;
                DU[I + 2] = (DUA[I + 1] + DUB[I + 1] - Z * DUC[I + 1]) * SD +
                            (DUA[I + 2] + DUB[I + 2] - Z * DUC[I + 2]) * CD;
                if (I < 9) continue;
                DU[9] = DU[9] + DUC[0];
                DU[10] = DU[10] + DUC[1] * CD - DUC[2] * SD;
                DU[11] = DU[11] - DUC[1] * SD - DUC[2] * CD;
            }
            if (J + K != 3) for (int k = 0; k < 12; ++k) U[k] = U[k] + DU[k];
            else            for (int k = 0; k < 12; ++k) U[k] = U[k] - DU[k];
        }
    }
    *iret = (double)(IRET >= 1.0);
}

/* Vectorised entry with the reference's signature: U is n x 12, column-major
 * (column k holds UX,UY,UZ,UXX,UYX,UZX,UXY,UYY,UZY,UXZ,UYZ,UZZ for k=0..11). */
void orc_dc3d(double alpha, int64_t n, const double *X, const double *Y, const double *Z,
              double depth, double dip, double al1, double al2, double aw1, double aw2,
              double disl1, double disl2, double disl3, double *U, double *iret) {
    double u[12];
    for (int64_t i = 0; i < n; ++i) {
        dc3d_point(alpha, X[i], Y[i], Z[i], depth, dip, al1, al2, aw1, aw2, disl1, disl2, disl3,
                   u, &iret[i]);
        for (int k = 0; k < 12; ++k) U[i + k * n] = u[k];
    }
}

/* Julia's 3x3 product, left-to-right sums (LinearAlgebra matmul3x3). */
static void mm3(const double A[3][3], const double B[3][3], double C[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

/* StiffnessMatrix_ByParts_Calculation_Rec, Functions_BuildInputFile.jl:386-512.
 * seg tables are column-major (element i, column c at seg[i + c*ld]) with the
 * reference's Input_Segment columns: 0-2 centre (z = +depth), 3 strike length,
 * 4 dip length, 5 strike deg, 6 dip deg, 7 rake deg.  KS/KN are nr x ns,
 * column-major with leading dimension ldk. */
void orc_stiffness_rect(const double *src, int64_t ns, int64_t lds,
                        const double *rcv, int64_t nr, int64_t ldr,
                        double mu, double nu, double *KS, double *KN, int64_t ldk) {
    double lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
    double alpha = (lam + mu) / (lam + 2.0 * mu);
    for (int64_t s = 0; s < ns; ++s) {
        double sx = src[s], sy = src[s + lds], sz = src[s + 2 * lds];
        double L = src[s + 3 * lds], W = src[s + 4 * lds];
        double sstr = src[s + 5 * lds], sdip = src[s + 6 * lds], srake = src[s + 7 * lds];
        double disl1 = orc_cosd(srake), disl2 = orc_sind(srake), disl3 = 0.0;
        double depth = sz;
        double al1 = L / 2.0, al2 = L / 2.0, aw1 = W / 2.0, aw2 = W / 2.0;
        double cs = orc_cosd(-sstr), sn = orc_sind(-sstr);
        for (int64_t r = 0; r < nr; ++r) {
            double xd = rcv[r] - sx, yd = rcv[r + ldr] - sy;
            double X = xd * cs - yd * sn;
            double Y = xd * sn + yd * cs;
            double Z = -rcv[r + 2 * ldr];
            double u[12], iret;
            dc3d_point(alpha, X, Y, Z, depth, sdip, al1, al2, aw1, aw2, disl1, disl2, disl3, u, &iret);
            double UXX = u[3], UYX = u[4], UZX = u[5], UXY = u[6], UYY = u[7], UZY = u[8];
            double UXZ = u[9], UYZ = u[10], UZZ = u[11];
            double sxx = lam * (UXX + UYY + UZZ) + 2.0 * mu * UXX;
            double syy = lam * (UXX + UYY + UZZ) + 2.0 * mu * UYY;
            double szz = lam * (UXX + UYY + UZZ) + 2.0 * mu * UZZ;
            double sxy = (UXY + UYX) * mu;
            double sxz = (UXZ + UZX) * mu;
            double syz = (UYZ + UZY) * mu;
            double S[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
            double rel = rcv[r + 5 * ldr] - sstr;
            double rdip = rcv[r + 6 * ldr];
            double Rs[3][3] = {{orc_cosd(-rel), -orc_sind(-rel), 0.0},
                               {orc_sind(-rel), orc_cosd(-rel), 0.0},
                               {0.0, 0.0, 1.0}};
            double Rd[3][3] = {{1.0, 0.0, 0.0},
                               {0.0, orc_cosd(-rdip), -orc_sind(-rdip)},
                               {0.0, orc_sind(-rdip), orc_cosd(-rdip)}};
            double R[3][3], RS[3][3], Rt[3][3], Sr[3][3];
            mm3(Rd, Rs, R);
            mm3(R, S, RS);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Rt[i][j] = R[j][i];
            mm3(RS, Rt, Sr);
            double rrake = rcv[r + 7 * ldr];
            KN[r + s * ldk] = -Sr[2][2];
            KS[r + s * ldk] = orc_cosd(rrake) * Sr[0][2] + orc_sind(rrake) * Sr[1][2];
        }
    }
}
