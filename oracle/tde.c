/* The angular-dislocation and free-surface-correction strain expressions
 * below follow the published solution of Nikkhoo & Walter (2015), as carried
 * in the reference's scripts/Functions_TDstressHS.jl, whose notice reads:
 * 
 *   Copyright (c) 2014 Mehdi Nikkhoo
 * 
 *   Permission is hereby granted, free of charge, to any person obtaining a
 *   copy of this software and associated documentation files
 *   (the "Software"), to deal in the Software without restriction, including
 *   without limitation the rights to use, copy, modify, merge, publish,
 *   distribute, sublicense, and/or sell copies of the Software, and to permit
 *   persons to whom the Software is furnished to do so, subject to the
 *   following conditions:
 * 
 *   The above copyright notice and this permission notice shall be included
 *   in all copies or substantial portions of the Software.
 * 
 *   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
 *   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
 *   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
 *   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
 *   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
 *   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
 *   USE OR OTHER DEALINGS IN THE SOFTWARE.
 */
/* ORACLE — test infrastructure only (see okada.c header).
 *
 * CPU FP64 restatement of the triangular-dislocation half-space stress of
 * Nikkhoo & Walter (2015) as the reference's scripts/Functions_TDstressHS.jl
 * states it (TDstressHS :2-130, TDstressFS :133-286, TDstress_HarFunc
 * :289-335, TensTrans :338-362, CoordTrans :365-390, trimodefinder
 * :393-430, TDSetupS :433-464, AngSetupFSC_S :467-574, AngDisStrain
 * :577-651, AngDisStrainFSC :654-1034), evaluated point by point, plus the
 * triangle assembly of scripts/Functions_BuildInputFile.jl
 * (RotVerts_UnitVectors :338-377, StiffnessMatrix_ByParts_Calculation_Tri
 * :551-594).  Unlike the reference, TDstressHS does not flip the caller's
 * vertex z in place (the reference's caller passes row copies, so this is
 * not observable there).
 * Build with -ffp-contract=off. */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "degtrig.h"
#include "oracle.h"

#define PI 3.141592653589793
static inline double sq(double x) { return x * x; }
static inline double cube(double x) { return x * x * x; }

static void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double norm3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* TensTrans (:338-362): A is 3x3 column-major (A[0..8] = Julia A[1..9]). */
static void tens_trans(const double T1[6], const double A[9], double T2[6]) {
    const double xx = T1[0], yy = T1[1], zz = T1[2], xy = T1[3], xz = T1[4], yz = T1[5];
    T2[0] = sq(A[0]) * xx + 2 * A[0] * A[3] * xy + 2 * A[0] * A[6] * xz + 2 * A[3] * A[6] * yz + sq(A[3]) * yy +
            sq(A[6]) * zz;
    T2[1] = sq(A[1]) * xx + 2 * A[1] * A[4] * xy + 2 * A[1] * A[7] * xz + 2 * A[4] * A[7] * yz + sq(A[4]) * yy +
            sq(A[7]) * zz;
    T2[2] = sq(A[2]) * xx + 2 * A[2] * A[5] * xy + 2 * A[2] * A[8] * xz + 2 * A[5] * A[8] * yz + sq(A[5]) * yy +
            sq(A[8]) * zz;
    T2[3] = A[0] * A[1] * xx + (A[0] * A[4] + A[1] * A[3]) * xy + (A[0] * A[7] + A[1] * A[6]) * xz +
            (A[7] * A[3] + A[6] * A[4]) * yz + A[4] * A[3] * yy + A[6] * A[7] * zz;
    T2[4] = A[0] * A[2] * xx + (A[0] * A[5] + A[2] * A[3]) * xy + (A[0] * A[8] + A[2] * A[6]) * xz +
            (A[8] * A[3] + A[6] * A[5]) * yz + A[5] * A[3] * yy + A[6] * A[8] * zz;
    T2[5] = A[1] * A[2] * xx + (A[2] * A[4] + A[1] * A[5]) * xy + (A[2] * A[7] + A[1] * A[8]) * xz +
            (A[7] * A[5] + A[8] * A[4]) * yz + A[4] * A[5] * yy + A[7] * A[8] * zz;
}

/* CoordTrans (:365-390): r = A * x with A column-major. */
static void coord_trans(double x1, double x2, double x3, const double A[9], double r[3]) {
    r[0] = A[0] * x1 + A[3] * x2 + A[6] * x3;
    r[1] = A[1] * x1 + A[4] * x2 + A[7] * x3;
    r[2] = A[2] * x1 + A[5] * x2 + A[8] * x3;
}

/* trimodefinder (:393-430) for one point; x,y are the TDCS y,z coordinates,
 * z the TDCS x coordinate; p1..p3 are (y,z) of the vertices. */
static int trimode(double x, double y, double z, const double p1[2], const double p2[2], const double p3[2]) {
    const double den = (p2[1] - p3[1]) * (p1[0] - p3[0]) + (p3[0] - p2[0]) * (p1[1] - p3[1]);
    const double a = ((p2[1] - p3[1]) * (x - p3[0]) + (p3[0] - p2[0]) * (y - p3[1])) / den;
    const double b = ((p3[1] - p1[1]) * (x - p3[0]) + (p1[0] - p3[0]) * (y - p3[1])) / den;
    const double c = 1 - a - b;
    int m = 1;
    if (a <= 0 && b > c && c > a) m = -1;
    if (b <= 0 && c > a && a > b) m = -1;
    if (c <= 0 && a > b && b > c) m = -1;
    if (a == 0 && b >= 0 && c >= 0) m = 0;
    if (a >= 0 && b == 0 && c >= 0) m = 0;
    if (a >= 0 && b >= 0 && c == 0) m = 0;
    if (m == 0 && z != 0) m = 1;
    return m;
}

/* AngDisStrain (:577-651): strains of an angular dislocation, full space. */
static void ang_dis_strain(double x, double y, double z, double alpha, double bx, double by, double bz, double nu,
                           double e[6]) {
    const double sinA = sin(alpha), cosA = cos(alpha);
    const double eta = y * cosA - z * sinA;
    const double zeta = y * sinA + z * cosA;
    const double x2 = sq(x), y2 = sq(y), z2 = sq(z);
    const double r2 = x2 + y2 + z2;
    const double r = sqrt(r2);
    const double r3 = r * r2;
    const double rz = r * (r - z);
    const double r2z2 = r2 * sq(r - z);
    const double r3z = r3 * (r - z);
    const double W = zeta - r;
    const double W2 = sq(W);
    const double Wr = W * r;
    const double W2r = W2 * r;
    const double Wr3 = W * r3;
    const double W2r2 = W2 * r2;
    const double C = (r * cosA - z) / Wr;
    const double S = (r * sinA - y) / Wr;
    const double rFi_rx = (eta / r / (r - zeta) - y / r / (r - z)) / 4 / PI;
    const double rFi_ry = (x / r / (r - z) - cosA * x / r / (r - zeta)) / 4 / PI;
    const double rFi_rz = (sinA * x / r / (r - zeta)) / 4 / PI;
    const double k = 1.0 / 8 / PI / (1 - nu);
    (void)k;
    e[0] = bx * (rFi_rx) +
           bx / 8 / PI / (1 - nu) * (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz - x2 * y / r2z2 - x2 * y / r3z) -
           by * x / 8 / PI / (1 - nu) *
               (((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3) * cosA + (2 * nu + 1) / rz - x2 / r2z2 - x2 / r3z) +
           bz * x * sinA / 8 / PI / (1 - nu) * ((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3);
    e[1] = by * (rFi_ry) +
           bx / 8 / PI / (1 - nu) *
               ((1 / Wr + sq(S) - y2 / Wr3) * eta + (2 * nu + 1) * y / rz - cube(y) / r2z2 - cube(y) / r3z -
                2 * nu * cosA * S) -
           by * x / 8 / PI / (1 - nu) * (1 / rz - y2 / r2z2 - y2 / r3z + (1 / Wr + sq(S) - y2 / Wr3) * cosA) +
           bz * x * sinA / 8 / PI / (1 - nu) * (1 / Wr + sq(S) - y2 / Wr3);
    e[2] = bz * (rFi_rz) +
           bx / 8 / PI / (1 - nu) * (eta / W / r + eta * sq(C) - eta * z2 / Wr3 + y * z / r3 + 2 * nu * sinA * C) -
           by * x / 8 / PI / (1 - nu) * ((1 / Wr + sq(C) - z2 / Wr3) * cosA + z / r3) +
           bz * x * sinA / 8 / PI / (1 - nu) * (1 / Wr + sq(C) - z2 / Wr3);
    e[3] = bx * (rFi_ry) / 2 + by * (rFi_rx) / 2 -
           bx / 8 / PI / (1 - nu) *
               (x * y2 / r2z2 - nu * x / rz + x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr +
                eta * x * y / Wr3) +
           by / 8 / PI / (1 - nu) *
               (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S + x2 * y * cosA / Wr3 +
                x2 * cosA * S / Wr) -
           bz * sinA / 8 / PI / (1 - nu) * (nu * S + x2 * S / Wr + x2 * y / Wr3);
    e[4] = bx * (rFi_rz) / 2 + bz * (rFi_rx) / 2 -
           bx / 8 / PI / (1 - nu) * (-x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr + eta * x * z / Wr3) +
           by / 8 / PI / (1 - nu) *
               (-x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 + x2 * cosA * C / Wr) -
           bz * sinA / 8 / PI / (1 - nu) * (nu * C + x2 * C / Wr + x2 * z / Wr3);
    e[5] = by * (rFi_rz) / 2 + bz * (rFi_ry) / 2 +
           bx / 8 / PI / (1 - nu) *
               (y2 / r3 - nu / r - nu * cosA * C + nu * sinA * S + eta * sinA * cosA / W2 -
                eta * (y * cosA + z * sinA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3) -
           by * x / 8 / PI / (1 - nu) *
               (y / r3 + sinA * sq(cosA) / W2 - cosA * (y * cosA + z * sinA) / W2r + y * z * cosA / W2r2 -
                y * z * cosA / Wr3) -
           bz * x * sinA / 8 / PI / (1 - nu) *
               (y * z / Wr3 - sinA * cosA / W2 + (y * cosA + z * sinA) / W2r - y * z / W2r2);
}

/* AngDisStrainFSC (:654-1034): harmonic free-surface correction strains of an
 * angular dislocation pair in a half-space. */
static void ang_dis_strain_fsc(double y1, double y2, double y3, double beta, double b1, double b2, double b3,
                               double nu, double a, double v[6]) {
    const double sinB = sin(beta), cosB = cos(beta), cotB = 1.0 / tan(beta);
    const double y3b = y3 + 2 * a;
    const double z1b = y1 * cosB + y3b * sinB;
    const double z3b = -y1 * sinB + y3b * cosB;
    const double rb2 = sq(y1) + sq(y2) + sq(y3b);
    const double rb = sqrt(rb2);
    const double rb3 = cube(rb), rb5 = rb3 * rb2;
    const double W1 = rb * cosB + y3b;
    const double W2 = cosB + a / rb;
    const double W3 = cosB + y3b / rb;
    const double W4 = nu + a / rb;
    const double W5 = 2 * nu + a / rb;
    const double W6 = rb + y3b;
    const double W7 = rb + z3b;
    const double W8 = y3 + a;
    const double W9 = 1 + a / rb / cosB;
    const double N1 = 1 - 2 * nu;
    const double cot2 = sq(cotB);
    const double K = 1.0 / 4 / PI / (1 - nu); /* the common factor 1/4 ... /pi/(1-nu) */
    const double rFib_ry2 = z1b / rb / (rb + z3b) - y1 / rb / (rb + y3b);
    const double rFib_ry1 = y2 / rb / (rb + y3b) - cosB * y2 / rb / (rb + z3b);
    const double rFib_ry3 = -sinB * y2 / rb / (rb + z3b);

    /* v11 */
    v[0] =
        b1 * K *
            ((-2 + 2 * nu) * N1 * rFib_ry1 * cot2 - N1 * y2 / sq(W6) * ((1 - W5) * cotB - y1 / W6 * W4) / rb * y1 +
             N1 * y2 / W6 * (a / rb3 * y1 * cotB - 1 / W6 * W4 + sq(y1) / sq(W6) * W4 / rb + sq(y1) / W6 * a / rb3) -
             N1 * y2 * cosB * cotB / sq(W7) * W2 * (y1 / rb - sinB) - N1 * y2 * cosB * cotB / W7 * a / rb3 * y1 -
             3 * a * y2 * W8 * cotB / rb5 * y1 -
             y2 * W8 / rb3 / W6 * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) * y1 -
             y2 * W8 / rb2 / sq(W6) * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) * y1 +
             y2 * W8 / rb / W6 *
                 (1 / W6 * W5 - sq(y1) / sq(W6) * W5 / rb - sq(y1) / W6 * a / rb3 + a / rb2 -
                  2 * a * sq(y1) / sq(rb2)) -
             y2 * W8 / rb3 / W7 *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) *
                 y1 -
             y2 * W8 / rb / sq(W7) *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) *
                 (y1 / rb - sinB) +
             y2 * W8 / rb / W7 *
                 (-cosB / sq(W7) * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) *
                      (y1 / rb - sinB) +
                  cosB / W7 *
                      (1 / rb * cosB * y1 * (N1 * cosB - a / rb) * cotB + W1 * a / rb3 * y1 * cotB +
                       (2 - 2 * nu) * (1 / rb * sinB * y1 - 1) * cosB) +
                  2 * a * y3b * cosB * cotB / sq(rb2) * y1)) +
        b2 * K *
            (N1 * (((2 - 2 * nu) * cot2 + nu) / rb * y1 / W6 -
                   ((2 - 2 * nu) * cot2 + 1) * cosB * (y1 / rb - sinB) / W7) -
             N1 / sq(W6) * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + sq(y1) / W6 * W4) / rb * y1 +
             N1 / W6 *
                 (-N1 * cotB + a * cotB / rb - a * sq(y1) * cotB / rb3 + 2 * y1 / W6 * W4 -
                  cube(y1) / sq(W6) * W4 / rb - cube(y1) / W6 * a / rb3) +
             N1 * cotB / sq(W7) * (z1b * cosB - a * (rb * sinB - y1) / rb / cosB) * (y1 / rb - sinB) -
             N1 * cotB / W7 *
                 (sq(cosB) - a * (1 / rb * sinB * y1 - 1) / rb / cosB + a * (rb * sinB - y1) / rb3 / cosB * y1) -
             a * W8 * cotB / rb3 + 3 * a * sq(y1) * W8 * cotB / rb5 -
             W8 / sq(W6) * (2 * nu + 1 / rb * (N1 * y1 * cotB + a) - sq(y1) / rb / W6 * W5 - a * sq(y1) / rb3) / rb *
                 y1 +
             W8 / W6 *
                 (-1 / rb3 * (N1 * y1 * cotB + a) * y1 + 1 / rb * N1 * cotB - 2 * y1 / rb / W6 * W5 +
                  cube(y1) / rb3 / W6 * W5 + cube(y1) / rb2 / sq(W6) * W5 + cube(y1) / sq(rb2) / W6 * a -
                  2 * a / rb3 * y1 + 3 * a * cube(y1) / rb5) -
             W8 * cotB / sq(W7) *
                 (-cosB * sinB + a * y1 * y3b / rb3 / cosB + (rb * sinB - y1) / rb * ((2 - 2 * nu) * cosB - W1 / W7 * W9)) *
                 (y1 / rb - sinB) +
             W8 * cotB / W7 *
                 (a * y3b / rb3 / cosB - 3 * a * sq(y1) * y3b / rb5 / cosB +
                  (1 / rb * sinB * y1 - 1) / rb * ((2 - 2 * nu) * cosB - W1 / W7 * W9) -
                  (rb * sinB - y1) / rb3 * ((2 - 2 * nu) * cosB - W1 / W7 * W9) * y1 +
                  (rb * sinB - y1) / rb *
                      (-1 / rb * cosB * y1 / W7 * W9 + W1 / sq(W7) * W9 * (y1 / rb - sinB) + W1 / W7 * a / rb3 / cosB * y1))) +
        b3 * K *
            (N1 * (-y2 / sq(W6) * (1 + a / rb) / rb * y1 - y2 / W6 * a / rb3 * y1 + y2 * cosB / sq(W7) * W2 * (y1 / rb - sinB) +
                   y2 * cosB / W7 * a / rb3 * y1) +
             y2 * W8 / rb3 * (a / rb2 + 1 / W6) * y1 - y2 * W8 / rb * (-2 * a / sq(rb2) * y1 - 1 / sq(W6) / rb * y1) -
             y2 * W8 * cosB / rb3 / W7 * (W1 / W7 * W2 + a * y3b / rb2) * y1 -
             y2 * W8 * cosB / rb / sq(W7) * (W1 / W7 * W2 + a * y3b / rb2) * (y1 / rb - sinB) +
             y2 * W8 * cosB / rb / W7 *
                 (1 / rb * cosB * y1 / W7 * W2 - W1 / sq(W7) * W2 * (y1 / rb - sinB) - W1 / W7 * a / rb3 * y1 -
                  2 * a * y3b / sq(rb2) * y1));

    /* v22 */
    v[1] =
        b1 * K *
            (N1 * (((2 - 2 * nu) * cot2 - nu) / rb * y2 / W6 - ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB / rb * y2 / W7) +
             N1 / sq(W6) * (y1 * cotB * (1 - W5) + nu * y3b - a + sq(y2) / W6 * W4) / rb * y2 -
             N1 / W6 * (a * y1 * cotB / rb3 * y2 + 2 * y2 / W6 * W4 - cube(y2) / sq(W6) * W4 / rb - cube(y2) / W6 * a / rb3) +
             N1 * z1b * cotB / sq(W7) * W2 / rb * y2 + N1 * z1b * cotB / W7 * a / rb3 * y2 +
             3 * a * y2 * W8 * cotB / rb5 * y1 -
             W8 / sq(W6) * (-2 * nu + 1 / rb * (N1 * y1 * cotB - a) + sq(y2) / rb / W6 * W5 + a * sq(y2) / rb3) / rb * y2 +
             W8 / W6 *
                 (-1 / rb3 * (N1 * y1 * cotB - a) * y2 + 2 * y2 / rb / W6 * W5 - cube(y2) / rb3 / W6 * W5 -
                  cube(y2) / rb2 / sq(W6) * W5 - cube(y2) / sq(rb2) / W6 * a + 2 * a / rb3 * y2 - 3 * a * cube(y2) / rb5) -
             W8 / sq(W7) *
                 (sq(cosB) - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / rb3 -
                  1 / rb / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1)) /
                 rb * y2 +
             W8 / W7 *
                 (1 / rb3 * (N1 * z1b * cotB + a * cosB) * y2 - 3 * a * y3b * z1b * cotB / rb5 * y2 +
                  1 / rb3 / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * y2 +
                  1 / rb2 / sq(W7) * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * y2 -
                  1 / rb / W7 * (2 * y2 * sq(cosB) + a * z1b * cotB / rb3 * W1 * y2 - a * z1b * cotB / rb2 * cosB * y2))) +
        b2 * K *
            ((2 - 2 * nu) * N1 * rFib_ry2 * cot2 + N1 / W6 * ((W5 - 1) * cotB + y1 / W6 * W4) -
             N1 * sq(y2) / sq(W6) * ((W5 - 1) * cotB + y1 / W6 * W4) / rb +
             N1 * y2 / W6 * (-a / rb3 * y2 * cotB - y1 / sq(W6) * W4 / rb * y2 - y2 / W6 * a / rb3 * y1) -
             N1 * cotB / W7 * W9 + N1 * sq(y2) * cotB / sq(W7) * W9 / rb + N1 * sq(y2) * cotB / W7 * a / rb3 / cosB -
             a * W8 * cotB / rb3 + 3 * a * sq(y2) * W8 * cotB / rb5 +
             W8 / rb / W6 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) -
             sq(y2) * W8 / rb3 / W6 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) -
             sq(y2) * W8 / rb2 / sq(W6) * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) +
             y2 * W8 / rb / W6 *
                 (2 * nu * y1 / sq(W6) / rb * y2 + a * y1 / rb3 * (1 / rb + 1 / W6) * y2 -
                  a * y1 / rb * (-1 / rb3 * y2 - 1 / sq(W6) / rb * y2)) +
             W8 * cotB / rb / W7 * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) -
             sq(y2) * W8 * cotB / rb3 / W7 * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) -
             sq(y2) * W8 * cotB / rb2 / sq(W7) * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) +
             y2 * W8 * cotB / rb / W7 *
                 (1 / rb * cosB * y2 / W7 * W9 - W1 / sq(W7) * W9 / rb * y2 - W1 / W7 * a / rb3 / cosB * y2 -
                  2 * a * y3b / sq(rb2) / cosB * y2)) +
        b3 * K *
            (N1 * (-sinB / rb * y2 / W7 + y2 / sq(W6) * (1 + a / rb) / rb * y1 + y2 / W6 * a / rb3 * y1 -
                   z1b / sq(W7) * W2 / rb * y2 - z1b / W7 * a / rb3 * y2) -
             y2 * W8 / rb3 * (a / rb2 + 1 / W6) * y1 + y1 * W8 / rb * (-2 * a / sq(rb2) * y2 - 1 / sq(W6) / rb * y2) +
             W8 / sq(W7) *
                 (sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / rb2) -
                  1 / rb / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1)) /
                 rb * y2 -
             W8 / W7 *
                 (sinB * a / rb3 * y2 - z1b / rb3 * (1 + a * y3b / rb2) * y2 - 2 * z1b / rb5 * a * y3b * y2 +
                  1 / rb3 / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * y2 +
                  1 / rb2 / sq(W7) * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * y2 -
                  1 / rb / W7 * (2 * y2 * cosB * sinB + a * z1b / rb3 * W1 * y2 - a * z1b / rb2 * cosB * y2)));

    /* v33 */
    v[2] =
        b1 * K *
            ((2 - 2 * nu) *
                 (N1 * rFib_ry3 * cotB - y2 / sq(W6) * W5 * (y3b / rb + 1) - 0.5 * y2 / W6 * a / rb3 * 2 * y3b +
                  y2 * cosB / sq(W7) * W2 * W3 + 0.5 * y2 * cosB / W7 * a / rb3 * 2 * y3b) +
             y2 / rb * (2 * nu / W6 + a / rb2) - 0.5 * y2 * W8 / rb3 * (2 * nu / W6 + a / rb2) * 2 * y3b +
             y2 * W8 / rb * (-2 * nu / sq(W6) * (y3b / rb + 1) - a / sq(rb2) * 2 * y3b) +
             y2 * cosB / rb / W7 * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) -
             0.5 * y2 * W8 * cosB / rb3 / W7 * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) * 2 * y3b -
             y2 * W8 * cosB / rb / sq(W7) * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) * W3 +
             y2 * W8 * cosB / rb / W7 *
                 (-(cosB * y3b / rb + 1) / W7 * W2 + W1 / sq(W7) * W2 * W3 + 0.5 * W1 / W7 * a / rb3 * 2 * y3b -
                  a / rb2 + a * y3b / sq(rb2) * 2 * y3b)) +
        b2 * K *
            ((-2 + 2 * nu) * N1 * cotB * ((y3b / rb + 1) / W6 - cosB * W3 / W7) +
             (2 - 2 * nu) * y1 / sq(W6) * W5 * (y3b / rb + 1) + 0.5 * (2 - 2 * nu) * y1 / W6 * a / rb3 * 2 * y3b +
             (2 - 2 * nu) * sinB / W7 * W2 - (2 - 2 * nu) * z1b / sq(W7) * W2 * W3 -
             0.5 * (2 - 2 * nu) * z1b / W7 * a / rb3 * 2 * y3b + 1 / rb * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb2) -
             0.5 * W8 / rb3 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb2) * 2 * y3b +
             W8 / rb * (2 * nu * y1 / sq(W6) * (y3b / rb + 1) + a * y1 / sq(rb2) * 2 * y3b) -
             1 / W7 *
                 (cosB * sinB + W1 * cotB / rb * ((2 - 2 * nu) * cosB - W1 / W7) +
                  a / rb * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7)) +
             W8 / sq(W7) *
                 (cosB * sinB + W1 * cotB / rb * ((2 - 2 * nu) * cosB - W1 / W7) +
                  a / rb * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7)) *
                 W3 -
             W8 / W7 *
                 ((cosB * y3b / rb + 1) * cotB / rb * ((2 - 2 * nu) * cosB - W1 / W7) -
                  0.5 * W1 * cotB / rb3 * ((2 - 2 * nu) * cosB - W1 / W7) * 2 * y3b +
                  W1 * cotB / rb * (-(cosB * y3b / rb + 1) / W7 + W1 / sq(W7) * W3) -
                  0.5 * a / rb3 * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7) * 2 * y3b +
                  a / rb *
                      (-z1b / rb2 - y3b * sinB / rb2 + y3b * z1b / sq(rb2) * 2 * y3b - sinB * W1 / rb / W7 -
                       z1b * (cosB * y3b / rb + 1) / rb / W7 + 0.5 * z1b * W1 / rb3 / W7 * 2 * y3b +
                       z1b * W1 / rb / sq(W7) * W3))) +
        b3 * K *
            ((2 - 2 * nu) * rFib_ry3 - (2 - 2 * nu) * y2 * sinB / sq(W7) * W2 * W3 -
             0.5 * (2 - 2 * nu) * y2 * sinB / W7 * a / rb3 * 2 * y3b +
             y2 * sinB / rb / W7 * (1 + W1 / W7 * W2 + a * y3b / rb2) -
             0.5 * y2 * W8 * sinB / rb3 / W7 * (1 + W1 / W7 * W2 + a * y3b / rb2) * 2 * y3b -
             y2 * W8 * sinB / rb / sq(W7) * (1 + W1 / W7 * W2 + a * y3b / rb2) * W3 +
             y2 * W8 * sinB / rb / W7 *
                 ((cosB * y3b / rb + 1) / W7 * W2 - W1 / sq(W7) * W2 * W3 - 0.5 * W1 / W7 * a / rb3 * 2 * y3b +
                  a / rb2 - a * y3b / sq(rb2) * 2 * y3b));

    /* v12 */
    v[3] =
        b1 / 2 * K *
            ((-2 + 2 * nu) * N1 * rFib_ry2 * cot2 + N1 / W6 * ((1 - W5) * cotB - y1 / W6 * W4) -
             N1 * sq(y2) / sq(W6) * ((1 - W5) * cotB - y1 / W6 * W4) / rb +
             N1 * y2 / W6 * (a / rb3 * y2 * cotB + y1 / sq(W6) * W4 / rb * y2 + y2 / W6 * a / rb3 * y1) +
             N1 * cosB * cotB / W7 * W2 - N1 * sq(y2) * cosB * cotB / sq(W7) * W2 / rb -
             N1 * sq(y2) * cosB * cotB / W7 * a / rb3 + a * W8 * cotB / rb3 - 3 * a * sq(y2) * W8 * cotB / rb5 +
             W8 / rb / W6 * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) -
             sq(y2) * W8 / rb3 / W6 * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) -
             sq(y2) * W8 / rb2 / sq(W6) * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) +
             y2 * W8 / rb / W6 * (-y1 / sq(W6) * W5 / rb * y2 - y2 / W6 * a / rb3 * y1 - 2 * a * y1 / sq(rb2) * y2) +
             W8 / rb / W7 *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) -
             sq(y2) * W8 / rb3 / W7 *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) -
             sq(y2) * W8 / rb2 / sq(W7) *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) +
             y2 * W8 / rb / W7 *
                 (-cosB / sq(W7) * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) / rb * y2 +
                  cosB / W7 *
                      (1 / rb * cosB * y2 * (N1 * cosB - a / rb) * cotB + W1 * a / rb3 * y2 * cotB +
                       (2 - 2 * nu) / rb * sinB * y2 * cosB) +
                  2 * a * y3b * cosB * cotB / sq(rb2) * y2)) +
        b2 / 2 * K *
            (N1 * (((2 - 2 * nu) * cot2 + nu) / rb * y2 / W6 - ((2 - 2 * nu) * cot2 + 1) * cosB / rb * y2 / W7) -
             N1 / sq(W6) * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + sq(y1) / W6 * W4) / rb * y2 +
             N1 / W6 * (-a * y1 * cotB / rb3 * y2 - sq(y1) / sq(W6) * W4 / rb * y2 - sq(y1) / W6 * a / rb3 * y2) +
             N1 * cotB / sq(W7) * (z1b * cosB - a * (rb * sinB - y1) / rb / cosB) / rb * y2 -
             N1 * cotB / W7 * (-a / rb2 * sinB * y2 / cosB + a * (rb * sinB - y1) / rb3 / cosB * y2) +
             3 * a * y2 * W8 * cotB / rb5 * y1 -
             W8 / sq(W6) * (2 * nu + 1 / rb * (N1 * y1 * cotB + a) - sq(y1) / rb / W6 * W5 - a * sq(y1) / rb3) / rb * y2 +
             W8 / W6 *
                 (-1 / rb3 * (N1 * y1 * cotB + a) * y2 + sq(y1) / rb3 / W6 * W5 * y2 + sq(y1) / rb2 / sq(W6) * W5 * y2 +
                  sq(y1) / sq(rb2) / W6 * a * y2 + 3 * a * sq(y1) / rb5 * y2) -
             W8 * cotB / sq(W7) *
                 (-cosB * sinB + a * y1 * y3b / rb3 / cosB + (rb * sinB - y1) / rb * ((2 - 2 * nu) * cosB - W1 / W7 * W9)) /
                 rb * y2 +
             W8 * cotB / W7 *
                 (-3 * a * y1 * y3b / rb5 / cosB * y2 + 1 / rb2 * sinB * y2 * ((2 - 2 * nu) * cosB - W1 / W7 * W9) -
                  (rb * sinB - y1) / rb3 * ((2 - 2 * nu) * cosB - W1 / W7 * W9) * y2 +
                  (rb * sinB - y1) / rb *
                      (-1 / rb * cosB * y2 / W7 * W9 + W1 / sq(W7) * W9 / rb * y2 + W1 / W7 * a / rb3 / cosB * y2))) +
        b3 / 2 * K *
            (N1 * (1 / W6 * (1 + a / rb) - sq(y2) / sq(W6) * (1 + a / rb) / rb - sq(y2) / W6 * a / rb3 - cosB / W7 * W2 +
                   sq(y2) * cosB / sq(W7) * W2 / rb + sq(y2) * cosB / W7 * a / rb3) -
             W8 / rb * (a / rb2 + 1 / W6) + sq(y2) * W8 / rb3 * (a / rb2 + 1 / W6) -
             y2 * W8 / rb * (-2 * a / sq(rb2) * y2 - 1 / sq(W6) / rb * y2) +
             W8 * cosB / rb / W7 * (W1 / W7 * W2 + a * y3b / rb2) -
             sq(y2) * W8 * cosB / rb3 / W7 * (W1 / W7 * W2 + a * y3b / rb2) -
             sq(y2) * W8 * cosB / rb2 / sq(W7) * (W1 / W7 * W2 + a * y3b / rb2) +
             y2 * W8 * cosB / rb / W7 *
                 (1 / rb * cosB * y2 / W7 * W2 - W1 / sq(W7) * W2 / rb * y2 - W1 / W7 * a / rb3 * y2 -
                  2 * a * y3b / sq(rb2) * y2)) +
        b1 / 2 * K *
            (N1 * (((2 - 2 * nu) * cot2 - nu) / rb * y1 / W6 -
                   ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB * (y1 / rb - sinB) / W7) +
             N1 / sq(W6) * (y1 * cotB * (1 - W5) + nu * y3b - a + sq(y2) / W6 * W4) / rb * y1 -
             N1 / W6 * ((1 - W5) * cotB + a * sq(y1) * cotB / rb3 - sq(y2) / sq(W6) * W4 / rb * y1 - sq(y2) / W6 * a / rb3 * y1) -
             N1 * cosB * cotB / W7 * W2 + N1 * z1b * cotB / sq(W7) * W2 * (y1 / rb - sinB) +
             N1 * z1b * cotB / W7 * a / rb3 * y1 - a * W8 * cotB / rb3 + 3 * a * sq(y1) * W8 * cotB / rb5 -
             W8 / sq(W6) * (-2 * nu + 1 / rb * (N1 * y1 * cotB - a) + sq(y2) / rb / W6 * W5 + a * sq(y2) / rb3) / rb * y1 +
             W8 / W6 *
                 (-1 / rb3 * (N1 * y1 * cotB - a) * y1 + 1 / rb * N1 * cotB - sq(y2) / rb3 / W6 * W5 * y1 -
                  sq(y2) / rb2 / sq(W6) * W5 * y1 - sq(y2) / sq(rb2) / W6 * a * y1 - 3 * a * sq(y2) / rb5 * y1) -
             W8 / sq(W7) *
                 (sq(cosB) - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / rb3 -
                  1 / rb / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1)) *
                 (y1 / rb - sinB) +
             W8 / W7 *
                 (1 / rb3 * (N1 * z1b * cotB + a * cosB) * y1 - 1 / rb * N1 * cosB * cotB + a * y3b * cosB * cotB / rb3 -
                  3 * a * y3b * z1b * cotB / rb5 * y1 + 1 / rb3 / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * y1 +
                  1 / rb / sq(W7) * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * (y1 / rb - sinB) -
                  1 / rb / W7 * (-a * cosB * cotB / rb * W1 + a * z1b * cotB / rb3 * W1 * y1 - a * z1b * cotB / rb2 * cosB * y1))) +
        b2 / 2 * K *
            ((2 - 2 * nu) * N1 * rFib_ry1 * cot2 - N1 * y2 / sq(W6) * ((W5 - 1) * cotB + y1 / W6 * W4) / rb * y1 +
             N1 * y2 / W6 * (-a / rb3 * y1 * cotB + 1 / W6 * W4 - sq(y1) / sq(W6) * W4 / rb - sq(y1) / W6 * a / rb3) +
             N1 * y2 * cotB / sq(W7) * W9 * (y1 / rb - sinB) + N1 * y2 * cotB / W7 * a / rb3 / cosB * y1 +
             3 * a * y2 * W8 * cotB / rb5 * y1 -
             y2 * W8 / rb3 / W6 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) * y1 -
             y2 * W8 / rb2 / sq(W6) * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) * y1 +
             y2 * W8 / rb / W6 *
                 (-2 * nu / W6 + 2 * nu * sq(y1) / sq(W6) / rb - a / rb * (1 / rb + 1 / W6) +
                  a * sq(y1) / rb3 * (1 / rb + 1 / W6) - a * y1 / rb * (-1 / rb3 * y1 - 1 / sq(W6) / rb * y1)) -
             y2 * W8 * cotB / rb3 / W7 * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) * y1 -
             y2 * W8 * cotB / rb / sq(W7) * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) * (y1 / rb - sinB) +
             y2 * W8 * cotB / rb / W7 *
                 (1 / rb * cosB * y1 / W7 * W9 - W1 / sq(W7) * W9 * (y1 / rb - sinB) - W1 / W7 * a / rb3 / cosB * y1 -
                  2 * a * y3b / sq(rb2) / cosB * y1)) +
        b3 / 2 * K *
            (N1 * (-sinB * (y1 / rb - sinB) / W7 - 1 / W6 * (1 + a / rb) + sq(y1) / sq(W6) * (1 + a / rb) / rb +
                   sq(y1) / W6 * a / rb3 + cosB / W7 * W2 - z1b / sq(W7) * W2 * (y1 / rb - sinB) - z1b / W7 * a / rb3 * y1) +
             W8 / rb * (a / rb2 + 1 / W6) - sq(y1) * W8 / rb3 * (a / rb2 + 1 / W6) +
             y1 * W8 / rb * (-2 * a / sq(rb2) * y1 - 1 / sq(W6) / rb * y1) +
             W8 / sq(W7) *
                 (sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / rb2) -
                  1 / rb / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1)) *
                 (y1 / rb - sinB) -
             W8 / W7 *
                 (sinB * a / rb3 * y1 + cosB / rb * (1 + a * y3b / rb2) - z1b / rb3 * (1 + a * y3b / rb2) * y1 -
                  2 * z1b / rb5 * a * y3b * y1 + 1 / rb3 / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * y1 +
                  1 / rb / sq(W7) * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * (y1 / rb - sinB) -
                  1 / rb / W7 * (-a * cosB / rb * W1 + a * z1b / rb3 * W1 * y1 - a * z1b / rb2 * cosB * y1)));

    /* v13 */
    v[4] =
        b1 / 2 * K *
            ((-2 + 2 * nu) * N1 * rFib_ry3 * cot2 - N1 * y2 / sq(W6) * ((1 - W5) * cotB - y1 / W6 * W4) * (y3b / rb + 1) +
             N1 * y2 / W6 * (0.5 * a / rb3 * 2 * y3b * cotB + y1 / sq(W6) * W4 * (y3b / rb + 1) + 0.5 * y1 / W6 * a / rb3 * 2 * y3b) -
             N1 * y2 * cosB * cotB / sq(W7) * W2 * W3 - 0.5 * N1 * y2 * cosB * cotB / W7 * a / rb3 * 2 * y3b +
             a / rb3 * y2 * cotB - 1.5 * a * y2 * W8 * cotB / rb5 * 2 * y3b +
             y2 / rb / W6 * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) -
             0.5 * y2 * W8 / rb3 / W6 * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) * 2 * y3b -
             y2 * W8 / rb / sq(W6) * (-N1 * cotB + y1 / W6 * W5 + a * y1 / rb2) * (y3b / rb + 1) +
             y2 * W8 / rb / W6 * (-y1 / sq(W6) * W5 * (y3b / rb + 1) - 0.5 * y1 / W6 * a / rb3 * 2 * y3b - a * y1 / sq(rb2) * 2 * y3b) +
             y2 / rb / W7 *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) -
             0.5 * y2 * W8 / rb3 / W7 *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) *
                 2 * y3b -
             y2 * W8 / rb / sq(W7) *
                 (cosB / W7 * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) -
                  a * y3b * cosB * cotB / rb2) *
                 W3 +
             y2 * W8 / rb / W7 *
                 (-cosB / sq(W7) * (W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB) * W3 +
                  cosB / W7 *
                      ((cosB * y3b / rb + 1) * (N1 * cosB - a / rb) * cotB + 0.5 * W1 * a / rb3 * 2 * y3b * cotB +
                       0.5 * (2 - 2 * nu) / rb * sinB * 2 * y3b * cosB) -
                  a * cosB * cotB / rb2 + a * y3b * cosB * cotB / sq(rb2) * 2 * y3b)) +
        b2 / 2 * K *
            (N1 * (((2 - 2 * nu) * cot2 + nu) * (y3b / rb + 1) / W6 - ((2 - 2 * nu) * cot2 + 1) * cosB * W3 / W7) -
             N1 / sq(W6) * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + sq(y1) / W6 * W4) * (y3b / rb + 1) +
             N1 / W6 * (nu - 0.5 * a * y1 * cotB / rb3 * 2 * y3b - sq(y1) / sq(W6) * W4 * (y3b / rb + 1) -
                        0.5 * sq(y1) / W6 * a / rb3 * 2 * y3b) +
             N1 * cotB / sq(W7) * (z1b * cosB - a * (rb * sinB - y1) / rb / cosB) * W3 -
             N1 * cotB / W7 * (cosB * sinB - 0.5 * a / rb2 * sinB * 2 * y3b / cosB + 0.5 * a * (rb * sinB - y1) / rb3 / cosB * 2 * y3b) -
             a / rb3 * y1 * cotB + 1.5 * a * y1 * W8 * cotB / rb5 * 2 * y3b +
             1 / W6 * (2 * nu + 1 / rb * (N1 * y1 * cotB + a) - sq(y1) / rb / W6 * W5 - a * sq(y1) / rb3) -
             W8 / sq(W6) * (2 * nu + 1 / rb * (N1 * y1 * cotB + a) - sq(y1) / rb / W6 * W5 - a * sq(y1) / rb3) * (y3b / rb + 1) +
             W8 / W6 *
                 (-0.5 / rb3 * (N1 * y1 * cotB + a) * 2 * y3b + 0.5 * sq(y1) / rb3 / W6 * W5 * 2 * y3b +
                  sq(y1) / rb / sq(W6) * W5 * (y3b / rb + 1) + 0.5 * sq(y1) / sq(rb2) / W6 * a * 2 * y3b +
                  1.5 * a * sq(y1) / rb5 * 2 * y3b) +
             cotB / W7 * (-cosB * sinB + a * y1 * y3b / rb3 / cosB + (rb * sinB - y1) / rb * ((2 - 2 * nu) * cosB - W1 / W7 * W9)) -
             W8 * cotB / sq(W7) *
                 (-cosB * sinB + a * y1 * y3b / rb3 / cosB + (rb * sinB - y1) / rb * ((2 - 2 * nu) * cosB - W1 / W7 * W9)) * W3 +
             W8 * cotB / W7 *
                 (a / rb3 / cosB * y1 - 1.5 * a * y1 * y3b / rb5 / cosB * 2 * y3b +
                  0.5 / rb2 * sinB * 2 * y3b * ((2 - 2 * nu) * cosB - W1 / W7 * W9) -
                  0.5 * (rb * sinB - y1) / rb3 * ((2 - 2 * nu) * cosB - W1 / W7 * W9) * 2 * y3b +
                  (rb * sinB - y1) / rb *
                      (-(cosB * y3b / rb + 1) / W7 * W9 + W1 / sq(W7) * W9 * W3 + 0.5 * W1 / W7 * a / rb3 / cosB * 2 * y3b))) +
        b3 / 2 * K *
            (N1 * (-y2 / sq(W6) * (1 + a / rb) * (y3b / rb + 1) - 0.5 * y2 / W6 * a / rb3 * 2 * y3b +
                   y2 * cosB / sq(W7) * W2 * W3 + 0.5 * y2 * cosB / W7 * a / rb3 * 2 * y3b) -
             y2 / rb * (a / rb2 + 1 / W6) + 0.5 * y2 * W8 / rb3 * (a / rb2 + 1 / W6) * 2 * y3b -
             y2 * W8 / rb * (-a / sq(rb2) * 2 * y3b - 1 / sq(W6) * (y3b / rb + 1)) +
             y2 * cosB / rb / W7 * (W1 / W7 * W2 + a * y3b / rb2) -
             0.5 * y2 * W8 * cosB / rb3 / W7 * (W1 / W7 * W2 + a * y3b / rb2) * 2 * y3b -
             y2 * W8 * cosB / rb / sq(W7) * (W1 / W7 * W2 + a * y3b / rb2) * W3 +
             y2 * W8 * cosB / rb / W7 *
                 ((cosB * y3b / rb + 1) / W7 * W2 - W1 / sq(W7) * W2 * W3 - 0.5 * W1 / W7 * a / rb3 * 2 * y3b + a / rb2 -
                  a * y3b / sq(rb2) * 2 * y3b)) +
        b1 / 2 * K *
            ((2 - 2 * nu) *
                 (N1 * rFib_ry1 * cotB - y1 / sq(W6) * W5 / rb * y2 - y2 / W6 * a / rb3 * y1 +
                  y2 * cosB / sq(W7) * W2 * (y1 / rb - sinB) + y2 * cosB / W7 * a / rb3 * y1) -
             y2 * W8 / rb3 * (2 * nu / W6 + a / rb2) * y1 + y2 * W8 / rb * (-2 * nu / sq(W6) / rb * y1 - 2 * a / sq(rb2) * y1) -
             y2 * W8 * cosB / rb3 / W7 * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) * y1 -
             y2 * W8 * cosB / rb / sq(W7) * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) * (y1 / rb - sinB) +
             y2 * W8 * cosB / rb / W7 *
                 (-1 / rb * cosB * y1 / W7 * W2 + W1 / sq(W7) * W2 * (y1 / rb - sinB) + W1 / W7 * a / rb3 * y1 +
                  2 * a * y3b / sq(rb2) * y1)) +
        b2 / 2 * K *
            ((-2 + 2 * nu) * N1 * cotB * (1 / rb * y1 / W6 - cosB * (y1 / rb - sinB) / W7) - (2 - 2 * nu) / W6 * W5 +
             (2 - 2 * nu) * sq(y1) / sq(W6) * W5 / rb + (2 - 2 * nu) * sq(y1) / W6 * a / rb3 +
             (2 - 2 * nu) * cosB / W7 * W2 - (2 - 2 * nu) * z1b / sq(W7) * W2 * (y1 / rb - sinB) -
             (2 - 2 * nu) * z1b / W7 * a / rb3 * y1 - W8 / rb3 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb2) * y1 +
             W8 / rb * (-2 * nu / W6 + 2 * nu * sq(y1) / sq(W6) / rb - a / rb2 + 2 * a * sq(y1) / sq(rb2)) +
             W8 / sq(W7) *
                 (cosB * sinB + W1 * cotB / rb * ((2 - 2 * nu) * cosB - W1 / W7) +
                  a / rb * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7)) *
                 (y1 / rb - sinB) -
             W8 / W7 *
                 (1 / rb2 * cosB * y1 * cotB * ((2 - 2 * nu) * cosB - W1 / W7) -
                  W1 * cotB / rb3 * ((2 - 2 * nu) * cosB - W1 / W7) * y1 +
                  W1 * cotB / rb * (-1 / rb * cosB * y1 / W7 + W1 / sq(W7) * (y1 / rb - sinB)) -
                  a / rb3 * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7) * y1 +
                  a / rb *
                      (-y3b * cosB / rb2 + 2 * y3b * z1b / sq(rb2) * y1 - cosB * W1 / rb / W7 - z1b / rb2 * cosB * y1 / W7 +
                       z1b * W1 / rb3 / W7 * y1 + z1b * W1 / rb / sq(W7) * (y1 / rb - sinB)))) +
        b3 / 2 * K *
            ((2 - 2 * nu) * rFib_ry1 - (2 - 2 * nu) * y2 * sinB / sq(W7) * W2 * (y1 / rb - sinB) -
             (2 - 2 * nu) * y2 * sinB / W7 * a / rb3 * y1 -
             y2 * W8 * sinB / rb3 / W7 * (1 + W1 / W7 * W2 + a * y3b / rb2) * y1 -
             y2 * W8 * sinB / rb / sq(W7) * (1 + W1 / W7 * W2 + a * y3b / rb2) * (y1 / rb - sinB) +
             y2 * W8 * sinB / rb / W7 *
                 (1 / rb * cosB * y1 / W7 * W2 - W1 / sq(W7) * W2 * (y1 / rb - sinB) - W1 / W7 * a / rb3 * y1 -
                  2 * a * y3b / sq(rb2) * y1));

    /* v23 */
    v[5] =
        b1 / 2 * K *
            (N1 * (((2 - 2 * nu) * cot2 - nu) * (y3b / rb + 1) / W6 - ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB * W3 / W7) +
             N1 / sq(W6) * (y1 * cotB * (1 - W5) + nu * y3b - a + sq(y2) / W6 * W4) * (y3b / rb + 1) -
             N1 / W6 * (0.5 * a * y1 * cotB / rb3 * 2 * y3b + nu - sq(y2) / sq(W6) * W4 * (y3b / rb + 1) -
                        0.5 * sq(y2) / W6 * a / rb3 * 2 * y3b) -
             N1 * sinB * cotB / W7 * W2 + N1 * z1b * cotB / sq(W7) * W2 * W3 + 0.5 * N1 * z1b * cotB / W7 * a / rb3 * 2 * y3b -
             a / rb3 * y1 * cotB + 1.5 * a * y1 * W8 * cotB / rb5 * 2 * y3b +
             1 / W6 * (-2 * nu + 1 / rb * (N1 * y1 * cotB - a) + sq(y2) / rb / W6 * W5 + a * sq(y2) / rb3) -
             W8 / sq(W6) * (-2 * nu + 1 / rb * (N1 * y1 * cotB - a) + sq(y2) / rb / W6 * W5 + a * sq(y2) / rb3) * (y3b / rb + 1) +
             W8 / W6 *
                 (-0.5 / rb3 * (N1 * y1 * cotB - a) * 2 * y3b - 0.5 * sq(y2) / rb3 / W6 * W5 * 2 * y3b -
                  sq(y2) / rb / sq(W6) * W5 * (y3b / rb + 1) - 0.5 * sq(y2) / sq(rb2) / W6 * a * 2 * y3b -
                  1.5 * a * sq(y2) / rb5 * 2 * y3b) +
             1 / W7 *
                 (sq(cosB) - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / rb3 -
                  1 / rb / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1)) -
             W8 / sq(W7) *
                 (sq(cosB) - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / rb3 -
                  1 / rb / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1)) *
                 W3 +
             W8 / W7 *
                 (0.5 / rb3 * (N1 * z1b * cotB + a * cosB) * 2 * y3b - 1 / rb * N1 * sinB * cotB + a * z1b * cotB / rb3 +
                  a * y3b * sinB * cotB / rb3 - 1.5 * a * y3b * z1b * cotB / rb5 * 2 * y3b +
                  0.5 / rb3 / W7 * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * 2 * y3b +
                  1 / rb / sq(W7) * (sq(y2) * sq(cosB) - a * z1b * cotB / rb * W1) * W3 -
                  1 / rb / W7 *
                      (-a * sinB * cotB / rb * W1 + 0.5 * a * z1b * cotB / rb3 * W1 * 2 * y3b -
                       a * z1b * cotB / rb * (cosB * y3b / rb + 1)))) +
        b2 / 2 * K *
            ((2 - 2 * nu) * N1 * rFib_ry3 * cot2 - N1 * y2 / sq(W6) * ((W5 - 1) * cotB + y1 / W6 * W4) * (y3b / rb + 1) +
             N1 * y2 / W6 * (-0.5 * a / rb3 * 2 * y3b * cotB - y1 / sq(W6) * W4 * (y3b / rb + 1) - 0.5 * y1 / W6 * a / rb3 * 2 * y3b) +
             N1 * y2 * cotB / sq(W7) * W9 * W3 + 0.5 * N1 * y2 * cotB / W7 * a / rb3 / cosB * 2 * y3b -
             a / rb3 * y2 * cotB + 1.5 * a * y2 * W8 * cotB / rb5 * 2 * y3b +
             y2 / rb / W6 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) -
             0.5 * y2 * W8 / rb3 / W6 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) * 2 * y3b -
             y2 * W8 / rb / sq(W6) * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6)) * (y3b / rb + 1) +
             y2 * W8 / rb / W6 *
                 (2 * nu * y1 / sq(W6) * (y3b / rb + 1) + 0.5 * a * y1 / rb3 * (1 / rb + 1 / W6) * 2 * y3b -
                  a * y1 / rb * (-0.5 / rb3 * 2 * y3b - 1 / sq(W6) * (y3b / rb + 1))) +
             y2 * cotB / rb / W7 * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) -
             0.5 * y2 * W8 * cotB / rb3 / W7 * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) * 2 * y3b -
             y2 * W8 * cotB / rb / sq(W7) * ((-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB) * W3 +
             y2 * W8 * cotB / rb / W7 *
                 ((cosB * y3b / rb + 1) / W7 * W9 - W1 / sq(W7) * W9 * W3 - 0.5 * W1 / W7 * a / rb3 / cosB * 2 * y3b +
                  a / rb2 / cosB - a * y3b / sq(rb2) / cosB * 2 * y3b)) +
        b3 / 2 * K *
            (N1 * (-sinB * W3 / W7 + y1 / sq(W6) * (1 + a / rb) * (y3b / rb + 1) + 0.5 * y1 / W6 * a / rb3 * 2 * y3b +
                   sinB / W7 * W2 - z1b / sq(W7) * W2 * W3 - 0.5 * z1b / W7 * a / rb3 * 2 * y3b) +
             y1 / rb * (a / rb2 + 1 / W6) - 0.5 * y1 * W8 / rb3 * (a / rb2 + 1 / W6) * 2 * y3b +
             y1 * W8 / rb * (-a / sq(rb2) * 2 * y3b - 1 / sq(W6) * (y3b / rb + 1)) -
             1 / W7 *
                 (sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / rb2) -
                  1 / rb / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1)) +
             W8 / sq(W7) *
                 (sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / rb2) -
                  1 / rb / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1)) *
                 W3 -
             W8 / W7 *
                 (0.5 * sinB * a / rb3 * 2 * y3b + sinB / rb * (1 + a * y3b / rb2) -
                  0.5 * z1b / rb3 * (1 + a * y3b / rb2) * 2 * y3b + z1b / rb * (a / rb2 - a * y3b / sq(rb2) * 2 * y3b) +
                  0.5 / rb3 / W7 * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * 2 * y3b +
                  1 / rb / sq(W7) * (sq(y2) * cosB * sinB - a * z1b / rb * W1) * W3 -
                  1 / rb / W7 *
                      (-a * sinB / rb * W1 + 0.5 * a * z1b / rb3 * W1 * 2 * y3b - a * z1b / rb * (cosB * y3b / rb + 1)))) +
        b1 / 2 * K *
            ((2 - 2 * nu) *
                 (N1 * rFib_ry2 * cotB + 1 / W6 * W5 - sq(y2) / sq(W6) * W5 / rb - sq(y2) / W6 * a / rb3 - cosB / W7 * W2 +
                  sq(y2) * cosB / sq(W7) * W2 / rb + sq(y2) * cosB / W7 * a / rb3) +
             W8 / rb * (2 * nu / W6 + a / rb2) - sq(y2) * W8 / rb3 * (2 * nu / W6 + a / rb2) +
             y2 * W8 / rb * (-2 * nu / sq(W6) / rb * y2 - 2 * a / sq(rb2) * y2) +
             W8 * cosB / rb / W7 * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) -
             sq(y2) * W8 * cosB / rb3 / W7 * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) -
             sq(y2) * W8 * cosB / rb2 / sq(W7) * (1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2) +
             y2 * W8 * cosB / rb / W7 *
                 (-1 / rb * cosB * y2 / W7 * W2 + W1 / sq(W7) * W2 / rb * y2 + W1 / W7 * a / rb3 * y2 +
                  2 * a * y3b / sq(rb2) * y2)) +
        b2 / 2 * K *
            ((-2 + 2 * nu) * N1 * cotB * (1 / rb * y2 / W6 - cosB / rb * y2 / W7) + (2 - 2 * nu) * y1 / sq(W6) * W5 / rb * y2 +
             (2 - 2 * nu) * y1 / W6 * a / rb3 * y2 - (2 - 2 * nu) * z1b / sq(W7) * W2 / rb * y2 -
             (2 - 2 * nu) * z1b / W7 * a / rb3 * y2 - W8 / rb3 * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb2) * y2 +
             W8 / rb * (2 * nu * y1 / sq(W6) / rb * y2 + 2 * a * y1 / sq(rb2) * y2) +
             W8 / sq(W7) *
                 (cosB * sinB + W1 * cotB / rb * ((2 - 2 * nu) * cosB - W1 / W7) +
                  a / rb * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7)) /
                 rb * y2 -
             W8 / W7 *
                 (1 / rb2 * cosB * y2 * cotB * ((2 - 2 * nu) * cosB - W1 / W7) -
                  W1 * cotB / rb3 * ((2 - 2 * nu) * cosB - W1 / W7) * y2 +
                  W1 * cotB / rb * (-cosB / rb * y2 / W7 + W1 / sq(W7) / rb * y2) -
                  a / rb3 * (sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7) * y2 +
                  a / rb * (2 * y3b * z1b / sq(rb2) * y2 - z1b / rb2 * cosB * y2 / W7 + z1b * W1 / rb3 / W7 * y2 +
                            z1b * W1 / rb2 / sq(W7) * y2))) +
        b3 / 2 * K *
            ((2 - 2 * nu) * rFib_ry2 + (2 - 2 * nu) * sinB / W7 * W2 - (2 - 2 * nu) * sq(y2) * sinB / sq(W7) * W2 / rb -
             (2 - 2 * nu) * sq(y2) * sinB / W7 * a / rb3 + W8 * sinB / rb / W7 * (1 + W1 / W7 * W2 + a * y3b / rb2) -
             sq(y2) * W8 * sinB / rb3 / W7 * (1 + W1 / W7 * W2 + a * y3b / rb2) -
             sq(y2) * W8 * sinB / rb2 / sq(W7) * (1 + W1 / W7 * W2 + a * y3b / rb2) +
             y2 * W8 * sinB / rb / W7 *
                 (1 / rb * cosB * y2 / W7 * W2 - W1 / sq(W7) * W2 / rb * y2 - W1 / W7 * a / rb3 * y2 -
                  2 * a * y3b / sq(rb2) * y2));
}

/* TDSetupS (:433-464) for one point. */
static void td_setup_s(double x, double y, double z, double alpha, double bx, double by, double bz, double nu,
                       const double tv[3], const double sv[3], double e[6]) {
    /* A = [sv3 -sv2; sv2 sv3] (row-major) */
    const double y1 = sv[2] * (y - tv[1]) - sv[1] * (z - tv[2]);
    const double z1 = sv[1] * (y - tv[1]) + sv[2] * (z - tv[2]);
    const double by1 = sv[2] * by - sv[1] * bz;
    const double bz1 = sv[1] * by + sv[2] * bz;
    double ea[6];
    ang_dis_strain(x, y1, z1, -PI + alpha, bx, by1, bz1, nu, ea);
    /* B = [1 0 0; 0 sv3 sv2; 0 -sv2 sv3], column-major */
    const double B[9] = {1.0, 0.0, 0.0, 0.0, sv[2], -sv[1], 0.0, sv[1], sv[2]};
    tens_trans(ea, B, e);
}

/* TDstressFS (:133-286) for one point: strain (6) in EFCS.  zsign selects the
 * horizontal-element strike correction for image sources (P1z > 0). */
static void td_strain_fs(double X, double Y, double Z, const double P1[3], const double P2[3], const double P3[3],
                         double Ss, double Ds, double Ts, double nu, double E[6]) {
    const double bx = Ts, by = Ss, bz = Ds;
    double u[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    double w[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    double Vnorm[3], Vstrike[3], Vdip[3];
    cross3(u, w, Vnorm);
    double nn = norm3(Vnorm);
    for (int k = 0; k < 3; ++k) Vnorm[k] /= nn;
    const double eY[3] = {0, 1, 0}, eZ[3] = {0, 0, 1};
    cross3(eZ, Vnorm, Vstrike);
    if (norm3(Vstrike) == 0) {
        for (int k = 0; k < 3; ++k) Vstrike[k] = eY[k] * Vnorm[2];
        if (P1[2] > 0)
            for (int k = 0; k < 3; ++k) Vstrike[k] = -Vstrike[k];
    }
    double ns = norm3(Vstrike);
    for (int k = 0; k < 3; ++k) Vstrike[k] /= ns;
    cross3(Vnorm, Vstrike, Vdip);
    /* A = [Vnorm Vstrike Vdip]' : rows Vnorm, Vstrike, Vdip -> column-major */
    const double A[9] = {Vnorm[0], Vstrike[0], Vdip[0], Vnorm[1], Vstrike[1], Vdip[1], Vnorm[2], Vstrike[2], Vdip[2]};
    double r[3], p1[3], p2[3] = {0, 0, 0}, p3[3];
    coord_trans(X - P2[0], Y - P2[1], Z - P2[2], A, r);
    const double x = r[0], y = r[1], z = r[2];
    coord_trans(P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2], A, p1);
    coord_trans(P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2], A, p3);
    double e12[3], e13[3], e23[3];
    double d12[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double d13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double d23[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    double n12 = norm3(d12), n13 = norm3(d13), n23 = norm3(d23);
    for (int k = 0; k < 3; ++k) {
        e12[k] = d12[k] / n12;
        e13[k] = d13[k] / n13;
        e23[k] = d23[k] / n23;
    }
    const double Aa = acos(dot3(e12, e13));
    const double me12[3] = {-e12[0], -e12[1], -e12[2]};
    const double Ba = acos(dot3(me12, e23));
    const double Ca = acos(dot3(e23, e13));
    const double q1[2] = {p1[1], p1[2]}, q2[2] = {p2[1], p2[2]}, q3[2] = {p3[1], p3[2]};
    const int m = trimode(y, z, x, q1, q2, q3);
    double e[6];
    if (m == 1) {
        double s1[6], s2[6], s3[6];
        const double me13[3] = {-e13[0], -e13[1], -e13[2]};
        td_setup_s(x, y, z, Aa, bx, by, bz, nu, p1, me13, s1);
        td_setup_s(x, y, z, Ba, bx, by, bz, nu, p2, e12, s2);
        td_setup_s(x, y, z, Ca, bx, by, bz, nu, p3, e23, s3);
        for (int k = 0; k < 6; ++k) e[k] = s1[k] + s2[k] + s3[k];
    } else if (m == -1) {
        double s1[6], s2[6], s3[6];
        const double me23[3] = {-e23[0], -e23[1], -e23[2]};
        td_setup_s(x, y, z, Aa, bx, by, bz, nu, p1, e13, s1);
        td_setup_s(x, y, z, Ba, bx, by, bz, nu, p2, me12, s2);
        td_setup_s(x, y, z, Ca, bx, by, bz, nu, p3, me23, s3);
        for (int k = 0; k < 6; ++k) e[k] = s1[k] + s2[k] + s3[k];
    } else {
        for (int k = 0; k < 6; ++k) e[k] = NAN;
    }
    /* TensTrans with [Vnorm Vstrike Vdip] (columns) */
    const double At[9] = {Vnorm[0], Vnorm[1], Vnorm[2], Vstrike[0], Vstrike[1], Vstrike[2], Vdip[0], Vdip[1], Vdip[2]};
    tens_trans(e, At, E);
}

/* AngSetupFSC_S (:467-574) strain part for one point. */
static void ang_setup_fsc(double X, double Y, double Z, double bX, double bY, double bZ, const double PA[3],
                          const double PB[3], double nu, double E[6]) {
    const double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    const double beta = acos(-sv[2] / norm3(sv));
    if (fabs(beta) < 1e-16 || fabs(PI - beta) < 1e-16) {
        for (int k = 0; k < 6; ++k) E[k] = 0.0;
        return;
    }
    double ey1[3] = {sv[0], sv[1], 0.0};
    double n1 = norm3(ey1);
    for (int k = 0; k < 3; ++k) ey1[k] /= n1;
    const double ey3[3] = {0, 0, -1};
    double ey2[3];
    cross3(ey3, ey1, ey2);
    /* A = [ey1 ey2 ey3] (columns), column-major */
    const double A[9] = {ey1[0], ey1[1], ey1[2], ey2[0], ey2[1], ey2[2], ey3[0], ey3[1], ey3[2]};
    double yA[3], yAB[3], bb[3];
    coord_trans(X - PA[0], Y - PA[1], Z - PA[2], A, yA);
    coord_trans(sv[0], sv[1], sv[2], A, yAB);
    const double yB[3] = {yA[0] - yAB[0], yA[1] - yAB[1], yA[2] - yAB[2]};
    coord_trans(bX, bY, bZ, A, bb);
    double vA[6], vB[6];
    if (beta * yA[0] >= 0) { /* configuration I */
        ang_dis_strain_fsc(-yA[0], -yA[1], yA[2], PI - beta, -bb[0], -bb[1], bb[2], nu, -PA[2], vA);
        vA[4] = -vA[4];
        vA[5] = -vA[5];
        ang_dis_strain_fsc(-yB[0], -yB[1], yB[2], PI - beta, -bb[0], -bb[1], bb[2], nu, -PB[2], vB);
        vB[4] = -vB[4];
        vB[5] = -vB[5];
    } else { /* configuration II */
        ang_dis_strain_fsc(yA[0], yA[1], yA[2], beta, bb[0], bb[1], bb[2], nu, -PA[2], vA);
        ang_dis_strain_fsc(yB[0], yB[1], yB[2], beta, bb[0], bb[1], bb[2], nu, -PB[2], vB);
    }
    double v[6];
    for (int k = 0; k < 6; ++k) v[k] = vB[k] - vA[k];
    /* TensTrans(v, A') with A' column-major */
    const double At[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
    tens_trans(v, At, E);
}

/* TDstress_HarFunc (:289-335) strain for one point. */
static void td_strain_harfunc(double X, double Y, double Z, const double P1[3], const double P2[3], const double P3[3],
                              double Ss, double Ds, double Ts, double nu, double E[6]) {
    const double bx = Ts, by = Ss, bz = Ds;
    double u[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    double w[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    double Vnorm[3], Vstrike[3], Vdip[3];
    cross3(u, w, Vnorm);
    double nn = norm3(Vnorm);
    for (int k = 0; k < 3; ++k) Vnorm[k] /= nn;
    const double eY[3] = {0, 1, 0}, eZ[3] = {0, 0, 1};
    cross3(eZ, Vnorm, Vstrike);
    if (norm3(Vstrike) == 0)
        for (int k = 0; k < 3; ++k) Vstrike[k] = eY[k] * Vnorm[2];
    double ns = norm3(Vstrike);
    for (int k = 0; k < 3; ++k) Vstrike[k] /= ns;
    cross3(Vnorm, Vstrike, Vdip);
    const double A[9] = {Vnorm[0], Vnorm[1], Vnorm[2], Vstrike[0], Vstrike[1], Vstrike[2], Vdip[0], Vdip[1], Vdip[2]};
    double b[3];
    coord_trans(bx, by, bz, A, b);
    double e1[6], e2[6], e3[6];
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P1, P2, nu, e1);
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P2, P3, nu, e2);
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P3, P1, nu, e3);
    for (int k = 0; k < 6; ++k) E[k] = e1[k] + e2[k] + e3[k];
}

static void strain_to_stress(const double E[6], double mu, double lam, double S[6]) {
    const double tr = E[0] + E[1] + E[2];
    S[0] = 2 * mu * E[0] + lam * tr;
    S[1] = 2 * mu * E[1] + lam * tr;
    S[2] = 2 * mu * E[2] + lam * tr;
    S[3] = 2 * mu * E[3];
    S[4] = 2 * mu * E[4];
    S[5] = 2 * mu * E[5];
}

/* TDstressHS (:2-130) at one point. */
static void td_stress_hs_point(double X, double Y, double Z, const double P1[3], const double P2[3], const double P3[3],
                               double Ss, double Ds, double Ts, double mu, double lam, double S[6], double E[6]) {
    const double nu = 1 / (1 + lam / mu) / 2;
    double Ems[6], Efsc[6], Eis[6];
    td_strain_fs(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, Ems);
    td_strain_harfunc(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, Efsc);
    const double Q1[3] = {P1[0], P1[1], -P1[2]}, Q2[3] = {P2[0], P2[1], -P2[2]}, Q3[3] = {P3[0], P3[1], -P3[2]};
    td_strain_fs(X, Y, Z, Q1, Q2, Q3, Ss, Ds, Ts, nu, Eis);
    if (Q1[2] == 0 && Q2[2] == 0 && Q3[2] == 0) {
        Eis[4] = -Eis[4];
        Eis[5] = -Eis[5];
    }
    double Sms[6], Sis[6], Sfsc[6];
    strain_to_stress(Ems, mu, lam, Sms);
    strain_to_stress(Eis, mu, lam, Sis);
    strain_to_stress(Efsc, mu, lam, Sfsc);
    for (int k = 0; k < 6; ++k) {
        S[k] = Sms[k] + Sis[k] + Sfsc[k];
        E[k] = Ems[k] + Eis[k] + Efsc[k];
    }
}

/* Stress, Strain: n x 6 column-major. */
void orc_tdstress_hs(int64_t n, const double *X, const double *Y, const double *Z, const double *P1, const double *P2,
                     const double *P3, double Ss, double Ds, double Ts, double mu, double lambda, double *Stress,
                     double *Strain) {
    for (int64_t i = 0; i < n; ++i) {
        double S[6], E[6];
        td_stress_hs_point(X[i], Y[i], Z[i], P1, P2, P3, Ss, Ds, Ts, mu, lambda, S, E);
        for (int k = 0; k < 6; ++k) {
            Stress[i + k * n] = S[k];
            Strain[i + k * n] = E[k];
        }
    }
}

/* RotVerts_UnitVectors (Functions_BuildInputFile.jl:338-377).  P1..P3 and the
 * outputs are n x 3 column-major; P1/P2 are swapped in place where the normal
 * has y <= 0. */
void orc_rotverts_unitvectors(int64_t n, double *P1, double *P2, double *P3, const double *rake, double *nrm,
                              double *strike, double *dip, double *slip) {
    for (int64_t i = 0; i < n; ++i) {
        double p1[3], p2[3], p3[3], u[3], w[3], c[3];
        for (int k = 0; k < 3; ++k) {
            p1[k] = P1[i + k * n];
            p2[k] = P2[i + k * n];
            p3[k] = P3[i + k * n];
        }
        for (int k = 0; k < 3; ++k) { u[k] = p2[k] - p1[k]; w[k] = p3[k] - p1[k]; }
        cross3(u, w, c);
        double cn = norm3(c);
        if (c[1] / cn <= 0) {
            for (int k = 0; k < 3; ++k) { double t = p1[k]; p1[k] = p2[k]; p2[k] = t; }
        }
        for (int k = 0; k < 3; ++k) {
            P1[i + k * n] = p1[k];
            P2[i + k * n] = p2[k];
            P3[i + k * n] = p3[k];
        }
        for (int k = 0; k < 3; ++k) { u[k] = p2[k] - p1[k]; w[k] = p3[k] - p1[k]; }
        cross3(u, w, c);
        cn = norm3(c);
        double N[3] = {c[0] / cn, c[1] / cn, c[2] / cn};
        const double ez[3] = {0, 0, 1};
        double s[3], d[3];
        cross3(N, ez, s);
        double sn = norm3(s);
        for (int k = 0; k < 3; ++k) s[k] /= sn;
        cross3(s, N, d);
        double dn = norm3(d);
        for (int k = 0; k < 3; ++k) d[k] /= dn;
        const double cr = orc_cosd(rake[i]), sr = orc_sind(rake[i]);
        for (int k = 0; k < 3; ++k) {
            nrm[i + k * n] = N[k];
            strike[i + k * n] = s[k];
            dip[i + k * n] = d[k];
            slip[i + k * n] = s[k] * cr + d[k] * sr;
        }
    }
}

/* StiffnessMatrix_ByParts_Calculation_Tri (Functions_BuildInputFile.jl:551-594).
 * Inputs n x 3 column-major; KS/KN nr x ns column-major (ldk). */
void orc_stiffness_tri(int64_t ns, const double *P1, const double *P2, const double *P3, const double *rake,
                       int64_t nr, const double *ctr, const double *nrm, const double *slip, double mu, double lambda,
                       double *KS, double *KN, int64_t ldk) {
    for (int64_t s = 0; s < ns; ++s) {
        const double p1[3] = {P1[s], P1[s + ns], P1[s + 2 * ns]};
        const double p2[3] = {P2[s], P2[s + ns], P2[s + 2 * ns]};
        const double p3[3] = {P3[s], P3[s + ns], P3[s + 2 * ns]};
        const double Ss = -orc_cosd(rake[s]), Ds = orc_sind(rake[s]), Ts = 0.0;
        for (int64_t r = 0; r < nr; ++r) {
            double S[6], E[6];
            td_stress_hs_point(ctr[r], ctr[r + nr], ctr[r + 2 * nr], p1, p2, p3, Ss, Ds, Ts, mu, lambda, S, E);
            const double n0 = nrm[r], n1 = nrm[r + nr], n2 = nrm[r + 2 * nr];
            /* Stress_i * n with Stress_i = [xx xy xz; xy yy yz; xz yz zz] */
            const double t0 = S[0] * n0 + S[3] * n1 + S[4] * n2;
            const double t1 = S[3] * n0 + S[1] * n1 + S[5] * n2;
            const double t2 = S[4] * n0 + S[5] * n1 + S[2] * n2;
            KN[r + s * ldk] = t0 * n0 + t1 * n1 + t2 * n2;
            KS[r + s * ldk] = t0 * slip[r] + t1 * slip[r + nr] + t2 * slip[r + 2 * nr];
        }
    }
}
