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
// Triangular dislocation in an elastic half-space (Nikkhoo & Walter 2015):
// device routines behind qdfn_tde_stress_hs / qdfn_tde_stiffness_tri.
// Semantics follow the reference's scripts/Functions_TDstressHS.jl
// (TDstressHS :2-130 = main full-space dislocation + image dislocation +
// harmonic free-surface correction; trimodefinder's configuration choice and
// NaN on the triangle's edges; AngSetupFSC_S's configuration choice by the
// sign of beta*y1).  Organisation for the GPU:
//   * everything that depends on the source triangle only (TDCS basis and
//     vertices of the main and image triangles, angles, the angular-dislocation
//     frames of the three edges, slip vectors) is computed once per source by
//     k_tri_prepare into a TriSource record that a workgroup reads uniformly;
//   * the per-pair work is the six angular-dislocation strains, the six
//     free-surface-correction strains and the tensor transforms, in FP64,
//     written as loops over one inlined copy of each strain expression: the
//     code stays small enough for short branches and is inlined into the
//     kernels (no device-function call, so no return address to keep);
//   * the free-surface-correction expressions are evaluated with their
//     repeated brackets computed once (T*, U*, V* temporaries below).
#pragma once
#include <hip/hip_runtime.h>

namespace qdfn {

#define QDFN_PI 3.141592653589793

struct TdFrame {          // one triangle (main or image) in its own TDCS
    double A[9];          // rows Vnorm, Vstrike, Vdip (EFCS -> TDCS), row-major
    double P2[3];         // origin of the TDCS (second vertex), EFCS
    double tv[3][2];      // (y, z) of the vertices A = p1, B = p2 = 0, C = p3 in TDCS
    double sinA[3], cosA[3];  // sin, cos of (angle - pi) at A, B, C (TDSetupS :440)
    double sv[3][2];      // (y, z) side directions at A, B, C for configuration II:
                          // e13, -e12, -e23 (configuration I negates all three)
    int flip_xz;          // image of a surface triangle: negate xz, yz
    int pad;
};

struct FscEdge {          // one side PA->PB for the free-surface correction
    double sinB, cosB, cotB;      // beta
    double sinB2, cosB2, cotB2;   // pi - beta
    double PA[3], PB3;            // PA (EFCS) and PB's z
    double ey1x, ey1y;            // A = [ey1 ey2 ey3], ey2 = (ey1y, -ey1x, 0), ey3 = (0, 0, -1)
    double yAB[3];                // side vector in ADCS
    double bb[3];                 // slip vector in ADCS
    double beta;
    int degenerate, pad;
};

struct TriSource {
    TdFrame main, image;
    FscEdge edge[3];
    double bx, by, bz;            // Ts, Ss, Ds
    double nu;
};

__device__ __host__ inline double dsq(double x) { return x * x; }

// AngDisStrain (:577-651): strains of an angular dislocation in a full space.
__device__ __forceinline__ void ang_strain(double x, double y, double z, double sinA, double cosA, double bx, double by,
                                  double bz, double nu, double e[6]) {
    const double eta = y * cosA - z * sinA;
    const double zeta = y * sinA + z * cosA;
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    const double r = sqrt(r2);
    const double r3 = r * r2;
    const double rmz = r - z;
    const double rz = r * rmz;
    const double r2z2 = r2 * rmz * rmz;
    const double r3z = r3 * rmz;
    const double W = zeta - r;
    const double W2 = W * W;
    const double Wr = W * r, W2r = W2 * r, Wr3 = W * r3, W2r2 = W2 * r2;
    const double C = (r * cosA - z) / Wr;
    const double S = (r * sinA - y) / Wr;
    const double rmzeta = r - zeta;
    const double k4 = 1.0 / 4 / QDFN_PI;
    const double Fx = (eta / r / rmzeta - y / r / rmz) * k4;
    const double Fy = (x / r / rmz - cosA * x / r / rmzeta) * k4;
    const double Fz = (sinA * x / r / rmzeta) * k4;
    const double k8 = 1.0 / 8 / QDFN_PI / (1 - nu);
    const double n21 = 2 * nu + 1;
    const double gx = n21 / Wr + x2 / W2r2 - x2 / Wr3;       // shared by exx's by- and bz-terms
    const double gy = 1 / Wr + S * S - y2 / Wr3;              // eyy
    const double gz = 1 / Wr + C * C - z2 / Wr3;              // ezz
    const double ycz = y * cosA + z * sinA;
    e[0] = bx * Fx + bx * k8 * (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz - x2 * y / r2z2 - x2 * y / r3z) -
           by * x * k8 * (gx * cosA + n21 / rz - x2 / r2z2 - x2 / r3z) + bz * x * sinA * k8 * gx;
    e[1] = by * Fy +
           bx * k8 * (gy * eta + n21 * y / rz - y2 * y / r2z2 - y2 * y / r3z - 2 * nu * cosA * S) -
           by * x * k8 * (1 / rz - y2 / r2z2 - y2 / r3z + gy * cosA) + bz * x * sinA * k8 * gy;
    e[2] = bz * Fz + bx * k8 * (eta / W / r + eta * C * C - eta * z2 / Wr3 + y * z / r3 + 2 * nu * sinA * C) -
           by * x * k8 * (gz * cosA + z / r3) + bz * x * sinA * k8 * gz;
    e[3] = bx * Fy / 2 + by * Fx / 2 -
           bx * k8 * (x * y2 / r2z2 - nu * x / rz + x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr + eta * x * y / Wr3) +
           by * k8 * (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S + x2 * y * cosA / Wr3 + x2 * cosA * S / Wr) -
           bz * sinA * k8 * (nu * S + x2 * S / Wr + x2 * y / Wr3);
    e[4] = bx * Fz / 2 + bz * Fx / 2 - bx * k8 * (-x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr + eta * x * z / Wr3) +
           by * k8 * (-x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 + x2 * cosA * C / Wr) -
           bz * sinA * k8 * (nu * C + x2 * C / Wr + x2 * z / Wr3);
    e[5] = by * Fz / 2 + bz * Fy / 2 +
           bx * k8 * (y2 / r3 - nu / r - nu * cosA * C + nu * sinA * S + eta * sinA * cosA / W2 - eta * ycz / W2r +
                      eta * y * z / W2r2 - eta * y * z / Wr3) -
           by * x * k8 * (y / r3 + sinA * cosA * cosA / W2 - cosA * ycz / W2r + y * z * cosA / W2r2 - y * z * cosA / Wr3) -
           bz * x * sinA * k8 * (y * z / Wr3 - sinA * cosA / W2 + ycz / W2r - y * z / W2r2);
}

// AngDisStrainFSC (:654-1034): harmonic free-surface correction of an angular
// dislocation; sinB/cosB/cotB describe its angle, a its vertex depth.
__device__ __forceinline__ void ang_strain_fsc(double y1, double y2, double y3, double sinB, double cosB, double cotB,
                                      double b1, double b2, double b3, double nu, double a, double v[6]) {
    const double y3b = y3 + 2 * a;
    const double z1b = y1 * cosB + y3b * sinB;
    const double z3b = -y1 * sinB + y3b * cosB;
    const double rb2 = y1 * y1 + y2 * y2 + y3b * y3b;
    const double rb = sqrt(rb2);
    const double rb3 = rb * rb2, rb5 = rb3 * rb2, rb4 = rb2 * rb2;
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
    const double cot2 = cotB * cotB, cB2 = cosB * cosB;
    const double W6s = W6 * W6, W7s = W7 * W7;
    const double y1s = y1 * y1, y2s = y2 * y2, y1c = y1s * y1, y2c = y2s * y2;
    const double K = 1.0 / 4 / QDFN_PI / (1 - nu);
    const double rFib_ry2 = z1b / rb / (rb + z3b) - y1 / rb / (rb + y3b);
    const double rFib_ry1 = y2 / rb / (rb + y3b) - cosB * y2 / rb / (rb + z3b);
    const double rFib_ry3 = -sinB * y2 / rb / (rb + z3b);
    const double D7 = y1 / rb - sinB;
    const double R3 = y3b / rb + 1;
    const double C3 = cosB * y3b / rb + 1;
    const double T1 = (1 - W5) * cotB - y1 / W6 * W4;
    const double T2 = -N1 * cotB + y1 / W6 * W5 + a * y1 / rb2;
    const double T3i = W1 * (N1 * cosB - a / rb) * cotB + (2 - 2 * nu) * (rb * sinB - y1) * cosB;
    const double T3 = cosB / W7 * T3i - a * y3b * cosB * cotB / rb2;
    const double T4 = -N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + y1s / W6 * W4;
    const double T5 = z1b * cosB - a * (rb * sinB - y1) / rb / cosB;
    const double T6 = 2 * nu + 1 / rb * (N1 * y1 * cotB + a) - y1s / rb / W6 * W5 - a * y1s / rb3;
    const double T7 = (2 - 2 * nu) * cosB - W1 / W7 * W9;
    const double T8 = -cosB * sinB + a * y1 * y3b / rb3 / cosB + (rb * sinB - y1) / rb * T7;
    const double T9 = W1 / W7 * W2 + a * y3b / rb2;
    const double T10 = a / rb2 + 1 / W6;
    const double U1 = y1 * cotB * (1 - W5) + nu * y3b - a + y2s / W6 * W4;
    const double U2 = -2 * nu + 1 / rb * (N1 * y1 * cotB - a) + y2s / rb / W6 * W5 + a * y2s / rb3;
    const double U3i = y2s * cB2 - a * z1b * cotB / rb * W1;
    const double U3 = cB2 - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / rb3 - 1 / rb / W7 * U3i;
    const double U4 = (W5 - 1) * cotB + y1 / W6 * W4;
    const double U5 = N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6);
    const double U6 = (-2 + 2 * nu) * cosB + W1 / W7 * W9 + a * y3b / rb2 / cosB;
    const double U7i = y2s * cosB * sinB - a * z1b / rb * W1;
    const double U7 = sinB * (cosB - a / rb) + z1b / rb * (1 + a * y3b / rb2) - 1 / rb / W7 * U7i;
    const double V1 = 2 * nu / W6 + a / rb2;
    const double V2 = 1 - 2 * nu - W1 / W7 * W2 - a * y3b / rb2;
    const double V3 = N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb2;
    const double V4i = (2 - 2 * nu) * cosB - W1 / W7;
    const double V4s = sinB - y3b * z1b / rb2 - z1b * W1 / rb / W7;
    const double V4 = cosB * sinB + W1 * cotB / rb * V4i + a / rb * V4s;
    const double V5 = 1 + W1 / W7 * W2 + a * y3b / rb2;
    v[0] = b1 * K * ((-2 + 2 * nu) * N1 * rFib_ry1 * cot2 - N1 * y2 / W6s * T1 / rb * y1 + N1 * y2 / W6 * (a / rb3 * y1 * cotB - 1 / W6 * W4 + y1s / W6s * W4 / rb + y1s / W6 * a / rb3) - N1 * y2 * cosB * cotB / W7s * W2 * D7 - N1 * y2 * cosB * cotB / W7 * a / rb3 * y1 - 3 * a * y2 * W8 * cotB / rb5 * y1 - y2 * W8 / rb3 / W6 * T2 * y1 - y2 * W8 / rb2 / W6s * T2 * y1 + y2 * W8 / rb / W6 * (1 / W6 * W5 - y1s / W6s * W5 / rb - y1s / W6 * a / rb3 + a / rb2 - 2 * a * y1s / rb4) - y2 * W8 / rb3 / W7 * T3 * y1 - y2 * W8 / rb / W7s * T3 * D7 + y2 * W8 / rb / W7 * (-cosB / W7s * T3i * D7 + cosB / W7 * (1 / rb * cosB * y1 * (N1 * cosB - a / rb) * cotB + W1 * a / rb3 * y1 * cotB + (2 - 2 * nu) * (1 / rb * sinB * y1 - 1) * cosB) + 2 * a * y3b * cosB * cotB / rb4 * y1)) + b2 * K * (N1 * (((2 - 2 * nu) * cot2 + nu) / rb * y1 / W6 - ((2 - 2 * nu) * cot2 + 1) * cosB * D7 / W7) - N1 / W6s * T4 / rb * y1 + N1 / W6 * (-N1 * cotB + a * cotB / rb - a * y1s * cotB / rb3 + 2 * y1 / W6 * W4 - y1c / W6s * W4 / rb - y1c / W6 * a / rb3) + N1 * cotB / W7s * T5 * D7 - N1 * cotB / W7 * (cB2 - a * (1 / rb * sinB * y1 - 1) / rb / cosB + a * (rb * sinB - y1) / rb3 / cosB * y1) - a * W8 * cotB / rb3 + 3 * a * y1s * W8 * cotB / rb5 - W8 / W6s * T6 / rb * y1 + W8 / W6 * (-1 / rb3 * (N1 * y1 * cotB + a) * y1 + 1 / rb * N1 * cotB - 2 * y1 / rb / W6 * W5 + y1c / rb3 / W6 * W5 + y1c / rb2 / W6s * W5 + y1c / rb4 / W6 * a - 2 * a / rb3 * y1 + 3 * a * y1c / rb5) - W8 * cotB / W7s * T8 * D7 + W8 * cotB / W7 * (a * y3b / rb3 / cosB - 3 * a * y1s * y3b / rb5 / cosB + (1 / rb * sinB * y1 - 1) / rb * T7 - (rb * sinB - y1) / rb3 * T7 * y1 + (rb * sinB - y1) / rb * (-1 / rb * cosB * y1 / W7 * W9 + W1 / W7s * W9 * D7 + W1 / W7 * a / rb3 / cosB * y1))) + b3 * K * (N1 * (-y2 / W6s * (1 + a / rb) / rb * y1 - y2 / W6 * a / rb3 * y1 + y2 * cosB / W7s * W2 * D7 + y2 * cosB / W7 * a / rb3 * y1) + y2 * W8 / rb3 * T10 * y1 - y2 * W8 / rb * (-2 * a / rb4 * y1 - 1 / W6s / rb * y1) - y2 * W8 * cosB / rb3 / W7 * T9 * y1 - y2 * W8 * cosB / rb / W7s * T9 * D7 + y2 * W8 * cosB / rb / W7 * (1 / rb * cosB * y1 / W7 * W2 - W1 / W7s * W2 * D7 - W1 / W7 * a / rb3 * y1 - 2 * a * y3b / rb4 * y1));
    v[1] = b1 * K * (N1 * (((2 - 2 * nu) * cot2 - nu) / rb * y2 / W6 - ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB / rb * y2 / W7) + N1 / W6s * U1 / rb * y2 - N1 / W6 * (a * y1 * cotB / rb3 * y2 + 2 * y2 / W6 * W4 - y2c / W6s * W4 / rb - y2c / W6 * a / rb3) + N1 * z1b * cotB / W7s * W2 / rb * y2 + N1 * z1b * cotB / W7 * a / rb3 * y2 + 3 * a * y2 * W8 * cotB / rb5 * y1 - W8 / W6s * U2 / rb * y2 + W8 / W6 * (-1 / rb3 * (N1 * y1 * cotB - a) * y2 + 2 * y2 / rb / W6 * W5 - y2c / rb3 / W6 * W5 - y2c / rb2 / W6s * W5 - y2c / rb4 / W6 * a + 2 * a / rb3 * y2 - 3 * a * y2c / rb5) - W8 / W7s * U3 / rb * y2 + W8 / W7 * (1 / rb3 * (N1 * z1b * cotB + a * cosB) * y2 - 3 * a * y3b * z1b * cotB / rb5 * y2 + 1 / rb3 / W7 * U3i * y2 + 1 / rb2 / W7s * U3i * y2 - 1 / rb / W7 * (2 * y2 * cB2 + a * z1b * cotB / rb3 * W1 * y2 - a * z1b * cotB / rb2 * cosB * y2))) + b2 * K * ((2 - 2 * nu) * N1 * rFib_ry2 * cot2 + N1 / W6 * U4 - N1 * y2s / W6s * U4 / rb + N1 * y2 / W6 * (-a / rb3 * y2 * cotB - y1 / W6s * W4 / rb * y2 - y2 / W6 * a / rb3 * y1) - N1 * cotB / W7 * W9 + N1 * y2s * cotB / W7s * W9 / rb + N1 * y2s * cotB / W7 * a / rb3 / cosB - a * W8 * cotB / rb3 + 3 * a * y2s * W8 * cotB / rb5 + W8 / rb / W6 * U5 - y2s * W8 / rb3 / W6 * U5 - y2s * W8 / rb2 / W6s * U5 + y2 * W8 / rb / W6 * (2 * nu * y1 / W6s / rb * y2 + a * y1 / rb3 * (1 / rb + 1 / W6) * y2 - a * y1 / rb * (-1 / rb3 * y2 - 1 / W6s / rb * y2)) + W8 * cotB / rb / W7 * U6 - y2s * W8 * cotB / rb3 / W7 * U6 - y2s * W8 * cotB / rb2 / W7s * U6 + y2 * W8 * cotB / rb / W7 * (1 / rb * cosB * y2 / W7 * W9 - W1 / W7s * W9 / rb * y2 - W1 / W7 * a / rb3 / cosB * y2 - 2 * a * y3b / rb4 / cosB * y2)) + b3 * K * (N1 * (-sinB / rb * y2 / W7 + y2 / W6s * (1 + a / rb) / rb * y1 + y2 / W6 * a / rb3 * y1 - z1b / W7s * W2 / rb * y2 - z1b / W7 * a / rb3 * y2) - y2 * W8 / rb3 * T10 * y1 + y1 * W8 / rb * (-2 * a / rb4 * y2 - 1 / W6s / rb * y2) + W8 / W7s * U7 / rb * y2 - W8 / W7 * (sinB * a / rb3 * y2 - z1b / rb3 * (1 + a * y3b / rb2) * y2 - 2 * z1b / rb5 * a * y3b * y2 + 1 / rb3 / W7 * U7i * y2 + 1 / rb2 / W7s * U7i * y2 - 1 / rb / W7 * (2 * y2 * cosB * sinB + a * z1b / rb3 * W1 * y2 - a * z1b / rb2 * cosB * y2)));
    v[2] = b1 * K * ((2 - 2 * nu) * (N1 * rFib_ry3 * cotB - y2 / W6s * W5 * R3 - 0.5 * y2 / W6 * a / rb3 * 2 * y3b + y2 * cosB / W7s * W2 * W3 + 0.5 * y2 * cosB / W7 * a / rb3 * 2 * y3b) + y2 / rb * V1 - 0.5 * y2 * W8 / rb3 * V1 * 2 * y3b + y2 * W8 / rb * (-2 * nu / W6s * R3 - a / rb4 * 2 * y3b) + y2 * cosB / rb / W7 * V2 - 0.5 * y2 * W8 * cosB / rb3 / W7 * V2 * 2 * y3b - y2 * W8 * cosB / rb / W7s * V2 * W3 + y2 * W8 * cosB / rb / W7 * (-C3 / W7 * W2 + W1 / W7s * W2 * W3 + 0.5 * W1 / W7 * a / rb3 * 2 * y3b - a / rb2 + a * y3b / rb4 * 2 * y3b)) + b2 * K * ((-2 + 2 * nu) * N1 * cotB * (R3 / W6 - cosB * W3 / W7) + (2 - 2 * nu) * y1 / W6s * W5 * R3 + 0.5 * (2 - 2 * nu) * y1 / W6 * a / rb3 * 2 * y3b + (2 - 2 * nu) * sinB / W7 * W2 - (2 - 2 * nu) * z1b / W7s * W2 * W3 - 0.5 * (2 - 2 * nu) * z1b / W7 * a / rb3 * 2 * y3b + 1 / rb * V3 - 0.5 * W8 / rb3 * V3 * 2 * y3b + W8 / rb * (2 * nu * y1 / W6s * R3 + a * y1 / rb4 * 2 * y3b) - 1 / W7 * V4 + W8 / W7s * V4 * W3 - W8 / W7 * (C3 * cotB / rb * V4i - 0.5 * W1 * cotB / rb3 * V4i * 2 * y3b + W1 * cotB / rb * (-C3 / W7 + W1 / W7s * W3) - 0.5 * a / rb3 * V4s * 2 * y3b + a / rb * (-z1b / rb2 - y3b * sinB / rb2 + y3b * z1b / rb4 * 2 * y3b - sinB * W1 / rb / W7 - z1b * C3 / rb / W7 + 0.5 * z1b * W1 / rb3 / W7 * 2 * y3b + z1b * W1 / rb / W7s * W3))) + b3 * K * ((2 - 2 * nu) * rFib_ry3 - (2 - 2 * nu) * y2 * sinB / W7s * W2 * W3 - 0.5 * (2 - 2 * nu) * y2 * sinB / W7 * a / rb3 * 2 * y3b + y2 * sinB / rb / W7 * V5 - 0.5 * y2 * W8 * sinB / rb3 / W7 * V5 * 2 * y3b - y2 * W8 * sinB / rb / W7s * V5 * W3 + y2 * W8 * sinB / rb / W7 * (C3 / W7 * W2 - W1 / W7s * W2 * W3 - 0.5 * W1 / W7 * a / rb3 * 2 * y3b + a / rb2 - a * y3b / rb4 * 2 * y3b));
    v[3] = b1 / 2 * K * ((-2 + 2 * nu) * N1 * rFib_ry2 * cot2 + N1 / W6 * T1 - N1 * y2s / W6s * T1 / rb + N1 * y2 / W6 * (a / rb3 * y2 * cotB + y1 / W6s * W4 / rb * y2 + y2 / W6 * a / rb3 * y1) + N1 * cosB * cotB / W7 * W2 - N1 * y2s * cosB * cotB / W7s * W2 / rb - N1 * y2s * cosB * cotB / W7 * a / rb3 + a * W8 * cotB / rb3 - 3 * a * y2s * W8 * cotB / rb5 + W8 / rb / W6 * T2 - y2s * W8 / rb3 / W6 * T2 - y2s * W8 / rb2 / W6s * T2 + y2 * W8 / rb / W6 * (-y1 / W6s * W5 / rb * y2 - y2 / W6 * a / rb3 * y1 - 2 * a * y1 / rb4 * y2) + W8 / rb / W7 * T3 - y2s * W8 / rb3 / W7 * T3 - y2s * W8 / rb2 / W7s * T3 + y2 * W8 / rb / W7 * (-cosB / W7s * T3i / rb * y2 + cosB / W7 * (1 / rb * cosB * y2 * (N1 * cosB - a / rb) * cotB + W1 * a / rb3 * y2 * cotB + (2 - 2 * nu) / rb * sinB * y2 * cosB) + 2 * a * y3b * cosB * cotB / rb4 * y2)) + b2 / 2 * K * (N1 * (((2 - 2 * nu) * cot2 + nu) / rb * y2 / W6 - ((2 - 2 * nu) * cot2 + 1) * cosB / rb * y2 / W7) - N1 / W6s * T4 / rb * y2 + N1 / W6 * (-a * y1 * cotB / rb3 * y2 - y1s / W6s * W4 / rb * y2 - y1s / W6 * a / rb3 * y2) + N1 * cotB / W7s * T5 / rb * y2 - N1 * cotB / W7 * (-a / rb2 * sinB * y2 / cosB + a * (rb * sinB - y1) / rb3 / cosB * y2) + 3 * a * y2 * W8 * cotB / rb5 * y1 - W8 / W6s * T6 / rb * y2 + W8 / W6 * (-1 / rb3 * (N1 * y1 * cotB + a) * y2 + y1s / rb3 / W6 * W5 * y2 + y1s / rb2 / W6s * W5 * y2 + y1s / rb4 / W6 * a * y2 + 3 * a * y1s / rb5 * y2) - W8 * cotB / W7s * T8 / rb * y2 + W8 * cotB / W7 * (-3 * a * y1 * y3b / rb5 / cosB * y2 + 1 / rb2 * sinB * y2 * T7 - (rb * sinB - y1) / rb3 * T7 * y2 + (rb * sinB - y1) / rb * (-1 / rb * cosB * y2 / W7 * W9 + W1 / W7s * W9 / rb * y2 + W1 / W7 * a / rb3 / cosB * y2))) + b3 / 2 * K * (N1 * (1 / W6 * (1 + a / rb) - y2s / W6s * (1 + a / rb) / rb - y2s / W6 * a / rb3 - cosB / W7 * W2 + y2s * cosB / W7s * W2 / rb + y2s * cosB / W7 * a / rb3) - W8 / rb * T10 + y2s * W8 / rb3 * T10 - y2 * W8 / rb * (-2 * a / rb4 * y2 - 1 / W6s / rb * y2) + W8 * cosB / rb / W7 * T9 - y2s * W8 * cosB / rb3 / W7 * T9 - y2s * W8 * cosB / rb2 / W7s * T9 + y2 * W8 * cosB / rb / W7 * (1 / rb * cosB * y2 / W7 * W2 - W1 / W7s * W2 / rb * y2 - W1 / W7 * a / rb3 * y2 - 2 * a * y3b / rb4 * y2)) + b1 / 2 * K * (N1 * (((2 - 2 * nu) * cot2 - nu) / rb * y1 / W6 - ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB * D7 / W7) + N1 / W6s * U1 / rb * y1 - N1 / W6 * ((1 - W5) * cotB + a * y1s * cotB / rb3 - y2s / W6s * W4 / rb * y1 - y2s / W6 * a / rb3 * y1) - N1 * cosB * cotB / W7 * W2 + N1 * z1b * cotB / W7s * W2 * D7 + N1 * z1b * cotB / W7 * a / rb3 * y1 - a * W8 * cotB / rb3 + 3 * a * y1s * W8 * cotB / rb5 - W8 / W6s * U2 / rb * y1 + W8 / W6 * (-1 / rb3 * (N1 * y1 * cotB - a) * y1 + 1 / rb * N1 * cotB - y2s / rb3 / W6 * W5 * y1 - y2s / rb2 / W6s * W5 * y1 - y2s / rb4 / W6 * a * y1 - 3 * a * y2s / rb5 * y1) - W8 / W7s * U3 * D7 + W8 / W7 * (1 / rb3 * (N1 * z1b * cotB + a * cosB) * y1 - 1 / rb * N1 * cosB * cotB + a * y3b * cosB * cotB / rb3 - 3 * a * y3b * z1b * cotB / rb5 * y1 + 1 / rb3 / W7 * U3i * y1 + 1 / rb / W7s * U3i * D7 - 1 / rb / W7 * (-a * cosB * cotB / rb * W1 + a * z1b * cotB / rb3 * W1 * y1 - a * z1b * cotB / rb2 * cosB * y1))) + b2 / 2 * K * ((2 - 2 * nu) * N1 * rFib_ry1 * cot2 - N1 * y2 / W6s * U4 / rb * y1 + N1 * y2 / W6 * (-a / rb3 * y1 * cotB + 1 / W6 * W4 - y1s / W6s * W4 / rb - y1s / W6 * a / rb3) + N1 * y2 * cotB / W7s * W9 * D7 + N1 * y2 * cotB / W7 * a / rb3 / cosB * y1 + 3 * a * y2 * W8 * cotB / rb5 * y1 - y2 * W8 / rb3 / W6 * U5 * y1 - y2 * W8 / rb2 / W6s * U5 * y1 + y2 * W8 / rb / W6 * (-2 * nu / W6 + 2 * nu * y1s / W6s / rb - a / rb * (1 / rb + 1 / W6) + a * y1s / rb3 * (1 / rb + 1 / W6) - a * y1 / rb * (-1 / rb3 * y1 - 1 / W6s / rb * y1)) - y2 * W8 * cotB / rb3 / W7 * U6 * y1 - y2 * W8 * cotB / rb / W7s * U6 * D7 + y2 * W8 * cotB / rb / W7 * (1 / rb * cosB * y1 / W7 * W9 - W1 / W7s * W9 * D7 - W1 / W7 * a / rb3 / cosB * y1 - 2 * a * y3b / rb4 / cosB * y1)) + b3 / 2 * K * (N1 * (-sinB * D7 / W7 - 1 / W6 * (1 + a / rb) + y1s / W6s * (1 + a / rb) / rb + y1s / W6 * a / rb3 + cosB / W7 * W2 - z1b / W7s * W2 * D7 - z1b / W7 * a / rb3 * y1) + W8 / rb * T10 - y1s * W8 / rb3 * T10 + y1 * W8 / rb * (-2 * a / rb4 * y1 - 1 / W6s / rb * y1) + W8 / W7s * U7 * D7 - W8 / W7 * (sinB * a / rb3 * y1 + cosB / rb * (1 + a * y3b / rb2) - z1b / rb3 * (1 + a * y3b / rb2) * y1 - 2 * z1b / rb5 * a * y3b * y1 + 1 / rb3 / W7 * U7i * y1 + 1 / rb / W7s * U7i * D7 - 1 / rb / W7 * (-a * cosB / rb * W1 + a * z1b / rb3 * W1 * y1 - a * z1b / rb2 * cosB * y1)));
    v[4] = b1 / 2 * K * ((-2 + 2 * nu) * N1 * rFib_ry3 * cot2 - N1 * y2 / W6s * T1 * R3 + N1 * y2 / W6 * (0.5 * a / rb3 * 2 * y3b * cotB + y1 / W6s * W4 * R3 + 0.5 * y1 / W6 * a / rb3 * 2 * y3b) - N1 * y2 * cosB * cotB / W7s * W2 * W3 - 0.5 * N1 * y2 * cosB * cotB / W7 * a / rb3 * 2 * y3b + a / rb3 * y2 * cotB - 1.5 * a * y2 * W8 * cotB / rb5 * 2 * y3b + y2 / rb / W6 * T2 - 0.5 * y2 * W8 / rb3 / W6 * T2 * 2 * y3b - y2 * W8 / rb / W6s * T2 * R3 + y2 * W8 / rb / W6 * (-y1 / W6s * W5 * R3 - 0.5 * y1 / W6 * a / rb3 * 2 * y3b - a * y1 / rb4 * 2 * y3b) + y2 / rb / W7 * T3 - 0.5 * y2 * W8 / rb3 / W7 * T3 * 2 * y3b - y2 * W8 / rb / W7s * T3 * W3 + y2 * W8 / rb / W7 * (-cosB / W7s * T3i * W3 + cosB / W7 * (C3 * (N1 * cosB - a / rb) * cotB + 0.5 * W1 * a / rb3 * 2 * y3b * cotB + 0.5 * (2 - 2 * nu) / rb * sinB * 2 * y3b * cosB) - a * cosB * cotB / rb2 + a * y3b * cosB * cotB / rb4 * 2 * y3b)) + b2 / 2 * K * (N1 * (((2 - 2 * nu) * cot2 + nu) * R3 / W6 - ((2 - 2 * nu) * cot2 + 1) * cosB * W3 / W7) - N1 / W6s * T4 * R3 + N1 / W6 * (nu - 0.5 * a * y1 * cotB / rb3 * 2 * y3b - y1s / W6s * W4 * R3 - 0.5 * y1s / W6 * a / rb3 * 2 * y3b) + N1 * cotB / W7s * T5 * W3 - N1 * cotB / W7 * (cosB * sinB - 0.5 * a / rb2 * sinB * 2 * y3b / cosB + 0.5 * a * (rb * sinB - y1) / rb3 / cosB * 2 * y3b) - a / rb3 * y1 * cotB + 1.5 * a * y1 * W8 * cotB / rb5 * 2 * y3b + 1 / W6 * T6 - W8 / W6s * T6 * R3 + W8 / W6 * (-0.5 / rb3 * (N1 * y1 * cotB + a) * 2 * y3b + 0.5 * y1s / rb3 / W6 * W5 * 2 * y3b + y1s / rb / W6s * W5 * R3 + 0.5 * y1s / rb4 / W6 * a * 2 * y3b + 1.5 * a * y1s / rb5 * 2 * y3b) + cotB / W7 * T8 - W8 * cotB / W7s * T8 * W3 + W8 * cotB / W7 * (a / rb3 / cosB * y1 - 1.5 * a * y1 * y3b / rb5 / cosB * 2 * y3b + 0.5 / rb2 * sinB * 2 * y3b * T7 - 0.5 * (rb * sinB - y1) / rb3 * T7 * 2 * y3b + (rb * sinB - y1) / rb * (-C3 / W7 * W9 + W1 / W7s * W9 * W3 + 0.5 * W1 / W7 * a / rb3 / cosB * 2 * y3b))) + b3 / 2 * K * (N1 * (-y2 / W6s * (1 + a / rb) * R3 - 0.5 * y2 / W6 * a / rb3 * 2 * y3b + y2 * cosB / W7s * W2 * W3 + 0.5 * y2 * cosB / W7 * a / rb3 * 2 * y3b) - y2 / rb * T10 + 0.5 * y2 * W8 / rb3 * T10 * 2 * y3b - y2 * W8 / rb * (-a / rb4 * 2 * y3b - 1 / W6s * R3) + y2 * cosB / rb / W7 * T9 - 0.5 * y2 * W8 * cosB / rb3 / W7 * T9 * 2 * y3b - y2 * W8 * cosB / rb / W7s * T9 * W3 + y2 * W8 * cosB / rb / W7 * (C3 / W7 * W2 - W1 / W7s * W2 * W3 - 0.5 * W1 / W7 * a / rb3 * 2 * y3b + a / rb2 - a * y3b / rb4 * 2 * y3b)) + b1 / 2 * K * ((2 - 2 * nu) * (N1 * rFib_ry1 * cotB - y1 / W6s * W5 / rb * y2 - y2 / W6 * a / rb3 * y1 + y2 * cosB / W7s * W2 * D7 + y2 * cosB / W7 * a / rb3 * y1) - y2 * W8 / rb3 * V1 * y1 + y2 * W8 / rb * (-2 * nu / W6s / rb * y1 - 2 * a / rb4 * y1) - y2 * W8 * cosB / rb3 / W7 * V2 * y1 - y2 * W8 * cosB / rb / W7s * V2 * D7 + y2 * W8 * cosB / rb / W7 * (-1 / rb * cosB * y1 / W7 * W2 + W1 / W7s * W2 * D7 + W1 / W7 * a / rb3 * y1 + 2 * a * y3b / rb4 * y1)) + b2 / 2 * K * ((-2 + 2 * nu) * N1 * cotB * (1 / rb * y1 / W6 - cosB * D7 / W7) - (2 - 2 * nu) / W6 * W5 + (2 - 2 * nu) * y1s / W6s * W5 / rb + (2 - 2 * nu) * y1s / W6 * a / rb3 + (2 - 2 * nu) * cosB / W7 * W2 - (2 - 2 * nu) * z1b / W7s * W2 * D7 - (2 - 2 * nu) * z1b / W7 * a / rb3 * y1 - W8 / rb3 * V3 * y1 + W8 / rb * (-2 * nu / W6 + 2 * nu * y1s / W6s / rb - a / rb2 + 2 * a * y1s / rb4) + W8 / W7s * V4 * D7 - W8 / W7 * (1 / rb2 * cosB * y1 * cotB * V4i - W1 * cotB / rb3 * V4i * y1 + W1 * cotB / rb * (-1 / rb * cosB * y1 / W7 + W1 / W7s * D7) - a / rb3 * V4s * y1 + a / rb * (-y3b * cosB / rb2 + 2 * y3b * z1b / rb4 * y1 - cosB * W1 / rb / W7 - z1b / rb2 * cosB * y1 / W7 + z1b * W1 / rb3 / W7 * y1 + z1b * W1 / rb / W7s * D7))) + b3 / 2 * K * ((2 - 2 * nu) * rFib_ry1 - (2 - 2 * nu) * y2 * sinB / W7s * W2 * D7 - (2 - 2 * nu) * y2 * sinB / W7 * a / rb3 * y1 - y2 * W8 * sinB / rb3 / W7 * V5 * y1 - y2 * W8 * sinB / rb / W7s * V5 * D7 + y2 * W8 * sinB / rb / W7 * (1 / rb * cosB * y1 / W7 * W2 - W1 / W7s * W2 * D7 - W1 / W7 * a / rb3 * y1 - 2 * a * y3b / rb4 * y1));
    v[5] = b1 / 2 * K * (N1 * (((2 - 2 * nu) * cot2 - nu) * R3 / W6 - ((2 - 2 * nu) * cot2 + 1 - 2 * nu) * cosB * W3 / W7) + N1 / W6s * U1 * R3 - N1 / W6 * (0.5 * a * y1 * cotB / rb3 * 2 * y3b + nu - y2s / W6s * W4 * R3 - 0.5 * y2s / W6 * a / rb3 * 2 * y3b) - N1 * sinB * cotB / W7 * W2 + N1 * z1b * cotB / W7s * W2 * W3 + 0.5 * N1 * z1b * cotB / W7 * a / rb3 * 2 * y3b - a / rb3 * y1 * cotB + 1.5 * a * y1 * W8 * cotB / rb5 * 2 * y3b + 1 / W6 * U2 - W8 / W6s * U2 * R3 + W8 / W6 * (-0.5 / rb3 * (N1 * y1 * cotB - a) * 2 * y3b - 0.5 * y2s / rb3 / W6 * W5 * 2 * y3b - y2s / rb / W6s * W5 * R3 - 0.5 * y2s / rb4 / W6 * a * 2 * y3b - 1.5 * a * y2s / rb5 * 2 * y3b) + 1 / W7 * U3 - W8 / W7s * U3 * W3 + W8 / W7 * (0.5 / rb3 * (N1 * z1b * cotB + a * cosB) * 2 * y3b - 1 / rb * N1 * sinB * cotB + a * z1b * cotB / rb3 + a * y3b * sinB * cotB / rb3 - 1.5 * a * y3b * z1b * cotB / rb5 * 2 * y3b + 0.5 / rb3 / W7 * U3i * 2 * y3b + 1 / rb / W7s * U3i * W3 - 1 / rb / W7 * (-a * sinB * cotB / rb * W1 + 0.5 * a * z1b * cotB / rb3 * W1 * 2 * y3b - a * z1b * cotB / rb * C3))) + b2 / 2 * K * ((2 - 2 * nu) * N1 * rFib_ry3 * cot2 - N1 * y2 / W6s * U4 * R3 + N1 * y2 / W6 * (-0.5 * a / rb3 * 2 * y3b * cotB - y1 / W6s * W4 * R3 - 0.5 * y1 / W6 * a / rb3 * 2 * y3b) + N1 * y2 * cotB / W7s * W9 * W3 + 0.5 * N1 * y2 * cotB / W7 * a / rb3 / cosB * 2 * y3b - a / rb3 * y2 * cotB + 1.5 * a * y2 * W8 * cotB / rb5 * 2 * y3b + y2 / rb / W6 * U5 - 0.5 * y2 * W8 / rb3 / W6 * U5 * 2 * y3b - y2 * W8 / rb / W6s * U5 * R3 + y2 * W8 / rb / W6 * (2 * nu * y1 / W6s * R3 + 0.5 * a * y1 / rb3 * (1 / rb + 1 / W6) * 2 * y3b - a * y1 / rb * (-0.5 / rb3 * 2 * y3b - 1 / W6s * R3)) + y2 * cotB / rb / W7 * U6 - 0.5 * y2 * W8 * cotB / rb3 / W7 * U6 * 2 * y3b - y2 * W8 * cotB / rb / W7s * U6 * W3 + y2 * W8 * cotB / rb / W7 * (C3 / W7 * W9 - W1 / W7s * W9 * W3 - 0.5 * W1 / W7 * a / rb3 / cosB * 2 * y3b + a / rb2 / cosB - a * y3b / rb4 / cosB * 2 * y3b)) + b3 / 2 * K * (N1 * (-sinB * W3 / W7 + y1 / W6s * (1 + a / rb) * R3 + 0.5 * y1 / W6 * a / rb3 * 2 * y3b + sinB / W7 * W2 - z1b / W7s * W2 * W3 - 0.5 * z1b / W7 * a / rb3 * 2 * y3b) + y1 / rb * T10 - 0.5 * y1 * W8 / rb3 * T10 * 2 * y3b + y1 * W8 / rb * (-a / rb4 * 2 * y3b - 1 / W6s * R3) - 1 / W7 * U7 + W8 / W7s * U7 * W3 - W8 / W7 * (0.5 * sinB * a / rb3 * 2 * y3b + sinB / rb * (1 + a * y3b / rb2) - 0.5 * z1b / rb3 * (1 + a * y3b / rb2) * 2 * y3b + z1b / rb * (a / rb2 - a * y3b / rb4 * 2 * y3b) + 0.5 / rb3 / W7 * U7i * 2 * y3b + 1 / rb / W7s * U7i * W3 - 1 / rb / W7 * (-a * sinB / rb * W1 + 0.5 * a * z1b / rb3 * W1 * 2 * y3b - a * z1b / rb * C3))) + b1 / 2 * K * ((2 - 2 * nu) * (N1 * rFib_ry2 * cotB + 1 / W6 * W5 - y2s / W6s * W5 / rb - y2s / W6 * a / rb3 - cosB / W7 * W2 + y2s * cosB / W7s * W2 / rb + y2s * cosB / W7 * a / rb3) + W8 / rb * V1 - y2s * W8 / rb3 * V1 + y2 * W8 / rb * (-2 * nu / W6s / rb * y2 - 2 * a / rb4 * y2) + W8 * cosB / rb / W7 * V2 - y2s * W8 * cosB / rb3 / W7 * V2 - y2s * W8 * cosB / rb2 / W7s * V2 + y2 * W8 * cosB / rb / W7 * (-1 / rb * cosB * y2 / W7 * W2 + W1 / W7s * W2 / rb * y2 + W1 / W7 * a / rb3 * y2 + 2 * a * y3b / rb4 * y2)) + b2 / 2 * K * ((-2 + 2 * nu) * N1 * cotB * (1 / rb * y2 / W6 - cosB / rb * y2 / W7) + (2 - 2 * nu) * y1 / W6s * W5 / rb * y2 + (2 - 2 * nu) * y1 / W6 * a / rb3 * y2 - (2 - 2 * nu) * z1b / W7s * W2 / rb * y2 - (2 - 2 * nu) * z1b / W7 * a / rb3 * y2 - W8 / rb3 * V3 * y2 + W8 / rb * (2 * nu * y1 / W6s / rb * y2 + 2 * a * y1 / rb4 * y2) + W8 / W7s * V4 / rb * y2 - W8 / W7 * (1 / rb2 * cosB * y2 * cotB * V4i - W1 * cotB / rb3 * V4i * y2 + W1 * cotB / rb * (-cosB / rb * y2 / W7 + W1 / W7s / rb * y2) - a / rb3 * V4s * y2 + a / rb * (2 * y3b * z1b / rb4 * y2 - z1b / rb2 * cosB * y2 / W7 + z1b * W1 / rb3 / W7 * y2 + z1b * W1 / rb2 / W7s * y2))) + b3 / 2 * K * ((2 - 2 * nu) * rFib_ry2 + (2 - 2 * nu) * sinB / W7 * W2 - (2 - 2 * nu) * y2s * sinB / W7s * W2 / rb - (2 - 2 * nu) * y2s * sinB / W7 * a / rb3 + W8 * sinB / rb / W7 * V5 - y2s * W8 * sinB / rb3 / W7 * V5 - y2s * W8 * sinB / rb2 / W7s * V5 + y2 * W8 * sinB / rb / W7 * (1 / rb * cosB * y2 / W7 * W2 - W1 / W7s * W2 / rb * y2 - W1 / W7 * a / rb3 * y2 - 2 * a * y3b / rb4 * y2));
}

}  // namespace qdfn
