/* ORACLE — test infrastructure only (see okada.c header).
 *
 * Forward-mode dual numbers for a second build of the loop oracle
 * (sim.c, solvers.c, hmat.c compiled as C++ with this header forced in):
 * every `double` of those files becomes qd_dual, a value v and its derivative
 * d along one direction, so one run of the restated main() returns the
 * derivative of every output along that direction — the stand-in, here, for
 * a ForwardDiff pass through the reference's Julia code, which cannot run.
 * Comparisons, branches and argmins act on values: the derivative is that of
 * the discrete path the run takes, as the device adjoint's is.  The secant
 * iterations are differentiated as iterated (their converged derivative
 * equals the implicit-function one to about the convergence tolerance). */
#ifndef DUAL_ORACLE_HPP
#define DUAL_ORACLE_HPP
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <cmath>

struct qd_dual {
    double v, d;
    qd_dual() = default;
    qd_dual(double x) : v(x), d(0.0) {}
    qd_dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    qd_dual& operator+=(qd_dual o) { d += o.d; v += o.v; return *this; }
    qd_dual& operator-=(qd_dual o) { d -= o.d; v -= o.v; return *this; }
    qd_dual& operator*=(qd_dual o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    qd_dual& operator/=(qd_dual o) { const double q = v / o.v; d = (d - q * o.d) / o.v; v = q; return *this; }
};
inline qd_dual operator+(qd_dual a, qd_dual b) { return qd_dual(a.v + b.v, a.d + b.d); }
inline qd_dual operator-(qd_dual a, qd_dual b) { return qd_dual(a.v - b.v, a.d - b.d); }
inline qd_dual operator-(qd_dual a) { return qd_dual(-a.v, -a.d); }
inline qd_dual operator+(qd_dual a) { return a; }
inline qd_dual operator*(qd_dual a, qd_dual b) { return qd_dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline qd_dual operator/(qd_dual a, qd_dual b) { const double q = a.v / b.v; return qd_dual(q, (a.d - q * b.d) / b.v); }
inline bool operator<(qd_dual a, qd_dual b) { return a.v < b.v; }
inline bool operator>(qd_dual a, qd_dual b) { return a.v > b.v; }
inline bool operator<=(qd_dual a, qd_dual b) { return a.v <= b.v; }
inline bool operator>=(qd_dual a, qd_dual b) { return a.v >= b.v; }
inline bool operator==(qd_dual a, qd_dual b) { return a.v == b.v; }
inline bool operator!=(qd_dual a, qd_dual b) { return a.v != b.v; }
inline qd_dual exp(qd_dual a) { const double e = std::exp(a.v); return qd_dual(e, e * a.d); }
inline qd_dual log(qd_dual a) { return qd_dual(std::log(a.v), a.d / a.v); }
inline qd_dual sqrt(qd_dual a) { const double s = std::sqrt(a.v); return qd_dual(s, a.d / (2.0 * s)); }
inline qd_dual sin(qd_dual a) { return qd_dual(std::sin(a.v), std::cos(a.v) * a.d); }
inline qd_dual cos(qd_dual a) { return qd_dual(std::cos(a.v), -std::sin(a.v) * a.d); }
inline qd_dual fabs(qd_dual a) { return a.v < 0.0 ? -a : (a.v > 0.0 ? a : qd_dual(a.v, std::fabs(a.d))); }
inline bool isnan(qd_dual a) { return std::isnan(a.v); }

#define double qd_dual
#endif
