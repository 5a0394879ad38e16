/* ORACLE (test infrastructure only).  Degree-argument sine/cosine with the
 * argument-reduction rules of Julia's Base `sind`/`cosd`, which the reference
 * calls throughout (e.g. Functions_BuildInputFile.jl:430-431, :450-451,
 * :480-488, :497).  What matters for parity is that multiples of 90 degrees
 * give exact 0 / +-1 and that reduction happens in degrees before the
 * conversion to radians.  The final conversion here is a plain double
 * product (Julia uses an extended-precision product), so results may differ
 * from Julia by about one ulp away from the exact points. */
#ifndef QDFN_ORACLE_DEGTRIG_H
#define QDFN_ORACLE_DEGTRIG_H
#include <math.h>

static inline double orc_d2r(double x) { return x * (3.141592653589793 / 180.0); }

static inline double orc_sind(double x) {
    if (isinf(x) || isnan(x)) return NAN;
    double rx = copysign(fmod(x, 360.0), x);
    double arx = fabs(rx);
    if (rx == 0.0) return rx;
    if (arx < 45.0) return sin(orc_d2r(rx));
    if (arx <= 135.0) return copysign(cos(orc_d2r(90.0 - arx)), rx);
    if (arx == 180.0) return copysign(0.0, rx);
    if (arx < 225.0) return sin(orc_d2r((180.0 - arx) * (rx > 0 ? 1.0 : -1.0)));
    if (arx <= 315.0) return -copysign(cos(orc_d2r(270.0 - arx)), rx);
    return sin(orc_d2r(rx - copysign(360.0, rx)));
}

static inline double orc_cosd(double x) {
    if (isinf(x) || isnan(x)) return NAN;
    double rx = fabs(fmod(x, 360.0));
    if (rx <= 45.0) return cos(orc_d2r(rx));
    if (rx < 135.0) return sin(orc_d2r(90.0 - rx));
    if (rx <= 225.0) return -cos(orc_d2r(rx - 180.0));
    if (rx < 315.0) return sin(orc_d2r(rx - 270.0));
    return cos(orc_d2r(360.0 - rx));
}
#endif
