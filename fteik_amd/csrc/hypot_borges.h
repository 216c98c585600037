// hypot_borges.h -- sqrt(x*x + y*y) with the one-step correction of
// C. F. Borges, "An Improved Algorithm for hypot(a,b)", ACM TOMS 47(1),
// 2021 (arXiv:1904.09481), unfused form (no FMA).
//
// The reference's 2D analytic time is SZERO*HYPOT(dx, dz)
// (localSolver.F90:46); its Fortran runtime calls the C library's hypot,
// which on glibc >= 2.35 evaluates this same corrected formula.  Writing it
// out here gives the device the identical rounding.  Arguments are grid
// distances in metres: no scaling for overflow or underflow is needed.
// Compiles for host and device so a host build can check it against the C
// library.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define FTEIK_HD __host__ __device__
#else
#define FTEIK_HD
#endif

namespace fteik {

FTEIK_HD inline double hypot_borges(double x, double y)
{
    double ax = fabs(x), ay = fabs(y);
    if (ax < ay) { const double t = ax; ax = ay; ay = t; }
    if (ay == 0.0) return ax;
    double h = sqrt(ax*ax + ay*ay);
    double t1, t2;
    if (h <= 2.0*ay) {
        const double d = h - ay;
        t1 = ax*(2.0*d - ax);
        t2 = (d - 2.0*(ax - ay))*d;
    } else {
        const double d = h - ax;
        t1 = 2.0*d*(ax - 2.0*ay);
        t2 = (4.0*d - ay)*ay + d*d;
    }
    return h - (t1 + t2)/(2.0*h);
}

} // namespace fteik
