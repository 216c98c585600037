// stencils.h -- device-side local solvers of the factored eikonal fast
// sweeping method (3D and 2D), written for gfx950 in plain HIP.
//
// Operation order follows the reference exactly so the FP64 path reproduces
// its rounding (the library is compiled with -ffp-contract=off):
//   3D iteration stencil   localSolver.F90:384-534
//   3D init stencil        localSolver.F90:538-751
//   3D analytic helpers    localSolver.F90:775-851
//   2D iteration stencil   localSolver.F90:123-158
//   2D init stencil        localSolver.F90:195-289
// The templates are instantiated for double (parity path) and float (the
// single-precision sweeps of the mixed-precision mode).
#pragma once
#include <hip/hip_runtime.h>
#include "hypot_borges.h"

namespace fteik {

constexpr double kHuge = 99999.0;

// Per-source solver constants (localSolver.F90:878-931 in 3D,
// localSolver.F90:299-352 in 2D).  Filled on the device by the source-init
// kernels so every value is computed in the same arithmetic as the sweeps.
template <typename R>
struct Consts {
    R dz, dx, dy;
    R dzi, dxi, dyi, dz2i, dx2i, dy2i;
    R dsum_nine, dsumi, epsSolver;
    R dz2dx2, dz2dy2, dx2dy2;
    R dz2i_dx2i, dz2i_dy2i, dx2i_dy2i;
    R dz2i_p_dx2i, dz2i_p_dy2i, dx2i_p_dy2i;
    R dz2i_p_dy2i_inv, dz2i_p_dx2i_inv, dx2i_p_dy2i_inv;
    R szero, szero2, zsa, xsa, ysa;
    int zsi, xsi, ysi, pad;
};

// Minimum and maximum of travel times and slownesses.  The operands are
// never NaN (every square root is guarded) and never -0, so the hardware
// min/max (one v_min_f64/v_max_f64 instead of a compare and two selects)
// returns exactly what the reference's MIN/MAX return.
__device__ __forceinline__ double rmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double rmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float rmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double rsqrt_exact(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float rsqrt_exact(float x) { return __builtin_sqrtf(x); }

// Minimum / maximum of operands read from memory.  The compiler precedes a
// minnum of a loaded value with a canonicalizing max (it cannot rule out a
// signalling NaN); the operands here are never NaN, so the hardware
// instruction alone returns the same value.
__device__ __forceinline__ double lmin(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double lmax(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float lmin(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float lmax(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Correctly rounded square root of a positive normal x >= 2^-767 (the
// arguments of the 2D operators of the iteration stencil wherever their
// gates are open): the reciprocal square-root estimate refined by
// Newton-Raphson steps in fused multiply-adds (Markstein's method), the
// sequence the compiler's general square root runs on that range, without
// its rescaling of tiny arguments and its zero/infinity selection.
__device__ __forceinline__ double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x*y;
    double h = y*0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float sqrt_pos(float x) { return __builtin_sqrtf(x); }

// localSolver.F90:775-792
template <typename R>
__device__ __forceinline__ R tana3d(int i, int j, int k, const Consts<R>& c)
{
    R dz = (R(i) - c.zsa)*c.dz;
    R dx = (R(j) - c.xsa)*c.dx;
    R dy = (R(k) - c.ysa)*c.dy;
    R d0 = dz*dz + dx*dx + dy*dy;
    return c.szero*rsqrt_exact(d0);
}

// localSolver.F90:819-851
template <typename R>
__device__ __forceinline__ void tanad3d(int i, int j, int k, const Consts<R>& c,
                                        R& t0c, R& tzc, R& txc, R& tyc)
{
    R dz = (R(i) - c.zsa)*c.dz;
    R dx = (R(j) - c.xsa)*c.dx;
    R dy = (R(k) - c.ysa)*c.dy;
    R d0 = dz*dz + dx*dx + dy*dy;
    R d0i = R(0);
    t0c = R(0);
    if (d0 > R(0)) {
        R sq = rsqrt_exact(d0);
        t0c = c.szero*sq;
        d0i = c.szero/sq;
    }
    tzc = d0i*dz;
    txc = d0i*dx;
    tyc = d0i*dy;
}

// Upwind neighbour times of one node, in the reference's order
// (autoCode.f90:10-17): tv, te, tn, tev, ten, tnv, tnve.
template <typename R>
struct Nbr { R tv, te, tn, tev, ten, tnv, tnve; };

// 7 min-reduced slownesses (autoCode.f90:439-445).
template <typename R>
struct Slw { R s1, s2, s3, s4, s5, s6, s7; };

// localSolver.F90:384-534 for one node; tupd_in is the node's current time.
// Every operator is evaluated unconditionally and the reference's gates
// only select its value (or HUGE): the four square roots are independent,
// so the compiler can interleave them instead of branching around each.
// Where a gate is closed the value computed under it is discarded (it may
// be NaN); where it is open the value is computed in the reference's
// operation order, so results are bitwise those of the gated form.
template <typename R>
__device__ __forceinline__ R noinit3d(const Consts<R>& c, const Nbr<R>& t,
                                      const Slw<R>& s, R tupd_in)
{
    const R huge = R(kHuge);
    R t12min = t.tv + c.dz*s.s1;
    t12min = rmin(t12min, t.te + c.dx*s.s2);
    t12min = rmin(t12min, t.tn + c.dy*s.s3);
    // ZX plane
    const R temtv = t.te - t.tv;
    R t1_2d;
    {
        const R sref2 = s.s4*s.s4;
        const R ta = t.tev + temtv;
        const R tb = t.tev - temtv;
        const R tab = ta - tb;
        const R tab2 = tab*tab;
        const R four_sref2 = R(4)*sref2;
        t1_2d = ((tb*c.dz2i + ta*c.dx2i)
              + sqrt_pos(four_sref2*c.dz2i_p_dx2i - c.dz2i_dx2i*tab2))
              *c.dz2i_p_dx2i_inv;
    }
    const bool g1 = temtv < c.dz*s.s4 && -temtv < c.dx*s.s4;
    // ZY plane
    const R tvmtn = t.tv - t.tn;
    R t2_2d;
    {
        const R sref2 = s.s5*s.s5;
        const R ta = tvmtn + t.tnv;
        const R tb = -tvmtn + t.tnv;
        const R tab = ta - tb;
        const R tab2 = tab*tab;
        const R four_sref2 = R(4)*sref2;
        t2_2d = ((ta*c.dz2i + tb*c.dy2i)
              + sqrt_pos(four_sref2*c.dz2i_p_dy2i - c.dz2i_dy2i*tab2))
              *c.dz2i_p_dy2i_inv;
    }
    const bool g2 = tvmtn < c.dy*s.s5 && -tvmtn < c.dz*s.s5;
    // XY plane
    const R temtn = t.te - t.tn;
    R t3_2d;
    {
        const R sref2 = s.s6*s.s6;
        const R ta = temtn + t.ten;
        const R tb = -temtn + t.ten;
        const R tab = ta - tb;
        const R tab2 = tab*tab;
        const R four_sref2 = R(4)*sref2;
        t3_2d = ((ta*c.dx2i + tb*c.dy2i)
              + sqrt_pos(four_sref2*c.dx2i_p_dy2i - c.dx2i_dy2i*tab2))
              *c.dx2i_p_dy2i_inv;
    }
    const bool g3 = temtn < c.dy*s.s6 && -temtn < c.dx*s.s6;
    // 3D operator
    R t3d_v, t2, t3;
    {
        const R sref2 = s.s7*s.s7;
        const R ta = t.te + R(0.5)*(-t.tn + t.ten - t.tv + t.tev) - t.tnv + t.tnve;
        const R tb = t.tv + R(0.5)*(-t.tn + t.tnv - t.te + t.tev) - t.ten + t.tnve;
        const R tc = t.tn + R(0.5)*(-t.te + t.ten - t.tv + t.tnv) - t.tev + t.tnve;
        const R tab = ta - tb;
        const R tbc = tb - tc;
        const R tac = ta - tc;
        t2 = sref2*c.dsum_nine;
        const R tab2 = tab*tab;
        const R tbc2 = tbc*tbc;
        const R tac2 = tac*tac;
        t3 = c.dz2dx2*tab2 + c.dz2dy2*tbc2 + c.dx2dy2*tac2;
        const R t1 = tb*c.dz2i + ta*c.dx2i + tc*c.dy2i;
        t3d_v = (t1 + rsqrt_exact(t2 - t3))*c.dsumi;
    }
    // the reference's order of minimisation
    t12min = rmin(t12min, g1 ? t1_2d : huge);
    t12min = rmin(t12min, g2 ? t2_2d : huge);
    t12min = rmin(t12min, g3 ? t3_2d : huge);
    const R tmax = lmax(lmax(t.tv, t.te), t.tn);
    const R t3d = (t12min > tmax && t2 >= t3) ? t3d_v : huge;
    return lmin(tupd_in, rmin(t3d, t12min));
}

// localSolver.F90:538-751 for one node.  (i, j, k) are 1-based, sgnt* are
// the sweep's time signs, sgnr*_d*i = sgnt/d (autoCode.f90:1097-1099).
template <typename R>
__device__ __forceinline__ R init3d(const Consts<R>& c, const Nbr<R>& t,
                                    const Slw<R>& s, R tt0, int i, int j,
                                    int k, int sgntz, int sgntx, int sgnty)
{
    const R huge = R(kHuge);
    const R sgnrz_dzi = R(sgntz)/c.dz;
    const R sgnrx_dxi = R(sgntx)/c.dx;
    const R sgnry_dyi = R(sgnty)/c.dy;
    const R tv = t.tv, te = t.te, tn = t.tn, tev = t.tev, ten = t.ten,
            tnv = t.tnv, tnve = t.tnve;
    R t1d = tv + c.dz*s.s1;
    t1d = rmin(t1d, te + c.dx*s.s2);
    t1d = rmin(t1d, tn + c.dy*s.s3);
    R tmin = rmin(rmin(tv, te), tn);
    R tmax = rmax(rmax(tv, te), tn);
    bool lcart = (tmin > c.epsSolver);
    R t0c = R(0), tzc = R(0), txc = R(0), tyc = R(0);
    R tauv = huge, taue = huge, taun = huge, tauev = huge, tauen = huge,
      taunv = huge, taunve = huge;
    if (!lcart) {
        tanad3d(i, j, k, c, t0c, tzc, txc, tyc);
        tauv   = tv   - tana3d(i - sgntz, j,         k,         c);
        taue   = te   - tana3d(i,         j - sgntx, k,         c);
        taun   = tn   - tana3d(i,         j,         k - sgnty, c);
        tauev  = tev  - tana3d(i - sgntz, j - sgntx, k,         c);
        tauen  = ten  - tana3d(i,         j - sgntx, k - sgnty, c);
        taunv  = tnv  - tana3d(i - sgntz, j,         k - sgnty, c);
        taunve = tnve - tana3d(i - sgntz, j - sgntx, k - sgnty, c);
    }
    bool ldoZX = (tv < te + c.dx*s.s4) && (te < tv + c.dz*s.s4);
    bool ldoZY = (tv < tn + c.dy*s.s5) && (tn < tv + c.dz*s.s5);
    bool ldoXY = (te < tn + c.dy*s.s6) && (tn < te + c.dx*s.s6);
    bool ldoZXCart = ldoZX && (lcart || k != c.ysi);
    bool ldoZYCart = ldoZY && (lcart || j != c.xsi);
    bool ldoXYCart = ldoXY && (lcart || i != c.zsi);
    bool ldoZXSphere = ldoZX && !ldoZXCart;
    bool ldoZYSphere = ldoZY && !ldoZYCart;
    bool ldoXYSphere = ldoXY && !ldoXYCart;
    // ZX plane
    R t1_2d = huge;
    if (ldoZXCart) {
        R sref2 = s.s4*s.s4;
        R ta = tev + te - tv;
        R tb = tev - te + tv;
        R tab = ta - tb;
        R tab2 = tab*tab;
        R four_sref2 = R(4)*sref2;
        t1_2d = ((tb*c.dz2i + ta*c.dx2i)
              + rsqrt_exact(four_sref2*c.dz2i_p_dx2i - c.dz2i_dx2i*tab2))
              *c.dz2i_p_dx2i_inv;
    }
    if (ldoZXSphere) {
        R sref2 = s.s4*s.s4;
        R ta = tauev + taue - tauv;
        R tb = tauev - taue + tauv;
        R apoly = c.dz2i_p_dx2i;
        R bpoly = R(4)*(sgnrx_dxi*txc + sgnrz_dzi*tzc)
                - R(2)*(ta*c.dx2i + tb*c.dz2i);
        R cpoly = ((ta*ta)*c.dx2i) + ((tb*tb)*c.dz2i)
                - R(4)*(sgnrx_dxi*txc*ta + sgnrz_dzi*tzc*tb)
                + R(4)*(c.szero2 - sref2 + tyc*tyc);
        R dpoly = bpoly*bpoly - R(4)*apoly*cpoly;
        if (dpoly >= R(0)) t1_2d = (rsqrt_exact(dpoly) - bpoly)/(R(2)*apoly) + t0c;
        if (t1_2d < tv || t1_2d < te) t1_2d = huge;
    }
    R t1d_t2d_min = rmin(t1d, t1_2d);
    // ZY plane
    R t2_2d = huge;
    if (ldoZYCart) {
        R sref2 = s.s5*s.s5;
        R ta = tv - tn + tnv;
        R tb = tn - tv + tnv;
        R tab = ta - tb;
        R tab2 = tab*tab;
        R four_sref2 = R(4)*sref2;
        t2_2d = ((ta*c.dz2i + tb*c.dy2i)
              + rsqrt_exact(four_sref2*c.dz2i_p_dy2i - c.dz2i_dy2i*tab2))
              *c.dz2i_p_dy2i_inv;
    }
    if (ldoZYSphere) {
        R sref2 = s.s5*s.s5;
        R ta = tauv - taun + taunv;
        R tb = taun - tauv + taunv;
        R apoly = c.dz2i_p_dy2i;
        R bpoly = R(4)*(sgnry_dyi*tyc + sgnrz_dzi*tzc)
                - R(2)*(ta*c.dz2i + tb*c.dy2i);
        R cpoly = ((ta*ta)*c.dz2i) + ((tb*tb)*c.dy2i)
                - R(4)*(sgnrz_dzi*tzc*ta + sgnry_dyi*tyc*tb)
                + R(4)*(c.szero2 - sref2 + txc*txc);
        R dpoly = bpoly*bpoly - R(4)*apoly*cpoly;
        if (dpoly >= R(0)) t2_2d = (rsqrt_exact(dpoly) - bpoly)/(R(2)*apoly) + t0c;
        if (t2_2d < tv || t2_2d < tn) t2_2d = huge;
    }
    t1d_t2d_min = rmin(t1d_t2d_min, t2_2d);
    // XY plane
    R t3_2d = huge;
    if (ldoXYCart) {
        R sref2 = s.s6*s.s6;
        R ta = te - tn + ten;
        R tb = tn - te + ten;
        R tab = ta - tb;
        R tab2 = tab*tab;
        R four_sref2 = R(4)*sref2;
        t3_2d = ((ta*c.dx2i + tb*c.dy2i)
              + rsqrt_exact(four_sref2*c.dx2i_p_dy2i - c.dx2i_dy2i*tab2))
              *c.dx2i_p_dy2i_inv;
    }
    if (ldoXYSphere) {
        R sref2 = s.s6*s.s6;
        R ta = taue - taun + tauen;
        R tb = taun - taue + tauen;
        R apoly = c.dx2i_p_dy2i;
        R bpoly = R(4)*(sgnry_dyi*tyc + sgnrx_dxi*txc)
                - R(2)*(ta*c.dx2i + tb*c.dy2i);
        R cpoly = ((ta*ta)*c.dx2i) + ((tb*tb)*c.dy2i)
                - R(4)*(sgnrx_dxi*txc*ta + sgnry_dyi*tyc*tb)
                + R(4)*(c.szero2 - sref2 + tzc*tzc);
        R dpoly = bpoly*bpoly - R(4)*apoly*cpoly;
        if (dpoly >= R(0)) t3_2d = (rsqrt_exact(dpoly) - bpoly)/(R(2)*apoly) + t0c;
        if (t3_2d < te || t3_2d < tn) t3_2d = huge;
    }
    t1d_t2d_min = rmin(t1d_t2d_min, t3_2d);
    // 3D operators
    R t3d = huge;
    bool ldo3D = (t1d_t2d_min > tmax);
    if (ldo3D && lcart) {
        R sref2 = s.s7*s.s7;
        R ta = te + R(0.5)*(-tn + ten - tv + tev) - tnv + tnve;
        R tb = tv + R(0.5)*(-tn + tnv - te + tev) - ten + tnve;
        R tc = tn + R(0.5)*(-te + ten - tv + tnv) - tev + tnve;
        R tab = ta - tb;
        R tbc = tb - tc;
        R tac = ta - tc;
        R t2 = sref2*c.dsum_nine;
        R tab2 = tab*tab;
        R tbc2 = tbc*tbc;
        R tac2 = tac*tac;
        R t3 = c.dz2dx2*tab2 + c.dz2dy2*tbc2 + c.dx2dy2*tac2;
        if (t2 >= t3) {
            R t1 = tb*c.dz2i + ta*c.dx2i + tc*c.dy2i;
            t3d = (t1 + rsqrt_exact(t2 - t3))*c.dsumi;
        }
    }
    if (ldo3D && !lcart) {
        R sref2 = s.s7*s.s7;
        R ta = taue + R(0.5)*(-taun + tauen - tauv + tauev) - taunv + taunve;
        R tb = tauv + R(0.5)*(-taun + taunv - taue + tauev) - tauen + taunve;
        R tc = taun + R(0.5)*(-taue + tauen - tauv + taunv) - tauev + taunve;
        R apoly = c.dx2i + c.dz2i + c.dy2i;
        R bpoly = R(6)*(sgnrz_dzi*tzc + sgnry_dyi*tyc + sgnrx_dxi*txc)
                - R(2)*(ta*c.dx2i + tb*c.dz2i + tc*c.dy2i);
        R cpoly = ((ta*ta)*c.dx2i) + ((tb*tb)*c.dz2i) + ((tc*tc)*c.dy2i)
                - R(6)*(sgnrx_dxi*txc*ta + sgnrz_dzi*tzc*tb + sgnry_dyi*tyc*tc)
                + R(9)*(c.szero2 - sref2);
        R dpoly = bpoly*bpoly - R(4)*apoly*cpoly;
        if (dpoly >= R(0)) t3d = (rsqrt_exact(dpoly) - bpoly)/(R(2)*apoly) + t0c;
        if (t3d < te || t3d < tn || t3d < tv) t3d = huge;
    }
    R tupd = rmin(t1d_t2d_min, t3d);
    return rmin(tt0, tupd);
}

// ------------------------------------------------------------------ 2D
// localSolver.F90:37-50 (HYPOT(diffx, diffz))
__device__ __forceinline__ double tana2d(int i, int j, const Consts<double>& c)
{
    double dz = (double(i) - c.zsa)*c.dz;
    double dx = (double(j) - c.xsa)*c.dx;
    return c.szero*hypot_borges(dx, dz);
}

// localSolver.F90:74-112
__device__ __forceinline__ void tanad2d(int i, int j, const Consts<double>& c,
                                        double& t0c, double& tzc, double& txc)
{
    double dz = (double(i) - c.zsa)*c.dz;
    double dx = (double(j) - c.xsa)*c.dx;
    double d0 = dz*dz + dx*dx, d0i = 0.0;
    t0c = 0.0;
    if (d0 > 0.0) {
        double sq = __builtin_sqrt(d0);
        t0c = c.szero*sq;
        d0i = c.szero/sq;
    }
    tzc = d0i*dz;
    txc = d0i*dx;
}

// localSolver.F90:123-158: tv, te, tev upwind; s1 (V), s2 (WE), s3 (cell).
__device__ __forceinline__ double noinit2d(const Consts<double>& c, double tv,
                                           double te, double tev, double s1,
                                           double s2, double s3)
{
    double temtv = te - tv;
    double t12min = rmin(tv + c.dz*s1, te + c.dx*s2);
    double t1_2d = kHuge;
    if (temtv < c.dz*s3 && -temtv < c.dx*s3) {
        double sref2 = s3*s3;
        double ta = tev + temtv;
        double tb = tev - temtv;
        double tab = ta - tb;
        double tab2 = tab*tab;
        double four_sref2 = 4.0*sref2;
        t1_2d = ((tb*c.dz2i + ta*c.dx2i)
              + __builtin_sqrt(four_sref2*c.dz2i_p_dx2i - c.dz2i_dx2i*tab2))
              *c.dz2i_p_dx2i_inv;
    }
    return rmin(t12min, t1_2d);
}

// localSolver.F90:195-289 (iz, ix 1-based).  eps is in grid points.
__device__ __forceinline__ double init2d(const Consts<double>& c, int sgntz,
                                         int sgntx, int iz, int ix, double tv,
                                         double te, double tev, double t,
                                         double sref1, double sref2,
                                         double sref3)
{
    const double huge = kHuge;
    const double sgnrz = double(sgntz), sgnrx = double(sgntx);
    double t1d = rmin(tv + c.dz*sref1, te + c.dx*sref2);
    double t0c, tzc, txc;
    tanad2d(iz, ix, c, t0c, tzc, txc);
    double tauv  = tv  - tana2d(iz - sgntz, ix, c);
    double taue  = te  - tana2d(iz, ix - sgntx, c);
    double tauev = tev - tana2d(iz - sgntz, ix - sgntx, c);
    double t1 = huge, t2 = huge, t3 = huge;
    int aiz = iz - c.zsi; if (aiz < 0) aiz = -aiz;
    int aix = ix - c.xsi; if (aix < 0) aix = -aix;
    if (double(aiz) > c.epsSolver || double(aix) > c.epsSolver) {
        if (tv <= te + c.dx*sref3 && te <= tv + c.dz*sref3 &&
            te >= tev && tv >= tev) {
            double ta = tev + te - tv;
            double tb = tev - te + tv;
            t1 = ((tb*c.dz2i + ta*c.dx2i)
               + __builtin_sqrt(4.0*sref3*sref3*(c.dz2i + c.dx2i)
                                - c.dz2i*c.dx2i*(ta - tb)*(ta - tb)))
               /(c.dz2i + c.dx2i);
        } else if ((te - tev) <= c.dz*c.dz*sref3/__builtin_sqrt(c.dx*c.dx + c.dz*c.dz)
                   && te > tev) {
            double q = (te - tev)/c.dz;
            t2 = te + c.dx*__builtin_sqrt(sref3*sref3 - q*q);
        } else if ((tv - tev) <= c.dx*c.dx*sref3/__builtin_sqrt(c.dx*c.dx + c.dz*c.dz)
                   && tv > tev) {
            double q = (tv - tev)/c.dx;
            t3 = tv + c.dz*__builtin_sqrt(sref3*sref3 - q*q);
        }
    } else {
        if (tv < te + c.dx*sref3 && te < tv + c.dz*sref3 &&
            te >= tev && tv >= tev) {
            double ta = tauev + taue - tauv;
            double tb = tauev - taue + tauv;
            double apoly = c.dz2i + c.dx2i;
            double bpoly = 4.0*(sgnrx*txc*c.dxi + sgnrz*tzc*c.dzi)
                         - 2.0*(ta*c.dx2i + tb*c.dz2i);
            double cpoly = (ta*ta*c.dx2i) + (tb*tb*c.dz2i)
                         - 4.0*(sgnrx*txc*c.dxi*ta + sgnrz*tzc*c.dzi*tb)
                         + 4.0*(c.szero*c.szero - sref3*sref3);
            double dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
            if (dpoly >= 0.0) t1 = 0.5*(__builtin_sqrt(dpoly) - bpoly)/apoly + t0c;
            if (t1 < tv || t1 < te) t1 = huge;
        }
    }
    double t2d = rmin(rmin(t1, t2), t3);
    return rmin(rmin(t, t1d), t2d);
}

} // namespace fteik
