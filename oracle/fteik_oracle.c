/*
 * fteik_oracle.c -- CPU restatement of the reference fteik algorithms.
 *
 * TEST INFRASTRUCTURE ONLY (see fteik_oracle.h).  Every function cites the
 * reference file:line it restates.  Floating-point expressions keep the
 * reference's evaluation order (Fortran evaluates same-precedence operators
 * left to right) and the file is compiled with -ffp-contract=off, so no FMA
 * contraction changes the rounding.
 *
 * Pinning (what checks each part; details in tests/):
 *   - local solvers, mesh constants, seeds, source/receiver/model helpers,
 *     the slowness/travel-time gathers, the 2D solver and the locator are
 *     compared bit-for-bit with the reference build in oracle/_ref and with
 *     the committed fixtures in tests/golden;
 *   - the 3D level-set driver (masks, sweep order, init pass) is pinned by
 *     brute-force level/mask checks, level-order invariance (level-
 *     synchronous schedule == lexicographic schedule, bitwise), agreement
 *     with the analytic solutions and with the original Noble-Gesret solver
 *     (refsrc/fteik.f90) in oracle/_ref.
 */
#include "fteik_oracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double dmin2(double a, double b) { return (a < b) ? a : b; }
static inline double dmax2(double a, double b) { return (a > b) ? a : b; }
static inline int imin2(int a, int b) { return (a < b) ? a : b; }
static inline int imax2(int a, int b) { return (a > b) ? a : b; }

/* ===================================================================== */
/*                                 Model                                 */
/* ===================================================================== */

/* model.F90:771-780 */
int or_grid2index(int i, int j, int k, int nz, int nzx)
{
    return (k - 1)*nzx + (j - 1)*nz + i;
}

/* model.F90:827-836 */
int or_velgrid2index(int i, int j, int k, int nzm1, int nzm1_nxm1)
{
    return (k - 1)*nzm1_nxm1 + (j - 1)*nzm1 + i;
}

/* model.F90:794-813 */
int or_index2grid(int igrd, int nz, int nx, int ny, int *i, int *j, int *k)
{
    int nzx = nz*nx, igrd1 = igrd - 1, ierr = 0;
    *k = igrd1/nzx;
    *j = (igrd1 - (*k)*nzx)/nz;
    *i = igrd1 - (*k)*nzx - (*j)*nz;
    *k += 1; *j += 1; *i += 1;
    if (*i < 1 || *i > nz) ierr++;
    if (*j < 1 || *j > nx) ierr++;
    if (*k < 1 || *k > ny) ierr++;
    return ierr;
}

/* model.F90:574-607: slow = 1/vel for cells in natural order. */
int or_velocity_slowness(int ncell, const double *vel, double *slow)
{
    int i;
    if (ncell < 1) return 1;
    for (i = 0; i < ncell; i++) if (vel[i] <= 0.0) return 1;
    for (i = 0; i < ncell; i++) slow[i] = 1.0/vel[i];
    return 0;
}

/* model.F90:469-559: cell velocity model in natural (0), XYZ (1) or ZYX (2)
 * ordering.  2D uses natural (0) or XZ (1). */
int or_cell_slowness(int is3d, int nz, int nx, int ny, int order, int nc,
                     const double *vel, double *slow)
{
    int ncell = (nz - 1)*(nx - 1)*(is3d ? (ny - 1) : 1);
    int nzm1 = nz - 1, nzm1_nxm1 = (nz - 1)*(nx - 1);
    int ix, iy, iz, icell, jcell, i;
    if (nc < ncell || nc < 1) return 1;
    for (i = 0; i < nc; i++) if (vel[i] <= 0.0) return 1;
    if (is3d && order == 1)
    {
        for (iy = 1; iy <= ny - 1; iy++)
            for (ix = 1; ix <= nx - 1; ix++)
                for (iz = 1; iz <= nz - 1; iz++)
                {
                    jcell = (iz - 1)*(nx - 1)*(ny - 1) + (iy - 1)*(nx - 1) + ix;
                    icell = or_velgrid2index(iz, ix, iy, nzm1, nzm1_nxm1);
                    slow[icell - 1] = vel[jcell - 1];
                }
    }
    else if (is3d && order == 2)
    {
        for (iy = 1; iy <= ny - 1; iy++)
            for (ix = 1; ix <= nx - 1; ix++)
                for (iz = 1; iz <= nz - 1; iz++)
                {
                    jcell = (ix - 1)*(nz - 1)*(ny - 1) + (iy - 1)*(nz - 1) + iz;
                    icell = or_velgrid2index(iz, ix, iy, nzm1, nzm1_nxm1);
                    slow[icell - 1] = vel[jcell - 1];
                }
    }
    else if (!is3d && order == 1)
    {
        for (ix = 1; ix <= nx - 1; ix++)
            for (iz = 1; iz <= nz - 1; iz++)
            {
                jcell = (iz - 1)*(nx - 1) + ix;
                icell = (ix - 1)*(nz - 1) + iz;
                slow[icell - 1] = vel[jcell - 1];
            }
    }
    else
    {
        for (i = 0; i < ncell; i++) slow[i] = 1.0/vel[i];
        return 0;
    }
    for (i = 0; i < ncell; i++) if (slow[i] <= 0.0) return 1;
    for (i = 0; i < ncell; i++) slow[i] = 1.0/slow[i];
    return 0;
}

/* model.F90:311-440: nodal velocity model averaged onto cells.  Keeps the
 * reference behaviour exactly, including the missing 1/8 in the XYZ branch
 * (model.F90:353) and the iy (not iy+1) used for nodes 5-8 of the natural
 * branch (model.F90:386-389). */
int or_nodal_slowness(int is3d, int nz, int nx, int ny, int order, int ng,
                      const double *vel, double *slow)
{
    int nyl = is3d ? ny : 1;
    int ngrd = nz*nx*nyl, nzx = nz*nx;
    int ncell = (nz - 1)*(nx - 1)*(is3d ? (ny - 1) : 1);
    int nzm1 = nz - 1, nzm1_nxm1 = (nz - 1)*(nx - 1);
    int ix, iy, iz, icell, i, n;
    int indx[8];
    double vmin, s;
    if (ng < ngrd || ng < 1) return 1;
    vmin = vel[0];
    for (i = 1; i < ng; i++) vmin = dmin2(vmin, vel[i]);
    if (vmin <= 0.0) return 1;
    for (i = 0; i < ncell; i++) slow[i] = 0.0;
    if (is3d)
    {
        for (iy = 1; iy <= ny - 1; iy++)
            for (ix = 1; ix <= nx - 1; ix++)
                for (iz = 1; iz <= nz - 1; iz++)
                {
                    if (order == 1)
                    {
                        indx[0] = (iz - 1)*nx*ny + (iy - 1)*nx + ix;
                        indx[1] = (iz - 1)*nx*ny + (iy - 1)*nx + ix + 1;
                        indx[2] = (iz - 1)*nx*ny + (iy    )*nx + ix;
                        indx[3] = (iz - 1)*nx*ny + (iy    )*nx + ix + 1;
                        indx[4] = (iz    )*nx*ny + (iy - 1)*nx + ix;
                        indx[5] = (iz    )*nx*ny + (iy - 1)*nx + ix + 1;
                        indx[6] = (iz    )*nx*ny + (iy    )*nx + ix;
                        indx[7] = (iz    )*nx*ny + (iy    )*nx + ix + 1;
                    }
                    else if (order == 2)
                    {
                        indx[0] = (ix - 1)*ny*nz + (iy - 1)*nz + iz;
                        indx[1] = (ix - 1)*ny*nz + (iy - 1)*nz + iz + 1;
                        indx[2] = (ix - 1)*ny*nz + (iy    )*nz + iz;
                        indx[3] = (ix - 1)*ny*nz + (iy    )*nz + iz + 1;
                        indx[4] = (ix    )*ny*nz + (iy - 1)*nz + iz;
                        indx[5] = (ix    )*ny*nz + (iy - 1)*nz + iz + 1;
                        indx[6] = (ix    )*ny*nz + (iy    )*nz + iz;
                        indx[7] = (ix    )*ny*nz + (iy    )*nz + iz + 1;
                    }
                    else
                    {
                        indx[0] = or_grid2index(iz,     ix,     iy, nz, nzx);
                        indx[1] = or_grid2index(iz + 1, ix,     iy, nz, nzx);
                        indx[2] = or_grid2index(iz,     ix + 1, iy, nz, nzx);
                        indx[3] = or_grid2index(iz + 1, ix + 1, iy, nz, nzx);
                        indx[4] = or_grid2index(iz,     ix,     iy, nz, nzx);
                        indx[5] = or_grid2index(iz + 1, ix,     iy, nz, nzx);
                        indx[6] = or_grid2index(iz,     ix + 1, iy, nz, nzx);
                        indx[7] = or_grid2index(iz + 1, ix + 1, iy, nz, nzx);
                    }
                    s = 0.0;
                    for (n = 0; n < 8; n++) s = s + vel[indx[n] - 1];
                    icell = or_velgrid2index(iz, ix, iy, nzm1, nzm1_nxm1);
                    slow[icell - 1] = (order == 1) ? s : s*0.125;
                }
    }
    else
    {
        for (iz = 1; iz <= nz - 1; iz++)
            for (ix = 1; ix <= nx - 1; ix++)
            {
                if (order == 1)
                {
                    indx[0] = (iz - 1)*nx + ix;
                    indx[1] = (iz - 1)*nx + ix + 1;
                    indx[2] = (iz    )*nx + ix;
                    indx[3] = (iz    )*nx + ix + 1;
                }
                else
                {
                    indx[0] = (ix - 1)*nz + iz;
                    indx[1] = (ix - 1)*nz + iz + 1;
                    indx[2] = (ix    )*nz + iz;
                    indx[3] = (ix    )*nz + iz + 1;
                }
                s = 0.0;
                for (n = 0; n < 4; n++) s = s + vel[indx[n] - 1];
                icell = (ix - 1)*(nz - 1) + iz;
                slow[icell - 1] = s*0.25;
            }
    }
    for (i = 0; i < ncell; i++) if (slow[i] <= 0.0) return 1;
    for (i = 0; i < ncell; i++) slow[i] = 1.0/slow[i];
    return 0;
}

/* ===================================================================== */
/*                                Source                                 */
/* ===================================================================== */

/* source.F90:136-190 and source.F90:429-464 (szero). */
int or_source_locate(int is3d, int nz, int nx, int ny,
                     double dz, double dx, double dy,
                     double z0, double x0, double y0,
                     double zs, double xs, double ys,
                     const double *slow, or_source *src)
{
    const double perturb = 0.0001;
    double zmax, xmax, ymax, zl, xl, yl, zsa, xsa, ysa;
    int zsi, xsi, ysi, nyl = is3d ? ny : 1;
    if (!is3d) { dy = 1.0; y0 = 0.0; ys = 0.0; }
    zmax = (double) (nz - 1)*dz;
    xmax = (double) (nx - 1)*dx;
    ymax = (double) (nyl - 1)*dy;
    zl = zs - z0; xl = xs - x0; yl = ys - y0;
    if (zl < 0.0 || zl > zmax || xl < 0.0 || xl > xmax ||
        yl < 0.0 || yl > ymax) return 1;
    zsa = (zl/dz) + 1.0;
    xsa = (xl/dx) + 1.0;
    ysa = (yl/dy) + 1.0;
    if (fabs(zsa - 1.0) < DBL_EPSILON) zsa = zsa + perturb;
    if (zsa >= (double) nz)            zsa = zsa - perturb;
    if (fabs(xsa - 1.0) < DBL_EPSILON) xsa = xsa + perturb;
    if (xsa >= (double) nx)            xsa = xsa - perturb;
    if (fabs(ysa - 1.0) < DBL_EPSILON) ysa = ysa + perturb;
    if (ysa >= (double) nyl)           ysa = ysa - perturb;
    zsi = (int) zsa;
    xsi = (int) xsa;
    ysi = (int) ysa;
    if (!is3d) ysi = 1;
    if (zsi < 1 || zsi > nz || xsi < 1 || xsi > nx ||
        (is3d && (ysi < 1 || ysi > ny))) return 1;
    if (!is3d) ysa = 1.0;
    src->zsa = zsa; src->xsa = xsa; src->ysa = ysa;
    src->zsi = zsi; src->xsi = xsi; src->ysi = ysi;
    src->szero = 0.0;
    if (slow)
    {
        src->szero = slow[or_velgrid2index(zsi, xsi, ysi, nz - 1,
                                           (nz - 1)*(nx - 1)) - 1];
    }
    return 0;
}

/* ===================================================================== */
/*                          3D local solver                              */
/* ===================================================================== */

/* localSolver.F90:878-931 (mesh constants, epsSolver). */
void or_consts3d_init(double dz, double dx, double dy, double eps,
                      const or_source *src, or_consts3d *c)
{
    double hmin = dmin2(dmin2(dz, dx), dy);
    memset(c, 0, sizeof(*c));
    c->dz = dz; c->dx = dx; c->dy = dy;
    c->szero = src->szero;
    c->szero2 = src->szero*src->szero;
    c->zsa = src->zsa; c->xsa = src->xsa; c->ysa = src->ysa;
    c->zsi = src->zsi; c->xsi = src->xsi; c->ysi = src->ysi;
    c->epsSolver = eps*hmin*src->szero;
    c->dz2 = dz*dz; c->dx2 = dx*dx; c->dy2 = dy*dy;
    c->dzi = 1.0/dz; c->dxi = 1.0/dx; c->dyi = 1.0/dy;
    c->dz2i = 1.0/c->dz2; c->dx2i = 1.0/c->dx2; c->dy2i = 1.0/c->dy2;
    c->dsum = c->dz2i + c->dx2i + c->dy2i;
    c->dsum_nine = c->dsum*9.0;
    c->dsumi = 1.0/c->dsum;
    c->dz2dx2 = 1.0/(c->dz2*c->dx2);
    c->dz2dy2 = 1.0/(c->dz2*c->dy2);
    c->dx2dy2 = 1.0/(c->dx2*c->dy2);
    c->dz2i_dx2i = c->dz2i*c->dx2i;
    c->dz2i_dy2i = c->dz2i*c->dy2i;
    c->dx2i_dy2i = c->dx2i*c->dy2i;
    c->dz2i_p_dy2i_inv = 1.0/(c->dz2i + c->dy2i);
    c->dz2i_p_dx2i_inv = 1.0/(c->dz2i + c->dx2i);
    c->dx2i_p_dy2i_inv = 1.0/(c->dx2i + c->dy2i);
    c->dz2i_p_dx2i = c->dz2i + c->dx2i;
    c->dz2i_p_dy2i = c->dz2i + c->dy2i;
    c->dx2i_p_dy2i = c->dx2i + c->dy2i;
}

/* localSolver.F90:775-792 */
double or_tana3d(int i, int j, int k, const or_consts3d *c)
{
    double diffz = ((double) i - c->zsa)*c->dz;
    double diffx = ((double) j - c->xsa)*c->dx;
    double diffy = ((double) k - c->ysa)*c->dy;
    double d0 = diffz*diffz + diffx*diffx + diffy*diffy;
    return c->szero*sqrt(d0);
}

/* localSolver.F90:819-851 */
void or_tanad3d(int i, int j, int k, const or_consts3d *c,
                double *tanad, double *tzc, double *txc, double *tyc)
{
    double diffz = ((double) i - c->zsa)*c->dz;
    double diffx = ((double) j - c->xsa)*c->dx;
    double diffy = ((double) k - c->ysa)*c->dy;
    double d0 = diffz*diffz + diffx*diffx + diffy*diffy;
    double d0i_szero = 0.0, sqrtd0;
    *tanad = 0.0;
    if (d0 > 0.0)
    {
        sqrtd0 = sqrt(d0);
        *tanad = c->szero*sqrtd0;
        d0i_szero = c->szero/sqrtd0;
    }
    *tzc = d0i_szero*diffz;
    *txc = d0i_szero*diffx;
    *tyc = d0i_szero*diffy;
}

/* localSolver.F90:933-956: 8 analytic seed times at the source cell corners
 * and their 0-based destinations in the natural-order field. */
void or_seeds3d(const or_consts3d *c, int nz, int nx, int *dest, double *ts8)
{
    static const int oz[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    static const int ox[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int oy[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    int n;
    for (n = 0; n < 8; n++)
    {
        ts8[n] = or_tana3d(c->zsi + oz[n], c->xsi + ox[n], c->ysi + oy[n], c);
        dest[n] = or_grid2index(c->zsi + oz[n], c->xsi + ox[n],
                                c->ysi + oy[n], nz, nz*nx) - 1;
    }
}

/* localSolver.F90:384-534 for one node.  t = [tv,te,tn,tev,ten,tnv,tnve],
 * s = [s1..s7]; tt0 is the node's current time (tupd on input). */
double or_noinit3d(const or_consts3d *c, const double *t, const double *s,
                   double tt0)
{
    const double tv = t[0], te = t[1], tn = t[2], tev = t[3], ten = t[4],
                 tnv = t[5], tnve = t[6];
    double t12min, t1_2d, t2_2d, t3_2d, t3d, tmax, sref2, ta, tb, tc, tab,
           tab2, tbc, tbc2, tac, tac2, four_sref2, temtv, tvmtn, temtn,
           t1, t2, t3;
    t12min = tv + c->dz*s[0];
    t12min = dmin2(t12min, te + c->dx*s[1]);
    t12min = dmin2(t12min, tn + c->dy*s[2]);
    /* ZX plane */
    t1_2d = OR_HUGE;
    temtv = te - tv;
    if (temtv < c->dz*s[3] && -temtv < c->dx*s[3])
    {
        sref2 = s[3]*s[3];
        ta = tev + temtv;
        tb = tev - temtv;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t1_2d = ((tb*c->dz2i + ta*c->dx2i)
              + sqrt(four_sref2*c->dz2i_p_dx2i - c->dz2i_dx2i*tab2))
              *c->dz2i_p_dx2i_inv;
    }
    t12min = dmin2(t12min, t1_2d);
    /* ZY plane */
    t2_2d = OR_HUGE;
    tvmtn = tv - tn;
    if (tvmtn < c->dy*s[4] && -tvmtn < c->dz*s[4])
    {
        sref2 = s[4]*s[4];
        ta = tvmtn + tnv;
        tb = -tvmtn + tnv;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t2_2d = ((ta*c->dz2i + tb*c->dy2i)
              + sqrt(four_sref2*c->dz2i_p_dy2i - c->dz2i_dy2i*tab2))
              *c->dz2i_p_dy2i_inv;
    }
    t12min = dmin2(t12min, t2_2d);
    /* XY plane */
    t3_2d = OR_HUGE;
    temtn = te - tn;
    if (temtn < c->dy*s[5] && -temtn < c->dx*s[5])
    {
        sref2 = s[5]*s[5];
        ta = temtn + ten;
        tb = -temtn + ten;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t3_2d = ((ta*c->dx2i + tb*c->dy2i)
              + sqrt(four_sref2*c->dx2i_p_dy2i - c->dx2i_dy2i*tab2))
              *c->dx2i_p_dy2i_inv;
    }
    t12min = dmin2(t12min, t3_2d);
    /* 3D operator */
    t3d = OR_HUGE;
    tmax = dmax2(dmax2(tv, te), tn);
    if (t12min > tmax)
    {
        sref2 = s[6]*s[6];
        ta = te + 0.5*(-tn + ten - tv + tev) - tnv + tnve;
        tb = tv + 0.5*(-tn + tnv - te + tev) - ten + tnve;
        tc = tn + 0.5*(-te + ten - tv + tnv) - tev + tnve;
        tab = ta - tb;
        tbc = tb - tc;
        tac = ta - tc;
        t2 = sref2*c->dsum_nine;
        tab2 = tab*tab;
        tbc2 = tbc*tbc;
        tac2 = tac*tac;
        t3 = c->dz2dx2*tab2 + c->dz2dy2*tbc2 + c->dx2dy2*tac2;
        if (t2 >= t3)
        {
            t1 = tb*c->dz2i + ta*c->dx2i + tc*c->dy2i;
            t3d = (t1 + sqrt(t2 - t3))*c->dsumi;
        }
    }
    t3d = dmin2(t3d, t12min);
    return dmin2(tt0, t3d);
}

/* Sweep sign tables: prefetch.py:4-8 (sgnt) and :42-44 (sgnv); the sweep
 * order 1..8 is the one of autoCode.f90:1028-1031 and its siblings. */
void or_sweep_signs(int sweep, int *sgntz, int *sgntx, int *sgnty,
                    int *sgnvz, int *sgnvx, int *sgnvy)
{
    static const int tz[8] = { 1,  1,  1,  1, -1, -1, -1, -1};
    static const int tx[8] = { 1, -1,  1, -1,  1, -1,  1, -1};
    static const int ty[8] = { 1,  1, -1, -1,  1,  1, -1, -1};
    int s = sweep - 1;
    *sgntz = tz[s]; *sgntx = tx[s]; *sgnty = ty[s];
    *sgnvz = tz[s] > 0 ? 1 : 0;
    *sgnvx = tx[s] > 0 ? 1 : 0;
    *sgnvy = ty[s] > 0 ? 1 : 0;
}

/* localSolver.F90:538-751 for one node.  sgnr*_d*i = sgnr/d (autoCode.f90:
 * 1097-1099), which equals +-1/d exactly. */
double or_init3d(const or_consts3d *c, const double *t, const double *s,
                 double tt0, int i, int j, int k, int sweep)
{
    const double tv = t[0], te = t[1], tn = t[2], tev = t[3], ten = t[4],
                 tnv = t[5], tnve = t[6];
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy;
    double sgnrz_dzi, sgnrx_dxi, sgnry_dyi;
    double t1d, tmin, tmax, t0c, tzc, txc, tyc, tauv, taue, taun, tauev,
           tauen, taunv, taunve, t1_2d, t2_2d, t3_2d, t1d_t2d_min, t3d,
           sref2, ta, tb, tc, tab, tab2, tbc, tbc2, tac, tac2, four_sref2,
           apoly, bpoly, cpoly, dpoly, t1, t2, t3;
    int lcart, ldoZX, ldoZY, ldoXY, ldoZXCart, ldoZYCart, ldoXYCart,
        ldoZXSphere, ldoZYSphere, ldoXYSphere, ldo3D;
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    sgnrz_dzi = (double) sgntz/c->dz;
    sgnrx_dxi = (double) sgntx/c->dx;
    sgnry_dyi = (double) sgnty/c->dy;
    t1d = tv + c->dz*s[0];
    t1d = dmin2(t1d, te + c->dx*s[1]);
    t1d = dmin2(t1d, tn + c->dy*s[2]);
    tmin = dmin2(dmin2(tv, te), tn);
    tmax = dmax2(dmax2(tv, te), tn);
    lcart = (tmin > c->epsSolver);
    t0c = 0.0; tzc = 0.0; txc = 0.0; tyc = 0.0;
    tauv = OR_HUGE; taue = OR_HUGE; taun = OR_HUGE; tauev = OR_HUGE;
    tauen = OR_HUGE; taunv = OR_HUGE; taunve = OR_HUGE;
    if (!lcart)
    {
        or_tanad3d(i, j, k, c, &t0c, &tzc, &txc, &tyc);
        tauv   = tv   - or_tana3d(i - sgntz, j,         k,         c);
        taue   = te   - or_tana3d(i,         j - sgntx, k,         c);
        taun   = tn   - or_tana3d(i,         j,         k - sgnty, c);
        tauev  = tev  - or_tana3d(i - sgntz, j - sgntx, k,         c);
        tauen  = ten  - or_tana3d(i,         j - sgntx, k - sgnty, c);
        taunv  = tnv  - or_tana3d(i - sgntz, j,         k - sgnty, c);
        taunve = tnve - or_tana3d(i - sgntz, j - sgntx, k - sgnty, c);
    }
    ldoZX = (tv < te + c->dx*s[3]) && (te < tv + c->dz*s[3]);
    ldoZY = (tv < tn + c->dy*s[4]) && (tn < tv + c->dz*s[4]);
    ldoXY = (te < tn + c->dy*s[5]) && (tn < te + c->dx*s[5]);
    ldoZXCart = ldoZX && (lcart || k != c->ysi);
    ldoZYCart = ldoZY && (lcart || j != c->xsi);
    ldoXYCart = ldoXY && (lcart || i != c->zsi);
    ldoZXSphere = ldoZX && !ldoZXCart;
    ldoZYSphere = ldoZY && !ldoZYCart;
    ldoXYSphere = ldoXY && !ldoXYCart;
    /* ZX plane */
    t1_2d = OR_HUGE;
    if (ldoZXCart)
    {
        sref2 = s[3]*s[3];
        ta = tev + te - tv;
        tb = tev - te + tv;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t1_2d = ((tb*c->dz2i + ta*c->dx2i)
              + sqrt(four_sref2*c->dz2i_p_dx2i - c->dz2i_dx2i*tab2))
              *c->dz2i_p_dx2i_inv;
    }
    if (ldoZXSphere)
    {
        sref2 = s[3]*s[3];
        ta = tauev + taue - tauv;
        tb = tauev - taue + tauv;
        apoly = c->dz2i_p_dx2i;
        bpoly = 4.0*(sgnrx_dxi*txc + sgnrz_dzi*tzc)
              - 2.0*(ta*c->dx2i + tb*c->dz2i);
        cpoly = ((ta*ta)*c->dx2i) + ((tb*tb)*c->dz2i)
              - 4.0*(sgnrx_dxi*txc*ta + sgnrz_dzi*tzc*tb)
              + 4.0*(c->szero2 - sref2 + tyc*tyc);
        dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
        if (dpoly >= 0.0) t1_2d = (sqrt(dpoly) - bpoly)/(2.0*apoly) + t0c;
        if (t1_2d < tv || t1_2d < te) t1_2d = OR_HUGE;
    }
    t1d_t2d_min = dmin2(t1d, t1_2d);
    /* ZY plane */
    t2_2d = OR_HUGE;
    if (ldoZYCart)
    {
        sref2 = s[4]*s[4];
        ta = tv - tn + tnv;
        tb = tn - tv + tnv;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t2_2d = ((ta*c->dz2i + tb*c->dy2i)
              + sqrt(four_sref2*c->dz2i_p_dy2i - c->dz2i_dy2i*tab2))
              *c->dz2i_p_dy2i_inv;
    }
    if (ldoZYSphere)
    {
        sref2 = s[4]*s[4];
        ta = tauv - taun + taunv;
        tb = taun - tauv + taunv;
        apoly = c->dz2i_p_dy2i;
        bpoly = 4.0*(sgnry_dyi*tyc + sgnrz_dzi*tzc)
              - 2.0*(ta*c->dz2i + tb*c->dy2i);
        cpoly = ((ta*ta)*c->dz2i) + ((tb*tb)*c->dy2i)
              - 4.0*(sgnrz_dzi*tzc*ta + sgnry_dyi*tyc*tb)
              + 4.0*(c->szero2 - sref2 + txc*txc);
        dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
        if (dpoly >= 0.0) t2_2d = (sqrt(dpoly) - bpoly)/(2.0*apoly) + t0c;
        if (t2_2d < tv || t2_2d < tn) t2_2d = OR_HUGE;
    }
    t1d_t2d_min = dmin2(t1d_t2d_min, t2_2d);
    /* XY plane */
    t3_2d = OR_HUGE;
    if (ldoXYCart)
    {
        sref2 = s[5]*s[5];
        ta = te - tn + ten;
        tb = tn - te + ten;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t3_2d = ((ta*c->dx2i + tb*c->dy2i)
              + sqrt(four_sref2*c->dx2i_p_dy2i - c->dx2i_dy2i*tab2))
              *c->dx2i_p_dy2i_inv;
    }
    if (ldoXYSphere)
    {
        sref2 = s[5]*s[5];
        ta = taue - taun + tauen;
        tb = taun - taue + tauen;
        apoly = c->dx2i_p_dy2i;
        bpoly = 4.0*(sgnry_dyi*tyc + sgnrx_dxi*txc)
              - 2.0*(ta*c->dx2i + tb*c->dy2i);
        cpoly = ((ta*ta)*c->dx2i) + ((tb*tb)*c->dy2i)
              - 4.0*(sgnrx_dxi*txc*ta + sgnry_dyi*tyc*tb)
              + 4.0*(c->szero2 - sref2 + tzc*tzc);
        dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
        if (dpoly >= 0.0) t3_2d = (sqrt(dpoly) - bpoly)/(2.0*apoly) + t0c;
        if (t3_2d < te || t3_2d < tn) t3_2d = OR_HUGE;
    }
    t1d_t2d_min = dmin2(t1d_t2d_min, t3_2d);
    /* 3D operators */
    t3d = OR_HUGE;
    ldo3D = (t1d_t2d_min > tmax);
    if (ldo3D && lcart)
    {
        sref2 = s[6]*s[6];
        ta = te + 0.5*(-tn + ten - tv + tev) - tnv + tnve;
        tb = tv + 0.5*(-tn + tnv - te + tev) - ten + tnve;
        tc = tn + 0.5*(-te + ten - tv + tnv) - tev + tnve;
        tab = ta - tb;
        tbc = tb - tc;
        tac = ta - tc;
        t2 = sref2*c->dsum_nine;
        tab2 = tab*tab;
        tbc2 = tbc*tbc;
        tac2 = tac*tac;
        t3 = c->dz2dx2*tab2 + c->dz2dy2*tbc2 + c->dx2dy2*tac2;
        if (t2 >= t3)
        {
            t1 = tb*c->dz2i + ta*c->dx2i + tc*c->dy2i;
            t3d = (t1 + sqrt(t2 - t3))*c->dsumi;
        }
    }
    if (ldo3D && !lcart)
    {
        sref2 = s[6]*s[6];
        ta = taue + 0.5*(-taun + tauen - tauv + tauev) - taunv + taunve;
        tb = tauv + 0.5*(-taun + taunv - taue + tauev) - tauen + taunve;
        tc = taun + 0.5*(-taue + tauen - tauv + taunv) - tauev + taunve;
        apoly = c->dx2i + c->dz2i + c->dy2i;
        bpoly = 6.0*(sgnrz_dzi*tzc + sgnry_dyi*tyc + sgnrx_dxi*txc)
              - 2.0*(ta*c->dx2i + tb*c->dz2i + tc*c->dy2i);
        cpoly = ((ta*ta)*c->dx2i) + ((tb*tb)*c->dz2i) + ((tc*tc)*c->dy2i)
              - 6.0*(sgnrx_dxi*txc*ta + sgnrz_dzi*tzc*tb + sgnry_dyi*tyc*tc)
              + 9.0*(c->szero2 - sref2);
        dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
        if (dpoly >= 0.0) t3d = (sqrt(dpoly) - bpoly)/(2.0*apoly) + t0c;
        if (t3d < te || t3d < tn || t3d < tv) t3d = OR_HUGE;
    }
    t3d = dmin2(t1d_t2d_min, t3d);
    return dmin2(tt0, t3d);
}

/* ===================================================================== */
/*                        Levels, masks, gathers                         */
/* ===================================================================== */

/* solver3D.F90:493-565.  lim = [z1, z2, x1, x2, y1, y2] (1-based). */
int or_sweep_limits(int sweep, int linit, int nz, int nx, int ny,
                    int zsi, int xsi, int ysi, int *lim)
{
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy;
    if (sweep < 1 || sweep > 8) return 1;
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    if (sgntz > 0) { lim[0] = linit ? imax2(2, zsi) : 2; lim[1] = nz; }
    else           { lim[0] = linit ? zsi + 1 : nz - 1;  lim[1] = 1; }
    if (sgntx > 0) { lim[2] = linit ? imax2(2, xsi) : 2; lim[3] = nx; }
    else           { lim[2] = linit ? xsi + 1 : nx - 1;  lim[3] = 1; }
    if (sgnty > 0) { lim[4] = linit ? imax2(2, ysi) : 2; lim[5] = ny; }
    else           { lim[4] = linit ? ysi + 1 : ny - 1;  lim[5] = 1; }
    return 0;
}

/* graph3d.F90:122-353: level pointers (1-based, nLevels+1 entries) and the
 * (iz, ix, iy, igrd) quadruples of the given sweep, each level sorted by the
 * linear grid index. */
int or_levels3d(int sweep, int nz, int nx, int ny, int *levelPtr, int *ijkv)
{
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy;
    int nLevels = nx + ny + nz - 2, il, j, k, fi, fj, fk, i, n = 0;
    if (nz < 3 || nx < 3 || ny < 3 || sweep < 1 || sweep > 8) return 1;
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    levelPtr[0] = 1;
    for (il = 1; il <= nLevels; il++)
    {
        /* actual (k, j) ascending gives ascending linear index */
        for (k = 1; k <= ny; k++)
        {
            fk = (sgnty > 0) ? k : ny + 1 - k;
            for (j = 1; j <= nx; j++)
            {
                fj = (sgntx > 0) ? j : nx + 1 - j;
                fi = il + 2 - fj - fk;
                if (fi < 1 || fi > nz) continue;
                i = (sgntz > 0) ? fi : nz + 1 - fi;
                ijkv[4*n + 0] = i;
                ijkv[4*n + 1] = j;
                ijkv[4*n + 2] = k;
                ijkv[4*n + 3] = or_grid2index(i, j, k, nz, nz*nx);
                n++;
            }
        }
        levelPtr[il] = n + 1;
    }
    return 0;
}

/* autoCode.f90:389-447 (sweep 1; sweeps 2-8 differ only in sgnv, see
 * prefetch.py:38-153): 19 cell slownesses min-reduced to 7. */
void or_gather_slow3d(int sweep, int i, int j, int k, int nz, int nx, int ny,
                      const double *slow, double *s7)
{
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy;
    int nzm1 = nz - 1, nzm1_nxm1 = (nz - 1)*(nx - 1);
    int i1, j1, k1, im, jm, km, ip, jp, kp;
#define SL(a, b, c) slow[or_velgrid2index((a), (b), (c), nzm1, nzm1_nxm1) - 1]
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    i1 = i - sgnvz; j1 = j - sgnvx; k1 = k - sgnvy;
    im = imax2(i - 1, 1); jm = imax2(j - 1, 1); km = imax2(k - 1, 1);
    ip = imin2(i, nz - 1); jp = imin2(j, nx - 1); kp = imin2(k, ny - 1);
    s7[0] = dmin2(dmin2(dmin2(SL(i1, jm, km), SL(i1, jm, kp)),
                        SL(i1, jp, km)), SL(i1, jp, kp));
    s7[1] = dmin2(dmin2(dmin2(SL(im, j1, km), SL(ip, j1, km)),
                        SL(im, j1, kp)), SL(ip, j1, kp));
    s7[2] = dmin2(dmin2(dmin2(SL(im, jm, k1), SL(im, jp, k1)),
                        SL(ip, jm, k1)), SL(ip, jp, k1));
    s7[3] = dmin2(SL(i1, j1, km), SL(i1, j1, kp));
    s7[4] = dmin2(SL(i1, jm, k1), SL(i1, jp, k1));
    s7[5] = dmin2(SL(im, j1, k1), SL(ip, j1, k1));
    s7[6] = SL(i1, j1, k1);
#undef SL
}

/* autoCode.f90:23-46 (sweep 1; sweeps 2-8 differ only in sgnt). */
void or_gather_tt3d(int sweep, int i, int j, int k, int nz, int nx,
                    const double *tt, double *t7, double *tt8)
{
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy, nzx = nz*nx;
#define TT(a, b, c) tt[or_grid2index((a), (b), (c), nz, nzx) - 1]
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    t7[0] = TT(i - sgntz, j,         k);
    t7[1] = TT(i,         j - sgntx, k);
    t7[2] = TT(i,         j,         k - sgnty);
    t7[3] = TT(i - sgntz, j - sgntx, k);
    t7[4] = TT(i,         j - sgntx, k - sgnty);
    t7[5] = TT(i - sgntz, j,         k - sgnty);
    t7[6] = TT(i - sgntz, j - sgntx, k - sgnty);
    *tt8  = TT(i, j, k);
#undef TT
}

/* ===================================================================== */
/*                              3D solver                                */
/* ===================================================================== */

static inline void node_update3d(int sweep, int linit, int i, int j, int k,
                                 int nz, int nx, int ny, const double *slow,
                                 const or_consts3d *c, double *tt)
{
    double s7[7], t7[7], tt8;
    or_gather_slow3d(sweep, i, j, k, nz, nx, ny, slow, s7);
    or_gather_tt3d(sweep, i, j, k, nz, nx, tt, t7, &tt8);
    tt[or_grid2index(i, j, k, nz, nz*nx) - 1]
        = linit ? or_init3d(c, t7, s7, tt8, i, j, k, sweep)
                : or_noinit3d(c, t7, s7, tt8);
}

/* autoCode.f90:1001-1316 (fteik_evaluateSweep1LS64fF and siblings): one
 * sweep over the masked box.  schedule 0 visits the nodes lexicographically
 * in the sweep direction; schedule 1 visits them level by level (flipped
 * iz+ix+iy = const) with OpenMP over each level and a barrier per level,
 * like the reference.  Both are topological orders of the same dependency
 * graph, so they give bitwise-identical fields. */
int or_sweep3d(int sweep, int linit, int schedule, int nz, int nx, int ny,
               const double *slow, const or_consts3d *c, double *tt)
{
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy, lim[6];
    int zlo, zhi, xlo, xhi, ylo, yhi;
    if (or_sweep_limits(sweep, linit, nz, nx, ny, c->zsi, c->xsi, c->ysi,
                        lim) != 0) return 1;
    or_sweep_signs(sweep, &sgntz, &sgntx, &sgnty, &sgnvz, &sgnvx, &sgnvy);
    zlo = imin2(lim[0], lim[1]); zhi = imax2(lim[0], lim[1]);
    xlo = imin2(lim[2], lim[3]); xhi = imax2(lim[2], lim[3]);
    ylo = imin2(lim[4], lim[5]); yhi = imax2(lim[4], lim[5]);
    if (schedule == 0)
    {
        int a, b, d, i, j, k;
        for (a = 0; a <= yhi - ylo; a++)
        {
            k = (sgnty > 0) ? ylo + a : yhi - a;
            for (b = 0; b <= xhi - xlo; b++)
            {
                j = (sgntx > 0) ? xlo + b : xhi - b;
                for (d = 0; d <= zhi - zlo; d++)
                {
                    i = (sgntz > 0) ? zlo + d : zhi - d;
                    node_update3d(sweep, linit, i, j, k, nz, nx, ny, slow, c,
                                  tt);
                }
            }
        }
        return 0;
    }
    else
    {
        /* flipped 0-based coordinates within the box */
        int lz = zhi - zlo + 1, lx = xhi - xlo + 1, ly = yhi - ylo + 1;
        int nlev = lz + lx + ly - 2;
        #pragma omp parallel
        {
            int L, fk, fj, fi, i, j, k;
            for (L = 0; L < nlev; L++)
            {
                #pragma omp for schedule(static)
                for (fk = 0; fk < ly; fk++)
                {
                    int fjlo = imax2(0, L - fk - (lz - 1));
                    int fjhi = imin2(lx - 1, L - fk);
                    k = (sgnty > 0) ? ylo + fk : yhi - fk;
                    for (fj = fjlo; fj <= fjhi; fj++)
                    {
                        fi = L - fk - fj;
                        j = (sgntx > 0) ? xlo + fj : xhi - fj;
                        i = (sgntz > 0) ? zlo + fi : zhi - fi;
                        node_update3d(sweep, linit, i, j, k, nz, nx, ny, slow,
                                      c, tt);
                    }
                }
                /* implicit barrier of the omp for: one barrier per level */
            }
        }
        return 0;
    }
}

/* solver3D.F90:1156-1243: seeds, 8 init sweeps, nsweep x 8 sweeps. */
int or_solve3d(int nz, int nx, int ny, double dz, double dx, double dy,
               const double *slow, const or_source *src, int nsweep,
               double eps, int schedule, double *tt)
{
    or_consts3d c;
    int dest[8], n, s, kiter;
    double ts8[8];
    long ngrd = (long) nz*nx*ny, g;
    or_consts3d_init(dz, dx, dy, eps, src, &c);
    for (g = 0; g < ngrd; g++) tt[g] = OR_HUGE;
    or_seeds3d(&c, nz, nx, dest, ts8);
    for (n = 0; n < 8; n++) tt[dest[n]] = ts8[n];
    for (s = 1; s <= 8; s++)
        if (or_sweep3d(s, 1, schedule, nz, nx, ny, slow, &c, tt)) return 1;
    for (kiter = 0; kiter < nsweep; kiter++)
        for (s = 1; s <= 8; s++)
            if (or_sweep3d(s, 0, schedule, nz, nx, ny, slow, &c, tt))
                return 1;
    return 0;
}

/* solver3D.F90:844-897 */
int or_field_reorder3d(int nz, int nx, int ny, int order, const double *tt,
                       double *out)
{
    int ix, iy, iz, nzx = nz*nx;
    long indx, jndx, ngrd = (long) nz*nx*ny;
    if (order == 1)
    {
        for (iz = 1; iz <= nz; iz++)
            for (iy = 1; iy <= ny; iy++)
                for (ix = 1; ix <= nx; ix++)
                {
                    indx = (long) (iy - 1)*nzx + (ix - 1)*nz + iz;
                    jndx = (long) (iz - 1)*nx*ny + (iy - 1)*nx + ix;
                    out[jndx - 1] = tt[indx - 1];
                }
    }
    else if (order == 2)
    {
        for (ix = 1; ix <= nx; ix++)
            for (iy = 1; iy <= ny; iy++)
                for (iz = 1; iz <= nz; iz++)
                {
                    indx = (long) (iy - 1)*nzx + (ix - 1)*nz + iz;
                    jndx = (long) (ix - 1)*nz*ny + (iy - 1)*nz + iz;
                    out[jndx - 1] = tt[indx - 1];
                }
    }
    else
    {
        for (indx = 0; indx < ngrd; indx++) out[indx] = tt[indx];
    }
    return 0;
}

/* solver3D.F90:915-980 (XYZ and ZYX branches).  The natural-order branch of
 * the reference copies nodal values into the cell-sized buffer
 * (solver3D.F90:974, a non-conforming assignment); here the natural order
 * returns the cell averages in z-fastest cell order instead. */
int or_field_cell3d(int nz, int nx, int ny, int order, const double *tt,
                    double *out)
{
    int ix, iy, iz, n, nzx = nz*nx;
    long indx[8], jndx;
    double s;
    for (iy = 1; iy <= ny - 1; iy++)
        for (ix = 1; ix <= nx - 1; ix++)
            for (iz = 1; iz <= nz - 1; iz++)
            {
                indx[0] = (long) (iy - 1)*nzx + (ix - 1)*nz + iz;
                indx[1] = (long) (iy - 1)*nzx + (ix - 1)*nz + iz + 1;
                indx[2] = (long) (iy - 1)*nzx + (ix - 0)*nz + iz;
                indx[3] = (long) (iy - 1)*nzx + (ix - 0)*nz + iz + 1;
                indx[4] = (long) (iy - 0)*nzx + (ix - 1)*nz + iz;
                indx[5] = (long) (iy - 0)*nzx + (ix - 1)*nz + iz + 1;
                indx[6] = (long) (iy - 0)*nzx + (ix - 0)*nz + iz;
                indx[7] = (long) (iy - 0)*nzx + (ix - 0)*nz + iz + 1;
                s = 0.0;
                for (n = 0; n < 8; n++) s = s + tt[indx[n] - 1];
                if (order == 1)
                    jndx = (long) (iz - 1)*(nx - 1)*(ny - 1)
                         + (iy - 1)*(nx - 1) + ix;
                else if (order == 2)
                    jndx = (long) (ix - 1)*(nz - 1)*(ny - 1)
                         + (iy - 1)*(nz - 1) + iz;
                else
                    jndx = (long) (iy - 1)*(nz - 1)*(nx - 1)
                         + (ix - 1)*(nz - 1) + iz;
                out[jndx - 1] = 0.125*s;
            }
    return 0;
}

/* ===================================================================== */
/*                              Receivers                                */
/* ===================================================================== */

/* receiver.F90:103-134.  ri = 1-based base node [3*nrec] (z, x, y);
 * rd = offsets in meters [3*nrec].  Returns the number of out-of-bounds
 * coordinates (the reference counts them in ierr and carries on). */
int or_receivers_locate(int is3d, int nz, int nx, int ny, double dz,
                        double dx, double dy, double z0, double x0, double y0,
                        int nrec, const double *z, const double *x,
                        const double *y, int *ri, double *rd)
{
    int irec, iz, ix, iy, ierr = 0, nyl = is3d ? ny : 1;
    double zr, xr, yr, yy;
    if (!is3d) { dy = 1.0; y0 = 0.0; }
    for (irec = 0; irec < nrec; irec++)
    {
        yy = is3d ? y[irec] : y0;
        if (z[irec] < z0 || z[irec] > z0 + dz*(double) (nz - 1)) ierr++;
        if (x[irec] < x0 || x[irec] > x0 + dx*(double) (nx - 1)) ierr++;
        if (yy < y0 || yy > y0 + dy*(double) (nyl - 1)) ierr++;
        zr = (z[irec] - z0)/dz;
        xr = (x[irec] - x0)/dx;
        yr = (yy - y0)/dy;
        iz = (int) (zr + 1.0);
        ix = (int) (xr + 1.0);
        iy = (int) (yr + 1.0);
        if (!is3d) iy = 1;
        ri[3*irec + 0] = imax2(1, imin2(nz - 1, iz));
        ri[3*irec + 1] = imax2(1, imin2(nx - 1, ix));
        ri[3*irec + 2] = imax2(1, imin2(nyl - 1, iy));
        rd[3*irec + 0] = (z[irec] - (z0 + dz*(double) (ri[3*irec + 0] - 1)));
        rd[3*irec + 1] = (x[irec] - (x0 + dx*(double) (ri[3*irec + 1] - 1)));
        rd[3*irec + 2] = (yy - (y0 + dy*(double) (ri[3*irec + 2] - 1)));
        if (!is3d) rd[3*irec + 2] = 0.0;
    }
    return ierr;
}

/* receiver.F90:246-306: trilinear (3D) / bilinear (2D) interpolation. */
void or_receivers_interp(int is3d, int nz, int nx, double dz, double dx,
                         double dy, int nrec, const int *ri, const double *rd,
                         const double *tt, double *ttr)
{
    int i, zi, xi, yi, nzx = nz*nx;
    double zd, xd, yd, omz, omx, c00, c01, c10, c11, c0, c1;
    for (i = 0; i < nrec; i++)
    {
        zi = ri[3*i]; xi = ri[3*i + 1]; yi = ri[3*i + 2];
        if (is3d)
        {
#define T3(a, b, c) tt[or_grid2index((a), (b), (c), nz, nzx) - 1]
            double t000 = T3(zi, xi, yi), t100 = T3(zi + 1, xi, yi);
            double t010 = T3(zi, xi + 1, yi), t110 = T3(zi + 1, xi + 1, yi);
            double t001 = T3(zi, xi, yi + 1), t101 = T3(zi + 1, xi, yi + 1);
            double t011 = T3(zi, xi + 1, yi + 1);
            double t111 = T3(zi + 1, xi + 1, yi + 1);
#undef T3
            zd = rd[3*i]/dz;
            xd = rd[3*i + 1]/dx;
            yd = rd[3*i + 2]/dy;
            omz = 1.0 - zd;
            c00 = t000*omz + t100*zd;
            c01 = t001*omz + t101*zd;
            c10 = t010*omz + t110*zd;
            c11 = t011*omz + t111*zd;
            omx = 1.0 - xd;
            c0 = c00*omx + c10*xd;
            c1 = c01*omx + c11*xd;
            ttr[i] = c0*(1.0 - yd) + c1*yd;
        }
        else
        {
            double t00 = tt[(xi - 1)*nz + zi - 1];
            double t10 = tt[(xi - 1)*nz + zi];
            double t01 = tt[xi*nz + zi - 1];
            double t11 = tt[xi*nz + zi];
            zd = rd[3*i]/dz;
            xd = rd[3*i + 1]/dx;
            omz = 1.0 - zd;
            omx = 1.0 - xd;
            ttr[i] = t00*omz*omx + t10*zd*omx + t01*omz*xd + t11*zd*xd;
        }
    }
}

/* ===================================================================== */
/*                              2D solver                                */
/* ===================================================================== */

typedef struct
{
    double dz, dx, dzi, dxi, dz2i, dx2i, dz2i_dx2i, dz2i_p_dx2i,
           dz2i_p_dx2i_inv, szero, zsa, xsa, eps;
    int zsi, xsi;
} or_consts2d;

/* localSolver.F90:299-352 */
static void consts2d_init(double dz, double dx, double eps,
                          const or_source *src, or_consts2d *c)
{
    double dz2 = dz*dz, dx2 = dx*dx;
    c->dz = dz; c->dx = dx;
    c->dzi = 1.0/dz; c->dxi = 1.0/dx;
    c->dz2i = 1.0/dz2; c->dx2i = 1.0/dx2;
    c->dz2i_dx2i = c->dz2i*c->dx2i;
    c->dz2i_p_dx2i = c->dz2i + c->dx2i;
    c->dz2i_p_dx2i_inv = 1.0/(c->dz2i + c->dx2i);
    c->szero = src->szero; c->zsa = src->zsa; c->xsa = src->xsa;
    c->zsi = src->zsi; c->xsi = src->xsi; c->eps = eps;
}

/* localSolver.F90:37-50 */
static double tana2d(int i, int j, const or_consts2d *c)
{
    double diffz = ((double) i - c->zsa)*c->dz;
    double diffx = ((double) j - c->xsa)*c->dx;
    return c->szero*hypot(diffx, diffz);
}

/* localSolver.F90:74-112 */
static void tanad2d(int i, int j, const or_consts2d *c, double *tanad,
                    double *tzc, double *txc)
{
    double diffz = ((double) i - c->zsa)*c->dz;
    double diffx = ((double) j - c->xsa)*c->dx;
    double d0 = diffz*diffz + diffx*diffx, d0i_szero = 0.0, sqrtd0;
    *tanad = 0.0;
    if (d0 > 0.0)
    {
        sqrtd0 = sqrt(d0);
        *tanad = c->szero*sqrtd0;
        d0i_szero = c->szero/sqrtd0;
    }
    *tzc = d0i_szero*diffz;
    *txc = d0i_szero*diffx;
}

/* localSolver.F90:160-193 */
static void signs2d(int sweep, int *sgntz, int *sgntx, int *sgnvz, int *sgnvx)
{
    *sgntz = (sweep <= 2) ? 1 : -1;
    *sgntx = (sweep == 1 || sweep == 3) ? 1 : -1;
    *sgnvz = (*sgntz > 0) ? 1 : 0;
    *sgnvx = (*sgntx > 0) ? 1 : 0;
}

/* localSolver.F90:123-158 for one node.  t4 = [tv, te, tev, tt],
 * s3 = [sV, sWE, scell]. */
double or_noinit2d(const double *t4, const double *s3, double dz, double dx)
{
    double dz2i = 1.0/(dz*dz), dx2i = 1.0/(dx*dx);
    double dz2i_dx2i = dz2i*dx2i, dz2i_p_dx2i = dz2i + dx2i;
    double dz2i_p_dx2i_inv = 1.0/(dz2i + dx2i);
    double tv = t4[0], te = t4[1], tev = t4[2];
    double temtv = te - tv, t12min, t1_2d, sref2, ta, tb, tab, tab2,
           four_sref2;
    t12min = dmin2(tv + dz*s3[0], te + dx*s3[1]);
    t1_2d = OR_HUGE;
    if (temtv < dz*s3[2] && -temtv < dx*s3[2])
    {
        sref2 = s3[2]*s3[2];
        ta = tev + temtv;
        tb = tev - temtv;
        tab = ta - tb;
        tab2 = tab*tab;
        four_sref2 = 4.0*sref2;
        t1_2d = ((tb*dz2i + ta*dx2i)
              + sqrt(four_sref2*dz2i_p_dx2i - dz2i_dx2i*tab2))
              *dz2i_p_dx2i_inv;
    }
    return dmin2(t12min, t1_2d);
}

/* solver2D.F90:474-570 (slowness) shared by the init and update stencils. */
static void gather_slow2d(int sweep, int nz, int nx, int iz, int ix,
                          const double *slow, double *s3)
{
    int sgntz, sgntx, sgnvz, sgnvx, izc, ixc;
    signs2d(sweep, &sgntz, &sgntx, &sgnvz, &sgnvx);
    izc = iz - sgnvz;
    ixc = ix - sgnvx;
    s3[0] = dmin2(slow[(imax2(ix - 1, 1) - 1)*(nz - 1) + izc - 1],
                  slow[(imin2(ix, nx - 1) - 1)*(nz - 1) + izc - 1]);
    s3[1] = dmin2(slow[(ixc - 1)*(nz - 1) + imax2(iz - 1, 1) - 1],
                  slow[(ixc - 1)*(nz - 1) + imin2(iz, nz - 1) - 1]);
    s3[2] = slow[(ixc - 1)*(nz - 1) + izc - 1];
}

/* localSolver.F90:195-289 */
static double init2d_c(int sweep, int nz, int nx, int iz, int ix,
                       const or_consts2d *c, const double *slow,
                       const double *tt)
{
    int sgntz, sgntx, sgnvz, sgnvx;
    double sgnrz, sgnrx, tv, te, tev, t, s3[3], sref1, sref2, sref3, t1d,
           t0c, tzc, txc, tauv, taue, tauev, t1, t2, t3, ta, tb, apoly, bpoly,
           cpoly, dpoly, t2d, q;
    signs2d(sweep, &sgntz, &sgntx, &sgnvz, &sgnvx);
    (void) nx;
    sgnrz = (double) sgntz;
    sgnrx = (double) sgntx;
    tv  = tt[(ix - 1)*nz + iz - sgntz - 1];
    te  = tt[(ix - sgntx - 1)*nz + iz - 1];
    tev = tt[(ix - sgntx - 1)*nz + iz - sgntz - 1];
    t   = tt[(ix - 1)*nz + iz - 1];
    gather_slow2d(sweep, nz, nx, iz, ix, slow, s3);
    sref1 = s3[0]; sref2 = s3[1]; sref3 = s3[2];
    t1d = dmin2(tv + c->dz*sref1, te + c->dx*sref2);
    tanad2d(iz, ix, c, &t0c, &tzc, &txc);
    tauv  = tv  - tana2d(iz - sgntz, ix, c);
    taue  = te  - tana2d(iz, ix - sgntx, c);
    tauev = tev - tana2d(iz - sgntz, ix - sgntx, c);
    t1 = OR_HUGE; t2 = OR_HUGE; t3 = OR_HUGE;
    if ((double) abs(iz - c->zsi) > c->eps || (double) abs(ix - c->xsi) > c->eps)
    {
        if (tv <= te + c->dx*sref3 && te <= tv + c->dz*sref3 &&
            te >= tev && tv >= tev)
        {
            ta = tev + te - tv;
            tb = tev - te + tv;
            t1 = ((tb*c->dz2i + ta*c->dx2i)
               + sqrt(4.0*sref3*sref3*(c->dz2i + c->dx2i)
                      - c->dz2i*c->dx2i*(ta - tb)*(ta - tb)))
               /(c->dz2i + c->dx2i);
        }
        else if ((te - tev) <= c->dz*c->dz*sref3/sqrt(c->dx*c->dx + c->dz*c->dz)
                 && te > tev)
        {
            q = (te - tev)/c->dz;
            t2 = te + c->dx*sqrt(sref3*sref3 - q*q);
        }
        else if ((tv - tev) <= c->dx*c->dx*sref3/sqrt(c->dx*c->dx + c->dz*c->dz)
                 && tv > tev)
        {
            q = (tv - tev)/c->dx;
            t3 = tv + c->dz*sqrt(sref3*sref3 - q*q);
        }
    }
    else
    {
        if (tv < te + c->dx*sref3 && te < tv + c->dz*sref3 &&
            te >= tev && tv >= tev)
        {
            ta = tauev + taue - tauv;
            tb = tauev - taue + tauv;
            apoly = c->dz2i + c->dx2i;
            bpoly = 4.0*(sgnrx*txc*c->dxi + sgnrz*tzc*c->dzi)
                  - 2.0*(ta*c->dx2i + tb*c->dz2i);
            cpoly = (ta*ta*c->dx2i) + (tb*tb*c->dz2i)
                  - 4.0*(sgnrx*txc*c->dxi*ta + sgnrz*tzc*c->dzi*tb)
                  + 4.0*(c->szero*c->szero - sref3*sref3);
            dpoly = bpoly*bpoly - 4.0*apoly*cpoly;
            if (dpoly >= 0.0) t1 = 0.5*(sqrt(dpoly) - bpoly)/apoly + t0c;
            if (t1 < tv || t1 < te) t1 = OR_HUGE;
        }
    }
    t2d = dmin2(dmin2(t1, t2), t3);
    return dmin2(dmin2(t, t1d), t2d);
}

double or_init2d(int sweep, int nz, int nx, int iz, int ix, double dz,
                 double dx, double eps, const or_source *src,
                 const double *slow, const double *tt)
{
    or_consts2d c;
    consts2d_init(dz, dx, eps, src, &c);
    return init2d_c(sweep, nz, nx, iz, ix, &c, slow, tt);
}

/* solver2D.F90:782-978 (with the lexicographic init sweeps of :825-867 and
 * the level loop of :897-925, which is equivalent to visiting the masked
 * nodes lexicographically in the sweep direction).  The convergence test of
 * :935-946 never triggers (tdiff starts at FTEIK_HUGE), so exactly nsweep
 * iterations run. */
int or_solve2d(int nz, int nx, double dz, double dx, const double *slow,
               const or_source *src, int nsweep, double eps, double *tt)
{
    or_consts2d c;
    int zsi = src->zsi, xsi = src->xsi, kiter, sweep, ix, iz, a, b;
    long g, ngrd = (long) nz*nx;
    double s3[3], t4[4], tupd;
    consts2d_init(dz, dx, eps, src, &c);
    for (g = 0; g < ngrd; g++) tt[g] = OR_HUGE;
    tt[(xsi - 1)*nz + zsi - 1] = tana2d(zsi,     xsi,     &c);
    tt[(xsi - 1)*nz + zsi]     = tana2d(zsi + 1, xsi,     &c);
    tt[xsi*nz + zsi - 1]       = tana2d(zsi,     xsi + 1, &c);
    tt[xsi*nz + zsi]           = tana2d(zsi + 1, xsi + 1, &c);
    /* initialization sweeps (solver2D.F90:825-867) */
    for (ix = imax2(2, xsi); ix <= nx; ix++)
        for (iz = imax2(2, zsi); iz <= nz; iz++)
        {
            tupd = init2d_c(1, nz, nx, iz, ix, &c, slow, tt);
            tt[(ix - 1)*nz + iz - 1] = dmin2(tt[(ix - 1)*nz + iz - 1], tupd);
        }
    for (ix = xsi + 1; ix >= 1; ix--)
        for (iz = imax2(2, zsi); iz <= nz; iz++)
        {
            tupd = init2d_c(2, nz, nx, iz, ix, &c, slow, tt);
            tt[(ix - 1)*nz + iz - 1] = dmin2(tt[(ix - 1)*nz + iz - 1], tupd);
        }
    for (ix = imax2(2, xsi); ix <= nx; ix++)
        for (iz = zsi + 1; iz >= 1; iz--)
        {
            tupd = init2d_c(3, nz, nx, iz, ix, &c, slow, tt);
            tt[(ix - 1)*nz + iz - 1] = dmin2(tt[(ix - 1)*nz + iz - 1], tupd);
        }
    for (ix = xsi + 1; ix >= 1; ix--)
        for (iz = zsi + 1; iz >= 1; iz--)
        {
            tupd = init2d_c(4, nz, nx, iz, ix, &c, slow, tt);
            tt[(ix - 1)*nz + iz - 1] = dmin2(tt[(ix - 1)*nz + iz - 1], tupd);
        }
    /* Gauss-Seidel iterations (solver2D.F90:897-947) */
    for (kiter = 0; kiter < nsweep; kiter++)
    {
        for (sweep = 1; sweep <= 4; sweep++)
        {
            int sgntz, sgntx, sgnvz, sgnvx;
            signs2d(sweep, &sgntz, &sgntx, &sgnvz, &sgnvx);
            for (a = 0; a < nx - 1; a++)
            {
                ix = (sgntx > 0) ? 2 + a : nx - 1 - a;
                for (b = 0; b < nz - 1; b++)
                {
                    iz = (sgntz > 0) ? 2 + b : nz - 1 - b;
                    gather_slow2d(sweep, nz, nx, iz, ix, slow, s3);
                    t4[0] = tt[(ix - 1)*nz + iz - sgntz - 1];
                    t4[1] = tt[(ix - sgntx - 1)*nz + iz - 1];
                    t4[2] = tt[(ix - sgntx - 1)*nz + iz - sgntz - 1];
                    t4[3] = tt[(ix - 1)*nz + iz - 1];
                    tupd = or_noinit2d(t4, s3, dz, dx);
                    tt[(ix - 1)*nz + iz - 1] = dmin2(t4[3], tupd);
                }
            }
        }
    }
    return 0;
}

/* ===================================================================== */
/*                              Analytic                                 */
/* ===================================================================== */

/* analytic.F90:618-662 with analytic.F90:844-851.  Output is the module's
 * internal x-fastest [nz][ny][nx] layout (analytic.F90:645). */
int or_analytic_constant(int nz, int nx, int ny, double dz, double dx,
                         double dy, double z0, double x0, double y0,
                         double zs, double xs, double ys, double vconst,
                         double *tt_xfast)
{
    int i, j, k;
    double x, y, z, d;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
            {
                x = x0 + (double) (float) (i - 1)*dx;
                y = y0 + (double) (float) (j - 1)*dy;
                z = z0 + (double) (float) (k - 1)*dz;
                d = sqrt((x - xs)*(x - xs) + (y - ys)*(y - ys)
                         + (z - zs)*(z - zs));
                tt_xfast[(long) (k - 1)*nx*ny + (j - 1)*nx + i - 1] = d/vconst;
            }
    return 0;
}

/* analytic.F90:853-890 for one point. */
double or_analytic_gradient_point(int nz, double dz, double z0, double zs,
                                  double xs, double ys, double vtop,
                                  double vbottom, double z, double x,
                                  double y)
{
    double ztop = z0, zbot = z0 + (double) (float) (nz - 1)*dz;
    double vel0 = vtop + (vbottom - vtop)/(zbot - ztop)*zs;
    double vel1 = vbottom, slow0 = 1.0/vel0;
    double gx = 0.0, gy = 0.0, gz = (vel1 - vel0)/(zbot - zs);
    double absG2 = gx*gx + gy*gy + gz*gz, absG = sqrt(absG2);
    double ddx = x - xs, ddy = y - ys, ddz = z - zs;
    double dist2 = ddx*ddx + ddy*ddy + ddz*ddz;
    double vel = vel0 + gx*ddx + gy*ddy + gz*ddz;
    double arg = 1.0 + 0.5/vel*slow0*absG2*dist2;
    return acosh(arg)/dmax2(1.0e-10, absG);
}

/* analytic.F90:456-528 (job == 1) */
int or_analytic_gradient(int nz, int nx, int ny, double dz, double dx,
                         double dy, double z0, double x0, double y0,
                         double zs, double xs, double ys, double vtop,
                         double vbottom, double *tt_xfast)
{
    int i, j, k;
    double x, y, z;
    if (nz < 2) return 1;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
            {
                x = x0 + (double) (float) (i - 1)*dx;
                y = y0 + (double) (float) (j - 1)*dy;
                z = z0 + (double) (float) (k - 1)*dz;
                tt_xfast[(long) (k - 1)*nx*ny + (j - 1)*nx + i - 1]
                    = or_analytic_gradient_point(nz, dz, z0, zs, xs, ys, vtop,
                                                 vbottom, z, x, y);
            }
    return 0;
}

/* ===================================================================== */
/*                               Locator                                 */
/* ===================================================================== */

/* memory.F90:194-211 */
int or_padlength32(int alignment, int n)
{
    int xmod = (n*4) % alignment;
    return (xmod != 0) ? (alignment - xmod)/4 : 0;
}

/* locate.F90:431-497 for one event.  obs2tf1 is 1-based.  Output arrays
 * have lntf entries (padded with zeros / skip=1). */
void or_locate_observation(int nttimes, const int *obs2tf1,
                           const double *picks, const double *wts, int lntf,
                           double *tepoch, float *obs, float *wt,
                           unsigned char *skip)
{
    int i, indx;
    double tmax = picks[0];
    for (i = 1; i < nttimes; i++) tmax = dmax2(tmax, picks[i]);
    *tepoch = tmax;
    for (i = 0; i < lntf; i++) { obs[i] = 0.0f; wt[i] = 0.0f; skip[i] = 1; }
    for (i = 0; i < nttimes; i++)
    {
        indx = obs2tf1[i] - 1;
        if (wts[i] > 0.0)
        {
            obs[indx] = (float) (picks[i] - tmax);
            wt[indx] = (float) wts[i];
            skip[indx] = 0;
        }
    }
}

/* locate.F90:620-744.  t0work carries over between calls, as the
 * reference's t0Work does (it is only zeroed by locate_initialize,
 * locate.F90:212); objwork is zeroed for the grid points. */
void or_locate_objective(int ntf, int ngrd, int ldgrd, const float *tables,
                         const float *obs, const float *wt,
                         const unsigned char *skip, const float *statics,
                         int lntf, float *t0work, float *objwork,
                         int nthreads)
{
    const int blockSize = 512;
    const float sqrtHalf = sqrtf(0.5f);
    float xnorm, wsum = 0.0f;
    int i, igrd;
    for (igrd = 0; igrd < ngrd; igrd++) objwork[igrd] = 0.0f;
    for (i = 0; i < lntf; i++) wsum = wsum + wt[i];
    xnorm = 1.0f/wsum;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void) nthreads;
#endif
    #pragma omp parallel for schedule(static)
    for (igrd = 0; igrd < ngrd; igrd += blockSize)
    {
        int grd2 = imin2(igrd + blockSize, ngrd), ii, jg;
        for (ii = 0; ii < ntf; ii++)
        {
            float w, toff, o, est, res;
            const float *tp;
            if (skip[ii]) continue;
            w = wt[ii]; toff = statics[ii]; o = obs[ii];
            tp = tables + (long) ii*ldgrd;
            for (jg = igrd; jg < grd2; jg++)
            {
                est = tp[jg] + toff;
                res = xnorm*(o - est);
                t0work[jg] = t0work[jg] + w*res;
            }
        }
        for (ii = 0; ii < ntf; ii++)
        {
            float w, toff, o, est, res;
            const float *tp;
            if (skip[ii]) continue;
            w = sqrtHalf*wt[ii]; toff = statics[ii]; o = obs[ii];
            tp = tables + (long) ii*ldgrd;
            for (jg = igrd; jg < grd2; jg++)
            {
                est = tp[jg] + t0work[jg] + toff;
                res = w*(o - est);
                objwork[jg] = objwork[jg] + res*res;
            }
        }
    }
}

/* locate.F90:112-126.  Returns the 1-based optimum index.  With
 * accumulate_t0 == 0 the origin-time work array is zeroed first (fresh
 * state per event); with pad_argmin != 0 the first-minimum search runs over
 * max(ngrd, 512) entries whose tail is zero, as the reference's MINLOC does
 * (locate.F90:122, :200-203).  t0work/objwork need max(ngrd,512) entries. */
int or_locate_event(int ntf, int ngrd, int ldgrd, const float *tables,
                    const float *obs, const float *wt,
                    const unsigned char *skip, int lntf, double tepoch,
                    int accumulate_t0, int pad_argmin, float *t0work,
                    float *objwork, int nthreads, double *t0opt,
                    double *objopt)
{
    int i, n = ngrd, opt = 0;
    float *statics = (float *) calloc((size_t) (ntf > 0 ? ntf : 1),
                                      sizeof(float));
    if (!accumulate_t0) for (i = 0; i < ngrd; i++) t0work[i] = 0.0f;
    or_locate_objective(ntf, ngrd, ldgrd, tables, obs, wt, skip, statics,
                        lntf, t0work, objwork, nthreads);
    free(statics);
    if (pad_argmin && ngrd < 512)
    {
        for (i = ngrd; i < 512; i++) { objwork[i] = 0.0f; t0work[i] = 0.0f; }
        n = 512;
    }
    for (i = 1; i < n; i++) if (objwork[i] < objwork[opt]) opt = i;
    *t0opt = (double) t0work[opt] + tepoch;
    *objopt = (double) objwork[opt];
    return opt + 1;
}
