/*
 * fteik_oracle_rays.c -- CPU statement of the travel-time gradients and of
 * receiver-to-source ray tracing.  TEST INFRASTRUCTURE ONLY (see
 * fteik_oracle.h): the checker the HIP path is compared against.
 *
 * Gradients restate the reference bit for bit in operation order:
 *   3D  solver3D.F90:1249-1321  forward differences, backward on the last
 *                               z, x and y faces; [x y z] per node;
 *   2D  solver2D.F90:984-1032   backward differences inside, forward on the
 *                               first row and column, backward on the last.
 *
 * Ray tracing states the method rays.F90 documents (its 3D tracer,
 * rays.F90:442-447, is an empty stub that fails, and its 2D tracer,
 * :449-669, stops after the first ray): march from the receiver down the
 * travel-time gradient, dx/ds = -grad T / u (rays.F90:846-863), with the
 * gradient interpolated (bi/trilinearly) from the nodes of the cell holding
 * the point and u the cell's slowness; steps of ds = min(h)/2 by Euler,
 * midpoint or classical RK4 (the RK4 the comment at rays.F90:690-694
 * states); positions clamped into the grid (rays.F90:812-843); the ray ends
 * once a step lands within sqrt(sum h^2 / 2) of the source (the true
 * position, where the field has its minimum: the reference's 2D tracer aims
 * at the source cell's first node, rays.F90:588-594, which rays to a source
 * inside a cell stall short of), where it is snapped; its time accumulates
 * |step| * slowness of the cell the step starts in (rays.F90:596-598).
 */
#include <math.h>
#include "fteik_oracle.h"

/* ------------------------------------------------------------ gradients */
void or_fd_gradient3d(int nz, int nx, int ny, double dz, double dx,
                      double dy, const double *tt, double *tgrad)
{
    const long nzx = (long) nz*nx, ngrd = nzx*ny;
    for (long m = 0; m < 3*ngrd; m++) tgrad[m] = 0.0;
    for (int iy = 1; iy <= ny; iy++)
        for (int ix = 1; ix <= nx; ix++)
            for (int iz = 1; iz <= nz - 1; iz++) {
                long m1 = or_grid2index(iz, ix, iy, nz, (int) nzx);
                long m2 = or_grid2index(iz + 1, ix, iy, nz, (int) nzx);
                tgrad[3*(m1 - 1) + 2] = (tt[m2 - 1] - tt[m1 - 1])/dz;
            }
    for (int iy = 1; iy <= ny; iy++)
        for (int ix = 1; ix <= nx - 1; ix++)
            for (int iz = 1; iz <= nz; iz++) {
                long m1 = or_grid2index(iz, ix, iy, nz, (int) nzx);
                long m2 = or_grid2index(iz, ix + 1, iy, nz, (int) nzx);
                tgrad[3*(m1 - 1) + 0] = (tt[m2 - 1] - tt[m1 - 1])/dx;
            }
    for (int iy = 1; iy <= ny - 1; iy++)
        for (int ix = 1; ix <= nx; ix++)
            for (int iz = 1; iz <= nz; iz++) {
                long m1 = or_grid2index(iz, ix, iy, nz, (int) nzx);
                long m2 = or_grid2index(iz, ix, iy + 1, nz, (int) nzx);
                tgrad[3*(m1 - 1) + 1] = (tt[m2 - 1] - tt[m1 - 1])/dy;
            }
    for (int iy = 1; iy <= ny; iy++)
        for (int ix = 1; ix <= nx; ix++) {
            long m1 = or_grid2index(nz - 1, ix, iy, nz, (int) nzx);
            long m2 = or_grid2index(nz, ix, iy, nz, (int) nzx);
            tgrad[3*(m2 - 1) + 2] = (tt[m2 - 1] - tt[m1 - 1])/dz;
        }
    for (int iy = 1; iy <= ny; iy++)
        for (int iz = 1; iz <= nz; iz++) {
            long m1 = or_grid2index(iz, nx - 1, iy, nz, (int) nzx);
            long m2 = or_grid2index(iz, nx, iy, nz, (int) nzx);
            tgrad[3*(m2 - 1) + 0] = (tt[m2 - 1] - tt[m1 - 1])/dx;
        }
    for (int ix = 1; ix <= nx; ix++)
        for (int iz = 1; iz <= nz; iz++) {
            long m1 = or_grid2index(iz, ix, ny - 1, nz, (int) nzx);
            long m2 = or_grid2index(iz, ix, ny, nz, (int) nzx);
            tgrad[3*(m2 - 1) + 1] = (tt[m2 - 1] - tt[m1 - 1])/dy;
        }
}

void or_fd_gradient2d(int nz, int nx, double dz, double dx, const double *tt,
                      double *tgradx, double *tgradz)
{
    for (long m = 0; m < (long) nz*nx; m++) { tgradx[m] = 0.0; tgradz[m] = 0.0; }
    for (int ix = 1; ix <= nx; ix++)
        for (int iz = 2; iz <= nz - 1; iz++) {
            long m1 = (long) (ix - 1)*nz + iz - 1, m2 = (long) (ix - 1)*nz + iz;
            tgradz[m2 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dz;
        }
    for (int ix = 2; ix <= nx - 1; ix++)
        for (int iz = 1; iz <= nz; iz++) {
            long m1 = (long) (ix - 2)*nz + iz, m2 = (long) (ix - 1)*nz + iz;
            tgradx[m2 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dx;
        }
    for (int ix = 1; ix <= nx; ix++) {
        long m1 = (long) (ix - 1)*nz + 1, m2 = (long) (ix - 1)*nz + 2;
        tgradz[m1 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dz;
        m1 = (long) (ix - 1)*nz + nz - 1; m2 = (long) (ix - 1)*nz + nz;
        tgradz[m2 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dz;
    }
    for (int iz = 1; iz <= nz; iz++) {
        long m1 = iz, m2 = (long) nz + iz;
        tgradx[m1 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dx;
        m1 = (long) (nx - 2)*nz + iz; m2 = (long) (nx - 1)*nz + iz;
        tgradx[m2 - 1] = (tt[m2 - 1] - tt[m1 - 1])/dx;
    }
}

/* ------------------------------------------------------------------ rays */
typedef struct {
    int is3d, nz, nx, ny;
    double dz, dx, dy;
    const double *slow;     /* cells, natural order */
    const double *g3;       /* 3D: [x y z] per node */
    const double *gx, *gz;  /* 2D: per node */
} ray_grid;

static int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

static double lerp(double a, double b, double w)
{
    return a + w*(b - a);
}

/* Cell holding p (grid-relative metres, x y z), 1-based indices. */
static void cell_of(const ray_grid *G, const double *p, int *ix, int *iy,
                    int *iz)
{
    *ix = clampi((int) (p[0]/G->dx + 1.0), 1, G->nx - 1);
    *iz = clampi((int) (p[2]/G->dz + 1.0), 1, G->nz - 1);
    *iy = G->is3d ? clampi((int) (p[1]/G->dy + 1.0), 1, G->ny - 1) : 1;
}

static double cell_slowness(const ray_grid *G, int ix, int iy, int iz)
{
    const long c = (long) (iy - 1)*(G->nz - 1)*(G->nx - 1)
                 + (long) (ix - 1)*(G->nz - 1) + (iz - 1);
    return G->slow[c];
}

/* dx/ds = -grad T / u at p. */
static void ray_rhs(const ray_grid *G, const double *p, double *f)
{
    int ix, iy, iz;
    cell_of(G, p, &ix, &iy, &iz);
    const double wx = (p[0] - (double) (ix - 1)*G->dx)/G->dx;
    const double wz = (p[2] - (double) (iz - 1)*G->dz)/G->dz;
    const double u = cell_slowness(G, ix, iy, iz);
    if (G->is3d) {
        const double wy = (p[1] - (double) (iy - 1)*G->dy)/G->dy;
        const long nzx = (long) G->nz*G->nx;
        const long n000 = (long) (iy - 1)*nzx + (long) (ix - 1)*G->nz + (iz - 1);
        for (int d = 0; d < 3; d++) {
            /* corners (z, x, y) -> node n000 + dz + dx*nz + dy*nzx */
            const double *g = G->g3 + d;
            double c00 = lerp(g[3*n000], g[3*(n000 + 1)], wz);
            double c10 = lerp(g[3*(n000 + G->nz)], g[3*(n000 + G->nz + 1)], wz);
            double c01 = lerp(g[3*(n000 + nzx)], g[3*(n000 + nzx + 1)], wz);
            double c11 = lerp(g[3*(n000 + nzx + G->nz)],
                              g[3*(n000 + nzx + G->nz + 1)], wz);
            double c0 = lerp(c00, c10, wx), c1 = lerp(c01, c11, wx);
            f[d] = -lerp(c0, c1, wy)/u;
        }
    } else {
        const long n00 = (long) (ix - 1)*G->nz + (iz - 1);
        const double *gg[2] = {G->gx, G->gz};
        double v[2];
        for (int d = 0; d < 2; d++) {
            const double *g = gg[d];
            double c0 = lerp(g[n00], g[n00 + 1], wz);
            double c1 = lerp(g[n00 + G->nz], g[n00 + G->nz + 1], wz);
            v[d] = lerp(c0, c1, wx);
        }
        f[0] = -v[0]/u;
        f[1] = 0.0;
        f[2] = -v[1]/u;
    }
}

/* p + s*f, clamped into the grid (rays.F90:812-843). */
static void advance(const ray_grid *G, const double *p, double s,
                    const double *f, double *q)
{
    const double eps = 2.220446049250313e-16;
    const double hi[3] = {(double) (G->nx - 1)*G->dx,
                          (double) (G->ny - 1)*G->dy,
                          (double) (G->nz - 1)*G->dz};
    for (int d = 0; d < 3; d++) {
        if (d == 1 && !G->is3d) { q[1] = 0.0; continue; }
        double v = p[d] + s*f[d];
        v = v < eps ? eps : v;
        v = v > hi[d] - eps ? hi[d] - eps : v;
        q[d] = v;
    }
}

static void ray_step(const ray_grid *G, int integrator, double ds,
                     const double *p, double *q)
{
    double k1[3], k2[3], k3[3], k4[3], t[3];
    ray_rhs(G, p, k1);
    if (integrator == 2) {                       /* midpoint */
        advance(G, p, 0.5*ds, k1, t);
        ray_rhs(G, t, k2);
        advance(G, p, ds, k2, q);
    } else if (integrator == 3) {                /* classical RK4 */
        advance(G, p, 0.5*ds, k1, t);
        ray_rhs(G, t, k2);
        advance(G, p, 0.5*ds, k2, t);
        ray_rhs(G, t, k3);
        advance(G, p, ds, k3, t);
        ray_rhs(G, t, k4);
        double k[3];
        for (int d = 0; d < 3; d++)
            k[d] = (k1[d] + 2.0*(k2[d] + k3[d]) + k4[d])/6.0;
        advance(G, p, ds, k, q);
    } else {                                     /* Euler */
        advance(G, p, ds, k1, q);
    }
}

/* One ray from (xr, yr, zr) to the source (xs, ys, zs), both
 * grid-relative metres.  pts receives up to maxpts points (x y z);
 * returns 0 when the ray reached the source, 1 otherwise. */
static int trace_one(const ray_grid *G, int integrator, double xs, double ys,
                     double zs, double xr, double yr, double zr, int maxpts,
                     double *pts, int *npts, double *time)
{
    const double ds = G->is3d
        ? fmin(fmin(G->dx, G->dy), G->dz)/2.0 : fmin(G->dx, G->dz)/2.0;
    const double r2 = G->is3d
        ? (G->dx*G->dx + G->dy*G->dy + G->dz*G->dz)/2.0
        : (G->dx*G->dx + G->dz*G->dz)/2.0;
    double p[3] = {xr, G->is3d ? yr : 0.0, zr}, q[3];
    double t = 0.0;
    int n = 1;
    pts[0] = p[0]; pts[1] = p[1]; pts[2] = p[2];
    int done = 0;
    while (!done && n < maxpts) {
        ray_step(G, integrator, ds, p, q);
        const double ex = q[0] - xs, ey = G->is3d ? q[1] - ys : 0.0,
                     ez = q[2] - zs;
        if (ex*ex + ey*ey + ez*ez < r2) {
            q[0] = xs; q[1] = G->is3d ? ys : 0.0; q[2] = zs;
            done = 1;
        }
        int ix, iy, iz;
        cell_of(G, p, &ix, &iy, &iz);
        const double sx = q[0] - p[0], sy = q[1] - p[1], sz = q[2] - p[2];
        t = t + sqrt(sx*sx + sy*sy + sz*sz)*cell_slowness(G, ix, iy, iz);
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        pts[3*n] = p[0]; pts[3*n + 1] = p[1]; pts[3*n + 2] = p[2];
        n++;
    }
    *npts = n;
    *time = t;
    return done ? 0 : 1;
}

int or_trace_ray3d(int nz, int nx, int ny, double dz, double dx, double dy,
                   const double *slow, const double *tgrad, int integrator,
                   double xs, double ys, double zs, double xr, double yr,
                   double zr, int maxpts, double *pts, int *npts,
                   double *time)
{
    ray_grid G = {1, nz, nx, ny, dz, dx, dy, slow, tgrad, 0, 0};
    return trace_one(&G, integrator, xs, ys, zs, xr, yr, zr, maxpts, pts,
                     npts, time);
}

int or_trace_ray2d(int nz, int nx, double dz, double dx, const double *slow,
                   const double *tgradx, const double *tgradz, int integrator,
                   double xs, double zs, double xr, double zr, int maxpts,
                   double *pts, int *npts, double *time)
{
    ray_grid G = {0, nz, nx, 1, dz, dx, 1.0, slow, 0, tgradx, tgradz};
    return trace_one(&G, integrator, xs, 0.0, zs, xr, 0.0, zr, maxpts, pts,
                     npts, time);
}
