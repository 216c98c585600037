// rays.hip -- travel-time gradients by finite differences and
// receiver-to-source ray tracing on gfx950.
//
// Gradients (one thread per node, FP64, the reference's differences):
//   3D  solver3D.F90:1249-1321: forward differences along z, x and y,
//       backward on the last face of each axis; stored [x y z] per node;
//   2D  solver2D.F90:984-1032: backward differences, forward on the first
//       row / column.
//
// Rays (one thread per ray): from the receiver down the travel-time
// gradient, dx/ds = -grad T / u, the gradient interpolated from the eight
// (four in 2D) nodes of the cell holding the point and u the cell's
// slowness, in steps of ds = min(h)/2 by Euler, midpoint or classical RK4,
// positions clamped into the grid, until a step lands within
// sqrt(sum h^2 / 2) of the source, where the ray is snapped.  The ray
// time accumulates |step| times the slowness of the cell each step starts
// in.  Compiled without contraction, so a ray is bitwise the same as the
// same arithmetic on the host.
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

__global__ void k_fd_gradient3d(double* __restrict__ g,
                                const double* __restrict__ tt, int nz, int nx,
                                int ny, double dz, double dx, double dy)
{
    const long nzx = (long) nz*nx, ngrd = nzx*ny;
    const long m = (long) blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= ngrd) return;
    const int iz = (int) (m % nz);
    const int ix = (int) ((m / nz) % nx);
    const int iy = (int) (m / nzx);
    const double t = tt[m];
    const double gz = (iz < nz - 1) ? (tt[m + 1] - t)/dz : (t - tt[m - 1])/dz;
    const double gx = (ix < nx - 1) ? (tt[m + nz] - t)/dx
                                    : (t - tt[m - nz])/dx;
    const double gy = (iy < ny - 1) ? (tt[m + nzx] - t)/dy
                                    : (t - tt[m - nzx])/dy;
    g[3*m] = gx;
    g[3*m + 1] = gy;
    g[3*m + 2] = gz;
}

__global__ void k_fd_gradient2d(double* __restrict__ gx,
                                double* __restrict__ gz,
                                const double* __restrict__ tt, int nz, int nx,
                                double dz, double dx)
{
    const long m = (long) blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= (long) nz*nx) return;
    const int iz = (int) (m % nz), ix = (int) (m / nz);
    const double t = tt[m];
    gz[m] = (iz == 0) ? (tt[m + 1] - t)/dz : (t - tt[m - 1])/dz;
    gx[m] = (ix == 0) ? (tt[m + nz] - t)/dx : (t - tt[m - nz])/dx;
}

struct Grid {
    int is3d, nz, nx, ny;
    double dz, dx, dy;
    const double* slow;
    const double* g3;
    const double* gx;
    const double* gz;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ double lerp(double a, double b, double w)
{
    return a + w*(b - a);
}

__device__ __forceinline__ void cell_of(const Grid& G, const double* p,
                                        int& ix, int& iy, int& iz)
{
    ix = clampi((int) (p[0]/G.dx + 1.0), 1, G.nx - 1);
    iz = clampi((int) (p[2]/G.dz + 1.0), 1, G.nz - 1);
    iy = G.is3d ? clampi((int) (p[1]/G.dy + 1.0), 1, G.ny - 1) : 1;
}

__device__ __forceinline__ double cell_slow(const Grid& G, int ix, int iy,
                                            int iz)
{
    const long c = (long) (iy - 1)*(G.nz - 1)*(G.nx - 1)
                 + (long) (ix - 1)*(G.nz - 1) + (iz - 1);
    return G.slow[c];
}

__device__ void rhs(const Grid& G, const double* p, double* f)
{
    int ix, iy, iz;
    cell_of(G, p, ix, iy, iz);
    const double wx = (p[0] - (double) (ix - 1)*G.dx)/G.dx;
    const double wz = (p[2] - (double) (iz - 1)*G.dz)/G.dz;
    const double u = cell_slow(G, ix, iy, iz);
    if (G.is3d) {
        const double wy = (p[1] - (double) (iy - 1)*G.dy)/G.dy;
        const long nzx = (long) G.nz*G.nx;
        const long n000 = (long) (iy - 1)*nzx + (long) (ix - 1)*G.nz
                        + (iz - 1);
        for (int d = 0; d < 3; d++) {
            const double* g = G.g3 + d;
            const double c00 = lerp(g[3*n000], g[3*(n000 + 1)], wz);
            const double c10 = lerp(g[3*(n000 + G.nz)],
                                    g[3*(n000 + G.nz + 1)], wz);
            const double c01 = lerp(g[3*(n000 + nzx)],
                                    g[3*(n000 + nzx + 1)], wz);
            const double c11 = lerp(g[3*(n000 + nzx + G.nz)],
                                    g[3*(n000 + nzx + G.nz + 1)], wz);
            const double c0 = lerp(c00, c10, wx), c1 = lerp(c01, c11, wx);
            f[d] = -lerp(c0, c1, wy)/u;
        }
    } else {
        const long n00 = (long) (ix - 1)*G.nz + (iz - 1);
        double v[2];
        for (int d = 0; d < 2; d++) {
            const double* g = d == 0 ? G.gx : G.gz;
            const double c0 = lerp(g[n00], g[n00 + 1], wz);
            const double c1 = lerp(g[n00 + G.nz], g[n00 + G.nz + 1], wz);
            v[d] = lerp(c0, c1, wx);
        }
        f[0] = -v[0]/u;
        f[1] = 0.0;
        f[2] = -v[1]/u;
    }
}

__device__ void advance(const Grid& G, const double* p, double s,
                        const double* f, double* q)
{
    const double eps = 2.220446049250313e-16;
    const double hi[3] = {(double) (G.nx - 1)*G.dx, (double) (G.ny - 1)*G.dy,
                          (double) (G.nz - 1)*G.dz};
    for (int d = 0; d < 3; d++) {
        if (d == 1 && !G.is3d) { q[1] = 0.0; continue; }
        double v = p[d] + s*f[d];
        v = v < eps ? eps : v;
        v = v > hi[d] - eps ? hi[d] - eps : v;
        q[d] = v;
    }
}

__device__ void step(const Grid& G, int integrator, double ds,
                     const double* p, double* q)
{
    double k1[3], k2[3], k3[3], k4[3], t[3];
    rhs(G, p, k1);
    if (integrator == 2) {
        advance(G, p, 0.5*ds, k1, t);
        rhs(G, t, k2);
        advance(G, p, ds, k2, q);
    } else if (integrator == 3) {
        advance(G, p, 0.5*ds, k1, t);
        rhs(G, t, k2);
        advance(G, p, 0.5*ds, k2, t);
        rhs(G, t, k3);
        advance(G, p, ds, k3, t);
        rhs(G, t, k4);
        double k[3];
        for (int d = 0; d < 3; d++)
            k[d] = (k1[d] + 2.0*(k2[d] + k3[d]) + k4[d])/6.0;
        advance(G, p, ds, k, q);
    } else {
        advance(G, p, ds, k1, q);
    }
}

// start: [nray][3] grid-relative (x y z); pts: [nray][maxpts][3].
__global__ void k_trace(Grid G, int integrator, double ds, double r2,
                        double xs, double ys, double zs,
                        const double* __restrict__ start, int nray,
                        int maxpts, double* __restrict__ pts,
                        int* __restrict__ npts, double* __restrict__ time,
                        int* __restrict__ status)
{
    const int r = blockIdx.x*blockDim.x + threadIdx.x;
    if (r >= nray) return;
    double* out = pts + (long) r*maxpts*3;
    double p[3] = {start[3*r], G.is3d ? start[3*r + 1] : 0.0, start[3*r + 2]};
    double q[3];
    double t = 0.0;
    int n = 1;
    out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
    bool done = false;
    while (!done && n < maxpts) {
        step(G, integrator, ds, p, q);
        const double ex = q[0] - xs, ey = G.is3d ? q[1] - ys : 0.0,
                     ez = q[2] - zs;
        if (ex*ex + ey*ey + ez*ez < r2) {
            q[0] = xs; q[1] = G.is3d ? ys : 0.0; q[2] = zs;
            done = true;
        }
        int ix, iy, iz;
        cell_of(G, p, ix, iy, iz);
        const double sx = q[0] - p[0], sy = q[1] - p[1], sz = q[2] - p[2];
        t = t + sqrt(sx*sx + sy*sy + sz*sz)*cell_slow(G, ix, iy, iz);
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        out[3*n] = p[0]; out[3*n + 1] = p[1]; out[3*n + 2] = p[2];
        n++;
    }
    npts[r] = n;
    time[r] = t;
    status[r] = done ? 0 : 1;
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : 1; }

} // namespace

extern "C" {

int fteik_k_fd_gradient3d(double* g, const double* tt, int nz, int nx,
                          int ny, double dz, double dx, double dy,
                          hipStream_t s)
{
    const long n = (long) nz*nx*ny;
    hipLaunchKernelGGL(k_fd_gradient3d, dim3((unsigned) ((n + 255)/256)),
                       dim3(256), 0, s, g, tt, nz, nx, ny, dz, dx, dy);
    return ok();
}

int fteik_k_fd_gradient2d(double* gx, double* gz, const double* tt, int nz,
                          int nx, double dz, double dx, hipStream_t s)
{
    const long n = (long) nz*nx;
    hipLaunchKernelGGL(k_fd_gradient2d, dim3((unsigned) ((n + 255)/256)),
                       dim3(256), 0, s, gx, gz, tt, nz, nx, dz, dx);
    return ok();
}

int fteik_k_trace_rays(const fteik::RayGrid* rg, const double* slow,
                       const double* g3, const double* gx, const double* gz,
                       int integrator, double xs, double ys, double zs,
                       const double* start, int nray, int maxpts,
                       double* pts, int* npts, double* time, int* status,
                       hipStream_t s)
{
    Grid G;
    G.is3d = rg->is3d; G.nz = rg->nz; G.nx = rg->nx; G.ny = rg->ny;
    G.dz = rg->dz; G.dx = rg->dx; G.dy = rg->dy;
    G.slow = slow; G.g3 = g3; G.gx = gx; G.gz = gz;
    const double ds = G.is3d ? fmin(fmin(G.dx, G.dy), G.dz)/2.0
                             : fmin(G.dx, G.dz)/2.0;
    const double r2 = G.is3d ? (G.dx*G.dx + G.dy*G.dy + G.dz*G.dz)/2.0
                             : (G.dx*G.dx + G.dz*G.dz)/2.0;
    hipLaunchKernelGGL(k_trace, dim3((nray + 63)/64), dim3(64), 0, s, G,
                       integrator, ds, r2, xs, ys, zs, start, nray, maxpts,
                       pts, npts, time, status);
    return ok();
}

} // extern "C"
