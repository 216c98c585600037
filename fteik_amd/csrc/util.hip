// util.hip -- element-wise and gather kernels around the solvers: velocity
// to slowness conversions, field re-orderings and cell averages, receiver
// interpolation, precision conversions and the closed-form (analytic)
// travel-time fields.  All grid indices follow the reference's z-fastest
// ("natural") layout.
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

constexpr int TPB = 256;

inline int nblocks(long n)
{
    long b = (n + TPB - 1)/TPB;
    return (int) (b < 65535*8 ? b : 65535*8);
}

#define GRID_STRIDE(i, n) \
    for (long i = (long) blockIdx.x*blockDim.x + threadIdx.x; i < (n); \
         i += (long) gridDim.x*blockDim.x)

// model.F90:469-559 (cell model): natural order is 1/vel, XYZ and ZYX are
// permuted first.  2D: order 1 is XZ.
__global__ void k_slow_cells(double* __restrict__ slow,
                             const double* __restrict__ vel, int nz, int nx,
                             int ny, int is3d, int order)
{
    const long ncz = nz - 1, ncx = nx - 1, ncy = is3d ? ny - 1 : 1;
    const long ncell = ncz*ncx*ncy;
    GRID_STRIDE(c, ncell) {
        long iz = c % ncz, ix = (c / ncz) % ncx, iy = c / (ncz*ncx);
        long src = c;
        if (is3d && order == 1) src = iz*ncx*ncy + iy*ncx + ix;
        else if (is3d && order == 2) src = ix*ncz*ncy + iy*ncz + iz;
        else if (!is3d && order == 1) src = iz*ncx + ix;
        slow[c] = 1.0/vel[src];
    }
}

// model.F90:311-440 (nodal model averaged onto cells), quirks included:
// the XYZ branch omits the 1/8 (model.F90:353) and the natural branch uses
// iy instead of iy+1 for nodes 5-8 (model.F90:386-389).
__global__ void k_slow_nodes(double* __restrict__ slow,
                             const double* __restrict__ vel, int nz, int nx,
                             int ny, int is3d, int order)
{
    const long ncz = nz - 1, ncx = nx - 1, ncy = is3d ? ny - 1 : 1;
    const long ncell = ncz*ncx*ncy;
    GRID_STRIDE(c, ncell) {
        const long iz = c % ncz + 1, ix = (c / ncz) % ncx + 1,
                   iy = c / (ncz*ncx) + 1;
        long ind[8];
        double s = 0.0;
        if (is3d) {
            if (order == 1) {
                ind[0] = (iz - 1)*nx*ny + (iy - 1)*nx + ix;
                ind[1] = (iz - 1)*nx*ny + (iy - 1)*nx + ix + 1;
                ind[2] = (iz - 1)*nx*ny + (iy    )*nx + ix;
                ind[3] = (iz - 1)*nx*ny + (iy    )*nx + ix + 1;
                ind[4] = (iz    )*nx*ny + (iy - 1)*nx + ix;
                ind[5] = (iz    )*nx*ny + (iy - 1)*nx + ix + 1;
                ind[6] = (iz    )*nx*ny + (iy    )*nx + ix;
                ind[7] = (iz    )*nx*ny + (iy    )*nx + ix + 1;
            } else if (order == 2) {
                ind[0] = (ix - 1)*ny*nz + (iy - 1)*nz + iz;
                ind[1] = (ix - 1)*ny*nz + (iy - 1)*nz + iz + 1;
                ind[2] = (ix - 1)*ny*nz + (iy    )*nz + iz;
                ind[3] = (ix - 1)*ny*nz + (iy    )*nz + iz + 1;
                ind[4] = (ix    )*ny*nz + (iy - 1)*nz + iz;
                ind[5] = (ix    )*ny*nz + (iy - 1)*nz + iz + 1;
                ind[6] = (ix    )*ny*nz + (iy    )*nz + iz;
                ind[7] = (ix    )*ny*nz + (iy    )*nz + iz + 1;
            } else {
                const long nzx = (long) nz*nx;
                ind[0] = (iy - 1)*nzx + (ix - 1)*nz + iz;
                ind[1] = (iy - 1)*nzx + (ix - 1)*nz + iz + 1;
                ind[2] = (iy - 1)*nzx + (ix    )*nz + iz;
                ind[3] = (iy - 1)*nzx + (ix    )*nz + iz + 1;
                ind[4] = ind[0]; ind[5] = ind[1]; ind[6] = ind[2];
                ind[7] = ind[3];
            }
            for (int n = 0; n < 8; n++) s = s + vel[ind[n] - 1];
            slow[c] = 1.0/((order == 1) ? s : s*0.125);
        } else {
            if (order == 1) {
                ind[0] = (iz - 1)*nx + ix;
                ind[1] = (iz - 1)*nx + ix + 1;
                ind[2] = (iz    )*nx + ix;
                ind[3] = (iz    )*nx + ix + 1;
            } else {
                ind[0] = (ix - 1)*nz + iz;
                ind[1] = (ix - 1)*nz + iz + 1;
                ind[2] = (ix    )*nz + iz;
                ind[3] = (ix    )*nz + iz + 1;
            }
            for (int n = 0; n < 4; n++) s = s + vel[ind[n] - 1];
            slow[c] = 1.0/(s*0.25);
        }
    }
}

// model.F90:746-748
__global__ void k_slow_vel32(double* __restrict__ slow,
                             const float* __restrict__ vel, long n)
{
    GRID_STRIDE(c, n) slow[c] = 1.0/(double) vel[c];
}

// solver3D.F90:863-892, solver2D.F90:1313-1327
__global__ void k_reorder(double* __restrict__ out,
                          const double* __restrict__ tt, int nz, int nx,
                          int ny, int is3d, int order)
{
    const long ngrd = (long) nz*nx*(is3d ? ny : 1);
    GRID_STRIDE(g, ngrd) {
        const long iz = g % nz, ix = (g / nz) % nx, iy = g / ((long) nz*nx);
        long o = g;
        if (is3d && order == 1) o = iz*nx*ny + iy*nx + ix;
        else if (is3d && order == 2) o = ix*nz*ny + iy*nz + iz;
        else if (!is3d && order == 1) o = iz*nx + ix;
        out[o] = tt[g];
    }
}

// solver3D.F90:934-969 (XYZ, ZYX) and solver2D.F90:1363-1374 (XZ); the
// natural order returns the averages in z-fastest cell order.
__global__ void k_cells(double* __restrict__ out,
                        const double* __restrict__ tt, int nz, int nx,
                        int ny, int is3d, int order)
{
    const long ncz = nz - 1, ncx = nx - 1, ncy = is3d ? ny - 1 : 1;
    const long ncell = ncz*ncx*ncy, nzx = (long) nz*nx;
    GRID_STRIDE(c, ncell) {
        const long iz = c % ncz, ix = (c / ncz) % ncx, iy = c / (ncz*ncx);
        const long b = iy*nzx + ix*nz + iz;
        double s;
        long o = c;
        if (is3d) {
            s = tt[b];
            s = s + tt[b + 1];
            s = s + tt[b + nz];
            s = s + tt[b + nz + 1];
            s = s + tt[b + nzx];
            s = s + tt[b + nzx + 1];
            s = s + tt[b + nzx + nz];
            s = s + tt[b + nzx + nz + 1];
            s = 0.125*s;
            if (order == 1) o = iz*ncx*ncy + iy*ncx + ix;
            else if (order == 2) o = ix*ncz*ncy + iy*ncz + iz;
        } else {
            s = tt[b];
            s = s + tt[b + 1];
            s = s + tt[b + nz];
            s = s + tt[b + nz + 1];
            s = 0.25*s;
            if (order == 1) o = iz*ncx + ix;
        }
        out[o] = s;
    }
}

// receiver.F90:246-306: trilinear (3D) / bilinear (2D) interpolation from
// the base node ri (1-based z, x, y) and offsets rd (meters).  blockIdx.y
// selects one of several fields fstride apart; its nrec times go to
// ttr + blockIdx.y*nrec.
__global__ void k_receivers(double* __restrict__ ttr,
                            const double* __restrict__ tt,
                            const int* __restrict__ ri,
                            const double* __restrict__ rd, int nrec, int nz,
                            int nx, double dz, double dx, double dy, int is3d,
                            long fstride)
{
    const int r = blockIdx.x*blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    tt += (long) blockIdx.y*fstride;
    ttr += (long) blockIdx.y*nrec;
    const long zi = ri[3*r], xi = ri[3*r + 1], yi = ri[3*r + 2];
    const long nzx = (long) nz*nx;
    if (is3d) {
        auto T = [&](long a, long b, long c) -> double {
            return tt[(c - 1)*nzx + (b - 1)*nz + a - 1];
        };
        const double t000 = T(zi, xi, yi), t100 = T(zi + 1, xi, yi);
        const double t010 = T(zi, xi + 1, yi), t110 = T(zi + 1, xi + 1, yi);
        const double t001 = T(zi, xi, yi + 1), t101 = T(zi + 1, xi, yi + 1);
        const double t011 = T(zi, xi + 1, yi + 1);
        const double t111 = T(zi + 1, xi + 1, yi + 1);
        const double zd = rd[3*r]/dz, xd = rd[3*r + 1]/dx, yd = rd[3*r + 2]/dy;
        const double omz = 1.0 - zd;
        const double c00 = t000*omz + t100*zd;
        const double c01 = t001*omz + t101*zd;
        const double c10 = t010*omz + t110*zd;
        const double c11 = t011*omz + t111*zd;
        const double omx = 1.0 - xd;
        const double c0 = c00*omx + c10*xd;
        const double c1 = c01*omx + c11*xd;
        ttr[r] = c0*(1.0 - yd) + c1*yd;
    } else {
        const double t00 = tt[(xi - 1)*nz + zi - 1];
        const double t10 = tt[(xi - 1)*nz + zi];
        const double t01 = tt[xi*nz + zi - 1];
        const double t11 = tt[xi*nz + zi];
        const double zd = rd[3*r]/dz, xd = rd[3*r + 1]/dx;
        const double omz = 1.0 - zd, omx = 1.0 - xd;
        ttr[r] = t00*omz*omx + t10*zd*omx + t01*omz*xd + t11*zd*xd;
    }
}

__global__ void k_d2f(float* __restrict__ out, const double* __restrict__ in,
                      long n)
{
    GRID_STRIDE(i, n) out[i] = (float) in[i];
}

__global__ void k_f2d(double* __restrict__ out, const float* __restrict__ in,
                      long n)
{
    GRID_STRIDE(i, n) out[i] = (double) in[i];
}

// analytic.F90:844-890 for one point.
__device__ double analytic_point(double x, double y, double z, double zs,
                                 double xs, double ys, double vconst,
                                 double vel0, double slow0, double gz,
                                 double absG2, double absG, int gradient)
{
    if (!gradient) {
        double d = sqrt((x - xs)*(x - xs) + (y - ys)*(y - ys)
                        + (z - zs)*(z - zs));
        return d/vconst;
    }
    const double ddx = x - xs, ddy = y - ys, ddz = z - zs;
    const double dist2 = ddx*ddx + ddy*ddy + ddz*ddz;
    const double vel = vel0 + 0.0*ddx + 0.0*ddy + gz*ddz;
    const double arg = 1.0 + 0.5/vel*slow0*absG2*dist2;
    return acosh(arg)/(absG > 1.0e-10 ? absG : 1.0e-10);
}

struct GradVars { double vel0, slow0, gz, absG2, absG; };

// analytic.F90:874-890
inline GradVars gradient_vars(int nz, double dz, double z0, double zs,
                              double vtop, double vbot)
{
    GradVars v;
    const double ztop = z0, zbot = z0 + (double) (float) (nz - 1)*dz;
    v.vel0 = vtop + (vbot - vtop)/(zbot - ztop)*zs;
    v.slow0 = 1.0/v.vel0;
    v.gz = (vbot - v.vel0)/(zbot - zs);
    v.absG2 = 0.0*0.0 + 0.0*0.0 + v.gz*v.gz;
    v.absG = sqrt(v.absG2);
    return v;
}

// analytic.F90:488-509 / 641-653: field in the module's x-fastest layout.
__global__ void k_analytic(double* __restrict__ out, int nz, int nx, int ny,
                           double dz, double dx, double dy, double z0,
                           double x0, double y0, double zs, double xs,
                           double ys, double vconst, GradVars v, int gradient)
{
    const long n = (long) nz*nx*ny;
    GRID_STRIDE(g, n) {
        const long i = g % nx, j = (g / nx) % ny, k = g / ((long) nx*ny);
        const double x = x0 + (double) (float) i*dx;
        const double y = y0 + (double) (float) j*dy;
        const double z = z0 + (double) (float) k*dz;
        out[g] = analytic_point(x, y, z, zs, xs, ys, vconst, v.vel0, v.slow0,
                                v.gz, v.absG2, v.absG, gradient);
    }
}

__global__ void k_analytic_points(double* __restrict__ out,
                                  const double* __restrict__ z,
                                  const double* __restrict__ x,
                                  const double* __restrict__ y, int n,
                                  double zs, double xs, double ys,
                                  double vconst, GradVars v, int gradient)
{
    const int r = blockIdx.x*blockDim.x + threadIdx.x;
    if (r >= n) return;
    out[r] = analytic_point(x[r], y[r], z[r], zs, xs, ys, vconst, v.vel0,
                            v.slow0, v.gz, v.absG2, v.absG, gradient);
}

// analytic.F90:705-731: x-fastest -> requested order.
__global__ void k_analytic_reorder(double* __restrict__ out,
                                   const double* __restrict__ a, int nz,
                                   int nx, int ny, int order)
{
    const long n = (long) nz*nx*ny;
    GRID_STRIDE(g, n) {
        const long ix = g % nx, iy = (g / nx) % ny, iz = g / ((long) nx*ny);
        long o;
        if (order == 1) o = g;
        else if (order == 2) o = ix*nz*ny + iy*nz + iz;
        else o = iy*nz*nx + ix*nz + iz;
        out[o] = a[g];
    }
}


// Mixed precision: keep the smaller of the FP64 field and the FP32 one.
// Promotes single-precision times, raised by the relative margin bias,
// into the double field, keeping the smaller value.  With the margin the
// promoted field stays an upper bound of the double-precision solution even
// where single-precision rounding undershot it, so the double sweeps that
// follow (which can only lower times) converge onto that solution.
__global__ void k_min_merge_f(double* __restrict__ tt,
                              const float* __restrict__ t32, long n,
                              double bias)
{
    GRID_STRIDE(i, n) {
        double v = (double) t32[i];
        v = v + v*bias;
        tt[i] = (v < tt[i]) ? v : tt[i];
    }
}

// graph3d.F90:122-353 in closed form: a node's level is its flipped
// iz+ix+iy, and inside a level nodes are sorted by linear index, i.e. by
// actual (iy, ix).  The rank of a node counts the lattice points of its
// level in the rows before it and in its row before it.
__global__ void k_graph_ijkv(int* __restrict__ ijkv,
                             const int* __restrict__ levelPtr, int nz, int nx,
                             int ny, int sz, int sx, int sy)
{
    const long ngrd = (long) nz*nx*ny;
    GRID_STRIDE(g, ngrd) {
        const int i = (int) (g % nz), j = (int) ((g / nz) % nx),
                  k = (int) (g / ((long) nz*nx));
        const int fi = sz > 0 ? i : nz - 1 - i;
        const int fj = sx > 0 ? j : nx - 1 - j;
        const int fk = sy > 0 ? k : ny - 1 - k;
        const int L = fi + fj + fk;
        long rank = 0;
        for (int kk = 0; kk < k; kk++) {
            const int fkk = sy > 0 ? kk : ny - 1 - kk;
            const int a = max(0, L - fkk - (nz - 1)), b = min(nx - 1, L - fkk);
            if (b >= a) rank += b - a + 1;
        }
        {
            const int a = max(0, L - fk - (nz - 1)), b = min(nx - 1, L - fk);
            int lo, hi;
            if (sx > 0) { lo = a; hi = min(b, j - 1); }
            else        { lo = nx - 1 - b; hi = min(nx - 1 - a, j - 1); }
            if (hi >= lo) rank += hi - lo + 1;
        }
        const long p = (long) levelPtr[L] - 1 + rank;
        ijkv[4*p]     = i + 1;
        ijkv[4*p + 1] = j + 1;
        ijkv[4*p + 2] = k + 1;
        ijkv[4*p + 3] = (int) g + 1;
    }
}

// locate.F90:748-782
__global__ void k_statics(float* __restrict__ ts, const float* __restrict__ tobs,
                          const float* __restrict__ tt,
                          const float* __restrict__ wts,
                          const float* __restrict__ t0, int ldo, int nobs,
                          int nev)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= nobs) return;
    float s = 0.0f;
    for (int iv = 0; iv < nev; iv++) {
        const long k = (long) iv*ldo + i;
        float test = tt[k] + t0[iv];
        float res = wts[k]*(tobs[k] - test);
        s = s + res;
    }
    ts[i] = s;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, long n,
                           unsigned long long v)
{
    GRID_STRIDE(i, n) a[i] = v;
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : 1; }

} // namespace

extern "C" {

int fteik_k_slowness_from_cells(double* slow, const double* vel, int nz,
                                int nx, int ny, int is3d, int order,
                                hipStream_t s)
{
    long n = (long) (nz - 1)*(nx - 1)*(is3d ? ny - 1 : 1);
    hipLaunchKernelGGL(k_slow_cells, dim3(nblocks(n)), dim3(TPB), 0, s, slow,
                       vel, nz, nx, ny, is3d, order);
    return ok();
}

int fteik_k_slowness_from_nodes(double* slow, const double* vel, int nz,
                                int nx, int ny, int is3d, int order,
                                hipStream_t s)
{
    long n = (long) (nz - 1)*(nx - 1)*(is3d ? ny - 1 : 1);
    hipLaunchKernelGGL(k_slow_nodes, dim3(nblocks(n)), dim3(TPB), 0, s, slow,
                       vel, nz, nx, ny, is3d, order);
    return ok();
}

int fteik_k_slowness_from_velocity32(double* slow, const float* vel, long n,
                                     hipStream_t s)
{
    hipLaunchKernelGGL(k_slow_vel32, dim3(nblocks(n)), dim3(TPB), 0, s, slow,
                       vel, n);
    return ok();
}

int fteik_k_field_reorder(double* out, const double* tt, int nz, int nx,
                          int ny, int is3d, int order, hipStream_t s)
{
    long n = (long) nz*nx*(is3d ? ny : 1);
    hipLaunchKernelGGL(k_reorder, dim3(nblocks(n)), dim3(TPB), 0, s, out, tt,
                       nz, nx, ny, is3d, order);
    return ok();
}

int fteik_k_field_cells(double* out, const double* tt, int nz, int nx,
                        int ny, int is3d, int order, hipStream_t s)
{
    long n = (long) (nz - 1)*(nx - 1)*(is3d ? ny - 1 : 1);
    hipLaunchKernelGGL(k_cells, dim3(nblocks(n)), dim3(TPB), 0, s, out, tt,
                       nz, nx, ny, is3d, order);
    return ok();
}

int fteik_k_receivers(double* ttr, const double* tt, const int* ri,
                      const double* rd, int nrec, int nz, int nx, double dz,
                      double dx, double dy, int is3d, hipStream_t s)
{
    return fteik_k_receivers_many(ttr, tt, 1, 0, ri, rd, nrec, nz, nx, dz,
                                  dx, dy, is3d, s);
}

int fteik_k_receivers_many(double* ttr, const double* tt, int nfield,
                           long fstride, const int* ri, const double* rd,
                           int nrec, int nz, int nx, double dz, double dx,
                           double dy, int is3d, hipStream_t s)
{
    if (nrec < 1 || nfield < 1) return 0;
    if (nfield > 65535) return 1;
    hipLaunchKernelGGL(k_receivers, dim3((nrec + TPB - 1)/TPB, nfield),
                       dim3(TPB), 0, s, ttr, tt, ri, rd, nrec, nz, nx, dz, dx,
                       dy, is3d, fstride);
    return ok();
}

int fteik_k_fill_u64(unsigned long long* a, long n, unsigned long long v,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(nblocks(n)), dim3(TPB), 0, s, a, n, v);
    return ok();
}

int fteik_k_d2f(float* out, const double* in, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_d2f, dim3(nblocks(n)), dim3(TPB), 0, s, out, in, n);
    return ok();
}

int fteik_k_f2d_up(double* out, const float* in, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f2d, dim3(nblocks(n)), dim3(TPB), 0, s, out, in, n);
    return ok();
}

int fteik_k_analytic(double* out, int nz, int nx, int ny, double dz,
                     double dx, double dy, double z0, double x0, double y0,
                     double zs, double xs, double ys, double vconst,
                     double vtop, double vbot, int gradient, hipStream_t s)
{
    GradVars v = gradient_vars(nz, dz, z0, zs, vtop, vbot);
    long n = (long) nz*nx*ny;
    hipLaunchKernelGGL(k_analytic, dim3(nblocks(n)), dim3(TPB), 0, s, out,
                       nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys, vconst,
                       v, gradient);
    return ok();
}

int fteik_k_analytic_points(double* out, const double* z, const double* x,
                            const double* y, int n, int nz, double dz,
                            double z0, double zs, double xs, double ys,
                            double vconst, double vtop, double vbot,
                            int gradient, hipStream_t s)
{
    if (n < 1) return 0;
    GradVars v = gradient_vars(nz, dz, z0, zs, vtop, vbot);
    hipLaunchKernelGGL(k_analytic_points, dim3((n + TPB - 1)/TPB), dim3(TPB),
                       0, s, out, z, x, y, n, zs, xs, ys, vconst, v, gradient);
    return ok();
}

int fteik_k_analytic_reorder(double* out, const double* a, int nz, int nx,
                             int ny, int order, hipStream_t s)
{
    long n = (long) nz*nx*ny;
    hipLaunchKernelGGL(k_analytic_reorder, dim3(nblocks(n)), dim3(TPB), 0, s,
                       out, a, nz, nx, ny, order);
    return ok();
}

int fteik_k_min_merge_f(double* tt, const float* t32, long n, double bias,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_min_merge_f, dim3(nblocks(n)), dim3(TPB), 0, s, tt,
                       t32, n, bias);
    return ok();
}

int fteik_k_graph_ijkv(int* ijkv, const int* levelPtr, int nz, int nx, int ny,
                       int sweep, hipStream_t s)
{
    static const int tz[8] = { 1,  1,  1,  1, -1, -1, -1, -1};
    static const int tx[8] = { 1, -1,  1, -1,  1, -1,  1, -1};
    static const int ty[8] = { 1,  1, -1, -1,  1,  1, -1, -1};
    long n = (long) nz*nx*ny;
    hipLaunchKernelGGL(k_graph_ijkv, dim3(nblocks(n)), dim3(TPB), 0, s, ijkv,
                       levelPtr, nz, nx, ny, tz[sweep - 1], tx[sweep - 1],
                       ty[sweep - 1]);
    return ok();
}

int fteik_k_statics(float* ts, const float* tobs, const float* tt,
                    const float* wts, const float* t0, int ldo, int nobs,
                    int nev, hipStream_t s)
{
    hipLaunchKernelGGL(k_statics, dim3((nobs + TPB - 1)/TPB), dim3(TPB), 0, s,
                       ts, tobs, tt, wts, t0, ldo, nobs, nev);
    return ok();
}

} // extern "C"
