// sweep2d.hip -- 2D level-set fast sweeping on gfx950.
//
// solver2D.F90:782-978: 4 initialization sweeps (lexicographic in the
// reference, :825-867) and nsweep x 4 Gauss-Seidel sweeps over the levels
// ix+iz (:897-925).  Every sweep is evaluated here level by level -- a
// topological order of the same dependency graph, so the results equal the
// reference's bit for bit -- by one workgroup per source: threads own x
// columns, one barrier per level.  Several sources (solveLSM) run as
// independent workgroups of one launch.
#include <hip/hip_runtime.h>
#include "stencils.h"
#include "kernels.h"

namespace fteik {

constexpr int T2 = 512;

// localSolver.F90:160-193
__device__ __forceinline__ void signs2d(int sweep, int& sgntz, int& sgntx)
{
    sgntz = (sweep <= 2) ? 1 : -1;
    sgntx = (sweep == 1 || sweep == 3) ? 1 : -1;
}

__global__ __launch_bounds__(T2)
void solve2d_kernel(double* __restrict__ ttall, const double* __restrict__ slow,
                    int nz, int nx, double dz, double dx, double eps,
                    int nsweep, long ld_tt, const double* __restrict__ zsa_v,
                    const double* __restrict__ xsa_v,
                    const int* __restrict__ zsi_v, const int* __restrict__ xsi_v)
{
    const int src = blockIdx.x;
    double* tt = ttall + (long) src*ld_tt;
    const long ngrd = (long) nz*nx;
    __shared__ Consts<double> cs;
    if (threadIdx.x == 0) {
        // localSolver.F90:299-345
        Consts<double> c;
        c.zsa = zsa_v[src]; c.xsa = xsa_v[src]; c.ysa = 1.0;
        c.zsi = zsi_v[src]; c.xsi = xsi_v[src]; c.ysi = 1;
        c.szero = slow[(long) (c.xsi - 1)*(nz - 1) + c.zsi - 1];
        c.szero2 = c.szero*c.szero;
        c.epsSolver = eps;
        c.dz = dz; c.dx = dx; c.dy = 1.0;
        const double dz2 = dz*dz, dx2 = dx*dx;
        c.dzi = 1.0/dz; c.dxi = 1.0/dx; c.dyi = 1.0;
        c.dz2i = 1.0/dz2; c.dx2i = 1.0/dx2; c.dy2i = 1.0;
        c.dz2i_dx2i = c.dz2i*c.dx2i;
        c.dz2i_p_dx2i = c.dz2i + c.dx2i;
        c.dz2i_p_dx2i_inv = 1.0/(c.dz2i + c.dx2i);
        cs = c;
    }
    __syncthreads();
    const Consts<double> c = cs;
    for (long g = threadIdx.x; g < ngrd; g += blockDim.x) tt[g] = kHuge;
    __syncthreads();
    if (threadIdx.x < 4) {
        const int i = c.zsi + (threadIdx.x & 1), j = c.xsi + (threadIdx.x >> 1);
        tt[(long) (j - 1)*nz + i - 1] = tana2d(i, j, c);
    }
    __syncthreads();
    const int zsi = c.zsi, xsi = c.xsi;
    const int nsw = 4 + 4*nsweep;
    for (int q = 0; q < nsw; q++) {
        const bool linit = q < 4;
        const int sweep = (q & 3) + 1;
        int sgntz, sgntx;
        signs2d(sweep, sgntz, sgntx);
        // actual 1-based box (solver2D.F90:826-866 init, :377-454 iterate)
        int z0, z1, x0, x1;
        if (sgntz > 0) { z0 = linit ? max(2, zsi) : 2; z1 = nz; }
        else           { z0 = 1; z1 = linit ? zsi + 1 : nz - 1; }
        if (sgntx > 0) { x0 = linit ? max(2, xsi) : 2; x1 = nx; }
        else           { x0 = 1; x1 = linit ? xsi + 1 : nx - 1; }
        // flipped 0-based box
        const int fz0 = (sgntz > 0) ? z0 - 1 : nz - z1;
        const int fz1 = (sgntz > 0) ? z1 - 1 : nz - z0;
        const int fx0 = (sgntx > 0) ? x0 - 1 : nx - x1;
        const int fx1 = (sgntx > 0) ? x1 - 1 : nx - x0;
        if (fz0 > fz1 || fx0 > fx1) continue;
        for (int L = fz0 + fx0; L <= fz1 + fx1; L++) {
            for (int fj = fx0 + (int) threadIdx.x; fj <= fx1; fj += blockDim.x) {
                const int fi = L - fj;
                if (fi < fz0 || fi > fz1) continue;
                const int iz = (sgntz > 0) ? fi + 1 : nz - fi;
                const int ix = (sgntx > 0) ? fj + 1 : nx - fj;
                const long me = (long) (ix - 1)*nz + iz - 1;
                const double tv  = tt[(long) (ix - 1)*nz + iz - sgntz - 1];
                const double te  = tt[(long) (ix - sgntx - 1)*nz + iz - 1];
                const double tev = tt[(long) (ix - sgntx - 1)*nz + iz - sgntz - 1];
                const double t   = tt[me];
                // solver2D.F90:491-502 / localSolver.F90:222-231
                const int izc = iz - (sgntz > 0 ? 1 : 0);
                const int ixc = ix - (sgntx > 0 ? 1 : 0);
                const double s1 = rmin(slow[(long) (max(ix - 1, 1) - 1)*(nz - 1) + izc - 1],
                                       slow[(long) (min(ix, nx - 1) - 1)*(nz - 1) + izc - 1]);
                const double s2 = rmin(slow[(long) (ixc - 1)*(nz - 1) + max(iz - 1, 1) - 1],
                                       slow[(long) (ixc - 1)*(nz - 1) + min(iz, nz - 1) - 1]);
                const double s3 = slow[(long) (ixc - 1)*(nz - 1) + izc - 1];
                double tupd;
                if (linit)
                    tupd = init2d(c, sgntz, sgntx, iz, ix, tv, te, tev, t, s1,
                                  s2, s3);
                else
                    tupd = noinit2d(c, tv, te, tev, s1, s2, s3);
                tt[me] = rmin(t, tupd);
            }
            __syncthreads();
        }
    }
}

} // namespace fteik

using namespace fteik;

extern "C" int fteik_k_solve2d(double* tt, const double* slow, int nz, int nx,
                               double dz, double dx, double eps, double zsa,
                               double xsa, int zsi, int xsi, int nsweep,
                               int nsrc_batch, int ld_tt, const double* zsa_v,
                               const double* xsa_v, const int* zsi_v,
                               const int* xsi_v, hipStream_t stream)
{
    (void) zsa; (void) xsa; (void) zsi; (void) xsi;
    hipLaunchKernelGGL(solve2d_kernel, dim3(nsrc_batch), dim3(T2), 0, stream,
                       tt, slow, nz, nx, dz, dx, eps, nsweep, (long) ld_tt,
                       zsa_v, xsa_v, zsi_v, xsi_v);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
