// stencil_eval.hip -- evaluates the device 3D local solvers on caller-given
// node inputs.  Diagnostic entry used by the parity tests to compare the
// GPU stencils one node at a time with the reference's outputs.
#include <hip/hip_runtime.h>
#include "stencils.h"
#include "kernels.h"

namespace {

// in: n rows of [tv te tn tev ten tnv tnve tt0 s1..s7]; node: n rows of
// [sweep i j k] (1-based).
__global__ void k_eval3d(const double* __restrict__ in,
                         const int* __restrict__ node,
                         const fteik::Consts<double>* __restrict__ cst, int n,
                         double* __restrict__ out_noinit,
                         double* __restrict__ out_init)
{
    static const int tz[8] = { 1,  1,  1,  1, -1, -1, -1, -1};
    static const int tx[8] = { 1, -1,  1, -1,  1, -1,  1, -1};
    static const int ty[8] = { 1,  1, -1, -1,  1,  1, -1, -1};
    const int r = blockIdx.x*blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double* p = in + 15L*r;
    fteik::Nbr<double> t{p[0], p[1], p[2], p[3], p[4], p[5], p[6]};
    fteik::Slw<double> s{p[8], p[9], p[10], p[11], p[12], p[13], p[14]};
    const int sw = node[4*r] - 1;
    out_noinit[r] = fteik::noinit3d(*cst, t, s, p[7]);
    out_init[r] = fteik::init3d(*cst, t, s, p[7], node[4*r + 1],
                                node[4*r + 2], node[4*r + 3], tz[sw], tx[sw],
                                ty[sw]);
}

// Latency of the iteration stencil on its dependency chain: each lane runs
// nsteps noinit3d evaluations, each one's neighbour times taken from the
// previous result (a diagonal plane wave, so the 1D, 2D and 3D operators
// all compete), as a column does from level to level in a sweep.  One
// lane's chain is the lower bound of a sweep step.
template <typename R>
__global__ void k_stencil_chain(const fteik::Consts<double>* __restrict__ cd,
                                int nsteps, R* __restrict__ out)
{
    const fteik::Consts<double>& a = *cd;
    fteik::Consts<R> c;
#define CV(f) c.f = (R) a.f
    CV(dz); CV(dx); CV(dy); CV(dzi); CV(dxi); CV(dyi); CV(dz2i); CV(dx2i);
    CV(dy2i); CV(dsum_nine); CV(dsumi); CV(epsSolver); CV(dz2dx2); CV(dz2dy2);
    CV(dx2dy2); CV(dz2i_dx2i); CV(dz2i_dy2i); CV(dx2i_dy2i); CV(dz2i_p_dx2i);
    CV(dz2i_p_dy2i); CV(dx2i_p_dy2i); CV(dz2i_p_dy2i_inv); CV(dz2i_p_dx2i_inv);
    CV(dx2i_p_dy2i_inv); CV(szero); CV(szero2); CV(zsa); CV(xsa); CV(ysa);
#undef CV
    c.zsi = a.zsi; c.xsi = a.xsi; c.ysi = a.ysi; c.pad = 0;
    const R h = c.dz, s0 = R(1)/R(3000);
    fteik::Slw<R> s{s0, s0, s0, s0, s0, s0, s0};
    R t = R(0.1) + R(1e-6)*(R) threadIdx.x;
    const R d1 = h*s0*R(0.57735), d2 = h*s0*R(0.81650);
    for (int k = 0; k < nsteps; k++) {
        fteik::Nbr<R> n{t, t + R(1e-7), t + R(2e-7), t + d1, t + d1 + R(1e-7),
                        t + d1 + R(2e-7), t + d2};
        t = fteik::noinit3d(c, n, s, t + R(1));
    }
    out[blockIdx.x*blockDim.x + threadIdx.x] = t;
}

// The stencil's shortened square root on caller-given arguments.
__global__ void k_sqrt_pos(const double* __restrict__ x,
                           double* __restrict__ y, long n)
{
    const long r = (long) blockIdx.x*blockDim.x + threadIdx.x;
    if (r < n) y[r] = fteik::sqrt_pos(x[r]);
}

} // namespace

extern "C" int fteik_k_sqrt_pos(const double* x, double* y, long n,
                                hipStream_t s)
{
    if (n < 1) return 0;
    hipLaunchKernelGGL(k_sqrt_pos, dim3((unsigned) ((n + 255)/256)),
                       dim3(256), 0, s, x, y, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Times one dependent stencil chain per lane of nwg workgroups of one wave.
extern "C" int fteik_k_time_stencil_chain(const void* cst, int nsteps,
                                          int nwg, int fp32, void* out,
                                          hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_stencil_chain<float>, dim3(nwg), dim3(64), 0, s,
                           (const fteik::Consts<double>*) cst, nsteps,
                           (float*) out);
    else
        hipLaunchKernelGGL(k_stencil_chain<double>, dim3(nwg), dim3(64), 0, s,
                           (const fteik::Consts<double>*) cst, nsteps,
                           (double*) out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

extern "C" int fteik_k_eval3d(const double* in, const int* node,
                              const void* cst, int n, double* out_noinit,
                              double* out_init, hipStream_t s)
{
    hipLaunchKernelGGL(k_eval3d, dim3((n + 127)/128), dim3(128), 0, s, in,
                       node, (const fteik::Consts<double>*) cst, n,
                       out_noinit, out_init);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
