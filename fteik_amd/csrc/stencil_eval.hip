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

} // namespace

extern "C" int fteik_k_eval3d(const double* in, const int* node,
                              const void* cst, int n, double* out_noinit,
                              double* out_init, hipStream_t s)
{
    hipLaunchKernelGGL(k_eval3d, dim3((n + 127)/128), dim3(128), 0, s, in,
                       node, (const fteik::Consts<double>*) cst, n,
                       out_noinit, out_init);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
