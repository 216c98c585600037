// surrogate.hip -- learned travel-time surrogate for the grid-search
// locator: an MLP 3 -> 256 -> 256 -> 256 -> 256 -> ntf (SiLU) mapping a
// normalised hypocentre to station travel times, trained on the GPU with
// hand-written BF16 MFMA / FP32-accumulate kernels, then used to propose
// candidate cells that the exact tables score.
//
// This has no counterpart in the reference (locate.F90 only searches the
// full grid); the exact full grid search stays the fallback and the
// reference result.
//
// Layout: activations are row-major [rows][features]; weights W_l are
// row-major [out][in] (FP32 master copy, BF16 copy for the MFMA GEMMs).
// The network predicts (T - T0) / scale, where T0 is the straight-ray time
// |x - station| / v0; the baseline takes most of the dynamic range.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "kernels.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short bf16_t;          // storage type

constexpr int TPB = 256;

__device__ __forceinline__ bf16_t f2bf(float x)
{
    // round to nearest even (inputs are finite)
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t) (u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h)
{
    return __uint_as_float(((unsigned) h) << 16);
}
__device__ __forceinline__ float sigm(float z) { return 1.0f/(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z*sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s*(1.0f + z*(1.0f - s));
}

#define GRID_STRIDE(i, n) \
    for (long i = (long) blockIdx.x*blockDim.x + threadIdx.x; i < (n); \
         i += (long) gridDim.x*blockDim.x)

inline int nblocks(long n)
{
    long b = (n + TPB - 1)/TPB;
    return (int) (b < 65536 ? (b > 0 ? b : 1) : 65536);
}
inline int ok() { return hipGetLastError() == hipSuccess ? 0 : 1; }

// ------------------------------------------------------------------ GEMM
// C[m][n] (+)= sum_k A(m,k) B(k,n) (+ bias[n]), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], BF16 operands, FP32 accumulation and output.
// Workgroup tile 64 x 64, four waves of 32 x 32, K in chunks of 32 staged
// through LDS as [row][k] so each lane reads its 8 consecutive k of the
// v_mfma_f32_32x32x16_bf16 operand with one 16-byte LDS read:
//   A/B operand: lane l holds row (l % 32), k = 8*(l / 32) .. +7;
//   C: lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
constexpr int GT = 64, GK = 32, GPAD = 8;

// Staging of one operand tile [64 rows][32 k] into LDS as [row][k].  The
// global reads follow the operand's layout: with k contiguous (sk == 1) a
// thread reads 8 consecutive k of one row as one 16-byte load; with rows
// contiguous (sr == 1, a transposed operand) it reads 8 consecutive rows
// of one k, so a wave's reads are contiguous either way.
__device__ __forceinline__ void stage_tile(const bf16_t* __restrict__ X,
                                           long sr, long sk, long r0, int R,
                                           int k0, int K,
                                           bf16_t (*Xs)[GK + GPAD], int t)
{
    if (sk == 1) {
        const int lr = t >> 2, lk = (t & 3)*8;
        const long r = r0 + lr;
        const int k = k0 + lk;
        const bf16_t* p = X + r*sr + k;
        if (r < R && k + 7 < K && (sr & 7) == 0 &&
            (((uintptr_t) p) & 15) == 0) {
            *(uint4*) &Xs[lr][lk] = *(const uint4*) p;
        } else {
            #pragma unroll
            for (int q = 0; q < 8; q++)
                Xs[lr][lk + q] = (r < R && k + q < K) ? p[q] : (bf16_t) 0;
        }
    } else if (sr == 1) {
        const int lk = t & 31, lr = (t >> 5)*8;
        const long r = r0 + lr;
        const int k = k0 + lk;
        const bf16_t* p = X + (long) k*sk + r;
        bf16_t v[8];
        if (k < K && r + 7 < R && (((uintptr_t) p) & 15) == 0) {
            const uint4 w = *(const uint4*) p;
            const bf16_t* wv = (const bf16_t*) &w;
            #pragma unroll
            for (int q = 0; q < 8; q++) v[q] = wv[q];
        } else {
            #pragma unroll
            for (int q = 0; q < 8; q++)
                v[q] = (k < K && r + q < R) ? p[q] : (bf16_t) 0;
        }
        #pragma unroll
        for (int q = 0; q < 8; q++) Xs[lr + q][lk] = v[q];
    } else {
        const int lr = t >> 2, lk = (t & 3)*8;
        const long r = r0 + lr;
        #pragma unroll
        for (int q = 0; q < 8; q++) {
            const int k = k0 + lk + q;
            Xs[lr][lk + q] = (r < R && k < K) ? X[r*sr + (long) k*sk]
                                              : (bf16_t) 0;
        }
    }
}

// One workgroup: the 64 x 64 tile (blockIdx.y, blockIdx.x) of C over the
// k range of split blockIdx.z (kper values each).  nsplit == 1 writes C
// (with bias, optionally accumulating); otherwise the partial sums go to
// P[z][M][N] for k_gemm_reduce.
__global__ __launch_bounds__(256)
void k_gemm_bf16(const bf16_t* __restrict__ A, long sam, long sak,
                 const bf16_t* __restrict__ B, long sbk, long sbn,
                 float* __restrict__ C, long ldc, int M, int N, int K,
                 const float* __restrict__ bias, int accumulate, int kper,
                 float* __restrict__ P)
{
    __shared__ __attribute__((aligned(16))) bf16_t As[GT][GK + GPAD];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[GT][GK + GPAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const long m0 = (long) blockIdx.y*GT, n0 = (long) blockIdx.x*GT;
    const int kb = blockIdx.z*kper, ke = min(K, kb + kper);
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += GK) {
        stage_tile(A, sam, sak, m0, M, k0, ke, As, t);
        stage_tile(B, sbn, sbk, n0, N, k0, ke, Bs, t);
        __syncthreads();
        #pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            const bf16x8 a = *(const bf16x8*) &As[wm*32 + (lane & 31)]
                                                  [kk + 8*(lane >> 5)];
            const bf16x8 b = *(const bf16x8*) &Bs[wn*32 + (lane & 31)]
                                                  [kk + 8*(lane >> 5)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const long n = n0 + wn*32 + (lane & 31);
    if (n >= N) return;
    if (P) {
        float* Pz = P + (long) blockIdx.z*M*N;
        #pragma unroll
        for (int r = 0; r < 16; r++) {
            const long m = m0 + wm*32 + 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
            if (m < M) Pz[m*N + n] = acc[r];
        }
        return;
    }
    const float bn = bias ? bias[n] : 0.0f;
    #pragma unroll
    for (int r = 0; r < 16; r++) {
        const long m = m0 + wm*32 + 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
        if (m < M) {
            float v = acc[r] + bn;
            if (accumulate) v += C[m*ldc + n];
            C[m*ldc + n] = v;
        }
    }
}

// C = sum of the nsplit partial products (in split order) + bias.
__global__ void k_gemm_reduce(const float* __restrict__ P, int nsplit, int M,
                              int N, float* __restrict__ C, long ldc,
                              const float* __restrict__ bias, int accumulate)
{
    GRID_STRIDE(i, (long) M*N) {
        const long m = i / N, n = i - m*N;
        float v = 0.0f;
        for (int z = 0; z < nsplit; z++) v += P[(long) z*M*N + i];
        v += bias ? bias[n] : 0.0f;
        if (accumulate) v += C[m*ldc + n];
        C[m*ldc + n] = v;
    }
}

// workspace of the split-K partial products (grown on demand)
float* g_split = nullptr;
size_t g_split_cap = 0;

// ------------------------------------------------------ elementwise parts
// First layer on the VALU (K = 3): z = W1 x + b1, h = silu(z).
__global__ void k_layer1(const float* __restrict__ x, const float* __restrict__ W1,
                         const float* __restrict__ b1, float* __restrict__ z,
                         bf16_t* __restrict__ h, long rows, int H)
{
    GRID_STRIDE(i, rows*H) {
        const long r = i / H;
        const int j = (int) (i - r*H);
        const float v = W1[3*j]*x[3*r] + W1[3*j + 1]*x[3*r + 1]
                      + W1[3*j + 2]*x[3*r + 2] + b1[j];
        z[i] = v;
        h[i] = f2bf(silu(v));
    }
}

__global__ void k_silu(const float* __restrict__ z, bf16_t* __restrict__ h,
                       long n)
{
    GRID_STRIDE(i, n) h[i] = f2bf(silu(z[i]));
}

// dz = dh * silu'(z), to BF16 for the next GEMMs (and FP32 for the bias
// gradient)
__global__ void k_silu_bwd(const float* __restrict__ dh,
                           const float* __restrict__ z,
                           bf16_t* __restrict__ dz, long n)
{
    GRID_STRIDE(i, n) dz[i] = f2bf(dh[i]*dsilu(z[i]));
}

// Mean squared error over rows x nout; dy = 2 (y - t) / (rows nout).
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ tgt,
                      bf16_t* __restrict__ dy, long n, float inv,
                      double* __restrict__ loss)
{
    __shared__ double red[TPB];
    double s = 0.0;
    GRID_STRIDE(i, n) {
        const float d = y[i] - tgt[i];
        dy[i] = f2bf(2.0f*d*inv);
        s += (double) d*d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = TPB/2; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]*(double) inv);
}

// Column sums of a BF16 [rows][cols] matrix (bias gradients).
__global__ void k_colsum(const bf16_t* __restrict__ d, long rows, int cols,
                         float* __restrict__ out)
{
    const int c = blockIdx.x;
    float s = 0.0f;
    for (long r = threadIdx.x; r < rows; r += blockDim.x)
        s += bf2f(d[r*cols + c]);
    __shared__ float red[TPB];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = TPB/2; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// dW1[j][d] = sum_r dz1[r][j] x[r][d]
__global__ void k_dw1(const bf16_t* __restrict__ dz, const float* __restrict__ x,
                      long rows, int H, float* __restrict__ dW1)
{
    const int j = blockIdx.x;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (long r = threadIdx.x; r < rows; r += blockDim.x) {
        const float g = bf2f(dz[r*H + j]);
        s0 += g*x[3*r]; s1 += g*x[3*r + 1]; s2 += g*x[3*r + 2];
    }
    __shared__ float red[3][TPB];
    red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1; red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int o = TPB/2; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o)
            for (int q = 0; q < 3; q++)
                red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) dW1[3*j + threadIdx.x] = red[threadIdx.x][0];
}

// Adam on the FP32 master parameters; refreshes the BF16 copies.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long n,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2, bf16_t* __restrict__ pb)
{
    GRID_STRIDE(i, n) {
        const float gi = g[i];
        const float mi = b1*m[i] + (1.0f - b1)*gi;
        const float vi = b2*v[i] + (1.0f - b2)*gi*gi;
        m[i] = mi; v[i] = vi;
        const float pi = p[i] - lr*(mi*c1)/(sqrtf(vi*c2) + eps);
        p[i] = pi;
        pb[i] = f2bf(pi);
    }
}

__global__ void k_f2bf(const float* __restrict__ a, bf16_t* __restrict__ b,
                       long n)
{
    GRID_STRIDE(i, n) b[i] = f2bf(a[i]);
}

// ------------------------------------------------------------ data
struct GeoArgs {
    int nz, nx, ny, ntf;
    float dz, dx, dy;            // spacing (m)
    float sx[64], sy[64], sz[64];  // stations in grid metres
    float v0, scale;             // straight-ray velocity, output scale (s)
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    // splitmix64 finaliser
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27))*0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// normalised coordinate of grid node c (natural order) in [-1, 1]^3
__device__ __forceinline__ void node_xyz(long c, const GeoArgs& g, float* p)
{
    const long nzx = (long) g.nz*g.nx;
    const int k = (int) (c / nzx);
    const long rem = c - (long) k*nzx;
    const int j = (int) (rem / g.nz), i = (int) (rem - (long) j*g.nz);
    p[0] = 2.0f*(float) j/(float) (g.nx - 1) - 1.0f;     // x
    p[1] = 2.0f*(float) k/(float) (g.ny - 1) - 1.0f;     // y
    p[2] = 2.0f*(float) i/(float) (g.nz - 1) - 1.0f;     // z
}

// straight-ray time to station s at normalised position p, and its
// gradient with respect to p
__device__ __forceinline__ float t0_ray(const float* p, int s, const GeoArgs& g,
                                        float* grad)
{
    const float hx = 0.5f*(float) (g.nx - 1)*g.dx;
    const float hy = 0.5f*(float) (g.ny - 1)*g.dy;
    const float hz = 0.5f*(float) (g.nz - 1)*g.dz;
    const float X = (p[0] + 1.0f)*hx - g.sx[s];
    const float Y = (p[1] + 1.0f)*hy - g.sy[s];
    const float Z = (p[2] + 1.0f)*hz - g.sz[s];
    const float d = sqrtf(X*X + Y*Y + Z*Z) + 1e-3f;
    if (grad) {
        grad[0] = X*hx/(d*g.v0); grad[1] = Y*hy/(d*g.v0);
        grad[2] = Z*hz/(d*g.v0);
    }
    return d/g.v0;
}

// Training samples: rows r of step `step` take cell hash(seed, index) of a
// pool of `pool` uniformly drawn cells; inputs and scaled residual targets.
__global__ void k_samples(const float* __restrict__ tab, long ldgrd, GeoArgs g,
                          long ngrd, unsigned long long seed, long pool,
                          long first, long rows, float* __restrict__ x,
                          float* __restrict__ tgt, int ldt)
{
    GRID_STRIDE(r, rows) {
        const long idx = (first + r) % pool;
        const long c = (long) (mix64(seed*0x100000001b3ull + (unsigned long long) idx)
                               % (unsigned long long) ngrd);
        float p[3];
        node_xyz(c, g, p);
        x[3*r] = p[0]; x[3*r + 1] = p[1]; x[3*r + 2] = p[2];
        for (int s = 0; s < ldt; s++) {
            float v = 0.0f;
            if (s < g.ntf)
                v = (tab[(long) s*ldgrd + c] - t0_ray(p, s, g, nullptr))/g.scale;
            tgt[r*ldt + s] = v;
        }
    }
}

// ------------------------------------------------------------ inversion
// Rows are laid out as 4 blocks of R: primal, then tangents along x, y, z.
__global__ void k_inv_layer1(const float* __restrict__ pos, long R,
                             const float* __restrict__ W1,
                             const float* __restrict__ b1,
                             float* __restrict__ zp, bf16_t* __restrict__ h,
                             int H)
{
    GRID_STRIDE(i, R*H) {
        const long r = i / H;
        const int j = (int) (i - r*H);
        const float* p = pos + 3*r;
        const float z = W1[3*j]*p[0] + W1[3*j + 1]*p[1] + W1[3*j + 2]*p[2]
                      + b1[j];
        const float ds = dsilu(z);
        zp[i] = z;
        h[i] = f2bf(silu(z));
        for (int d = 0; d < 3; d++)
            h[(long) (d + 1)*R*H + i] = f2bf(ds*W1[3*j + d]);
    }
}

// Hidden layer of the inversion network, fused: 4 row blocks (primal and
// three tangents) of R rows, H = 256.  z = h_in W^T on the MFMA units with
// FP32 accumulation in the same k order as k_gemm_bf16 (so z is bitwise
// the GEMM's), then h = silu(z + b) for the primal block and
// h_d = silu'(z_0 + b) z_d for the tangents, rounded to BF16: what
// k_gemm_bf16 followed by k_inv_act computes, without the FP32 round trip
// through memory.  One workgroup per 32 rows; wave d computes block d with
// its A rows held in registers over the whole K, the weights stream from
// L2 as 16-byte vectors, and the primal wave hands silu' to the tangent
// waves through LDS one 32-column tile at a time.
__global__ __launch_bounds__(256)
void k_inv_hidden(const bf16_t* __restrict__ hin, long R,
                  const bf16_t* __restrict__ W, const float* __restrict__ b,
                  bf16_t* __restrict__ hout)
{
    constexpr int H = 256, KS = H/16;
    __shared__ float ds[2][32][33];
    const int lane = threadIdx.x & 63, d = threadIdx.x >> 6;
    const long s0 = (long) blockIdx.x*32;
    const long arow = s0 + (lane & 31);
    const bool aok = arow < R;
    const bf16_t* ap = hin + ((long) d*R + (aok ? arow : 0))*H + 8*(lane >> 5);
    bf16x8 a[KS];
    #pragma unroll
    for (int kk = 0; kk < KS; kk++) {
        bf16x8 v = {};
        if (aok) v = *(const bf16x8*) (ap + 16*kk);
        a[kk] = v;
    }
    for (int ct = 0; ct < H/32; ct++) {
        const int n = 32*ct + (lane & 31);
        const bf16_t* wp = W + (long) n*H + 8*(lane >> 5);
        f32x16 acc = {};
        #pragma unroll
        for (int kk = 0; kk < KS; kk++) {
            const bf16x8 bv = *(const bf16x8*) (wp + 16*kk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk], bv, acc,
                                                          0, 0, 0);
        }
        const float bn = b[n];
        float (*dsb)[33] = ds[ct & 1];
        if (d == 0) {
            #pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
                const float z = acc[r] + bn;
                dsb[m][lane & 31] = dsilu(z);
                if (s0 + m < R) hout[(s0 + m)*H + n] = f2bf(silu(z));
            }
        }
        __syncthreads();
        if (d > 0) {
            #pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
                if (s0 + m < R)
                    hout[((long) d*R + s0 + m)*H + n] =
                        f2bf(dsb[m][lane & 31]*acc[r]);
            }
        }
    }
}

// After a hidden GEMM over all 4R rows (no bias): primal h = silu(z + b),
// tangents h_d = silu'(z + b) * z_d.
__global__ void k_inv_act(const float* __restrict__ zall, long R, int H,
                          const float* __restrict__ b, float* __restrict__ zp,
                          bf16_t* __restrict__ h)
{
    GRID_STRIDE(i, R*H) {
        const int j = (int) (i % H);
        const float z = zall[i] + b[j];
        const float ds = dsilu(z);
        zp[i] = z;
        h[i] = f2bf(silu(z));
        for (int d = 1; d < 4; d++)
            h[(long) d*R*H + i] = f2bf(ds*zall[(long) d*R*H + i]);
    }
}

// One damped Gauss-Newton step of the locator misfit
// 1/2 sum_i w_i^2 (r_i - t0)^2, r = obs - T(p), t0 the weighted mean of r,
// for every (event, start) row; y holds 4 blocks of R x ldy outputs.
__global__ void k_inv_gn(const float* __restrict__ y, long R, int ldy,
                         const float* __restrict__ b5, GeoArgs g,
                         const float* __restrict__ obs,
                         const float* __restrict__ wt, int lntf, int starts,
                         float* __restrict__ pos, float* __restrict__ misfit,
                         int apply)
{
    GRID_STRIDE(r, R) {
        const long e = r / starts;
        const float* o = obs + e*lntf;
        const float* w = wt + e*lntf;
        float* p = pos + 3*r;
        float res[64], J[64][3];
        float sw = 0.0f, mr = 0.0f, mj[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < g.ntf; s++) {
            float gr[3];
            const float t = t0_ray(p, s, g, gr)
                          + g.scale*(y[r*ldy + s] + b5[s]);
            res[s] = o[s] - t;
            for (int d = 0; d < 3; d++)
                J[s][d] = gr[d] + g.scale*y[(long) (d + 1)*R*ldy + r*ldy + s];
            const float ws = w[s];
            sw += ws; mr += ws*res[s];
            for (int d = 0; d < 3; d++) mj[d] += ws*J[s][d];
        }
        const float inv = sw > 0.0f ? 1.0f/sw : 0.0f;
        mr *= inv;
        for (int d = 0; d < 3; d++) mj[d] *= inv;
        float A[3][3] = {}, bb[3] = {}, f = 0.0f;
        for (int s = 0; s < g.ntf; s++) {
            const float w2 = w[s]*w[s];
            const float rc = res[s] - mr;
            float jc[3];
            for (int d = 0; d < 3; d++) jc[d] = J[s][d] - mj[d];
            f += 0.5f*w2*rc*rc;
            for (int a = 0; a < 3; a++) {
                bb[a] += w2*jc[a]*rc;
                for (int c = 0; c < 3; c++) A[a][c] += w2*jc[a]*jc[c];
            }
        }
        misfit[r] = f;
        if (!apply) continue;
        const float lam = 1e-3f*(A[0][0] + A[1][1] + A[2][2]) + 1e-12f;
        for (int a = 0; a < 3; a++) A[a][a] += lam;
        // 3 x 3 solve by Cramer's rule
        const float det = A[0][0]*(A[1][1]*A[2][2] - A[1][2]*A[2][1])
                        - A[0][1]*(A[1][0]*A[2][2] - A[1][2]*A[2][0])
                        + A[0][2]*(A[1][0]*A[2][1] - A[1][1]*A[2][0]);
        if (!(fabsf(det) > 0.0f)) continue;
        float dlt[3];
        for (int q = 0; q < 3; q++) {
            float M[3][3];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) M[a][c] = (c == q) ? bb[a] : A[a][c];
            dlt[q] = (M[0][0]*(M[1][1]*M[2][2] - M[1][2]*M[2][1])
                    - M[0][1]*(M[1][0]*M[2][2] - M[1][2]*M[2][0])
                    + M[0][2]*(M[1][0]*M[2][1] - M[1][1]*M[2][0]))/det;
        }
        // residual r = o - T: dr/dp = -J, so the step is +A^-1 J^T w r
        for (int d = 0; d < 3; d++) {
            float step = dlt[d];
            step = fminf(fmaxf(step, -0.25f), 0.25f);
            p[d] = fminf(fmaxf(p[d] + step, -1.0f), 1.0f);
        }
    }
}

// Exact scoring of the 4 x 4 x 4 cells around each event's best start, one
// wave per event (one lane per cell), with the locator's objective in its
// operation order.  If the best cell lies on a face of the cube that is not
// a face of the grid, the cube is re-centred on it and scored again (at
// most `rounds` times); events still on a face are flagged for the full
// grid search.
__device__ __forceinline__ unsigned long long cube_min(
    int i0, int j0, int k0, int lane, int e, const GeoArgs& g,
    const float* __restrict__ tab, long ldgrd, const float* __restrict__ obs,
    const float* __restrict__ wt, const unsigned char* __restrict__ skip,
    const float* __restrict__ xnorm, int lntf)
{
    const int di = lane & 3, dj = (lane >> 2) & 3, dk = lane >> 4;
    const long c = (long) (k0 + dk)*g.nz*g.nx + (long) (j0 + dj)*g.nz + i0 + di;
    const float* o = obs + (long) e*lntf;
    const float* w = wt + (long) e*lntf;
    const unsigned char* sk = skip + (long) e*lntf;
    const float sqrtHalf = 0.70710677f;
    float t0 = 0.0f;
    for (int s = 0; s < g.ntf; s++) {
        if (sk[s]) continue;
        float res = xnorm[e]*(o[s] - tab[(long) s*ldgrd + c]);
        t0 = t0 + w[s]*res;
    }
    float obj = 0.0f;
    for (int s = 0; s < g.ntf; s++) {
        if (sk[s]) continue;
        float ws = sqrtHalf*w[s];
        float est = tab[(long) s*ldgrd + c] + t0;
        float res = ws*(o[s] - est);
        obj = obj + res*res;
    }
    unsigned long long key = ((unsigned long long) __float_as_uint(obj) << 32)
                           | (unsigned long long) (unsigned) c;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long q = __shfl_xor(key, off, 64);
        key = q < key ? q : key;
    }
    return key;
}

__global__ void k_inv_score(const float* __restrict__ pos,
                            const float* __restrict__ misfit, int starts,
                            int nev, GeoArgs g, const float* __restrict__ tab,
                            long ldgrd, const float* __restrict__ obs,
                            const float* __restrict__ wt,
                            const unsigned char* __restrict__ skip,
                            const float* __restrict__ xnorm, int lntf,
                            int rounds, unsigned long long* __restrict__ best,
                            int* __restrict__ fallback)
{
    const int e = blockIdx.x*(blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= nev) return;
    int bs = 0;
    float bf = misfit[(long) e*starts];
    for (int s = 1; s < starts; s++) {
        const float f = misfit[(long) e*starts + s];
        if (f < bf) { bf = f; bs = s; }
    }
    const float* p = pos + 3*((long) e*starts + bs);
    int cj = (int) rintf((p[0] + 1.0f)*0.5f*(float) (g.nx - 1));
    int ck = (int) rintf((p[1] + 1.0f)*0.5f*(float) (g.ny - 1));
    int ci = (int) rintf((p[2] + 1.0f)*0.5f*(float) (g.nz - 1));
    unsigned long long key = ~0ull;
    bool face = true;
    for (int r = 0; r < rounds && face; r++) {
        const int i0 = min(max(ci - 1, 0), g.nz - 4);
        const int j0 = min(max(cj - 1, 0), g.nx - 4);
        const int k0 = min(max(ck - 1, 0), g.ny - 4);
        key = cube_min(i0, j0, k0, lane, e, g, tab, ldgrd, obs, wt, skip,
                       xnorm, lntf);
        const long wc = (long) (key & 0xffffffffull);
        const long nzx = (long) g.nz*g.nx;
        const int wk = (int) (wc / nzx), wj = (int) ((wc % nzx)/g.nz),
                  wi = (int) (wc % g.nz);
        face = (wi == i0 && i0 > 0) || (wi == i0 + 3 && i0 + 4 < g.nz)
            || (wj == j0 && j0 > 0) || (wj == j0 + 3 && j0 + 4 < g.nx)
            || (wk == k0 && k0 > 0) || (wk == k0 + 3 && k0 + 4 < g.ny);
        ci = wi; cj = wj; ck = wk;
    }
    if (lane == 0) {
        best[e] = key;
        fallback[e] = face ? 1 : 0;
    }
}

__global__ void k_inv_start(float* __restrict__ pos, long R, int starts)
{
    GRID_STRIDE(r, R) {
        const int s = (int) (r % starts);
        // starts on a 2 x 2 x 2 lattice (plus the centre when starts = 9)
        float q[3];
        if (s < 8) {
            q[0] = (s & 1) ? 0.5f : -0.5f;
            q[1] = (s & 2) ? 0.5f : -0.5f;
            q[2] = (s & 4) ? 0.5f : -0.5f;
        } else {
            q[0] = q[1] = q[2] = 0.0f;
        }
        pos[3*r] = q[0]; pos[3*r + 1] = q[1]; pos[3*r + 2] = q[2];
    }
}

} // namespace

extern "C" {

int fteik_k_gemm_bf16(const void* A, long sam, long sak, const void* B,
                      long sbk, long sbn, float* C, long ldc, int M, int N,
                      int K, const float* bias, int accumulate,
                      hipStream_t s)
{
    if (M < 1 || N < 1 || K < 1) return 0;
    const int tiles = ((N + GT - 1)/GT)*((M + GT - 1)/GT);
    // few output tiles over a long k (the weight gradients: 256 x 256 over
    // the batch rows): split k so the GPU's 256 CUs all work, then sum the
    // partial products in a fixed order (the result does not depend on
    // timing)
    int nsplit = 1;
    if (tiles < 256 && K >= 2048)
        nsplit = std::min(std::max(512/tiles, 1), K/256);
    const int kper = ((K + nsplit - 1)/nsplit + GK - 1)/GK*GK;
    nsplit = (K + kper - 1)/kper;
    float* P = nullptr;
    if (nsplit > 1) {
        const size_t need = (size_t) nsplit*M*N;
        if (need > g_split_cap) {
            if (g_split) hipFree(g_split);
            g_split = nullptr;
            g_split_cap = 0;
            if (hipMalloc((void**) &g_split, need*sizeof(float)) != hipSuccess)
                return 1;
            g_split_cap = need;
        }
        P = g_split;
    }
    dim3 grid((N + GT - 1)/GT, (M + GT - 1)/GT, nsplit);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s,
                       (const bf16_t*) A, sam, sak, (const bf16_t*) B, sbk,
                       sbn, C, ldc, M, N, K, bias, accumulate, kper, P);
    if (nsplit > 1)
        hipLaunchKernelGGL(k_gemm_reduce, dim3(nblocks((long) M*N)), dim3(TPB),
                           0, s, P, nsplit, M, N, C, ldc, bias, accumulate);
    return ok();
}

void fteik_k_sur_release(void)
{
    if (g_split) hipFree(g_split);
    g_split = nullptr;
    g_split_cap = 0;
}

int fteik_k_f2bf(const float* a, void* b, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f2bf, dim3(nblocks(n)), dim3(TPB), 0, s, a,
                       (bf16_t*) b, n);
    return ok();
}

int fteik_k_sur_layer1(const float* x, const float* W1, const float* b1,
                       float* z, void* h, long rows, int H, hipStream_t s)
{
    hipLaunchKernelGGL(k_layer1, dim3(nblocks(rows*H)), dim3(TPB), 0, s, x,
                       W1, b1, z, (bf16_t*) h, rows, H);
    return ok();
}

int fteik_k_sur_silu(const float* z, void* h, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_silu, dim3(nblocks(n)), dim3(TPB), 0, s, z,
                       (bf16_t*) h, n);
    return ok();
}

int fteik_k_sur_silu_bwd(const float* dh, const float* z, void* dz, long n,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_silu_bwd, dim3(nblocks(n)), dim3(TPB), 0, s, dh, z,
                       (bf16_t*) dz, n);
    return ok();
}

int fteik_k_sur_mse(const float* y, const float* tgt, void* dy, long n,
                    float inv, double* loss, hipStream_t s)
{
    hipLaunchKernelGGL(k_mse, dim3(nblocks(n) < 1024 ? nblocks(n) : 1024),
                       dim3(TPB), 0, s, y, tgt, (bf16_t*) dy, n, inv, loss);
    return ok();
}

int fteik_k_sur_colsum(const void* d, long rows, int cols, float* out,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3(cols), dim3(TPB), 0, s,
                       (const bf16_t*) d, rows, cols, out);
    return ok();
}

int fteik_k_sur_dw1(const void* dz, const float* x, long rows, int H,
                    float* dW1, hipStream_t s)
{
    hipLaunchKernelGGL(k_dw1, dim3(H), dim3(TPB), 0, s, (const bf16_t*) dz, x,
                       rows, H, dW1);
    return ok();
}

int fteik_k_sur_adam(float* p, const float* g, float* m, float* v, long n,
                     float lr, float b1, float b2, float eps, float c1,
                     float c2, void* pb, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(nblocks(n)), dim3(TPB), 0, s, p, g, m, v,
                       n, lr, b1, b2, eps, c1, c2, (bf16_t*) pb);
    return ok();
}

int fteik_k_sur_samples(const float* tab, long ldgrd, const void* geo,
                        long ngrd, unsigned long long seed, long pool,
                        long first, long rows, float* x, float* tgt, int ldt,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_samples, dim3(nblocks(rows)), dim3(TPB), 0, s, tab,
                       ldgrd, *(const GeoArgs*) geo, ngrd, seed, pool, first,
                       rows, x, tgt, ldt);
    return ok();
}

int fteik_k_sur_geo_size(void) { return (int) sizeof(GeoArgs); }

int fteik_k_inv_start(float* pos, long R, int starts, hipStream_t s)
{
    hipLaunchKernelGGL(k_inv_start, dim3(nblocks(R)), dim3(TPB), 0, s, pos, R,
                       starts);
    return ok();
}

int fteik_k_inv_hidden(const void* hin, long R, const void* W,
                       const float* b, void* hout, int H, hipStream_t s)
{
    if (H != 256 || R < 1) return 2;
    hipLaunchKernelGGL(k_inv_hidden, dim3((unsigned) ((R + 31)/32)),
                       dim3(256), 0, s, (const bf16_t*) hin, R,
                       (const bf16_t*) W, b, (bf16_t*) hout);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int fteik_k_inv_layer1(const float* pos, long R, const float* W1,
                       const float* b1, float* zp, void* h, int H,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_inv_layer1, dim3(nblocks(R*H)), dim3(TPB), 0, s, pos,
                       R, W1, b1, zp, (bf16_t*) h, H);
    return ok();
}

int fteik_k_inv_act(const float* zall, long R, int H, const float* b,
                    float* zp, void* h, hipStream_t s)
{
    hipLaunchKernelGGL(k_inv_act, dim3(nblocks(R*H)), dim3(TPB), 0, s, zall, R,
                       H, b, zp, (bf16_t*) h);
    return ok();
}

int fteik_k_inv_gn(const float* y, long R, int ldy, const float* b5,
                   const void* geo, const float* obs, const float* wt,
                   int lntf, int starts, float* pos, float* misfit, int apply,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_inv_gn, dim3(nblocks(R)), dim3(TPB), 0, s, y, R, ldy,
                       b5, *(const GeoArgs*) geo, obs, wt, lntf, starts, pos,
                       misfit, apply);
    return ok();
}

int fteik_k_inv_score(const float* pos, const float* misfit, int starts,
                      int nev, const void* geo, const float* tab, long ldgrd,
                      const float* obs, const float* wt,
                      const unsigned char* skip, const float* xnorm, int lntf,
                      unsigned long long* best, int* fallback, hipStream_t s)
{
    const int wpb = 4;
    hipLaunchKernelGGL(k_inv_score, dim3((nev + wpb - 1)/wpb), dim3(64*wpb),
                       0, s, pos, misfit, starts, nev,
                       *(const GeoArgs*) geo, tab, ldgrd, obs, wt, skip, xnorm,
                       lntf, 4, best, fallback);
    return ok();
}

} // extern "C"
