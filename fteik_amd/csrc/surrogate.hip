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
#include <type_traits>
#include "kernels.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x32 __attribute__((ext_vector_type(32)));
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
// logistic function with the hardware reciprocal (1 ulp; the network's
// activations are rounded to BF16 anyway)
__device__ __forceinline__ float sigm(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
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
// C[m][n] = sum_k A(m,k) B(k,n), then an epilogue; A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], BF16 operands, FP32 accumulation.
// Workgroup tile 64 x 64, four waves of 32 x 32 (v_mfma_f32_32x32x16_bf16):
//   A/B operand: lane l holds row (l % 32), k = 8*(l / 32) .. +7;
//   C: lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
// A workgroup stages up to GKW = 256 values of k of both operands into LDS
// as [row][k] at once (all its global loads in flight together, then one
// barrier) and runs its MFMAs in increasing k, 16 at a time; the network's
// layers have k <= 256, longer k loops over rounds.  The accumulation order
// is the same for every staging, so every epilogue sees the same sums.
constexpr int GT = 64, GK = 32, GKW = 256, LDK = GKW + 8;

// Epilogues:
//   EP_STORE     C = acc + bias (+ C when accumulating)
//   EP_SPLIT     P[z][m][n] = acc over the k range of split z
//   EP_SILU      z = acc + bias into C, silu(z) as BF16 into hout
enum { EP_STORE = 0, EP_SPLIT = 1, EP_SILU = 2 };

struct GemmArgs {
    const bf16_t* A; long sam, sak;
    const bf16_t* B; long sbk, sbn;
    int M, N, K, kper;
    float* C; long ldc;          // C and hout share the stride ldc
    const float* bias;
    int accumulate;
    float* P;
    bf16_t* hout;
};

// Thread t's 16-byte piece of an operand tile [64 rows][32 k] (rows r0..,
// k kc..kc+31), zero outside rows [0, R) and k [kc, ke).  The global reads
// follow the operand's layout: with k contiguous (sk == 1) 8 consecutive k
// of one row; with rows contiguous (sr == 1, a transposed operand) 8
// consecutive rows of one k; otherwise element by element.
__device__ __forceinline__ uint4 tile_fetch(const bf16_t* __restrict__ X,
                                            long sr, long sk, long r0, int R,
                                            int kc, int ke, int t)
{
    union { uint4 v; bf16_t h[8]; } u;
    if (sk == 1) {
        const long r = r0 + (t >> 2);
        const int k = kc + (t & 3)*8;
        const bf16_t* p = X + r*sr + k;
        if (r < R && k + 7 < ke && (((uintptr_t) p) & 15) == 0)
            return *(const uint4*) p;
        #pragma unroll
        for (int q = 0; q < 8; q++)
            u.h[q] = (r < R && k + q < ke) ? p[q] : (bf16_t) 0;
    } else if (sr == 1) {
        const long r = r0 + (t >> 5)*8;
        const int k = kc + (t & 31);
        const bf16_t* p = X + (long) k*sk + r;
        if (k < ke && r + 7 < R && (((uintptr_t) p) & 15) == 0)
            return *(const uint4*) p;
        #pragma unroll
        for (int q = 0; q < 8; q++)
            u.h[q] = (k < ke && r + q < R) ? p[q] : (bf16_t) 0;
    } else {
        const long r = r0 + (t >> 2);
        #pragma unroll
        for (int q = 0; q < 8; q++) {
            const int k = kc + (t & 3)*8 + q;
            u.h[q] = (r < R && k < ke) ? X[r*sr + (long) k*sk] : (bf16_t) 0;
        }
    }
    return u.v;
}

// Whether an operand's tiles can be loaded branch-free: 1 with k
// contiguous (rows 16-byte aligned, the k range a multiple of 8), 2 with
// rows contiguous (k strides 16-byte aligned, R a multiple of 8), 0
// otherwise (element-checked loads).  Rows past R are read as row R - 1
// (mode 1) or as the last 8 rows (mode 2): they only reach outputs past
// the matrix, which are not stored.  Values of k past the range are read
// at k0 and replaced by zeros.
__device__ __forceinline__ int full_tile(const bf16_t* X, long sr, long sk,
                                         int R, int ke)
{
    if ((((uintptr_t) X) & 15) != 0) return 0;
    if (sk == 1 && (sr & 7) == 0 && (ke & 7) == 0) return 1;
    if (sr == 1 && (sk & 7) == 0 && (R & 7) == 0 && R >= 8) return 2;
    return 0;
}

// Thread t's pieces of a round (chunks c = 0..GKW/GK-1 from k0) for a
// mode full_tile accepted.
__device__ __forceinline__ void full_round(const bf16_t* __restrict__ X,
                                           long sr, long sk, long r0, int R,
                                           int k0, int ke, int t, int mode,
                                           uint4* v)
{
    constexpr int NCH = GKW/GK;
    if (mode == 1) {
        const long r = min(r0 + (t >> 2), (long) R - 1);
        const bf16_t* p = X + r*sr;
        #pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int k = k0 + c*GK + (t & 3)*8;
            const bool ok = k < ke;
            const uint4 u = *(const uint4*) (p + (ok ? k : k0));
            v[c] = ok ? u : make_uint4(0, 0, 0, 0);
        }
    } else {
        const long r = min(r0 + (t >> 5)*8, (long) R - 8);
        #pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int k = k0 + c*GK + (t & 31);
            const bool ok = k < ke;
            const uint4 u = *(const uint4*) (X + (long) (ok ? k : k0)*sk + r);
            v[c] = ok ? u : make_uint4(0, 0, 0, 0);
        }
    }
}

// Stores tile_fetch's piece into the LDS tile [row][k] at k offset c*GK.
__device__ __forceinline__ void tile_put(bf16_t (*Xs)[LDK], long sr, long sk,
                                         int c, int t, uint4 v)
{
    if (sk == 1 || sr != 1) {
        *(uint4*) &Xs[t >> 2][c*GK + (t & 3)*8] = v;
    } else {
        union { uint4 v; bf16_t h[8]; } u;
        u.v = v;
        #pragma unroll
        for (int q = 0; q < 8; q++) Xs[(t >> 5)*8 + q][c*GK + (t & 31)] = u.h[q];
    }
}

// Tile of this workgroup.  Workgroups are dispatched to the GPU's 8 XCDs
// in turn (XCD = linear id mod 8), each XCD with its own L2: the tiles are
// dealt so that XCD q gets the q-th eighth of them in (column tile, row
// block, split) order, i.e. the column tiles of one row block (which read
// the same A rows) and the tiles of one split (the same k range) share an
// L2.
__device__ __forceinline__ void gemm_tile(int& bx, int& by, int& bz)
{
    const int nx = gridDim.x, ny = gridDim.y;
    const int T = nx*ny*gridDim.z;
    int L = blockIdx.x + nx*(blockIdx.y + ny*blockIdx.z);
    if ((T & 7) == 0) L = (L & 7)*(T >> 3) + (L >> 3);
    bx = L % nx;
    by = (L / nx) % ny;
    bz = L / (nx*ny);
}

// One workgroup: the 64 x 64 tile (by, bx) of C over the k range of split
// bz (kper values each).
template <int EP>
__device__ __forceinline__ void gemm_body(const GemmArgs& g, int bx, int by,
                                          int bz)
{
    __shared__ __attribute__((aligned(16))) bf16_t As[GT][LDK];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[GT][LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const long m0 = (long) by*GT, n0 = (long) bx*GT;
    const int kb = bz*g.kper, ke = min(g.K, kb + g.kper);
    constexpr int NCH = GKW/GK;
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += GKW) {
        const int nch = (min(GKW, ke - k0) + GK - 1)/GK;
        uint4 va[NCH], vb[NCH];
        // operands full_tile accepts load branch-free, every load of the
        // round in flight at once
        const int fa = full_tile(g.A, g.sam, g.sak, g.M, ke);
        const int fb = full_tile(g.B, g.sbn, g.sbk, g.N, ke);
        if (fa) {
            full_round(g.A, g.sam, g.sak, m0, g.M, k0, ke, t, fa, va);
        } else {
            #pragma unroll
            for (int c = 0; c < NCH; c++)
                if (c < nch)
                    va[c] = tile_fetch(g.A, g.sam, g.sak, m0, g.M, k0 + c*GK,
                                       ke, t);
        }
        if (fb) {
            full_round(g.B, g.sbn, g.sbk, n0, g.N, k0, ke, t, fb, vb);
        } else {
            #pragma unroll
            for (int c = 0; c < NCH; c++)
                if (c < nch)
                    vb[c] = tile_fetch(g.B, g.sbn, g.sbk, n0, g.N, k0 + c*GK,
                                       ke, t);
        }
        if (k0 > kb) __syncthreads();      // the last round's reads are done
        #pragma unroll
        for (int c = 0; c < NCH; c++)
            if (c < nch) {
                tile_put(As, g.sam, g.sak, c, t, va[c]);
                tile_put(Bs, g.sbn, g.sbk, c, t, vb[c]);
            }
        __syncthreads();
        const bf16_t* ar = &As[wm*32 + (lane & 31)][8*(lane >> 5)];
        const bf16_t* br = &Bs[wn*32 + (lane & 31)][8*(lane >> 5)];
        if (nch == NCH) {
            // operands of the next 4 k steps are read while the current 4
            // MFMAs run
            bf16x8 a0[4], b0[4];
            #pragma unroll
            for (int e = 0; e < 4; e++) {
                a0[e] = *(const bf16x8*) (ar + 16*e);
                b0[e] = *(const bf16x8*) (br + 16*e);
            }
            __builtin_amdgcn_iglp_opt(0);
            #pragma unroll
            for (int q = 0; q < GKW/64; q++) {
                bf16x8 a1[4], b1[4];
                if (q + 1 < GKW/64) {
                    #pragma unroll
                    for (int e = 0; e < 4; e++) {
                        a1[e] = *(const bf16x8*) (ar + 64*(q + 1) + 16*e);
                        b1[e] = *(const bf16x8*) (br + 64*(q + 1) + 16*e);
                    }
                }
                #pragma unroll
                for (int e = 0; e < 4; e++)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[e], b0[e],
                                                                  acc, 0, 0, 0);
                if (q + 1 < GKW/64) {
                    #pragma unroll
                    for (int e = 0; e < 4; e++) { a0[e] = a1[e]; b0[e] = b1[e]; }
                }
            }
        } else {
            for (int kk = 0; kk < nch*GK; kk += 16)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    *(const bf16x8*) (ar + kk), *(const bf16x8*) (br + kk),
                    acc, 0, 0, 0);
        }
    }
    const long n = n0 + wn*32 + (lane & 31);
    const bool nok = n < g.N;
    auto row = [&](int r) -> long {
        return m0 + wm*32 + 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
    };
    if (EP == EP_SPLIT) {
        if (!nok) return;
        float* Pz = g.P + (long) bz*g.M*g.N;
        #pragma unroll
        for (int r = 0; r < 16; r++) {
            const long m = row(r);
            if (m < g.M) Pz[m*g.N + n] = acc[r];
        }
    } else if (EP == EP_STORE || EP == EP_SILU) {
        if (!nok) return;
        const float bn = g.bias ? g.bias[n] : 0.0f;
        #pragma unroll
        for (int r = 0; r < 16; r++) {
            const long m = row(r);
            if (m < g.M) {
                float v = acc[r] + bn;
                if (EP == EP_STORE && g.accumulate) v += g.C[m*g.ldc + n];
                g.C[m*g.ldc + n] = v;
                if (EP == EP_SILU) g.hout[m*g.ldc + n] = f2bf(silu(v));
            }
        }
    }
}

template <int EP>
__global__ __launch_bounds__(256)
void k_gemm(GemmArgs g)
{
    int bx, by, bz;
    gemm_tile(bx, by, bz);
    gemm_body<EP>(g, bx, by, bz);
}

// Several split-K partial products in one launch: group q owns workgroups
// [start[q], start[q + 1]) of the linear id (after the XCD dealing of
// gemm_tile), laid out as its (column tile, row block, split) grid.
constexpr int kMaxGroup = 4;
struct GemmGroup {
    GemmArgs g[kMaxGroup];
    int nx[kMaxGroup], ny[kMaxGroup], start[kMaxGroup + 1];
    int ng;
};

__global__ __launch_bounds__(256)
void k_gemm_grouped(GemmGroup G)
{
    const int T = G.start[G.ng];
    int L = blockIdx.x;
    if ((T & 7) == 0) L = (L & 7)*(T >> 3) + (L >> 3);
    int q = 0;
    while (q + 1 < G.ng && L >= G.start[q + 1]) q++;
    const int loc = L - G.start[q], nx = G.nx[q], ny = G.ny[q];
    gemm_body<EP_SPLIT>(G.g[q], loc % nx, (loc / nx) % ny, loc / (nx*ny));
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

// Split of k for an M x N GEMM over K: few output tiles over a long k (the
// weight gradients: 256 x 256 over the batch rows) split k so the GPU's
// 256 CUs all work; the parts are summed in a fixed order, so the result
// does not depend on timing.
int gemm_split(int M, int N, int K, int* kper)
{
    const int tiles = ((N + GT - 1)/GT)*((M + GT - 1)/GT);
    int nsplit = 1;
    if (tiles < 256 && K >= 2048)
        nsplit = std::min(std::max(512/tiles, 1), K/256);
    const int kp = ((K + nsplit - 1)/nsplit + GK - 1)/GK*GK;
    *kper = kp;
    return (K + kp - 1)/kp;
}

template <int EP>
int launch_gemm(const GemmArgs& g, int nsplit, hipStream_t s)
{
    dim3 grid((g.N + GT - 1)/GT, (g.M + GT - 1)/GT, nsplit);
    hipLaunchKernelGGL(k_gemm<EP>, grid, dim3(256), 0, s, g);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

GemmArgs gemm_args(const void* A, long sam, long sak, const void* B,
                   long sbk, long sbn, int M, int N, int K)
{
    GemmArgs g = {};
    g.A = (const bf16_t*) A; g.sam = sam; g.sak = sak;
    g.B = (const bf16_t*) B; g.sbk = sbk; g.sbn = sbn;
    g.M = M; g.N = N; g.K = K; g.kper = K;
    return g;
}

// workspace of the split-K partial products of fteik_k_gemm_bf16 (grown
// on demand)
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

// Parameter gradients from their parts, then (adam != 0) the Adam step.
// Segment q covers parameters [off, off + len): gradient i is the sum of
// src[z*pstride + i - off] over its parts z in increasing order (split-K
// partial products, per-row-block column sums, per-block first-layer sums),
// so the result does not depend on timing.  The FP32 master parameters are
// updated and their BF16 copies refreshed, for weight segments also the
// transposed BF16 copy.
struct GradTable {
    FteikGradSeg s[kSurMaxSeg];
    int nseg;
};

__global__ void k_grad_adam(GradTable tab, float* __restrict__ g,
                            float* __restrict__ p, float* __restrict__ m,
                            float* __restrict__ v, bf16_t* __restrict__ pb,
                            long n, int adam, float lr, float b1, float b2,
                            float eps, float c1, float c2,
                            double* __restrict__ lclear)
{
    if (lclear && blockIdx.x == 0 && threadIdx.x == 0) *lclear = 0.0;
    GRID_STRIDE(i, n) {
        int qs = 0;
        while (qs + 1 < tab.nseg && i >= tab.s[qs + 1].off) qs++;
        const long o = i - tab.s[qs].off;
        const float* src = tab.s[qs].src;
        const long ps = tab.s[qs].pstride;
        const int np = tab.s[qs].nparts;
        float gi = 0.0f;
        int z = 0;
        for (; z + 32 <= np; z += 32) {
            float u[32];
            #pragma unroll
            for (int e = 0; e < 32; e++) u[e] = src[(long) (z + e)*ps + o];
            #pragma unroll
            for (int e = 0; e < 32; e++) gi += u[e];
        }
        for (; z + 8 <= np; z += 8) {
            float u[8];
            #pragma unroll
            for (int e = 0; e < 8; e++) u[e] = src[(long) (z + e)*ps + o];
            #pragma unroll
            for (int e = 0; e < 8; e++) gi += u[e];
        }
        for (; z < np; z++) gi += src[(long) z*ps + o];
        g[i] = gi;
        if (adam) {
            const float mi = b1*m[i] + (1.0f - b1)*gi;
            const float vi = b2*v[i] + (1.0f - b2)*gi*gi;
            m[i] = mi; v[i] = vi;
            const float pi = p[i] - lr*(mi*c1)/(sqrtf(vi*c2) + eps);
            p[i] = pi;
            const bf16_t q = f2bf(pi);
            pb[i] = q;
            if (tab.s[qs].tcopy) {
                const long rr = o / tab.s[qs].ncol;
                const long cc = o - rr*tab.s[qs].ncol;
                tab.s[qs].tcopy[cc*tab.s[qs].ldt + rr] = q;
            }
        }
    }
}

__global__ void k_f2bf(const float* __restrict__ a, bf16_t* __restrict__ b,
                       long n)
{
    GRID_STRIDE(i, n) b[i] = f2bf(a[i]);
}

// ----------------------------------------------------- fused training pass
// One workgroup per block of kMlpRows = 32 rows runs the whole network on
// them: layer 1 (3 -> 256) on the VALU, then the three hidden layers and
// the output layer as MFMA products of the block's activations (BF16, in
// LDS) with the weights (BF16, read from L2 as 16-byte vectors), the
// squared-error gradient, and back through the layers: dh = dz W from the
// transposed weight copies, dz = dh silu'(z).  Every output is one MFMA
// chain over k in steps of 16, in increasing k, as in k_gemm, so z, h, dy
// and dz are the values the separate kernels compute.  The pass writes
// what the weight gradients need (h_0..h_3, dz_1..dz_3, dy), the block's
// bias-gradient column sums and its first-layer weight-gradient sums.
constexpr int MR = kMlpRows, HU = 256, LDM = HU + 8;

// The pass is a chain of stages (3 hidden layers, the output layer, the
// output layer back, 3 hidden layers back), each an MFMA product of the
// block's 32 rows in LDS (A) with a weight matrix (B, 8 consecutive k per
// lane, read from L2), then an epilogue.  Loads and stores share one
// counter of at most 63 operations in flight, and a load issued after a
// store is waited for together with it, so:
//   * each stage issues the next stage's weight loads and epilogue inputs
//     right after its MFMAs, before its own stores;
//   * results leave through LDS as whole rows, 16 bytes a thread (z, h and
//     dz: 12 stores a thread per stage instead of 64), and the backward
//     pre-activations arrive the same way.
constexpr int LDZ = HU + 4;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_mlp_train(FteikMlpArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16_t hb[2][MR][LDM];
    __shared__ __attribute__((aligned(16))) float zb[2][MR][LDZ];
    __shared__ double lsum[2];
    __shared__ float xs[MR][3];       // the block's inputs
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int blk = blockIdx.x, nblk = gridDim.x;
    const long r0 = (long) blk*MR, rows = g.rows;
    const int T = g.T;
    const int lr = lane & 31, lk8 = 8*(lane >> 5);
    auto rowof = [&](int r) { return 8*(r >> 2) + 4*(lane >> 5) + (r & 3); };
    const int n0 = 64*w + lr, n1 = n0 + 32;       // this lane's columns
    const int no = 32*w + lr, noc = min(no, T - 1);   // output layer
    bf16x8 bw[2][HU/16];       // weights of the stage (two column tiles)
    float ec[2][16], en[2][16];  // epilogue inputs: this stage, next stage
    f32x32 zr;                 // pre-activation rows on their way to LDS
                               // (a vector value, so it stays in registers)
    // weights [out][in] (the forward layers): rows n0, n1 of W
    auto load_w = [&](const bf16_t* W, long ld) {
        const bf16_t* bp = W + (long) n0*ld + lk8;
        #pragma unroll
        for (int kk = 0; kk < HU/16; kk++) {
            bw[0][kk] = *(const bf16x8*) (bp + 16*kk);
            bw[1][kk] = *(const bf16x8*) (bp + 32*ld + 16*kk);
        }
    };
    auto load_bias = [&](const float* b) { en[0][0] = b[n0]; en[1][0] = b[n1]; };
    auto take = [&]() {
        #pragma unroll
        for (int c = 0; c < 2; c++)
            #pragma unroll
            for (int r = 0; r < 16; r++) ec[c][r] = en[c][r];
    };
    // whole rows of the block, 16 bytes a thread: FP32 rows of zb[b] to
    // dst, BF16 rows of hb[b] to dst; z rows from src into zr, then zb[b]
    auto put_z = [&](int b, float* dst) {
        #pragma unroll
        for (int i = 0; i < 8; i++) {
            const int idx = t + 256*i, m = idx >> 6, c4 = (idx & 63)*4;
            if (r0 + m < rows)
                *(float4*) (dst + (r0 + m)*HU + c4) = *(const float4*) &zb[b][m][c4];
        }
    };
    auto put_h = [&](int b, bf16_t* dst) {
        #pragma unroll
        for (int i = 0; i < 4; i++) {
            const int idx = t + 256*i, m = idx >> 5, c8 = (idx & 31)*8;
            if (r0 + m < rows)
                *(uint4*) (dst + (r0 + m)*HU + c8) = *(const uint4*) &hb[b][m][c8];
        }
    };
    auto get_z = [&](const float* src) {
        #pragma unroll
        for (int i = 0; i < 8; i++) {
            const int idx = t + 256*i, m = idx >> 6, c4 = (idx & 63)*4;
            const float4 v = *(const float4*) (src + min(r0 + m, rows - 1)*HU + c4);
            zr[4*i] = v.x; zr[4*i + 1] = v.y; zr[4*i + 2] = v.z; zr[4*i + 3] = v.w;
        }
    };
    auto park_z = [&](int b) {
        #pragma unroll
        for (int i = 0; i < 8; i++) {
            const int idx = t + 256*i, m = idx >> 6, c4 = (idx & 63)*4;
            *(float4*) &zb[b][m][c4] = make_float4(zr[4*i], zr[4*i + 1],
                                                   zr[4*i + 2], zr[4*i + 3]);
        }
    };
    // A = hb[cur], B = bw, nk steps of 16 (nk <= HU/16, wave-uniform)
    auto product = [&](int cur, int nk, f32x16& acc0, f32x16& acc1) {
        const bf16_t* ap = &hb[cur][lr][lk8];
        #pragma unroll
        for (int kk = 0; kk < HU/16; kk++) {
            if (kk < nk) {
                const bf16x8 a = *(const bf16x8*) (ap + 16*kk);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[0][kk],
                                                               acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[1][kk],
                                                               acc1, 0, 0, 0);
            }
        }
    };
    // ---- layer 1 (as k_layer1): z0 = W1 x + b1, h0 = silu(z0); the first
    // hidden layer's weights and biases are loaded before its stores
    // row r of the pass is row (first + r) % pool of x and tgt
    auto prow = [&](long r) { return (g.first + r) % g.pool; };
    if (t < MR*3) {
        const int m = t / 3;
        xs[m][t - 3*m] = g.x[3*prow(min(r0 + m, rows - 1)) + t - 3*m];
    }
    // thread t computes unit j = t of every row
    const float w10 = g.p[g.oW1 + 3*t], w11 = g.p[g.oW1 + 3*t + 1],
                w12 = g.p[g.oW1 + 3*t + 2], b1t = g.p[g.ob1 + t];
    load_w(g.pb + g.oW[0], HU);
    load_bias(g.p + g.ob[0]);
    __syncthreads();
    #pragma unroll 8
    for (int m = 0; m < MR; m++) {
        const float v = w10*xs[m][0] + w11*xs[m][1] + w12*xs[m][2] + b1t;
        zb[1][m][t] = v;
        hb[0][m][t] = f2bf(silu(v));
    }
    __syncthreads();
    put_z(1, g.z[0]);
    put_h(0, g.h[0]);
    int cur = 0;
    // ---- hidden layers: z_{l+1} = h_l W_l^T + b_l (stage staging buffer
    // zb[l & 1])
    #pragma unroll
    for (int l = 0; l < 3; l++) {
        f32x16 acc0 = {}, acc1 = {};
        product(cur, HU/16, acc0, acc1);
        take();
        if (l < 2) {
            load_w(g.pb + g.oW[l + 1], HU);
            load_bias(g.p + g.ob[l + 1]);
        } else {
            // the output layer: rows no of W5 (clamped; waves 2, 3 load
            // and multiply the last row and use nothing of it)
            const bf16_t* bp = g.pb + g.oW5 + (long) noc*HU + lk8;
            #pragma unroll
            for (int kk = 0; kk < HU/16; kk++) bw[0][kk] = *(const bf16x8*) (bp + 16*kk);
            en[0][0] = g.p[g.ob5 + noc];
            #pragma unroll
            for (int r = 0; r < 16; r++)
                en[1][r] = g.tgt[prow(min(r0 + rowof(r), rows - 1))*T + noc];
        }
        #pragma unroll
        for (int c = 0; c < 2; c++) {
            const int n = c ? n1 : n0;
            const float bn = ec[c][0];
            #pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = rowof(r);
                const float v = (c ? acc1[r] : acc0[r]) + bn;
                hb[cur ^ 1][m][n] = f2bf(silu(v));
                zb[l & 1][m][n] = v;
            }
        }
        __syncthreads();
        put_z(l & 1, g.z[l + 1]);
        put_h(cur ^ 1, g.h[l + 1]);
        cur ^= 1;
    }
    // ---- output layer and dy = 2 (y - tgt) / (rows T) (waves 0 and 1,
    // 32 columns each); dy also becomes the A operand of the way back
    {
        f32x16 acc = {}, unused = {};
        {
            const bf16_t* ap = &hb[cur][lr][lk8];
            #pragma unroll
            for (int kk = 0; kk < HU/16; kk++)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    *(const bf16x8*) (ap + 16*kk), bw[0][kk], acc, 0, 0, 0);
        }
        (void) unused;
        take();
        // next: the output layer back (W5 transposed, k = TP) and z3
        {
            const bf16_t* bp = g.pbT + g.oW5T + (long) n0*g.TP + lk8;
            #pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int k = min(kk, g.TP/16 - 1);
                bw[0][kk] = *(const bf16x8*) (bp + 16*k);
                bw[1][kk] = *(const bf16x8*) (bp + 32L*g.TP + 16*k);
            }
        }
        get_z(g.z[3]);
        double l2 = 0.0;
        if (w < 2) {
            const float bn = ec[0][0];
            float s = 0.0f;
            #pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = rowof(r);
                const long row = r0 + m;
                bf16_t q = 0;
                if (no < T && row < rows) {
                    const float d = (acc[r] + bn) - ec[1][r];
                    q = f2bf(2.0f*d*g.inv);
                    g.dy[row*T + no] = q;
                    s += bf2f(q);
                    l2 += (double) d*d;
                }
                hb[cur ^ 1][m][no] = q;
            }
            s += __shfl_xor(s, 32);        // lanes l and l + 32: same column
            if (lane < 32 && no < T) g.colp[3L*nblk*HU + (long) blk*T + no] = s;
            #pragma unroll
            for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o);
            if (lane == 0) lsum[w] = l2;
        }
        park_z(1);
        __syncthreads();
        if (t == 0) atomicAdd(g.loss, (lsum[0] + lsum[1])*(double) g.inv);
        cur ^= 1;
    }
    // ---- the way back: dh = dz_next W (B from the transposed copies), then
    // dz = dh silu'(z) into LDS, out as whole rows (when a weight gradient
    // needs it) and summed over the block's rows (bias-gradient parts).
    // Stage 0 is the output layer (k = TP), stages 1..3 the hidden layers
    // 2, 1, 0; stage st reads z from zb[(st + 1) & 1].
    #pragma unroll
    for (int st = 0; st < 4; st++) {
        const int l = 3 - st;                 // dz_l is made here
        f32x16 acc0 = {}, acc1 = {};
        product(cur, st == 0 ? g.TP/16 : HU/16, acc0, acc1);
        if (l > 0) {
            load_w(g.pbT + g.oWT[l - 1], HU);
            get_z(g.z[l - 1]);
        }
        const int zi = (st + 1) & 1;
        const int slot = (l == 3) ? 2 : (l >= 1 ? l - 1 : 4);
        #pragma unroll
        for (int c = 0; c < 2; c++) {
            const int n = c ? n1 : n0;
            float s = 0.0f;
            #pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = rowof(r);
                bf16_t q = 0;
                if (r0 + m < rows) {
                    q = f2bf(((c ? acc1[r] : acc0[r]) + 0.0f)*dsilu(zb[zi][m][n]));
                    s += bf2f(q);
                }
                hb[cur ^ 1][m][n] = q;
            }
            s += __shfl_xor(s, 32);
            if (lane < 32) g.colp[(long) slot*nblk*HU + (long) blk*HU + n] = s;
        }
        if (l > 0) park_z(st & 1);
        __syncthreads();
        if (l > 0) put_h(cur ^ 1, g.dz[l]);
        cur ^= 1;
    }
    // ---- first-layer weight-gradient sums of the block (thread = unit j)
    {
        const int j = t;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
        const int mr = (int) min((long) MR, rows - r0);
        for (int m = 0; m < mr; m++) {
            const float gz = bf2f(hb[cur][m][j]);
            s0 += gz*xs[m][0]; s1 += gz*xs[m][1]; s2 += gz*xs[m][2];
        }
        float* o = g.dw1p + (long) blk*3*HU + 3*j;
        o[0] = s0; o[1] = s1; o[2] = s2;
    }
}

// dst[c*ldt + r] = src[r*cols + c] (BF16)
__global__ void k_transpose_bf16(const bf16_t* __restrict__ src, int rows,
                                 int cols, bf16_t* __restrict__ dst, int ldt)
{
    GRID_STRIDE(i, (long) rows*cols) {
        const long r = i / cols, c = i - r*cols;
        dst[c*ldt + r] = src[i];
    }
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
// One thread per (row, station): the table reads of a row are independent
// gathers spread over the whole GPU.  Clears *loss for the step.
__global__ void k_samples(const float* __restrict__ tab, long ldgrd, GeoArgs g,
                          long ngrd, unsigned long long seed, long pool,
                          long first, long rows, float* __restrict__ x,
                          float* __restrict__ tgt, int ldt,
                          double* __restrict__ loss)
{
    if (loss && blockIdx.x == 0 && threadIdx.x == 0) *loss = 0.0;
    GRID_STRIDE(i, rows*ldt) {
        const long r = i / ldt;
        const int s = (int) (i - r*ldt);
        const long idx = (first + r) % pool;
        const long c = (long) (mix64(seed*0x100000001b3ull + (unsigned long long) idx)
                               % (unsigned long long) ngrd);
        float p[3];
        node_xyz(c, g, p);
        if (s == 0) { x[3*r] = p[0]; x[3*r + 1] = p[1]; x[3*r + 2] = p[2]; }
        float v = 0.0f;
        if (s < g.ntf)
            v = (tab[(long) s*ldgrd + c] - t0_ray(p, s, g, nullptr))/g.scale;
        tgt[r*ldt + s] = v;
    }
}

// ------------------------------------------------------------ inversion
// Rows are laid out as 4 blocks of R: primal, then tangents along x, y, z.
// zp (the primal pre-activations) may be null.
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
        if (zp) zp[i] = z;
        h[i] = f2bf(silu(z));
        for (int d = 0; d < 3; d++)
            h[(long) (d + 1)*R*H + i] = f2bf(ds*W1[3*j + d]);
    }
}

// Hidden layer of the inversion network, fused: 4 row blocks (primal and
// three tangents) of R rows, H = 256.  z = h_in W^T on the MFMA units with
// FP32 accumulation in the same k order as k_gemm (so z is bitwise
// the GEMM's), then h = silu(z + b) for the primal block and
// h_d = silu'(z_0 + b) z_d for the tangents, rounded to BF16: what
// k_gemm followed by k_inv_act computes, without the FP32 round trip
// through memory.  Persistent: each workgroup copies W into LDS once (rows
// padded to 264 elements so the 32 rows an MFMA operand load touches fall
// in different banks) and then takes 32-row blocks in turn.  Eight waves,
// two per SIMD (the weights fill most of the LDS, so a CU holds one
// workgroup): waves d and d + 4 compute block d, the first the column tiles
// 0-3 and the second 4-7, each with its A rows held in registers over the
// whole K and the next block's rows loading while it computes; a tile's
// MFMAs are issued before the previous tile's epilogue.  The primal waves
// hand silu' to the tangent waves through LDS one 32-column tile at a time.
// Streaming W from L2 for every wave instead read 1.3 GB of L2 per launch at
// 80,000 rows.
constexpr int kInvWStride = 264;
__global__ __launch_bounds__(512)
void k_inv_hidden(const bf16_t* __restrict__ hin, long R,
                  const bf16_t* __restrict__ W, const float* __restrict__ b,
                  bf16_t* __restrict__ hout)
{
    constexpr int H = 256, KS = H/16;
    __shared__ __attribute__((aligned(16))) bf16_t ws[H*kInvWStride];
    __shared__ float ds[2][2][32][33];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int d = w & 3, half = w >> 2;
    for (int c = threadIdx.x; c < H*H/8; c += 512) {
        const int n = c >> 5, k8 = (c & 31)*8;
        *(bf16x8*) (ws + n*kInvWStride + k8) = *(const bf16x8*) (W + n*H + k8);
    }
    const long nrb = (R + 31)/32;
    auto load_a = [&](long rb, bf16x8* a) {
        const long arow = min(rb*32 + (lane & 31), R - 1);
        const bf16_t* ap = hin + ((long) d*R + arow)*H + 8*(lane >> 5);
        #pragma unroll
        for (int kk = 0; kk < KS; kk++) a[kk] = *(const bf16x8*) (ap + 16*kk);
    };
    bf16x8 a[KS], an[KS];
    long rb = blockIdx.x;
    load_a(rb, a);
    __syncthreads();
    for (; rb < nrb; rb += gridDim.x) {
        load_a(rb + gridDim.x, an);
        const long s0 = rb*32;
        const bool full = s0 + 32 <= R;
        auto tile = [&](int ct, f32x16& acc) {
            const bf16_t* wp = ws + (32*ct + (lane & 31))*kInvWStride
                             + 8*(lane >> 5);
            acc = f32x16{};
            #pragma unroll
            for (int kk = 0; kk < KS; kk++) {
                const bf16x8 bv = *(const bf16x8*) (wp + 16*kk);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk], bv, acc,
                                                              0, 0, 0);
            }
        };
        f32x16 accs[2];
        tile(4*half, accs[0]);
        #pragma unroll
        for (int cl = 0; cl < 4; cl++) {
            const int ct = 4*half + cl;
            const int n = 32*ct + (lane & 31);
            if (cl + 1 < 4) tile(ct + 1, accs[(cl + 1) & 1]);
            const f32x16& acc = accs[cl & 1];
            const float bn = b[n];
            float (*dsb)[33] = ds[cl & 1][half];
            const int mlim = full ? 32 : (int) (R - s0);
            auto primal = [&](auto guarded) {
                bf16_t* hp = hout + s0*H + n;
                #pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int m = 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
                    const float z = acc[r] + bn;
                    dsb[m][lane & 31] = dsilu(z);
                    if (!guarded || m < mlim) hp[m*H] = f2bf(silu(z));
                }
            };
            auto tangent = [&](auto guarded) {
                bf16_t* hp = hout + ((long) d*R + s0)*H + n;
                #pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int m = 8*(r >> 2) + 4*(lane >> 5) + (r & 3);
                    if (!guarded || m < mlim)
                        hp[m*H] = f2bf(dsb[m][lane & 31]*acc[r]);
                }
            };
            if (d == 0) {
                if (full) primal(std::false_type{});
                else primal(std::true_type{});
            }
            __syncthreads();
            if (d > 0) {
                if (full) tangent(std::false_type{});
                else tangent(std::true_type{});
            }
        }
        #pragma unroll
        for (int kk = 0; kk < KS; kk++) a[kk] = an[kk];
    }
}

// Output layer of the inversion network (no bias: the Gauss-Newton kernel
// adds it), station-major so that the Gauss-Newton kernel's lanes, one row
// each, read consecutive words: y[s*ldy + q] = sum_k W5[s][k] h[q][k].  The
// weights (T <= 64 rows, padded with zero rows to 64 and to 264 elements
// per row) sit in LDS as the MFMA's A operand; each wave holds 32 rows of h
// over all of K as its B operand, so a result tile is 32 stations x 32 rows
// and consecutive lanes store consecutive rows.  Whole tiles are stored
// (y is allocated to ((T + 31)/32)*32 station rows and ldy >= NR rounded up
// to 32), so the stores need no branch; rows past NR load row NR - 1.
__global__ __launch_bounds__(256)
void k_inv_out(const bf16_t* __restrict__ hin, long NR,
               const bf16_t* __restrict__ W5, int T, float* __restrict__ y,
               long ldy)
{
    constexpr int H = 256, KS = H/16;
    __shared__ __attribute__((aligned(16))) bf16_t ws[64*kInvWStride];
    for (int c = threadIdx.x; c < 64*H/8; c += 256) {
        const int n = c >> 5, k8 = (c & 31)*8;
        bf16x8 v = {};
        if (n < T) v = *(const bf16x8*) (W5 + n*H + k8);
        *(bf16x8*) (ws + n*kInvWStride + k8) = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nst = (T + 31)/32;
    const long nblk = (NR + 31)/32;
    for (long bq = (long) blockIdx.x*4 + w; bq < nblk;
         bq += (long) gridDim.x*4) {
        const long q = min(bq*32 + (lane & 31), NR - 1);
        const bf16_t* hp = hin + q*H + 8*(lane >> 5);
        bf16x8 bh[KS];
        #pragma unroll
        for (int kk = 0; kk < KS; kk++) bh[kk] = *(const bf16x8*) (hp + 16*kk);
        for (int st = 0; st < nst; st++) {
            const bf16_t* wp = ws + (32*st + (lane & 31))*kInvWStride
                             + 8*(lane >> 5);
            f32x16 acc = {};
            #pragma unroll
            for (int kk = 0; kk < KS; kk++) {
                const bf16x8 av = *(const bf16x8*) (wp + 16*kk);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bh[kk], acc,
                                                              0, 0, 0);
            }
            float* yp = y + (long) (32*st + 4*(lane >> 5))*ldy + bq*32
                      + (lane & 31);
            #pragma unroll
            for (int r = 0; r < 16; r++)
                yp[(long) (8*(r >> 2) + (r & 3))*ldy] = acc[r];
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
// for every (event, start) row; y holds the network outputs station-major:
// output s of row r of block d (primal, tangents x, y, z) at
// y[s*ldy + d*R + r], ldy >= 4R.
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
        // station s's residual and Jacobian row; formed again by the second
        // pass (the same operations, so the same values) rather than kept in
        // per-thread arrays, which would live in scratch memory
        auto station = [&](int s, float& res, float* J) {
            float gr[3];
            const float* ys = y + (long) s*ldy + r;
            const float t = t0_ray(p, s, g, gr) + g.scale*(ys[0] + b5[s]);
            res = o[s] - t;
            for (int d = 0; d < 3; d++)
                J[d] = gr[d] + g.scale*ys[(long) (d + 1)*R];
        };
        float sw = 0.0f, mr = 0.0f, mj[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < g.ntf; s++) {
            float res, J[3];
            station(s, res, J);
            const float ws = w[s];
            sw += ws; mr += ws*res;
            for (int d = 0; d < 3; d++) mj[d] += ws*J[d];
        }
        const float inv = sw > 0.0f ? 1.0f/sw : 0.0f;
        mr *= inv;
        for (int d = 0; d < 3; d++) mj[d] *= inv;
        float A[3][3] = {}, bb[3] = {}, f = 0.0f;
        for (int s = 0; s < g.ntf; s++) {
            float res, J[3];
            station(s, res, J);
            const float w2 = w[s]*w[s];
            const float rc = res - mr;
            float jc[3];
            for (int d = 0; d < 3; d++) jc[d] = J[d] - mj[d];
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
    // the cell's times of every station (at most 64), all loads issued
    // before the first is used; station indices past ntf are clamped, so
    // the loads need no branch
    float T[64];
    #pragma unroll
    for (int s = 0; s < 64; s++)
        T[s] = tab[(long) min(s, g.ntf - 1)*ldgrd + c];
    float t0 = 0.0f;
    #pragma unroll
    for (int s = 0; s < 64; s++) {
        if (s >= g.ntf || sk[s]) continue;
        float res = xnorm[e]*(o[s] - T[s]);
        t0 = t0 + w[s]*res;
    }
    float obj = 0.0f;
    #pragma unroll
    for (int s = 0; s < 64; s++) {
        if (s >= g.ntf || sk[s]) continue;
        float ws = sqrtHalf*w[s];
        float est = T[s] + t0;
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
    GemmArgs g = gemm_args(A, sam, sak, B, sbk, sbn, M, N, K);
    const int nsplit = gemm_split(M, N, K, &g.kper);
    g.C = C; g.ldc = ldc; g.bias = bias; g.accumulate = accumulate;
    if (nsplit == 1) return launch_gemm<EP_STORE>(g, 1, s);
    const size_t need = (size_t) nsplit*M*N;
    if (need > g_split_cap) {
        if (g_split) hipFree(g_split);
        g_split = nullptr;
        g_split_cap = 0;
        if (hipMalloc((void**) &g_split, need*sizeof(float)) != hipSuccess)
            return 1;
        g_split_cap = need;
    }
    g.P = g_split;
    if (launch_gemm<EP_SPLIT>(g, nsplit, s)) return 1;
    hipLaunchKernelGGL(k_gemm_reduce, dim3(nblocks((long) M*N)), dim3(TPB),
                       0, s, g_split, nsplit, M, N, C, ldc, bias, accumulate);
    return ok();
}

int fteik_k_gemm_parts(int M, int N, int K)
{
    int kper;
    return gemm_split(M, N, K, &kper);
}

int fteik_k_gemm_partial(const void* A, long sam, long sak, const void* B,
                         long sbk, long sbn, int M, int N, int K, float* P,
                         hipStream_t s)
{
    if (M < 1 || N < 1 || K < 1) return 0;
    GemmArgs g = gemm_args(A, sam, sak, B, sbk, sbn, M, N, K);
    const int nsplit = gemm_split(M, N, K, &g.kper);
    g.P = P;
    return launch_gemm<EP_SPLIT>(g, nsplit, s);
}

int fteik_k_gemm_partials(int n, const FteikGemmPart* parts, hipStream_t s)
{
    if (n < 1 || n > kMaxGroup) return 2;
    GemmGroup G = {};
    G.ng = n;
    G.start[0] = 0;
    for (int q = 0; q < n; q++) {
        const FteikGemmPart& p = parts[q];
        if (p.M < 1 || p.N < 1 || p.K < 1) return 2;
        G.g[q] = gemm_args(p.A, p.sam, p.sak, p.B, p.sbk, p.sbn, p.M, p.N,
                           p.K);
        const int nsplit = gemm_split(p.M, p.N, p.K, &G.g[q].kper);
        G.g[q].P = p.P;
        G.nx[q] = (p.N + GT - 1)/GT;
        G.ny[q] = (p.M + GT - 1)/GT;
        G.start[q + 1] = G.start[q] + G.nx[q]*G.ny[q]*nsplit;
    }
    hipLaunchKernelGGL(k_gemm_grouped, dim3(G.start[n]), dim3(256), 0, s, G);
    return ok();
}

int fteik_k_gemm_silu(const void* A, long sam, long sak, const void* B,
                      long sbk, long sbn, int M, int N, int K,
                      const float* bias, float* z, void* h, long ld,
                      hipStream_t s)
{
    if (M < 1 || N < 1 || K < 1 || K > GKW) return 2;
    GemmArgs g = gemm_args(A, sam, sak, B, sbk, sbn, M, N, K);
    g.bias = bias; g.C = z; g.hout = (bf16_t*) h; g.ldc = ld;
    return launch_gemm<EP_SILU>(g, 1, s);
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

int fteik_k_sur_grad_adam(const FteikGradSeg* segs, int nseg, float* g,
                          float* p, float* m, float* v, void* pb, long n,
                          int adam, float lr, float b1, float b2, float eps,
                          float c1, float c2, double* lclear, hipStream_t s)
{
    if (nseg < 1 || nseg > kSurMaxSeg) return 2;
    GradTable tab = {};
    for (int q = 0; q < nseg; q++) tab.s[q] = segs[q];
    tab.nseg = nseg;
    hipLaunchKernelGGL(k_grad_adam, dim3(nblocks(n)), dim3(TPB), 0, s, tab,
                       g, p, m, v, (bf16_t*) pb, n, adam, lr, b1, b2, eps,
                       c1, c2, lclear);
    return ok();
}

int fteik_k_mlp_train(const FteikMlpArgs* a, hipStream_t s)
{
    if (a->rows < 1 || a->T < 1 || a->T > 64 || a->TP < a->T ||
        a->TP > 64 || (a->TP & 15) != 0)
        return 2;
    const long nblk = (a->rows + kMlpRows - 1)/kMlpRows;
    hipLaunchKernelGGL(k_mlp_train, dim3((unsigned) nblk), dim3(256), 0, s,
                       *a);
    return ok();
}

int fteik_k_sur_transpose(const void* src, int rows, int cols, void* dst,
                          int ldt, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose_bf16, dim3(nblocks((long) rows*cols)),
                       dim3(TPB), 0, s, (const bf16_t*) src, rows, cols,
                       (bf16_t*) dst, ldt);
    return ok();
}

int fteik_k_sur_samples(const float* tab, long ldgrd, const void* geo,
                        long ngrd, unsigned long long seed, long pool,
                        long first, long rows, float* x, float* tgt, int ldt,
                        double* loss, hipStream_t s)
{
    hipLaunchKernelGGL(k_samples, dim3(nblocks(rows*ldt)), dim3(TPB), 0, s,
                       tab, ldgrd, *(const GeoArgs*) geo, ngrd, seed, pool,
                       first, rows, x, tgt, ldt, loss);
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
    // persistent: one workgroup per CU (the weights fill most of its LDS)
    static int ncu = 0;
    if (ncu <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                  dev) != hipSuccess || ncu <= 0) ncu = 256;
    }
    const long nrb = (R + 31)/32;
    const unsigned nblk = (unsigned) (nrb < ncu ? nrb : ncu);
    hipLaunchKernelGGL(k_inv_hidden, dim3(nblk), dim3(512), 0, s,
                       (const bf16_t*) hin, R, (const bf16_t*) W, b,
                       (bf16_t*) hout);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int fteik_k_inv_out(const void* hin, long NR, const void* W5, int T,
                    float* y, long ldy, hipStream_t s)
{
    if (T < 1 || T > 64 || NR < 1 || ldy < NR || ldy % 32 != 0) return 2;
    static int ncu = 0;
    if (ncu <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                  dev) != hipSuccess || ncu <= 0) ncu = 256;
    }
    const long nwg = (NR + 127)/128;          // 4 waves of 32 rows each
    const unsigned nblk = (unsigned) (nwg < 4L*ncu ? nwg : 4L*ncu);
    hipLaunchKernelGGL(k_inv_out, dim3(nblk), dim3(256), 0, s,
                       (const bf16_t*) hin, NR, (const bf16_t*) W5, T, y, ldy);
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
