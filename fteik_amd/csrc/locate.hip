// locate.hip -- fused L2-misfit + argmin grid search on gfx950.
//
// locate.F90:620-744 (locate_computeL2ObjectiveFunction32f) followed by the
// MINLOC of locate.F90:122, in one pass over the travel-time tables:
//   pass 1  t0(c)  = sum_i w_i * (xnorm * (o_i - T_ic))          (:684-687)
//   pass 2  obj(c) = sum_i ((sqrt(1/2) w_i) * (o_i - (T_ic + t0)))^2 (:694-705)
// Each thread keeps its cell's station times in registers, so the tables
// stream from HBM once (the reference reads them twice).  The static
// corrections of the reference are identically zero (no setter exists,
// locate.F90:210) and T + 0 == T, so they are not added.  The argmin packs
// (objective bits << 32 | cell) into one 64-bit key: non-negative floats
// order like their bit patterns, so the smallest key is the first minimum,
// exactly MINLOC's tie-break.  Compiled without FMA contraction so every
// objective value equals the reference's bit for bit.
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

constexpr int TL = 256;
constexpr int MAXT = 64;      // tables held in registers
constexpr int CH = 16;        // stations per chunk of scalar operands

__device__ __forceinline__ unsigned long long make_key(float obj, int cell)
{
    return ((unsigned long long) __float_as_uint(obj) << 32)
         | (unsigned long long) (unsigned) cell;
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long k)
{
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(k, off, 64);
        k = (o < k) ? o : k;
    }
    return k;
}

__global__ void k_tables_d2f(float* __restrict__ t, const double* __restrict__ d,
                             int n)
{
    int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i < n) t[i] = (float) d[i];   // locate.F90:338 REAL(): round to nearest
}

// One event over all cells (locate_locateEvent).  t0work carries the
// reference's state between events when accumulate != 0 (locate.F90:212,
// :687: it is never reset between events).
__global__ __launch_bounds__(TL)
void k_locate_one(const float* __restrict__ tab, int ntf, int ngrd, long ldgrd,
                  const float* __restrict__ obs, const float* __restrict__ wt,
                  const unsigned char* __restrict__ skip, float xnorm,
                  float* __restrict__ t0work, float* __restrict__ objwork,
                  int accumulate, unsigned long long* __restrict__ best)
{
    __shared__ unsigned long long red[TL/64];
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    unsigned long long key = ~0ull;
    if (c < ngrd) {
        const float sqrtHalf = 0.70710677f;   // SQRT(0.5) in single precision
        float T[MAXT];
        float t0 = accumulate ? t0work[c] : 0.0f;
        #pragma unroll
        for (int i = 0; i < MAXT; i++) {
            if (i < ntf) {
                T[i] = tab[(long) i*ldgrd + c];
                if (!skip[i]) {
                    float res = xnorm*(obs[i] - T[i]);
                    t0 = t0 + wt[i]*res;
                }
            }
        }
        for (int i = MAXT; i < ntf; i++) {
            if (skip[i]) continue;
            float res = xnorm*(obs[i] - tab[(long) i*ldgrd + c]);
            t0 = t0 + wt[i]*res;
        }
        float obj = 0.0f;
        #pragma unroll
        for (int i = 0; i < MAXT; i++) {
            if (i < ntf && !skip[i]) {
                float w = sqrtHalf*wt[i];
                float est = T[i] + t0;
                float res = w*(obs[i] - est);
                obj = obj + res*res;
            }
        }
        for (int i = MAXT; i < ntf; i++) {
            if (skip[i]) continue;
            float w = sqrtHalf*wt[i];
            float est = tab[(long) i*ldgrd + c] + t0;
            float res = w*(obs[i] - est);
            obj = obj + res*res;
        }
        if (accumulate) t0work[c] = t0;
        if (objwork) objwork[c] = obj;
        if (!accumulate && t0work) t0work[c] = t0;
        key = make_key(obj, c);
    }
    key = wave_min(key);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long k = red[0];
        for (int i = 1; i < TL/64; i++) k = (red[i] < k) ? red[i] : k;
        atomicMin(best, k);
    }
}

// Many events, fresh origin-time state per event (the analytic origin-time
// removal of the grid search).  A thread holds CPT cells' tables in
// registers and loops over the events; event data are uniform (scalar
// loads).  best[e] receives the per-event key.
//
// Skipped picks arrive with obs = wt = wt2 = 0: their terms are +-0, and
// adding a zero leaves an FP32 sum unchanged, so the sums equal the
// reference's sums over the picks it keeps without a branch per pick.
// wt2 = sqrt(1/2) * wt, the product the reference forms per pick.  Each
// wave posts its minimum key with an atomic (no workgroup barrier per
// event).  blockIdx.y selects a group of epg consecutive events, so the
// launch can be cut into workgroups that fill whole rounds of the GPU
// (event_groups).
template <int NTF, int CPT, bool WGMIN>
__global__ __launch_bounds__(TL)
void k_locate_batch(const float* __restrict__ tab, int ngrd, long ldgrd,
                    const float* __restrict__ obs, const float* __restrict__ wt,
                    const float* __restrict__ wt2,
                    const float* __restrict__ xnorm, int lntf, int nev,
                    int epg, unsigned long long* __restrict__ best)
{
    __shared__ unsigned long long s_key[WGMIN ? TL/64 : 1];
    const long base = ((long) blockIdx.x*TL*CPT) + threadIdx.x;
    // events [e0, e1) of this workgroup's event group
    const int e0 = (int) blockIdx.y*epg, e1 = min(nev, e0 + epg);
    float T[CPT][NTF];
    #pragma unroll
    for (int q = 0; q < CPT; q++) {
        const long c = base + (long) q*TL;
        #pragma unroll
        for (int i = 0; i < NTF; i++)
            T[q][i] = (c < ngrd) ? tab[(long) i*ldgrd + c] : 0.0f;
    }
    for (int e = e0; e < e1; e++) {
        const float* o = obs + (long) e*lntf;
        const float* w = wt + (long) e*lntf;
        const float* w2 = wt2 + (long) e*lntf;
        const float xn = xnorm[e];
        float t0[CPT], ob[CPT];
        #pragma unroll
        for (int q = 0; q < CPT; q++) { t0[q] = 0.0f; ob[q] = 0.0f; }
        // The event's pick times and weights are scalar operands, loaded
        // CH stations at a time, one chunk ahead of the chunk in use: all
        // 3*NTF at once would not fit the scalar registers, and the values
        // spilled to vector lanes would cost a vector instruction per use.
        // The scheduling barriers keep the loads where they are placed.
        // The operations per (cell, station) and their order are those of
        // the reference.
        float cv[2][CH], cw[2][CH];
        auto fetch = [&](int pass, int c, int b) {
            const float* wsrc = pass ? w2 : w;
            #pragma unroll
            for (int j = 0; j < CH; j++) {
                if (c*CH + j < NTF) {
                    cv[b][j] = o[c*CH + j];
                    cw[b][j] = wsrc[c*CH + j];
                }
            }
        };
        constexpr int NCH = (NTF + CH - 1)/CH;
        fetch(0, 0, 0);
        #pragma unroll
        for (int k = 0; k < 2*NCH; k++) {
            const int pass = k / NCH, c = k - pass*NCH, b = k & 1;
            // the memory clobber keeps pass 2 from reusing pass 1's loads
            // of the pick times (which would keep all of them live)
            asm volatile("" ::: "memory");
            if (k + 1 < 2*NCH) fetch((k + 1)/NCH, (k + 1) % NCH, b ^ 1);
            __builtin_amdgcn_sched_barrier(0);
            #pragma unroll
            for (int j = 0; j < CH; j++) {
                const int i = c*CH + j;
                if (i >= NTF) break;
                const float oi = cv[b][j], wi = cw[b][j];
                #pragma unroll
                for (int q = 0; q < CPT; q++) {
                    if (pass == 0) {
                        float res = xn*(oi - T[q][i]);
                        t0[q] = t0[q] + wi*res;
                    } else {
                        float est = T[q][i] + t0[q];
                        float res = wi*(oi - est);
                        ob[q] = ob[q] + res*res;
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // selects, not a branch per cell: under a branch the compiler sinks
        // the cell's whole misfit computation into it
        unsigned long long key = ~0ull;
        #pragma unroll
        for (int q = 0; q < CPT; q++) {
            const long c = base + (long) q*TL;
            const unsigned long long k = make_key(ob[q], (int) c);
            key = (c < ngrd && k < key) ? k : key;
        }
        key = wave_min(key);
        if (WGMIN) {
            // few events: every workgroup would post to the same few
            // words, and atomics on one word serialize (one event over
            // 10,936 waves took 0.77 ms).  One post per workgroup, and none
            // when the event's minimum is already lower (the word only
            // decreases, so a stale read can only cause a needless post).
            if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = key;
            __syncthreads();
            if (threadIdx.x == 0) {
                unsigned long long k = s_key[0];
                #pragma unroll
                for (int q = 1; q < TL/64; q++) k = (s_key[q] < k) ? s_key[q] : k;
                if (k < __hip_atomic_load(best + e, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT))
                    atomicMin(best + e, k);
            }
            __syncthreads();
        } else if ((threadIdx.x & 63) == 0 && key != ~0ull) {
            atomicMin(best + e, key);
        }
    }
}

// locate.F90:684-688 at one cell per event (origin time of the optimum).
__global__ void k_t0_at(float* __restrict__ t0, const float* __restrict__ tab,
                        int ntf, long ldgrd, const float* __restrict__ obs,
                        const float* __restrict__ wt,
                        const unsigned char* __restrict__ skip,
                        const float* __restrict__ xnorm, int lntf,
                        const int* __restrict__ cells, int nev, int ngrd)
{
    const int e = blockIdx.x*blockDim.x + threadIdx.x;
    if (e >= nev) return;
    const long c = cells[e];
    if (c < 0 || c >= ngrd) { t0[e] = 0.0f; return; }
    const float* o = obs + (long) e*lntf;
    const float* w = wt + (long) e*lntf;
    const unsigned char* sk = skip + (long) e*lntf;
    float t = 0.0f;
    for (int i = 0; i < ntf; i++) {
        if (sk[i]) continue;
        float res = xnorm[e]*(o[i] - tab[(long) i*ldgrd + c]);
        t = t + w[i]*res;
    }
    t0[e] = t;
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : 1; }

// Event groups of a batched search.  Each workgroup searches its cells for
// the events of one group, so a launch is nblk x groups workgroups of equal
// length.  They run in rounds of as many as are resident at once, and a
// last round that is only partly full leaves CUs idle: 2734 cell blocks of
// 48 tables of 161 x 161 x 81 make 3.56 rounds of 768, 89 % of the GPU
// busy on average.  The group count minimizes rounds x events per group
// (each group reloads its cells' tables, so groups keep at least 256
// events).
int event_groups(const void* fn, int nblk, int nev)
{
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TL, 0)
            != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess) return 1;
    const long cap = (long) per_cu*ncu;
    if (cap <= 0 || nblk <= 0 || nev <= 0) return 1;
    int best = 1;
    double bt = 0.0;
    for (int g = 1; g <= 16; g++) {
        const long epg = (nev + g - 1)/g;
        if (g > 1 && epg < 256) break;
        const long rounds = ((long) nblk*g + cap - 1)/cap;
        const double t = (double) rounds*(double) epg;
        if (g == 1 || t < 0.999*bt) { bt = t; best = g; }
    }
    return best;
}

} // namespace

extern "C" {

int fteik_k_tables_d2f(float* tables, const double* tt, int ngrd,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_tables_d2f, dim3((ngrd + 255)/256), dim3(256), 0, s,
                       tables, tt, ngrd);
    return ok();
}

int fteik_k_locate(const float* tables, int ntf, int ngrd, long ldgrd,
                   const float* obs, const float* wt, const unsigned char* skip,
                   int lntf, float xnorm, float* t0work, float* objwork,
                   int accumulate, unsigned long long* best, hipStream_t s)
{
    (void) lntf;
    hipLaunchKernelGGL(k_locate_one, dim3((ngrd + TL - 1)/TL), dim3(TL), 0, s,
                       tables, ntf, ngrd, ldgrd, obs, wt, skip, xnorm, t0work,
                       objwork, accumulate, best);
    return ok();
}

int fteik_k_locate_batch(const float* tables, int ntf, int ngrd, long ldgrd,
                         const float* obs, const float* wt, const float* wt2,
                         const float* xnorm, int lntf, int nev,
                         unsigned long long* best, hipStream_t s)
{
    // cells per thread; measured on 48 tables of 161x161x81 and 10,000
    // events: 2 -> 58.7k, 3 -> 61.7k, 4 -> 53.6k events/s (4 costs occupancy)
    constexpr int CPT = 3;
    int nblk = (int) ((ngrd + (long) TL*CPT - 1)/((long) TL*CPT));
    // few events: the variant that posts one key per workgroup (see the
    // kernel's key posting)
    const bool few = nev <= 8;
    const void* fn = nullptr;
#define LB(N) case N: fn = few ? (const void*) k_locate_batch<N, CPT, true> \
                               : (const void*) k_locate_batch<N, CPT, false>; \
                      break;
    switch (ntf) {
        LB(1) LB(2) LB(3) LB(4) LB(5) LB(6) LB(7) LB(8) LB(10) LB(12) LB(16)
        LB(20) LB(24) LB(32) LB(40) LB(48) LB(56) LB(64)
        default: return 2;
    }
#undef LB
    int groups = event_groups(fn, nblk, nev);
    int epg = (nev + groups - 1)/groups;
    void* args[] = {(void*) &tables, (void*) &ngrd, (void*) &ldgrd,
                    (void*) &obs, (void*) &wt, (void*) &wt2, (void*) &xnorm,
                    (void*) &lntf, (void*) &nev, (void*) &epg, (void*) &best};
    if (hipLaunchKernel(fn, dim3(nblk, groups), dim3(TL), args, 0, s)
            != hipSuccess) return 1;
    return ok();
}

int fteik_k_locate_t0_at(float* t0, const float* tables, int ntf, long ldgrd,
                         const float* obs, const float* wt,
                         const unsigned char* skip, const float* xnorm,
                         int lntf, const int* cells, int nev, int ngrd,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_t0_at, dim3((nev + 127)/128), dim3(128), 0, s, t0,
                       tables, ntf, ldgrd, obs, wt, skip, xnorm, lntf, cells,
                       nev, ngrd);
    return ok();
}

} // extern "C"
