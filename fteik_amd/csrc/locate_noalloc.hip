// locate_noalloc.hip -- locate_locateEventsNoAlloc: the batched grid search
// of locate_locateEvents with every per-call device buffer (observations,
// weights, argmin keys, optimum cells, origin times) carved out of one
// file-scope device array instead of hipMalloc'd, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).
//
// The array lives once per GPU for the life of the process.  Its device
// address is looked up with hipGetSymbolAddress once per device (host code
// must never pass the array's own name to a kernel or copy: in host code it
// is a host-side stand-in).  Calls that overlap on one GPU share the array
// and race; the library is single-threaded like the reference.
#include "capi_internal.h"

using namespace fteik_capi;

namespace {

constexpr int kMaxEvents = 4096;   // events per call
constexpr int kMaxLntf = 64;       // padded observations per event
constexpr size_t kAlign = 256;

constexpr size_t al(size_t b) { return (b + kAlign - 1)/kAlign*kAlign; }

// obs, wt, wt2 (floats), skip (bytes), xnorm (floats), keys (u64), cells
// (ints), origin times (floats), each segment 256-byte aligned
constexpr size_t kScratchBytes =
    3*al((size_t) kMaxEvents*kMaxLntf*sizeof(float)) +
    al((size_t) kMaxEvents*kMaxLntf) + al(kMaxEvents*sizeof(float)) +
    al(kMaxEvents*sizeof(unsigned long long)) + al(kMaxEvents*sizeof(int)) +
    al(kMaxEvents*sizeof(float));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Host staging with the same layout: one copy moves every input.
static unsigned char h_pack[kScratchBytes];

constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];

// Byte offsets of the segments for n events of lntf observations; the
// inputs (obs .. keys) are contiguous from 0 to `in_end`, the outputs
// (keys, origin times) from `keys` to `out_end`.
struct Layout {
    size_t obs, wt, wt2, skip, xn, keys, t0, cells, in_end, out_end;
};

Layout layout(int n, int lntf)
{
    Layout L;
    const size_t m = (size_t) n*lntf;
    L.obs = 0;
    L.wt = L.obs + al(m*sizeof(float));
    L.wt2 = L.wt + al(m*sizeof(float));
    L.skip = L.wt2 + al(m*sizeof(float));
    L.xn = L.skip + al(m);
    L.keys = L.xn + al((size_t) n*sizeof(float));
    L.in_end = L.keys + (size_t) n*sizeof(unsigned long long);
    L.t0 = L.keys + al((size_t) n*sizeof(unsigned long long));
    L.out_end = L.t0 + (size_t) n*sizeof(float);
    L.cells = L.t0 + al((size_t) n*sizeof(float));
    return L;
}

__global__ void k_cells_from_keys(const unsigned long long* __restrict__ keys,
                                  int* __restrict__ cells, int n)
{
    const int e = blockIdx.x*blockDim.x + threadIdx.x;
    if (e < n) cells[e] = (int) (keys[e] & 0xffffffffull);
}

// Device address of this GPU's scratch array (looked up once per device).
unsigned char *scratch_ptr(const char *who)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        msg("%s: no current device", who);
        return nullptr;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            msg("%s: hipGetSymbolAddress failed (%s)", who,
                hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return (unsigned char *) g_scratch_ptr[dev];
}

} // namespace

extern "C" {

void locate_locateEventsNoAlloc(int n, const int *evnmbr, int *optIndx,
                                double *t0Opt, double *objOpt, int *ierr)
{
    static const char *who = "locate_locateEventsNoAlloc";
    *ierr = 0;
    if (!gl.init || n < 1) {
        msg("%s: locator not initialized or no events", who);
        *ierr = 1;
        return;
    }
    if (n > kMaxEvents || gl.lntf > kMaxLntf) {
        msg("%s: at most %d events of at most %d padded observations per "
            "call (got %d events, %d observations)", who, kMaxEvents,
            kMaxLntf, n, gl.lntf);
        *ierr = 1;
        return;
    }
    for (int e = 0; e < n; e++)
        if (evnmbr[e] < 1 || evnmbr[e] > gl.nEvents) {
            msg("%s: invalid event %d", who, evnmbr[e]);
            *ierr = 1;
            return;
        }
    unsigned char *d = scratch_ptr(who);
    if (!d) { *ierr = 1; return; }
    const int lntf = gl.lntf;
    const Layout Ly = layout(n, lntf);
    float *o = (float *) (h_pack + Ly.obs), *w = (float *) (h_pack + Ly.wt),
          *w2 = (float *) (h_pack + Ly.wt2), *xn = (float *) (h_pack + Ly.xn);
    unsigned char *sk = h_pack + Ly.skip;
    unsigned long long *keys = (unsigned long long *) (h_pack + Ly.keys);
    for (int e = 0; e < n; e++) {
        const size_t k = (size_t) (evnmbr[e] - 1)*lntf;
        memcpy(o + (size_t) e*lntf, &gl.obs[k], lntf*sizeof(float));
        memcpy(w + (size_t) e*lntf, &gl.wt[k], lntf*sizeof(float));
        memcpy(sk + (size_t) e*lntf, &gl.skip[k], (size_t) lntf);
        xn[e] = weight_norm(&gl.wt[k], lntf);
        keys[e] = ~0ull;
    }
    mask_skipped(o, w, w2, sk, (size_t) n*lntf);
    hipStream_t s = g_dev.stream;
    float *d_o = (float *) (d + Ly.obs), *d_w = (float *) (d + Ly.wt),
          *d_w2 = (float *) (d + Ly.wt2), *d_xn = (float *) (d + Ly.xn),
          *d_t0 = (float *) (d + Ly.t0);
    unsigned char *d_sk = d + Ly.skip;
    unsigned long long *d_keys = (unsigned long long *) (d + Ly.keys);
    int *d_cells = (int *) (d + Ly.cells);
    bool ok = hipMemcpyAsync(d, h_pack, Ly.in_end, hipMemcpyHostToDevice, s)
              == hipSuccess;
    if (ok) {
        const int rc = fteik_k_locate_batch(gl.d_tab, gl.ntf, gl.ngrd,
                                            gl.ldgrd, d_o, d_w, d_w2, d_xn,
                                            lntf, n, d_keys, s);
        if (rc == 2) {
            // table count without a specialised kernel: one pass per event
            for (int e = 0; e < n && ok; e++)
                ok = fteik_k_locate(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                    d_o + (size_t) e*lntf,
                                    d_w + (size_t) e*lntf,
                                    d_sk + (size_t) e*lntf, lntf, xn[e],
                                    nullptr, nullptr, 0, d_keys + e, s) == 0;
        } else {
            ok = rc == 0;
        }
    }
    if (ok) {
        hipLaunchKernelGGL(k_cells_from_keys, dim3((n + 255)/256), dim3(256),
                           0, s, d_keys, d_cells, n);
        ok = hipGetLastError() == hipSuccess;
    }
    // origin time at each optimum (locate.F90:684-688)
    ok = ok && fteik_k_locate_t0_at(d_t0, gl.d_tab, gl.ntf, gl.ldgrd, d_o,
                                    d_w, d_sk, d_xn, lntf, d_cells, n,
                                    gl.ngrd, s) == 0;
    ok = ok && hipMemcpyAsync(h_pack + Ly.keys, d + Ly.keys,
                              Ly.out_end - Ly.keys, hipMemcpyDeviceToHost, s)
               == hipSuccess
            && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    const float *t0 = (const float *) (h_pack + Ly.t0);
    for (int e = 0; e < n; e++) {
        const unsigned long long k = keys[e];
        unsigned bits = (unsigned) (k >> 32);
        float obj;
        memcpy(&obj, &bits, sizeof(obj));
        optIndx[e] = (int) (k & 0xffffffffull) + 1;
        objOpt[e] = (double) obj;
        t0Opt[e] = (double) t0[e] + gl.tepoch[(size_t) (evnmbr[e] - 1)];
    }
}

} // extern "C"
