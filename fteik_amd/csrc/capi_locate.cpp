// capi_locate.cpp -- the travel-time grid-search locator entry points
// (locate.F90).
#include "capi_internal.h"

using namespace fteik_capi;

namespace fteik_capi {

static void free_locator()
{
    dfree(gl.d_tab); dfree(gl.d_t0); dfree(gl.d_obs); dfree(gl.d_wt);
    dfree(gl.d_xn); dfree(gl.d_skip); dfree(gl.d_best); dfree(gl.d_tmp);
    int rc = gl.refCompat;
    gl = Locator();
    gl.refCompat = rc;
}

// memory.F90:194-211
int padlength32(int alignment, int n)
{
    long xmod = ((long) n*4) % alignment;
    return xmod != 0 ? (int) ((alignment - xmod)/4) : 0;
}

// Per-event arrays for the batched grid search: observations and weights
// of skipped picks zeroed (their terms then add +-0), and wt2 = sqrt(1/2) *
// wt, the product the reference forms per pick (locate.F90:694-706).
void mask_skipped(float *o, float *w, float *w2, const unsigned char *sk,
                  size_t n)
{
    const float sqrtHalf = 0.70710677f;   // SQRT(0.5) in single precision
    for (size_t i = 0; i < n; i++) {
        if (sk[i]) { o[i] = 0.0f; w[i] = 0.0f; }
        w2[i] = sqrtHalf*w[i];
    }
}

// float sum in index order (locate.F90:654 SUM over the padded weights)
float weight_norm(const float *w, int lntf)
{
    float s = 0.0f;
    for (int i = 0; i < lntf; i++) s = s + w[i];
    return 1.0f/s;
}

} // namespace fteik_capi

extern "C" {

void locate_finalize(void) { free_locator(); }

void locate_setNumberOfEvents(int nEvents, int *ierr)
{
    *ierr = 0;
    gl.nEvents = 0;
    if (nEvents < 1) {
        printf("locate_setNumberOfEventsf: nEvents=%4dmust be positive\n",
               nEvents);
        *ierr = 1;
        return;
    }
    gl.nEvents = nEvents;
}

void locate_setNumberOfTravelTimeFields(int ntf, int *ierr)
{
    *ierr = 0;
    gl.ntf = 0;
    if (ntf < 1) {
        printf("locate_setNumberOfTravelTimeFields: ntf=%4d must be positive\n",
               ntf);
        *ierr = 1;
        return;
    }
    gl.ntf = ntf;
}

void locate_setNumberOfGridPointsF(int ngrd, int *ierr)
{
    *ierr = 0;
    gl.ngrd = 0;
    if (ngrd < 1) {
        printf("locate_setNumberOfGridPointsF: Number of grid points=%8d must "
               "be positive\n", ngrd);
        *ierr = 1;
        return;
    }
    gl.ngrd = ngrd;
}

void locate_initialize(int nEvents, int ntf, int ngrd, int *ierr)
{
    *ierr = 0;
    locate_finalize();
    locate_setNumberOfEvents(nEvents, ierr);
    if (*ierr) { msg("locate_initializeF: Failed to set number of events");
                 return; }
    locate_setNumberOfTravelTimeFields(ntf, ierr);
    if (*ierr) { msg("locate_initializeF: Failed to set number of travel time "
                     "fields"); return; }
    locate_setNumberOfGridPointsF(ngrd, ierr);
    if (*ierr) { msg("locate_initializeF: Failed to set number of grid "
                     "points"); return; }
    if (!device_ok("locate_initialize")) { *ierr = 1; return; }
    gl.lntf = ntf + padlength32(64, ntf);
    gl.ldgrd = ngrd + padlength32(64, ngrd);
    gl.tepoch.assign((size_t) nEvents, 0.0);
    gl.obs.assign((size_t) nEvents*gl.lntf, 0.0f);
    gl.wt.assign((size_t) nEvents*gl.lntf, 0.0f);
    gl.skip.assign((size_t) nEvents*gl.lntf, 1);
    const long nwork = std::max(ngrd, 512);
    if (!dalloc(&gl.d_tab, (size_t) ntf*gl.ldgrd) ||
        !dalloc(&gl.d_t0, nwork) || !dalloc(&gl.d_obs, gl.lntf) ||
        !dalloc(&gl.d_wt, gl.lntf) || !dalloc(&gl.d_skip, gl.lntf) ||
        !dalloc(&gl.d_best, 1) || !dalloc(&gl.d_xn, 1)) {
        msg("locate_initializeF: out of GPU memory");
        free_locator();
        *ierr = 1;
        return;
    }
    hipMemsetAsync(gl.d_tab, 0, (size_t) ntf*gl.ldgrd*sizeof(float),
                   g_dev.stream);
    hipMemsetAsync(gl.d_t0, 0, nwork*sizeof(float), g_dev.stream);
    if (!sync_ok()) { *ierr = 1; return; }
    gl.init = true;
}

void locate_setTravelTimeField64f(int ngrdIn, int itf, const double *tt,
                                  int *ierr)
{
    *ierr = 0;
    if (!gl.init) {
        printf("locate_setTravelTimeField64fF: Locator not initialized\n");
        *ierr = 1;
        return;
    }
    if (itf < 0 || itf > gl.ntf) {
        printf("locate_setTravelTimeField64f: Field number=%4d must be in "
               "[1,%3d]\n", itf, gl.ntf);
        *ierr = 1;
        return;
    }
    if (ngrdIn != gl.ngrd) {
        printf("locate_setTravelTimeField64f: ngrdIn=%8d not equal ngrd=%8d\n",
               ngrdIn, gl.ngrd);
        *ierr = 1;
        return;
    }
    if (itf == 0) {
        // locate.F90:310 admits itf = 0, which addresses memory before the
        // table array; nothing sensible can be stored there.
        printf("locate_setTravelTimeField64f: Field number=%4d must be in "
               "[1,%3d]\n", itf, gl.ntf);
        *ierr = 1;
        return;
    }
    if (!gl.d_tmp && !dalloc(&gl.d_tmp, gl.ngrd)) { *ierr = 1; return; }
    if (!h2d(gl.d_tmp, tt, (size_t) gl.ngrd*sizeof(double)) ||
        fteik_k_tables_d2f(gl.d_tab + (long) (itf - 1)*gl.ldgrd, gl.d_tmp,
                           gl.ngrd, g_dev.stream) || !sync_ok()) *ierr = 1;
}

void locate_setTravelTimeField32f(int ngrdIn, int itf, const float *tt,
                                  int *ierr)
{
    *ierr = 0;
    if (!gl.init) {
        printf("locate_setTravelTimeField32f: Locator not initialized\n");
        *ierr = 1;
        return;
    }
    if (itf < 1 || itf > gl.ntf) {
        printf("locate_setTravelTimeField32f: Field number=%4d must be in "
               "[1,%3d]\n", itf, gl.ntf);
        *ierr = 1;
        return;
    }
    if (ngrdIn != gl.ngrd) {
        printf("locate_setTravelTimeField32f: ngrdIn=%8dnot equal ngrd=%8d\n",
               ngrdIn, gl.ngrd);
        *ierr = 1;
        return;
    }
    if (!h2d(gl.d_tab + (long) (itf - 1)*gl.ldgrd, tt,
             (size_t) gl.ngrd*sizeof(float))) *ierr = 1;
}

void locate_setObservation64f(int evnmbr, int nttimes, bool lhaveOT,
                              double t0, const int *obs2tf,
                              const double *picks, const double *wts,
                              int *ierr)
{
    (void) lhaveOT; (void) t0;   // stored but never used (locate.F90:478)
    *ierr = 0;
    if (!gl.init) {
        printf("locate_setObservation64f: Locator not initialized\n");
        *ierr = 1;
        return;
    }
    if (nttimes < 1 || nttimes > gl.ntf) {
        printf("locate_setObservation64f: nttimes=%4d must be in [1,%3d]\n",
               nttimes, gl.ntf);
        *ierr = 1;
        return;
    }
    int mn = obs2tf[0], mx = obs2tf[0];
    for (int i = 1; i < nttimes; i++) {
        mn = std::min(mn, obs2tf[i]);
        mx = std::max(mx, obs2tf[i]);
    }
    if (mn < 1) {
        printf("locate_setObservation64f: All values in obs2tf must be "
               "positive\n");
        *ierr = 1;
        return;
    }
    if (mx > gl.ntf) {
        printf("locate_setObservation64f: All values in obs2tf must be <%4d\n",
               gl.ntf);
        *ierr = 1;
        return;
    }
    if (evnmbr < 1 || evnmbr > gl.nEvents) {
        printf("locate_setObservation64f: Event number=%5d must be in range "
               "[1,%5d\n", evnmbr, gl.nEvents);
        *ierr = 1;
        return;
    }
    // locate.F90:475-495: reduce by the latest pick, REAL() to single
    double tmax = picks[0];
    for (int i = 1; i < nttimes; i++) tmax = (picks[i] > tmax) ? picks[i] : tmax;
    const size_t k = (size_t) (evnmbr - 1)*gl.lntf;
    gl.tepoch[(size_t) (evnmbr - 1)] = tmax;
    for (int i = 0; i < gl.lntf; i++) {
        gl.obs[k + i] = 0.0f; gl.wt[k + i] = 0.0f; gl.skip[k + i] = 1;
    }
    for (int i = 0; i < nttimes; i++) {
        const size_t j = k + (size_t) (obs2tf[i] - 1);
        if (wts[i] > 0.0) {
            gl.obs[j] = (float) (picks[i] - tmax);
            gl.wt[j] = (float) wts[i];
            gl.skip[j] = 0;
        }
    }
}

void locate_setObservation64fC(const int evnmbr, const int nttimes,
                               const bool lhaveOT, const double t0In,
                               const int *obs2tf, const double *pickTimes,
                               const double *wts, int *ierr)
{
    std::vector<int> o((size_t) std::max(nttimes, 1));
    for (int i = 0; i < nttimes; i++) o[(size_t) i] = obs2tf[i] + 1;
    locate_setObservation64f(evnmbr + 1, nttimes, lhaveOT, t0In, o.data(),
                             pickTimes, wts, ierr);
}

void locate_setReferenceCompatibility(int mode, int *ierr)
{
    *ierr = 0;
    if (mode != 0 && mode != 1) {
        msg("locate_setReferenceCompatibility: mode must be 0 or 1");
        *ierr = 1;
        return;
    }
    gl.refCompat = mode;
}


void locate_locateEvent(int evnmbr, int *optIndx, double *t0Opt,
                        double *objOpt)
{
    *optIndx = 0;
    *t0Opt = 0.0;
    *objOpt = 0.0;
    if (!gl.init || evnmbr < 1 || evnmbr > gl.nEvents) {
        msg("locate_locateEvent: locator not initialized or invalid event %d",
            evnmbr);
        return;
    }
    const size_t k = (size_t) (evnmbr - 1)*gl.lntf;
    const float xnorm = weight_norm(gl.wt.data() + k, gl.lntf);
    const unsigned long long init = ~0ull;
    bool ok = h2d(gl.d_obs, gl.obs.data() + k, gl.lntf*sizeof(float)) &&
              h2d(gl.d_wt, gl.wt.data() + k, gl.lntf*sizeof(float)) &&
              h2d(gl.d_skip, gl.skip.data() + k, (size_t) gl.lntf) &&
              h2d(gl.d_best, &init, sizeof(init));
    const int accumulate = gl.refCompat ? 1 : 0;
    ok = ok && fteik_k_locate(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd, gl.d_obs,
                              gl.d_wt, gl.d_skip, gl.lntf, xnorm, gl.d_t0,
                              nullptr, accumulate, gl.d_best,
                              g_dev.stream) == 0;
    unsigned long long best = 0;
    ok = ok && d2h(&best, gl.d_best, sizeof(best));
    if (!ok) {
        msg("locate_locateEvent: GPU error");
        return;
    }
    const int cell = (int) (best & 0xffffffffull);
    unsigned bits = (unsigned) (best >> 32);
    float obj;
    memcpy(&obj, &bits, sizeof(obj));
    if (gl.refCompat && gl.ngrd < 512 && obj > 0.0f) {
        // MINLOC over objWork(1:512), whose tail is zero (locate.F90:200-203)
        *optIndx = gl.ngrd + 1;
        *t0Opt = 0.0 + gl.tepoch[(size_t) (evnmbr - 1)];
        *objOpt = 0.0;
        return;
    }
    float t0 = 0.0f;
    if (!d2h(&t0, gl.d_t0 + cell, sizeof(float))) {
        msg("locate_locateEvent: GPU error");
        return;
    }
    *optIndx = cell + 1;
    *t0Opt = (double) t0 + gl.tepoch[(size_t) (evnmbr - 1)];
    *objOpt = (double) obj;
}

void locate_locateEvents(int n, const int *evnmbr, int *optIndx,
                         double *t0Opt, double *objOpt, int *ierr)
{
    *ierr = 0;
    if (!gl.init || n < 1) {
        msg("locate_locateEvents: locator not initialized or no events");
        *ierr = 1;
        return;
    }
    for (int e = 0; e < n; e++)
        if (evnmbr[e] < 1 || evnmbr[e] > gl.nEvents) {
            msg("locate_locateEvents: invalid event %d", evnmbr[e]);
            *ierr = 1;
            return;
        }
    std::vector<float> o((size_t) n*gl.lntf), w((size_t) n*gl.lntf), xn((size_t) n);
    std::vector<unsigned char> sk((size_t) n*gl.lntf);
    for (int e = 0; e < n; e++) {
        const size_t k = (size_t) (evnmbr[e] - 1)*gl.lntf;
        memcpy(&o[(size_t) e*gl.lntf], &gl.obs[k], gl.lntf*sizeof(float));
        memcpy(&w[(size_t) e*gl.lntf], &gl.wt[k], gl.lntf*sizeof(float));
        memcpy(&sk[(size_t) e*gl.lntf], &gl.skip[k], (size_t) gl.lntf);
        xn[(size_t) e] = weight_norm(&gl.wt[k], gl.lntf);
    }
    std::vector<float> w2(w.size());
    mask_skipped(o.data(), w.data(), w2.data(), sk.data(), o.size());
    float *d_o = nullptr, *d_w = nullptr, *d_w2 = nullptr, *d_xn = nullptr;
    unsigned char *d_sk = nullptr;
    unsigned long long *d_best = nullptr;
    std::vector<unsigned long long> best((size_t) n, ~0ull);
    bool ok = dalloc(&d_o, o.size()) && dalloc(&d_w, w.size()) &&
              dalloc(&d_w2, w2.size()) &&
              dalloc(&d_xn, xn.size()) && dalloc(&d_sk, sk.size()) &&
              dalloc(&d_best, best.size()) &&
              h2d(d_o, o.data(), o.size()*sizeof(float)) &&
              h2d(d_w, w.data(), w.size()*sizeof(float)) &&
              h2d(d_w2, w2.data(), w2.size()*sizeof(float)) &&
              h2d(d_xn, xn.data(), xn.size()*sizeof(float)) &&
              h2d(d_sk, sk.data(), sk.size()) &&
              h2d(d_best, best.data(), best.size()*sizeof(unsigned long long));
    if (ok) {
        int rc = fteik_k_locate_batch(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd, d_o,
                                      d_w, d_w2, d_xn, gl.lntf, n, d_best,
                                      g_dev.stream);
        if (rc == 2) {
            // table count without a specialised kernel: one pass per event
            for (int e = 0; e < n && ok; e++)
                ok = fteik_k_locate(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                    d_o + (size_t) e*gl.lntf,
                                    d_w + (size_t) e*gl.lntf,
                                    d_sk + (size_t) e*gl.lntf, gl.lntf,
                                    xn[(size_t) e], nullptr, nullptr, 0,
                                    d_best + e, g_dev.stream) == 0;
        } else {
            ok = rc == 0;
        }
    }
    ok = ok && d2h(best.data(), d_best, best.size()*sizeof(unsigned long long));
    dfree(d_best);
    std::vector<int> cells((size_t) n);
    for (int e = 0; e < n; e++) {
        const int cell = (int) (best[(size_t) e] & 0xffffffffull);
        unsigned bits = (unsigned) (best[(size_t) e] >> 32);
        float obj;
        memcpy(&obj, &bits, sizeof(obj));
        optIndx[e] = cell + 1;
        objOpt[e] = (double) obj;
        cells[(size_t) e] = cell;
    }
    // origin time at each optimum (locate.F90:684-688) on the device
    std::vector<float> t0v((size_t) n);
    int *d_c = nullptr;
    float *d_t0v = nullptr;
    ok = ok && dalloc(&d_c, (size_t) n) && dalloc(&d_t0v, (size_t) n) &&
         h2d(d_c, cells.data(), (size_t) n*sizeof(int)) &&
         fteik_k_locate_t0_at(d_t0v, gl.d_tab, gl.ntf, gl.ldgrd, d_o, d_w,
                              d_sk, d_xn, gl.lntf, d_c, n, gl.ngrd,
                              g_dev.stream) == 0 &&
         d2h(t0v.data(), d_t0v, (size_t) n*sizeof(float));
    dfree(d_o); dfree(d_w); dfree(d_w2); dfree(d_xn); dfree(d_sk); dfree(d_c);
    dfree(d_t0v);
    if (!ok) { msg("locate_locateEvents: GPU error"); *ierr = 1; return; }
    for (int e = 0; e < n; e++)
        t0Opt[e] = (double) t0v[(size_t) e]
                 + gl.tepoch[(size_t) (evnmbr[e] - 1)];
}

void locate_computeStatics32fF(const int *ldo, const int *nobs,
                               const int *nev, const float *tobs,
                               const float *tt, const float *t0,
                               const float *wts, float *tstatic)
{
    // locate.F90:748-782
    const int lo = *ldo, no = *nobs, ne = *nev;
    if (no < 1) return;
    if (!device_ok("locate_computeStatics32fF")) return;
    const size_t n = (size_t) std::max(1, ne)*lo;
    float *d = nullptr;
    bool ok = dalloc(&d, 3*n + ne + no) &&
              h2d(d, tobs, (size_t) ne*lo*sizeof(float)) &&
              h2d(d + n, tt, (size_t) ne*lo*sizeof(float)) &&
              h2d(d + 2*n, wts, (size_t) ne*lo*sizeof(float)) &&
              h2d(d + 3*n, t0, (size_t) ne*sizeof(float)) &&
              fteik_k_statics(d + 3*n + ne, d, d + n, d + 2*n, d + 3*n, lo,
                              no, ne, g_dev.stream) == 0 &&
              d2h(tstatic, d + 3*n + ne, (size_t) no*sizeof(float));
    dfree(d);
    if (!ok) msg("locate_computeStatics32fF: GPU error");
}

} // extern "C"
