// capi2d.cpp -- the 2D solver entry points (solver2D.F90).
#include "capi_internal.h"

using namespace fteik_capi;

namespace {

// ================================================================== 2D
struct Solver2D {
    bool init = false, haveModel = false, haveTimes = false;
    int nz = 0, nx = 0, nsweep = 1, verbose = 0, srcNumber = -1;
    double z0 = 0, x0 = 0, dz = 0, dx = 0, eps = 0, convTol = 0;
    std::vector<SrcInfo> src;
    std::vector<int> ri;
    std::vector<double> rd;
    std::vector<double> zr, xr;      // receiver positions (model coords)
    int nrec = 0;
    std::vector<double> ttimesRec;   // [nrec x nsrc]
    double *d_slow = nullptr, *d_tt = nullptr, *d_scr = nullptr;
    double *d_ttb = nullptr;         // batch fields for solveLSM
    long ttb_cap = 0;
    int *d_ri = nullptr;
    double *d_rd = nullptr, *d_ttr = nullptr;
    double *d_ttrb = nullptr;        // receiver times of a batch of sources
    int *d_err = nullptr;            // wait-limit flag of the pipelined sweeps
    double *d_scr2 = nullptr;        // their per-source scratch lines
    int scr2_cap = 0;
    int schedule = 0;                // 0 pipelined columns, 1 level by level
    long ttrb_cap = 0;
    double *d_sa = nullptr;          // zsa, xsa per source
    int *d_si = nullptr;             // zsi, xsi per source
    double *d_gx = nullptr, *d_gz = nullptr;   // gradient, on demand
    bool haveGrad = false;
};
Solver2D g2;

void free2d()
{
    dfree(g2.d_slow); dfree(g2.d_tt); dfree(g2.d_scr); dfree(g2.d_ttb);
    dfree(g2.d_ri); dfree(g2.d_rd); dfree(g2.d_ttr);
    dfree(g2.d_sa); dfree(g2.d_si); dfree(g2.d_gx); dfree(g2.d_gz);
    dfree(g2.d_ttrb); dfree(g2.d_err); dfree(g2.d_scr2);
    g2 = Solver2D();
}

} // namespace

extern "C" {

void fteik_solver2d_free(void)
{
    free2d();
    model_free();
}
void fteik_solver2d_setVerbosity(int verbose) { g2.verbose = verbose; }
void fteik_solver2d_setSchedule(int schedule, int *ierr)
{
    *ierr = 0;
    if (schedule != 0 && schedule != 1) {
        msg("fteik_solver2d_setSchedule: schedule must be 0 or 1");
        *ierr = 1;
        return;
    }
    g2.schedule = schedule;
}
void fteik_solver2d_setConvergenceTolerance(double convTol)
{
    g2.convTol = convTol;
}

void fteik_solver2d_setNumberOfSweeps(int nsweep, int *ierr)
{
    *ierr = 0;
    g2.nsweep = 1;
    if (nsweep < 0) {
        msg("fteik_solver2d_setNumberOfSweeps: nsweep=%d must be positive",
            g2.nsweep);
        *ierr = 1;
        return;
    }
    g2.nsweep = nsweep;
}

void fteik_solver2d_setSphereToCartEpsilon(double eps, int *ierr)
{
    *ierr = 0;
    g2.eps = 0.0;
    if (g2.nz < 1 || g2.nx < 1) {
        msg("fteik_solver2d_setSphereToCartEpsilon: Grid not yet set");
        *ierr = 1;
        return;
    }
    if (!check_eps("fteik_solver2d_setSphereToCartEpsilon", eps, g2.nz, g2.nx,
                   1, false)) {
        *ierr = 1;
        return;
    }
    g2.eps = eps;
}

void fteik_solver2d_initialize(int nz, int nx, double z0, double x0,
                               double dz, double dx, int nsweep, double eps,
                               double convTol, int verbose, int *ierr)
{
    *ierr = 0;
    if (verbose > 0) printf("fteik_solver2d_initialize64f: Initializing...\n");
    free2d();
    model_free();
    g2.verbose = verbose;
    g2.convTol = convTol;
    if (nz < 3 || nx < 3 || dz <= 0.0 || dx <= 0.0) {
        msg("fteik_solver2d_initialize64f: Error setting grid geometry");
        *ierr = 1;
        return;
    }
    g2.nz = nz; g2.nx = nx; g2.z0 = z0; g2.x0 = x0; g2.dz = dz; g2.dx = dx;
    fteik_solver2d_setNumberOfSweeps(nsweep, ierr);
    if (*ierr) {
        msg("fteik_solver2d_initialize64f: Failed to set max number of sweeps");
        return;
    }
    fteik_solver2d_setSphereToCartEpsilon(eps, ierr);
    if (*ierr) {
        msg("fteik_solver2d_initialize64f: Failed to set epsilon");
        return;
    }
    if (!device_ok("fteik_solver2d_initialize")) { *ierr = 1; return; }
    const long ngrd = (long) nz*nx, ncell = (long) (nz - 1)*(nx - 1);
    if (!dalloc(&g2.d_slow, ncell) || !dalloc(&g2.d_tt, ngrd) ||
        !dalloc(&g2.d_scr, ngrd) || !dalloc(&g2.d_sa, 2) ||
        !dalloc(&g2.d_si, 2)) {
        msg("fteik_solver2d_initialize: out of GPU memory");
        free2d();
        *ierr = 1;
        return;
    }
    model_set(false, nz, nx, 1);
    g2.init = true;
}

static void set_model2d(const char *who, int kind, int n, int order,
                        const double *vel, int *ierr)
{
    *ierr = 0;
    g2.haveModel = false;
    if (!g2.init) { msg("%s: solver not initialized", who); *ierr = 1; return; }
    if (!set_model(who, false, g2.nz, g2.nx, 1, kind, n,
                   (order == 1) ? 1 : 0, vel, g2.d_slow)) {
        msg("%s: Error setting velocity model", who);
        *ierr = 1;
        return;
    }
    g2.haveModel = true;
}

void fteik_solver2d_setVelocityModel64f(int ncell, const double *vel,
                                        int *ierr)
{
    set_model2d("fteik_solver2d_setVelocityModel64f", 0, ncell, 0, vel, ierr);
}

void fteik_solver2d_setCellVelocityModel64f(int nc, int order,
                                            const double *vel, int *ierr)
{
    set_model2d("fteik_solver2d_setCellVelocityModel64f", 1, nc, order, vel,
                ierr);
}

void fteik_solver2d_setNodalVelocityModel64f(int ng, int order,
                                             const double *vel, int *ierr)
{
    set_model2d("fteik_solver2d_setNodalVelocityModel64f", 2, ng, order, vel,
                ierr);
}

static void size_rec2d()
{
    if (g2.nrec > 0 && !g2.src.empty())
        g2.ttimesRec.assign((size_t) g2.nrec*g2.src.size(), 99999.0);
    else
        g2.ttimesRec.clear();
}

void fteik_solver2d_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, int *ierr)
{
    *ierr = 0;
    g2.src.clear();
    if (nsrc < 1 || g2.nz < 1) {
        msg("fteik_solver2d_setSources64f: Failed to set source");
        *ierr = 1;
        return;
    }
    std::vector<SrcInfo> v((size_t) nsrc);
    for (int i = 0; i < nsrc; i++) {
        if (locate_source(false, g2.nz, g2.nx, 1, g2.dz, g2.dx, 1.0, g2.z0,
                          g2.x0, 0.0, zsrc[i], xsrc[i], 0.0, i + 1,
                          &v[(size_t) i])) {
            msg("fteik_solver2d_setSources64f: Failed to set source");
            *ierr = 1;
            return;
        }
    }
    g2.src.swap(v);
    size_rec2d();
}

void fteik_solver2d_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, int *ierr)
{
    *ierr = 0;
    if (nrec < 1) {
        printf("fteik_solver2d_setReceivers64f: No receivers to set\n");
        return;
    }
    if (!g2.init) {
        msg("fteik_solver2d_setReceivers64f: Failed to set receivers");
        *ierr = 1;
        return;
    }
    int e = locate_receivers(false, g2.nz, g2.nx, 1, g2.dz, g2.dx, 1.0, g2.z0,
                             g2.x0, 0.0, nrec, zrec, xrec, nullptr, g2.ri,
                             g2.rd);
    dfree(g2.d_ri); dfree(g2.d_rd); dfree(g2.d_ttr);
    if (!dalloc(&g2.d_ri, 3*(size_t) nrec) ||
        !dalloc(&g2.d_rd, 3*(size_t) nrec) || !dalloc(&g2.d_ttr, nrec) ||
        !h2d(g2.d_ri, g2.ri.data(), 3*(size_t) nrec*sizeof(int)) ||
        !h2d(g2.d_rd, g2.rd.data(), 3*(size_t) nrec*sizeof(double))) {
        *ierr = 1;
        return;
    }
    g2.nrec = nrec;
    g2.zr.assign(zrec, zrec + nrec);
    g2.xr.assign(xrec, xrec + nrec);
    size_rec2d();
    if (e) {
        msg("fteik_solver2d_setReceivers64f: Failed to set receivers");
        *ierr = e;
    }
}

void fteik_solver2d_getNumberOfReceivers(int *nrec, int *ierr)
{
    *ierr = 0;
    *nrec = 0;
    if (g2.nrec < 1) {
        msg("fteik_receiver_getNumberOfReceivers: Module not initialized");
        *ierr = 1;
        return;
    }
    *nrec = g2.nrec;
}

// Solves sources [first, first+n) (0-based) in one launch; the last one's
// field stays in d_tt.  Receiver times go to ttimesRec.
static bool solve2d_range(int first, int n, const char *who)
{
    for (int s = first; s < first + n; s++) {
        const SrcInfo &si = g2.src[(size_t) s];
        if (si.zsi > g2.nz - 2 || si.xsi > g2.nx - 2) {
            msg("%s: source %d lies in the last cell along an axis; the "
                "reference reads outside the grid for such sources", who,
                s + 1);
            return false;
        }
    }
    const long ngrd = (long) g2.nz*g2.nx;
    if (n > 1 && g2.ttb_cap < n*ngrd) {
        dfree(g2.d_ttb);
        if (!dalloc(&g2.d_ttb, n*ngrd)) { msg("%s: out of GPU memory", who);
                                          return false; }
        g2.ttb_cap = n*ngrd;
    }
    dfree(g2.d_sa); dfree(g2.d_si);
    std::vector<double> sa(2*(size_t) n);
    std::vector<int> si(2*(size_t) n);
    for (int q = 0; q < n; q++) {
        sa[(size_t) q] = g2.src[(size_t) (first + q)].zsa;
        sa[(size_t) (n + q)] = g2.src[(size_t) (first + q)].xsa;
        si[(size_t) q] = g2.src[(size_t) (first + q)].zsi;
        si[(size_t) (n + q)] = g2.src[(size_t) (first + q)].xsi;
    }
    if (!dalloc(&g2.d_sa, 2*(size_t) n) || !dalloc(&g2.d_si, 2*(size_t) n) ||
        !h2d(g2.d_sa, sa.data(), sa.size()*sizeof(double)) ||
        !h2d(g2.d_si, si.data(), si.size()*sizeof(int))) return false;
    double *fields = (n > 1) ? g2.d_ttb : g2.d_tt;
    if (!g2.d_err && !dalloc(&g2.d_err, 1)) return false;
    if (g2.scr2_cap < n) {
        dfree(g2.d_scr2);
        g2.scr2_cap = 0;
        if (!dalloc(&g2.d_scr2, (size_t) n*kSolve2dScratch)) {
            msg("%s: out of GPU memory", who);
            return false;
        }
        g2.scr2_cap = n;
    }
    if (hipMemsetAsync(g2.d_err, 0, sizeof(int), g_dev.stream) != hipSuccess ||
        fteik_k_solve2d(fields, g2.d_slow, g2.nz, g2.nx, g2.dz, g2.dx, g2.eps,
                        0.0, 0.0, 0, 0, g2.nsweep, n, (int) ngrd, g2.d_sa,
                        g2.d_sa + n, g2.d_si, g2.d_si + n, g2.schedule,
                        g2.d_err, g2.d_scr2, g_dev.stream))
        return false;
    {
        int e = 0;
        if (!d2h(&e, g2.d_err, sizeof(int))) return false;
        if (e) {
            msg("%s: a sweep's wait limit was exceeded", who);
            return false;
        }
    }
    if (g2.nrec > 0) {
        // every source's receiver times in one launch and one copy
        const long need = (long) n*g2.nrec;
        if (g2.ttrb_cap < need) {
            dfree(g2.d_ttrb);
            g2.ttrb_cap = 0;
            if (!dalloc(&g2.d_ttrb, (size_t) need)) {
                msg("%s: out of GPU memory", who);
                return false;
            }
            g2.ttrb_cap = need;
        }
        for (int q = 0; q < n; q += 65535) {
            const int m = std::min(65535, n - q);
            if (fteik_k_receivers_many(g2.d_ttrb + (long) q*g2.nrec,
                                       fields + (long) q*ngrd, m, ngrd,
                                       g2.d_ri, g2.d_rd, g2.nrec, g2.nz,
                                       g2.nx, g2.dz, g2.dx, 1.0, 0,
                                       g_dev.stream)) return false;
        }
        if (!d2h(g2.ttimesRec.data() + (size_t) first*g2.nrec, g2.d_ttrb,
                 (size_t) need*sizeof(double))) return false;
    }
    if (n > 1 &&
        hipMemcpyAsync(g2.d_tt, fields + (long) (n - 1)*ngrd,
                       ngrd*sizeof(double), hipMemcpyDeviceToDevice,
                       g_dev.stream) != hipSuccess) return false;
    return sync_ok();
}

void fteik_solver2d_solveSourceLSM(int isrc, int *ierr)
{
    static const char *who = "fteik_solver2d_solveSourceLSM";
    *ierr = 0;
    g2.haveTimes = false;
    g2.haveGrad = false;
    g2.srcNumber = -1;
    if (!g2.init || !g2.haveModel) {
        msg("%s: Model not yet set", who);
        *ierr = 1;
        return;
    }
    const int nsrc = (int) g2.src.size();
    if (isrc < 1 || isrc > nsrc) {
        msg("%s: Source number =%d must be in range [1,%d]", who, isrc, nsrc);
        *ierr = 1;
        return;
    }
    if (!solve2d_range(isrc - 1, 1, who)) { *ierr = 1; return; }
    g2.haveTimes = true;
    g2.srcNumber = isrc;
}

void fteik_solver2d_solveSourceFSM(int isrc, int *ierr)
{
    // solver2D.F90:1042-1280: for every source the reference accepts
    // (zsi <= nz-2, xsi <= nx-2) the FSM visits the same masked nodes in a
    // lexicographic order; it does not extract receiver times.
    static const char *who = "fteik_solver2d_solveSourceFSM";
    *ierr = 0;
    g2.haveTimes = false;
    g2.haveGrad = false;
    if (!g2.init || !g2.haveModel) { msg("%s: Model not yet set", who);
                                     *ierr = 1; return; }
    const int nsrc = (int) g2.src.size();
    if (isrc < 1 || isrc > nsrc) {
        msg("%s: Source number =%d must be in range [1,%d]", who, isrc, nsrc);
        *ierr = 1;
        return;
    }
    int nrec = g2.nrec;
    g2.nrec = 0;
    bool ok = solve2d_range(isrc - 1, 1, who);
    g2.nrec = nrec;
    if (!ok) { *ierr = 1; return; }
    g2.haveTimes = true;
    g2.srcNumber = -1;
}

void fteik_solver2d_solveLSM(int *ierr)
{
    static const char *who = "fteik_solver2d_solveLSM";
    *ierr = 0;
    if (!g2.init || !g2.haveModel) { msg("%s: Model not yet set", who);
                                     *ierr = 1; return; }
    const int nsrc = (int) g2.src.size();
    if (nsrc < 1) {
        if (g2.verbose > 0) msg("%s: No sources", who);
        return;
    }
    // all sources in one launch (one workgroup per source), in chunks that
    // keep the batch of fields within a fixed memory budget
    const long ngrd = (long) g2.nz*g2.nx;
    const int chunk = (int) std::max(1L, std::min<long>(nsrc,
                                (1L << 28)/std::max(1L, ngrd)));
    g2.haveTimes = false;
    g2.haveGrad = false;
    for (int f = 0; f < nsrc; f += chunk) {
        int n = std::min(chunk, nsrc - f);
        if (!solve2d_range(f, n, who)) {
            msg("%s: Error solving source %d", who, f + 1);
            *ierr = 1;
            return;
        }
    }
    g2.haveTimes = true;
    g2.srcNumber = nsrc;
}

void fteik_solver2d_getTravelTimeField64f(int ngin, int order, double *tt,
                                          int *ierr)
{
    *ierr = 0;
    if (!g2.haveTimes) {
        msg("fteik_solver2d_getTravelTimeField64f: Travel times not yet "
            "computed");
        *ierr = 1;
        return;
    }
    const long ngrd = (long) g2.nz*g2.nx;
    if (ngin != ngrd) {
        msg("fteik_solver2d_getTravelTimeField64f: ngin=%d /= ngrd =%ld",
            ngin, ngrd);
        *ierr = 1;
        return;
    }
    const double *src = g2.d_tt;
    if (order == 1) {
        if (fteik_k_field_reorder(g2.d_scr, g2.d_tt, g2.nz, g2.nx, 1, 0, 1,
                                  g_dev.stream)) { *ierr = 1; return; }
        src = g2.d_scr;
    } else if (order != 0) {
        printf("fteik_solver2d_getTravelTimeField64f: Defaulting to natural "
               "order\n");
    }
    if (!d2h(tt, src, ngrd*sizeof(double))) *ierr = 1;
}

void fteik_solver2d_getTravelTimeFieldCell64f(int ncin, int order, double *tt,
                                              int *ierr)
{
    *ierr = 0;
    if (!g2.haveTimes) {
        msg("fteik_solver2d_getTravelTimeFieldCell64f: Travel times not yet "
            "computed");
        *ierr = 1;
        return;
    }
    const long ncell = (long) (g2.nz - 1)*(g2.nx - 1);
    if (ncin != ncell) {
        msg("fteik_solver2d_getTravelTimeFieldCell64f: ncin=%d /= ncell =%ld",
            ncin, ncell);
        *ierr = 1;
        return;
    }
    if (fteik_k_field_cells(g2.d_scr, g2.d_tt, g2.nz, g2.nx, 1, 0,
                            order == 1 ? 1 : 0, g_dev.stream) ||
        !d2h(tt, g2.d_scr, ncell*sizeof(double))) *ierr = 1;
}

void fteik_solver2d_getTravelTimes64f(int ldr, double *ttr, int *ierr)
{
    *ierr = 0;
    const int nsrc = (int) g2.src.size();
    if (g2.nrec < 1) {
        printf("fteik_solver2d_getTravelTimes64f: No receivers\n");
        return;
    }
    if (nsrc < 1) {
        msg("fteik_solver2d_getTravelTimes64f: No sources");
        *ierr = 1;
        return;
    }
    if (!g2.haveTimes) {
        msg("fteik_solver2d_getTravelTimes64f: Travel times not yet computed");
        *ierr = 1;
        return;
    }
    if (ldr < g2.nrec) {
        msg("fteik_solver2d_getTravelTimes64f: ldr = %d < nrec = %d", ldr,
            g2.nrec);
        *ierr = 1;
        return;
    }
    if (ldr > g2.nrec)
        for (long i = 0; i < (long) ldr*nsrc; i++) ttr[i] = 0.0;
    for (int s = 0; s < nsrc; s++)
        memcpy(ttr + (size_t) s*ldr, g2.ttimesRec.data() + (size_t) s*g2.nrec,
               (size_t) g2.nrec*sizeof(double));
}

// solver2D.F90:984-1032 on the device, once per solved field.
static bool gradient2d(const char *who)
{
    if (g2.haveGrad) return true;
    const long ngrd = (long) g2.nz*g2.nx;
    if ((!g2.d_gx && !dalloc(&g2.d_gx, (size_t) ngrd)) ||
        (!g2.d_gz && !dalloc(&g2.d_gz, (size_t) ngrd))) {
        msg("%s: out of GPU memory (gradient)", who);
        return false;
    }
    if (fteik_k_fd_gradient2d(g2.d_gx, g2.d_gz, g2.d_tt, g2.nz, g2.nx, g2.dz,
                              g2.dx, g_dev.stream) != 0 || !sync_ok()) {
        msg("%s: GPU error (gradient)", who);
        return false;
    }
    g2.haveGrad = true;
    return true;
}

// solver2D.F90:310-355: rays from the points to the solved source (the
// field of the last source solved by solveSourceLSM or solveLSM).
void fteik_solver2d_traceRaysToPoints(int np, const double *xp,
                                      const double *zp, int *ierr)
{
    static const char *who = "fteik_solver2d_traceRaysToPoints";
    *ierr = 0;
    if (!g2.haveTimes || g2.srcNumber < 1) {
        msg("%s: Travel time field not yet computed", who);
        *ierr = 1;
        return;
    }
    if (!gradient2d(who)) { *ierr = 1; return; }
    RayGeom G;
    G.is3d = false; G.nz = g2.nz; G.nx = g2.nx; G.ny = 1;
    G.dz = g2.dz; G.dx = g2.dx; G.dy = 1.0;
    G.z0 = g2.z0; G.x0 = g2.x0; G.y0 = 0.0;
    const SrcInfo &s = g2.src[(size_t) (g2.srcNumber - 1)];
    if (!trace_rays(G, g2.d_slow, nullptr, g2.d_gx, g2.d_gz, s.ztrue,
                    s.xtrue, 0.0, np, xp, nullptr, zp, who))
        *ierr = 1;
}

// solver2D.F90:283-302
void fteik_solver2d_traceRaysToReceivers(int *ierr)
{
    *ierr = 0;
    if (g2.nrec < 1) {
        msg("fteik_solver2d_traceRaysToReceivers: No receivers set");
        *ierr = 1;
        return;
    }
    fteik_solver2d_traceRaysToPoints(g2.nrec, g2.xr.data(), g2.zr.data(),
                                     ierr);
    if (*ierr != 0) {
        msg("fteik_solver2d_traceRaysToReceivers: Failed to compute rays");
        *ierr = 1;
    }
}

// The gradient of the solved field (solver2D.F90:984-1032), per node in
// natural order (z fastest).
void fteik_solver2d_getTravelTimeGradient64f(int ngin, double *tgradx,
                                             double *tgradz, int *ierr)
{
    static const char *who = "fteik_solver2d_getTravelTimeGradient64f";
    *ierr = 0;
    if (!g2.haveTimes) {
        msg("%s: Travel times not yet computed", who);
        *ierr = 1;
        return;
    }
    const long ngrd = (long) g2.nz*g2.nx;
    if (ngin != ngrd) {
        msg("%s: ngin=%d /= ngrd =%ld", who, ngin, ngrd);
        *ierr = 1;
        return;
    }
    if (!gradient2d(who) ||
        !d2h(tgradx, g2.d_gx, (size_t) ngrd*sizeof(double)) ||
        !d2h(tgradz, g2.d_gz, (size_t) ngrd*sizeof(double))) *ierr = 1;
}

} // extern "C"
