// capi.cpp -- the C ABI (include/fteik/fteik_amd.h): host-side state,
// argument checking and error reporting of the reference's BIND(C) entry
// points, with all arithmetic on fields, tables and models done by the HIP
// kernels behind kernels.h.  Host code here only validates inputs, computes
// scalar geometry (grid positions of sources and receivers), moves data and
// launches kernels.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <vector>
#include <algorithm>
#include "kernels.h"
#include "../../include/fteik/fteik_amd.h"

using fteik::SourceGeom;
using fteik::SweepLaunch;

namespace {

// ----------------------------------------------------------------- common
void msg(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
}

struct Device {
    bool ready = false;
    bool failed = false;
    hipStream_t stream = nullptr;
    int ncu = 0;
};
Device g_dev;

bool device_ok(const char *who)
{
    if (g_dev.ready) return true;
    if (g_dev.failed) { msg("%s: no usable GPU", who); return false; }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        g_dev.failed = true;
        msg("%s: no usable GPU (this library has no CPU path)", who);
        return false;
    }
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&g_dev.ncu, hipDeviceAttributeMultiprocessorCount,
                          dev);
    if (hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking)
        != hipSuccess) {
        g_dev.failed = true;
        msg("%s: could not create a HIP stream", who);
        return false;
    }
    g_dev.ready = true;
    return true;
}

template <typename T>
bool dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **) p, n*sizeof(T)) == hipSuccess;
}

template <typename T>
void dfree(T *&p)
{
    if (p) hipFree((void *) p);
    p = nullptr;
}

bool h2d(void *d, const void *h, size_t bytes)
{
    if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_dev.stream)
        != hipSuccess) return false;
    return hipStreamSynchronize(g_dev.stream) == hipSuccess;
}

bool d2h(void *h, const void *d, size_t bytes)
{
    if (hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_dev.stream)
        != hipSuccess) return false;
    return hipStreamSynchronize(g_dev.stream) == hipSuccess;
}

bool sync_ok()
{
    return hipStreamSynchronize(g_dev.stream) == hipSuccess;
}

double vmin(const double *v, long n)
{
    double m = v[0];
    for (long i = 1; i < n; i++) m = (v[i] < m) ? v[i] : m;
    return m;
}

// ------------------------------------------------ source / receiver host
// source.F90:136-190 (the grid position of a source; no field arithmetic).
struct SrcInfo {
    double ztrue, xtrue, ytrue;
    double zsa, xsa, ysa;
    int zsi, xsi, ysi;
};

int locate_source(bool is3d, int nz, int nx, int ny, double dz, double dx,
                  double dy, double z0, double x0, double y0, double zs,
                  double xs, double ys, int isrc, SrcInfo *o)
{
    const double perturb = 0.0001;
    if (!is3d) { dy = 1.0; y0 = 0.0; ys = 0.0; ny = 1; }
    const double zmax = (double) (nz - 1)*dz, xmax = (double) (nx - 1)*dx,
                 ymax = (double) (ny - 1)*dy;
    const double zl = zs - z0, xl = xs - x0, yl = ys - y0;
    if (zl < 0.0 || zl > zmax || xl < 0.0 || xl > xmax || yl < 0.0 ||
        yl > ymax) {
        if (zl < 0.0 || zl > zmax)
            msg("fteik_source_initialize64f: src=%d zsrc=%12.5E must be in "
                "range [%12.5E %12.5E]", isrc, zs, z0, zmax + z0);
        if (xl < 0.0 || xl > xmax)
            msg("fteik_source_initialize64f: src=%d xsrc=%12.5E must be in "
                "range [%12.5E %12.5E]", isrc, xs, x0, xmax + x0);
        if (yl < 0.0 || yl > ymax)
            msg("fteik_source_initialize64f: src=%d ysrc=%12.5E must be in "
                "range [%12.5E %12.5E]", isrc, ys, y0, ymax + y0);
        return 1;
    }
    double zsa = (zl/dz) + 1.0, xsa = (xl/dx) + 1.0, ysa = (yl/dy) + 1.0;
    if (fabs(zsa - 1.0) < DBL_EPSILON) zsa = zsa + perturb;
    if (zsa >= (double) nz)            zsa = zsa - perturb;
    if (fabs(xsa - 1.0) < DBL_EPSILON) xsa = xsa + perturb;
    if (xsa >= (double) nx)            xsa = xsa - perturb;
    if (fabs(ysa - 1.0) < DBL_EPSILON) ysa = ysa + perturb;
    if (ysa >= (double) ny)            ysa = ysa - perturb;
    o->zsi = (int) zsa; o->xsi = (int) xsa; o->ysi = is3d ? (int) ysa : 1;
    o->zsa = zsa; o->xsa = xsa; o->ysa = is3d ? ysa : 1.0;
    o->ztrue = zs; o->xtrue = xs; o->ytrue = ys;
    return 0;
}

// receiver.F90:103-134.  Returns the number of out-of-bounds coordinates
// (the reference reports them in ierr and keeps the receivers).
int locate_receivers(bool is3d, int nz, int nx, int ny, double dz, double dx,
                     double dy, double z0, double x0, double y0, int nrec,
                     const double *z, const double *x, const double *y,
                     std::vector<int> &ri, std::vector<double> &rd)
{
    int ierr = 0;
    if (!is3d) { dy = 1.0; y0 = 0.0; ny = 1; }
    ri.assign(3*(size_t) nrec, 1);
    rd.assign(3*(size_t) nrec, 0.0);
    for (int r = 0; r < nrec; r++) {
        const double yy = is3d ? y[r] : y0;
        if (z[r] < z0 || z[r] > z0 + dz*(double) (nz - 1)) {
            msg("fteik_receiver_initialize64f: z position = %12.5E is out of "
                "bounds[%12.5E%12.5E]", z[r], z0, z0 + dz*(double) (nz - 1));
            ierr++;
        }
        if (x[r] < x0 || x[r] > x0 + dx*(double) (nx - 1)) {
            msg("fteik_receiver_initialize64f: x position = %12.5E is out of "
                "bounds[%12.5E%12.5E]", x[r], x0, x0 + dx*(double) (nx - 1));
            ierr++;
        }
        if (yy < y0 || yy > y0 + dy*(double) (ny - 1)) {
            msg("fteik_receiver_initialize64f: y position = %12.5E is out of "
                "bounds[%12.5E%12.5E]", yy, y0, y0 + dy*(double) (ny - 1));
            ierr++;
        }
        const int iz = (int) ((z[r] - z0)/dz + 1.0);
        const int ix = (int) ((x[r] - x0)/dx + 1.0);
        const int iy = is3d ? (int) ((yy - y0)/dy + 1.0) : 1;
        const int zi = std::max(1, std::min(nz - 1, iz));
        const int xi = std::max(1, std::min(nx - 1, ix));
        const int yi = std::max(1, std::min(ny - 1, iy));
        ri[3*r] = zi; ri[3*r + 1] = xi; ri[3*r + 2] = yi;
        rd[3*r] = (z[r] - (z0 + dz*(double) (zi - 1)));
        rd[3*r + 1] = (x[r] - (x0 + dx*(double) (xi - 1)));
        rd[3*r + 2] = is3d ? (yy - (y0 + dy*(double) (yi - 1))) : 0.0;
    }
    return ierr;
}

// Velocity model upload + conversion shared by the 3D and 2D solvers
// (model.F90:311-607).  kind: 0 natural cells, 1 ordered cells, 2 nodes.
bool set_model(const char *who, bool is3d, int nz, int nx, int ny, int kind,
               int n, int order, const double *vel, double *d_slow)
{
    const long ncell = (long) (nz - 1)*(nx - 1)*(is3d ? ny - 1 : 1);
    const long ngrd = (long) nz*nx*(is3d ? ny : 1);
    if (kind == 0 && n != ncell) {
        msg("%s: ERROR - ncell /= nv %ld %d", who, ncell, n);
        return false;
    }
    if (kind == 1 && n < ncell) {
        msg("%s: nc should be at least %ld", who, ncell);
        return false;
    }
    if (kind == 2 && n < ngrd) {
        msg("%s: ng should be at least %ld", who, ngrd);
        return false;
    }
    if (n < 1) { msg("%s: no points in model %d", who, n); return false; }
    if (vmin(vel, n) <= 0.0) {
        msg("%s: All velocities must be positive", who);
        return false;
    }
    double *d_vel = nullptr;
    if (!dalloc(&d_vel, (size_t) n)) { msg("%s: out of GPU memory", who); return false; }
    bool ok = h2d(d_vel, vel, (size_t) n*sizeof(double));
    if (ok) {
        if (kind == 2)
            ok = fteik_k_slowness_from_nodes(d_slow, d_vel, nz, nx, ny,
                                             is3d ? 1 : 0, order,
                                             g_dev.stream) == 0;
        else
            ok = fteik_k_slowness_from_cells(d_slow, d_vel, nz, nx, ny,
                                             is3d ? 1 : 0,
                                             kind == 0 ? 0 : order,
                                             g_dev.stream) == 0;
        ok = ok && sync_ok();
    }
    dfree(d_vel);
    if (!ok) msg("%s: GPU error while setting the model", who);
    return ok;
}

// ================================================================== 3D
struct Solver3D {
    bool init = false, haveModel = false, haveTimes = false;
    int nz = 0, nx = 0, ny = 0, nsweep = 0, verbose = 0, srcNumber = -1;
    double z0 = 0, x0 = 0, y0 = 0, dz = 0, dx = 0, dy = 0, eps = 0,
           convTol = 0;
    int precision = 0, schedule = 0, cooperative = 1;
    int tile = 0, tile32 = 0;            // pipelined-sweep tile shapes
    std::vector<char> tileOK;            // tile usable on this grid
    std::vector<char> tileCoRes;         // whole grid of tiles co-resident
    double mixedBias = 1.0/64.0;         // 2^-6 relative promotion margin
    int maxCorr = 64;
    std::vector<SrcInfo> src;
    std::vector<int> ri;
    std::vector<double> rd;
    int nrec = 0;
    bool haveRec = false;
    double *d_slow = nullptr, *d_tt = nullptr, *d_scr = nullptr;
    float *d_slow32 = nullptr, *d_tt32 = nullptr;
    void *d_cst = nullptr, *d_cst32 = nullptr;
    unsigned long long *d_flags = nullptr;
    unsigned long long *d_mb = nullptr;  // tagged mailbox, 2 words per node
    // edge-major mailbox of single sweeps: x and y tile edges (at most
    // (n + 7)/8 + 1 tiles per axis), mbnt = nz + nx + ny steps each
    unsigned long long *d_mbx = nullptr, *d_mby = nullptr;
    int mbnt = 0;
    unsigned long long *d_colmod = nullptr, *d_lastproc = nullptr;
    int *d_plan = nullptr;                     // tile skipping state
    bool skipping = true;
    unsigned long long *d_prog = nullptr;      // sweep-chain progress words
    int chainTile = -1, chainTile32 = -1;      // mode-2 tiles for chains
    bool overlap = false;                      // chain iteration sweeps
    bool overlapForced = false;                // ... on every grid size
    int ncu_chain = 256;
    int *d_err = nullptr;
    double *d_maxdelta = nullptr;
    int *d_ri = nullptr;
    double *d_rd = nullptr, *d_ttr = nullptr;
    int max_flags = 0;
    unsigned long long serial = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, last_updates = 0.0;
    int last_corr = 0;
    bool cstReady = false;
};
Solver3D g3;

void free3d()
{
    dfree(g3.d_slow); dfree(g3.d_tt); dfree(g3.d_scr);
    dfree(g3.d_slow32); dfree(g3.d_tt32);
    dfree(g3.d_cst); dfree(g3.d_cst32);
    dfree(g3.d_flags); dfree(g3.d_err); dfree(g3.d_maxdelta);
    dfree(g3.d_mb); dfree(g3.d_mbx); dfree(g3.d_mby);
    dfree(g3.d_colmod); dfree(g3.d_lastproc); dfree(g3.d_plan);
    dfree(g3.d_prog);
    dfree(g3.d_ri); dfree(g3.d_rd); dfree(g3.d_ttr);
    if (g3.ev0) hipEventDestroy(g3.ev0);
    if (g3.ev1) hipEventDestroy(g3.ev1);
    g3 = Solver3D();
}

bool check_eps(const char *who, double eps, int nz, int nx, int ny, bool is3d)
{
    const int e = (int) eps;
    bool bad = e > nz || e > nx || (is3d && e > ny);
    if (bad) {
        if (e > nz) msg("%s: eps=%d bigger than nz =%d", who, e, nz);
        if (e > nx) msg("%s: eps=%d bigger than nx =%d", who, e, nx);
        if (is3d && e > ny) msg("%s: eps=%d bigger than ny =%d", who, e, ny);
    }
    return !bad;
}

SweepLaunch launch3d(bool fp32)
{
    SweepLaunch L;
    memset(&L, 0, sizeof(L));
    L.tt = fp32 ? (void *) g3.d_tt32 : (void *) g3.d_tt;
    L.slow = fp32 ? (const void *) g3.d_slow32 : (const void *) g3.d_slow;
    L.cst = fp32 ? g3.d_cst32 : g3.d_cst;
    L.flags = g3.d_flags;
    L.err = g3.d_err;
    L.maxdelta = nullptr;
    L.max_flags = g3.max_flags;
    L.nz = g3.nz; L.nx = g3.nx; L.ny = g3.ny;
    L.fp32 = fp32 ? 1 : 0;
    L.schedule = g3.schedule;
    // a cooperative launch when every tile is resident at once; mailbox
    // tiles also run without (a tile waits only on lower-numbered ones,
    // and dispatch is in order)
    {
        const int t = fp32 ? g3.tile32 : g3.tile;
        L.cooperative = (t >= 0 && t < (int) g3.tileCoRes.size() &&
                         g3.tileCoRes[(size_t) t]) ? 1 : 0;
    }
    L.tile = fp32 ? g3.tile32 : g3.tile;
    L.mb = g3.d_mb;
    // the edge-major mailbox, except for the 16-wide tiles without a store
    // wave (7, 8), which are faster with the per-node one (257^3 FP64 sweep
    // with 16x8 tiles: 1.14 ms per node, 1.21 ms per edge; 8x8 tiles: 1.11
    // and 1.03 ms)
    const bool edge = !(L.tile == 7 || L.tile == 8);
    L.mbx = edge ? g3.d_mbx : nullptr;
    L.mby = edge ? g3.d_mby : nullptr;
    L.mbnt = g3.mbnt;
    L.colmod = nullptr; L.lastproc = nullptr; L.plan = nullptr;
    L.prog = g3.d_prog;
    return L;
}

// Every tile dirty: columns changed at serial 1, never processed.
bool reset_skipping(SweepLaunch &L)
{
    if (!g3.skipping) return true;
    L.colmod = g3.d_colmod; L.lastproc = g3.d_lastproc; L.plan = g3.d_plan;
    return fteik_k_fill_u64(g3.d_colmod, (long) g3.nx*g3.ny, 1ull,
                            g_dev.stream) == 0 &&
           fteik_k_fill_u64(g3.d_lastproc, 8L*g3.max_flags, 0ull,
                            g_dev.stream) == 0;
}

// nsweep iterations of the 8 sweep directions.  With sweep chains, every
// iteration sweep of the range runs in one launch, each sweep overlapping
// the one before it; otherwise one launch per sweep.
bool run_chain(SweepLaunch L, int niter)
{
    const int tile = L.fp32 ? g3.chainTile32 : g3.chainTile;
    L.tile = tile;
    L.serial = g3.serial + 1;
    g3.serial += 8*(unsigned long long) niter;
    L.colmod = nullptr; L.lastproc = nullptr; L.plan = nullptr;
    return fteik_k_sweep3d_chain(&L, 0, 8*niter, g_dev.stream) == 0;
}

// Chains pay where one sweep leaves CUs idle (a sweep's tiles are active
// only while the wavefront crosses them): at most one tile per CU.  On
// larger grids every SIMD already holds a sweep's compute wave.  A chain
// cannot skip tiles, so once a solve converges early tile skipping is
// faster: chains are opt-in.
bool chain_ok(bool fp32)
{
    const int tile = fp32 ? g3.chainTile32 : g3.chainTile;
    if (!g3.overlap || g3.schedule != 0 || tile < 0) return false;
    int bx = 0, by = 0, hb = 0, mode = 0;
    fteik_k_tile_shape(tile, &bx, &by, &hb, &mode);
    const long ntile = (long) ((g3.nx - 1 + bx - 1)/bx)*((g3.ny - 1 + by - 1)/by);
    return g3.overlapForced || ntile <= g3.ncu_chain;
}

bool run_sweep(SweepLaunch &L, int sweep, bool linit)
{
    L.serial = ++g3.serial;
    int rc = fteik_k_sweep3d(&L, sweep, linit ? 1 : 0, g_dev.stream);
    if (rc != 0) {
        msg("fteik_solver3d_solveSourceLSM: sweep %d launch failed (%d)",
            sweep, rc);
        return false;
    }
    return true;
}

// Number of node updates of one sweep (for throughput reporting).
double sweep_updates(int sweep, bool linit, const SrcInfo &s)
{
    static const int tz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
    static const int tx[8] = {1, -1, 1, -1, 1, -1, 1, -1};
    static const int ty[8] = {1, 1, -1, -1, 1, 1, -1, -1};
    auto len = [&](int sg, int n, int si) -> double {
        int a0, a1;
        if (sg > 0) { a0 = linit ? std::max(2, si) : 2; a1 = n; }
        else        { a0 = 1; a1 = linit ? si + 1 : n - 1; }
        return (double) std::max(0, a1 - a0 + 1);
    };
    return len(tz[sweep - 1], g3.nz, s.zsi)*len(tx[sweep - 1], g3.nx, s.xsi)
         * len(ty[sweep - 1], g3.ny, s.ysi);
}

bool prepare_source3d(int isrc, const char *who)
{
    const SrcInfo &s = g3.src[isrc - 1];
    SourceGeom g;
    g.dz = g3.dz; g.dx = g3.dx; g.dy = g3.dy; g.eps = g3.eps;
    g.zsa = s.zsa; g.xsa = s.xsa; g.ysa = s.ysa;
    g.zsi = s.zsi; g.xsi = s.xsi; g.ysi = s.ysi;
    g.nz = g3.nz; g.nx = g3.nx; g.ny = g3.ny;
    const long ngrd = (long) g3.nz*g3.nx*g3.ny;
    if (fteik_k_source3d_init(g3.d_cst, g3.d_tt, g3.d_slow, &g, ngrd, 0,
                              g_dev.stream) != 0) {
        msg("%s: Error setting mesh constants", who);
        return false;
    }
    g3.cstReady = true;
    return true;
}

// ================================================================== 2D
struct Solver2D {
    bool init = false, haveModel = false, haveTimes = false;
    int nz = 0, nx = 0, nsweep = 1, verbose = 0, srcNumber = -1;
    double z0 = 0, x0 = 0, dz = 0, dx = 0, eps = 0, convTol = 0;
    std::vector<SrcInfo> src;
    std::vector<int> ri;
    std::vector<double> rd;
    int nrec = 0;
    std::vector<double> ttimesRec;   // [nrec x nsrc]
    double *d_slow = nullptr, *d_tt = nullptr, *d_scr = nullptr;
    double *d_ttb = nullptr;         // batch fields for solveLSM
    long ttb_cap = 0;
    int *d_ri = nullptr;
    double *d_rd = nullptr, *d_ttr = nullptr;
    double *d_sa = nullptr;          // zsa, xsa per source
    int *d_si = nullptr;             // zsi, xsi per source
};
Solver2D g2;

void free2d()
{
    dfree(g2.d_slow); dfree(g2.d_tt); dfree(g2.d_scr); dfree(g2.d_ttb);
    dfree(g2.d_ri); dfree(g2.d_rd); dfree(g2.d_ttr);
    dfree(g2.d_sa); dfree(g2.d_si);
    g2 = Solver2D();
}

// ============================================================== Locator
struct Locator {
    bool init = false;
    int nEvents = 0, ntf = 0, ngrd = 0, lntf = 0;
    long ldgrd = 0;
    int refCompat = 1;
    std::vector<double> tepoch;
    std::vector<float> obs, wt;
    std::vector<unsigned char> skip;
    float *d_tab = nullptr, *d_t0 = nullptr, *d_obs = nullptr,
          *d_wt = nullptr, *d_xn = nullptr;
    unsigned char *d_skip = nullptr;
    unsigned long long *d_best = nullptr;
    double *d_tmp = nullptr;
};
Locator gl;

void free_locator()
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

// ============================================================= Analytic
struct Analytic {
    bool init = false;
    int nz = 0, nx = 0, ny = 0, verbose = 0;
    double dz = 0, dx = 0, dy = 0, z0 = 0, x0 = 0, y0 = 0;
    double vconst = 5000.0, vtop = 5000.0, vbot = 5000.0;
    std::vector<double> zs, xs, ys;   // true source positions
    std::vector<double> zr, xr, yr;   // receivers
    double *d_tt = nullptr, *d_scr = nullptr, *d_pts = nullptr;
    bool haveField = false;
};
Analytic ga;

void free_analytic()
{
    dfree(ga.d_tt); dfree(ga.d_scr); dfree(ga.d_pts);
    ga = Analytic();
}

// =============================================================== Graph
struct Graph {
    bool init = false, made = false;
    int nz = 0, nx = 0, ny = 0, verbose = 0, nLevels = 0, maxLevelSize = 0;
    std::vector<int> levelPtr;
};
Graph gg;

} // namespace

// ====================================================================
//                         3D solver entry points
// ====================================================================
extern "C" {

void fteik_solver3d_free(void)
{
    free3d();
}

void fteik_solver3d_setVerbosity(int verbose) { g3.verbose = verbose; }

void fteik_solver3d_setConvergenceTolerance(double convTol)
{
    g3.convTol = convTol;
}

void fteik_solver3d_setNumberOfSweeps(int nsweep, int *ierr)
{
    *ierr = 0;
    g3.nsweep = 0;
    if (nsweep < 0) {
        msg("fteik_solver3d_setNumberOfSweeps: nsweep = %d must be positive",
            nsweep);
        *ierr = 1;
        return;
    }
    g3.nsweep = nsweep;
}

void fteik_solver3d_setSphereToCartEpsilon(double eps, int *ierr)
{
    *ierr = 0;
    g3.eps = 0.0;
    if (g3.nz < 1 || g3.nx < 1 || g3.ny < 1) {
        msg("fteik_solver3d_setSphereToCartEpsilon: Grid not yet set");
        *ierr = 1;
        return;
    }
    if (!check_eps("fteik_solver3d_setSphereToCartEpsilon", eps, g3.nz,
                   g3.nx, g3.ny, true)) {
        *ierr = 1;
        return;
    }
    g3.eps = eps;
}

void fteik_solver3d_initialize(int nz, int nx, int ny, double z0, double x0,
                               double y0, double dz, double dx, double dy,
                               int nsweep, double eps, double convTol,
                               int verbose, int *ierr)
{
    *ierr = 0;
    if (verbose > 0) printf("fteik_solver3d_initialize: Initializing...\n");
    free3d();
    g3.verbose = verbose;
    g3.convTol = convTol;
    if (nz < 3 || nx < 3 || ny < 3) {
        if (nz < 3) msg("fteik_model_setGridSize: ERROR nz is too small%d", nz);
        if (nx < 3) msg("fteik_model_setGridSize: ERROR nx is too small%d", nx);
        if (ny < 3) msg("fteik_model_setGridSize: ERROR ny is too small%d", ny);
        msg("fteik_solver3d_initialize: Error setting grid geometry");
        *ierr = 1;
        return;
    }
    if (dz <= 0.0 || dx <= 0.0 || dy <= 0.0) {
        msg("fteik_model_setGridSpacing: grid spacing must be positive");
        msg("fteik_solver3d_initialize: Error setting grid geometry");
        *ierr = 1;
        return;
    }
    g3.nz = nz; g3.nx = nx; g3.ny = ny;
    g3.z0 = z0; g3.x0 = x0; g3.y0 = y0;
    g3.dz = dz; g3.dx = dx; g3.dy = dy;
    fteik_solver3d_setNumberOfSweeps(nsweep, ierr);
    if (*ierr) {
        msg("fteik_solver3d_initialize: Failed to set max number of sweeps");
        return;
    }
    fteik_solver3d_setSphereToCartEpsilon(eps, ierr);
    if (*ierr) {
        msg("fteik_solver3d_initialize: Failed to set epsilon");
        return;
    }
    if (!device_ok("fteik_solver3d_initialize")) { *ierr = 1; return; }
    const long ngrd = (long) nz*nx*ny;
    const long ncell = (long) (nz - 1)*(nx - 1)*(ny - 1);
    g3.max_flags = fteik_k_max_flags(nx, ny);
    g3.mbnt = nz + nx + ny;
    const size_t mbxw = 2*(size_t) ((nx + 7)/8 + 1)*g3.mbnt*ny;
    const size_t mbyw = 2*(size_t) ((ny + 7)/8 + 1)*g3.mbnt*nx;
    bool ok = dalloc(&g3.d_slow, ncell) && dalloc(&g3.d_tt, ngrd) &&
              dalloc(&g3.d_scr, ngrd) &&
              dalloc((char **) &g3.d_cst, 512) &&
              dalloc((char **) &g3.d_cst32, 512) &&
              dalloc(&g3.d_flags, g3.max_flags) && dalloc(&g3.d_err, 1) &&
              dalloc(&g3.d_mb, 2*(size_t) ngrd) &&
              dalloc(&g3.d_mbx, mbxw) && dalloc(&g3.d_mby, mbyw) &&
              dalloc(&g3.d_colmod, (size_t) nx*ny) &&
              dalloc(&g3.d_lastproc, 8*(size_t) g3.max_flags) &&
              dalloc(&g3.d_plan, 1 + 3*(size_t) g3.max_flags) &&
              dalloc(&g3.d_prog, 2*(size_t) g3.max_flags) &&
              dalloc(&g3.d_maxdelta, 1);
    if (!ok) {
        msg("fteik_solver3d_initialize: out of GPU memory");
        free3d();
        *ierr = 1;
        return;
    }
    hipMemsetAsync(g3.d_flags, 0, g3.max_flags*sizeof(unsigned long long),
                   g_dev.stream);
    hipMemsetAsync(g3.d_err, 0, sizeof(int), g_dev.stream);
    hipMemsetAsync(g3.d_mb, 0, 2*(size_t) ngrd*sizeof(unsigned long long),
                   g_dev.stream);
    hipMemsetAsync(g3.d_mbx, 0, mbxw*sizeof(unsigned long long),
                   g_dev.stream);
    hipMemsetAsync(g3.d_mby, 0, mbyw*sizeof(unsigned long long),
                   g_dev.stream);
    hipMemsetAsync(g3.d_prog, 0, 2*(size_t) g3.max_flags*
                   sizeof(unsigned long long), g_dev.stream);
    // plan: count, order, skip flags and the dirty marks (all clear)
    hipMemsetAsync(g3.d_plan, 0, (1 + 3*(size_t) g3.max_flags)*sizeof(int),
                   g_dev.stream);
    hipEventCreate(&g3.ev0);
    hipEventCreate(&g3.ev1);
    // tile shapes whose whole grid of workgroups can be co-resident (FP64
    // and FP32 kernels); the first such one is the default
    const int ntile = fteik_k_num_tiles();
    g3.tileOK.assign((size_t) ntile, 0);
    g3.tileCoRes.assign((size_t) ntile, 0);
    for (int t = 0; t < ntile; t++) {
        g3.tileCoRes[(size_t) t] =
            (char) (fteik_k_pencil_fits(nx, ny, 0, t) &&
                    fteik_k_pencil_fits(nx, ny, 1, t));
        g3.tileOK[(size_t) t] = (char) (g3.tileCoRes[(size_t) t] ||
                                        fteik_k_tile_mode(t) >= 1);
    }
    // default: the first available shape in order of measured speed (FP64
    // sweep of a 257^3 grid: 8x8 tiles with a store wave 0.92 ms, launched
    // without co-residency; 8x8 without a store wave 1.03 ms; 16x8 with a
    // store wave 1.10 ms, without 1.14 ms); on smaller grids the 8x8 tiles
    // keep more CUs busy as well.
    static const int pref_large[] = {12, 9, 11, 8, 7, 10, 6, 3, 1, 0, 5, 2, 4};
    static const int pref_small[] = {12, 9, 11, 8, 10, 7, 6, 3, 1, 0, 5, 2, 4};
    const bool large = (long) (nx - 1)*(ny - 1) >= 128L*g_dev.ncu;
    int first = -1;
    for (int t : (large ? pref_large : pref_small))
        if (t < ntile && g3.tileOK[(size_t) t]) { first = t; break; }
    g3.cooperative = first >= 0;
    g3.tile = g3.tile32 = first >= 0 ? first : 0;
    // sweep chains: the first mode-2 tile with two sweeps' tiles resident
    g3.ncu_chain = g_dev.ncu > 0 ? g_dev.ncu : 256;
    g3.overlapForced = false;
    static const int chain_pref[] = {13, 11, 10, 12};
    g3.chainTile = g3.chainTile32 = -1;
    for (int t : chain_pref) {
        if (g3.chainTile < 0 && fteik_k_chain_fits(nx, ny, 0, t))
            g3.chainTile = t;
        if (g3.chainTile32 < 0 && fteik_k_chain_fits(nx, ny, 1, t))
            g3.chainTile32 = t;
    }
    if (!g3.cooperative) g3.schedule = 1;
    if (!sync_ok()) { msg("fteik_solver3d_initialize: GPU error"); *ierr = 1;
                      return; }
    g3.init = true;
}

void fteik_solver3d_setVelocityModel64f(int ncell, const double *vel,
                                        int *ierr)
{
    *ierr = 0;
    g3.haveModel = false;
    if (!g3.init) {
        msg("fteik_solver3d_setVelocityModel64f: solver not initialized");
        *ierr = 1;
        return;
    }
    if (!set_model("fteik_model_setVelocityModel64f", true, g3.nz, g3.nx,
                   g3.ny, 0, ncell, 0, vel, g3.d_slow)) {
        msg("fteik_solver3d_setVelocityModel64f: Error setting velocity model");
        *ierr = 1;
        return;
    }
    g3.haveModel = true;
}

void fteik_solver3d_setCellVelocityModel64f(int nc, int order,
                                            const double *vel, int *ierr)
{
    *ierr = 0;
    g3.haveModel = false;
    if (!g3.init) {
        msg("fteik_solver3d_setCellVelocityModel64f: solver not initialized");
        *ierr = 1;
        return;
    }
    if (order != 1 && order != 2 && order != 0 && g3.verbose > 0)
        printf("fteik_model_setCellVelocityModel64f: Defaulting ot zxy\n");
    if (!set_model("fteik_model_setCellVelocityModel64f", true, g3.nz, g3.nx,
                   g3.ny, 1, nc, (order == 1 || order == 2) ? order : 0, vel,
                   g3.d_slow)) {
        msg("fteik_solver3d_setCellVelocityModel64f: Error setting velocity "
            "model");
        *ierr = 1;
        return;
    }
    g3.haveModel = true;
}

void fteik_solver3d_setNodalVelocityModel64f(int ng, int order,
                                             const double *vel, int *ierr)
{
    *ierr = 0;
    g3.haveModel = false;
    if (!g3.init) {
        msg("fteik_solver3d_setNodalVelocityModel64f: solver not initialized");
        *ierr = 1;
        return;
    }
    if (!set_model("fteik_model_setNodalVelocityModel64f", true, g3.nz,
                   g3.nx, g3.ny, 2, ng, (order == 1 || order == 2) ? order : 0,
                   vel, g3.d_slow)) {
        msg("fteik_solver3d_setNodalVelocityModel64f: Error setting velocity "
            "model");
        *ierr = 1;
        return;
    }
    g3.haveModel = true;
}

void fteik_solver3d_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, const double *ysrc,
                                  int *ierr)
{
    *ierr = 0;
    g3.src.clear();
    if (nsrc < 1) {
        msg("fteik_source_initialize64fF: No sources");
        msg("fteik_solver3d_setSources64f: Failed to set source");
        *ierr = 1;
        return;
    }
    if (g3.nz < 1 || g3.dz <= 0.0) {
        msg("fteik_source_initialize64fF: Model not yet set");
        msg("fteik_solver3d_setSources64f: Failed to set source");
        *ierr = 1;
        return;
    }
    std::vector<SrcInfo> v((size_t) nsrc);
    for (int i = 0; i < nsrc; i++) {
        if (locate_source(true, g3.nz, g3.nx, g3.ny, g3.dz, g3.dx, g3.dy,
                          g3.z0, g3.x0, g3.y0, zsrc[i], xsrc[i], ysrc[i],
                          i + 1, &v[(size_t) i])) {
            msg("fteik_solver3d_setSources64f: Failed to set source");
            *ierr = 1;
            return;
        }
    }
    g3.src.swap(v);
}

void fteik_solver3d_getNumberOfSources(int *nsrc, int *ierr)
{
    *ierr = 0;
    *nsrc = 0;
    if (g3.src.empty()) {
        msg("fteik_source_getNumberOfSources: Never initialized");
        msg("fteik_solver3d_getNumberOfSources: No sources initialized!");
        *ierr = 1;
        return;
    }
    *nsrc = (int) g3.src.size();
}

void fteik_solver3d_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, const double *yrec,
                                    int *ierr)
{
    *ierr = 0;
    if (nrec < 1) {
        printf("fteik_solver3d_setReceivers64f: No receivers to set\n");
        return;
    }
    if (!g3.init) {
        msg("fteik_receiver_initialize64f: Grid not yet set");
        msg("fteik_solver3d_setReceivers64f: Failed to set receivers");
        *ierr = 1;
        return;
    }
    int e = locate_receivers(true, g3.nz, g3.nx, g3.ny, g3.dz, g3.dx, g3.dy,
                             g3.z0, g3.x0, g3.y0, nrec, zrec, xrec, yrec,
                             g3.ri, g3.rd);
    dfree(g3.d_ri); dfree(g3.d_rd); dfree(g3.d_ttr);
    if (!dalloc(&g3.d_ri, 3*(size_t) nrec) ||
        !dalloc(&g3.d_rd, 3*(size_t) nrec) || !dalloc(&g3.d_ttr, nrec) ||
        !h2d(g3.d_ri, g3.ri.data(), 3*(size_t) nrec*sizeof(int)) ||
        !h2d(g3.d_rd, g3.rd.data(), 3*(size_t) nrec*sizeof(double))) {
        msg("fteik_solver3d_setReceivers64f: GPU error");
        *ierr = 1;
        return;
    }
    g3.nrec = nrec;
    g3.haveRec = true;
    if (e) {
        msg("fteik_solver3d_setReceivers64f: Failed to set receivers");
        *ierr = e;
    }
}

void fteik_solver3d_getNumberOfReceivers(int *nrec, int *ierr)
{
    *ierr = 0;
    *nrec = 0;
    if (!g3.haveRec) {
        msg("fteik_receiver_getNumberOfReceivers: Module not initialized");
        *ierr = 1;
        return;
    }
    *nrec = g3.nrec;
}

void fteik_solver3d_setPrecision(int mode, int *ierr)
{
    *ierr = 0;
    if (mode != 0 && mode != 1) {
        msg("fteik_solver3d_setPrecision: mode must be 0 (FP64) or 1 (mixed)");
        *ierr = 1;
        return;
    }
    g3.precision = mode;
}

void fteik_solver3d_setSchedule(int schedule, int *ierr)
{
    *ierr = 0;
    if (schedule != 0 && schedule != 1) {
        msg("fteik_solver3d_setSchedule: schedule must be 0 or 1");
        *ierr = 1;
        return;
    }
    if (schedule == 0 && g3.init && !g3.cooperative) {
        msg("fteik_solver3d_setSchedule: column blocks of this grid cannot be "
            "co-resident; keeping the per-level schedule");
        *ierr = 1;
        return;
    }
    g3.schedule = schedule;
}

void fteik_solver3d_setMixedParameters(double bias, int maxCorrections,
                                       int *ierr)
{
    *ierr = 0;
    if (!(bias >= 0.0 && bias < 0.1) || maxCorrections < 1 ||
        maxCorrections > 64) {
        msg("fteik_solver3d_setMixedParameters: need 0 <= bias < 0.1 and "
            "1 <= maxCorrections <= 64");
        *ierr = 1;
        return;
    }
    g3.mixedBias = bias;
    g3.maxCorr = maxCorrections;
}

void fteik_solver3d_setTile(int tile, int tile32, int *ierr)
{
    *ierr = 0;
    const int n = (int) g3.tileOK.size();
    if (!g3.init || tile < 0 || tile >= n || tile32 < 0 || tile32 >= n ||
        !g3.tileOK[(size_t) tile] || !g3.tileOK[(size_t) tile32]) {
        msg("fteik_solver3d_setTile: tile shapes (%d, %d) are not available "
            "for this grid", tile, tile32);
        *ierr = 1;
        return;
    }
    g3.tile = tile;
    g3.tile32 = tile32;
}

void fteik_solver3d_getTile(int *tile, int *tile32, int *ntiles, int *ierr)
{
    *ierr = g3.init ? 0 : 1;
    *tile = g3.tile;
    *tile32 = g3.tile32;
    *ntiles = (int) g3.tileOK.size();
}

void fteik_solver3d_setTileSkipping(int on, int *ierr)
{
    *ierr = 0;
    g3.skipping = on != 0;
}

void fteik_solver3d_setSweepOverlap(int on, int *ierr)
{
    *ierr = 0;
    if (on < 0 || on > 2) {
        msg("fteik_solver3d_setSweepOverlap: on must be 0, 1 or 2");
        *ierr = 1;
        return;
    }
    g3.overlap = on != 0;
    g3.overlapForced = on == 2;
}

void fteik_solver3d_setChainTile(int tile, int tile32, int *ierr)
{
    *ierr = 0;
    if (!g3.init || !fteik_k_chain_fits(g3.nx, g3.ny, 0, tile) ||
        !fteik_k_chain_fits(g3.nx, g3.ny, 1, tile32)) {
        msg("fteik_solver3d_setChainTile: tile shapes (%d, %d) cannot chain "
            "sweeps on this grid", tile, tile32);
        *ierr = 1;
        return;
    }
    g3.chainTile = tile;
    g3.chainTile32 = tile32;
}

void fteik_solver3d_getSweepOverlap(int *on, int *tile, int *tile32,
                                    int *ierr)
{
    *ierr = g3.init ? 0 : 1;
    *on = chain_ok(false) ? 1 : 0;
    *tile = g3.chainTile;
    *tile32 = g3.chainTile32;
}

void fteik_solver3d_getLastSolveInfo(double *ms, int *corr, double *updates,
                                     int *ierr)
{
    *ierr = g3.haveTimes ? 0 : 1;
    *ms = g3.last_ms;
    *corr = g3.last_corr;
    *updates = g3.last_updates;
}

void fteik_amd_prepareSource3d(int isrc, int *ierr)
{
    *ierr = 0;
    if (!g3.init || !g3.haveModel || isrc < 1 ||
        isrc > (int) g3.src.size()) {
        msg("fteik_amd_prepareSource3d: solver, model or source not set");
        *ierr = 1;
        return;
    }
    if (!prepare_source3d(isrc, "fteik_amd_prepareSource3d") || !sync_ok())
        *ierr = 1;
}

void fteik_solver3d_solveSourceLSM(int isrc, int *ierr)
{
    static const char *who = "fteik_solver3d_solveSourceLSM";
    *ierr = 0;
    g3.haveTimes = false;
    g3.srcNumber = -1;
    if (!g3.init || !g3.haveModel) {
        msg("%s: Model not yet set", who);
        *ierr = 1;
        return;
    }
    const int nsrc = (int) g3.src.size();
    if (isrc < 1 || isrc > nsrc) {
        msg("%s: Source number = %d must be in range [1,%d]", who, isrc, nsrc);
        *ierr = 1;
        return;
    }
    const SrcInfo &s = g3.src[isrc - 1];
    if (s.zsi > g3.nz - 2 || s.xsi > g3.nx - 2 || s.ysi > g3.ny - 2) {
        // solver3D.F90:513-558 with zsi = nz-1 puts node nz in a
        // negative-direction box, whose upwind neighbour lies outside the
        // grid (the reference reads past the array there).
        msg("%s: source %d lies in the last cell along an axis "
            "(zsi,xsi,ysi)=(%d,%d,%d); the reference reads outside the grid "
            "for such sources", who, isrc, s.zsi, s.xsi, s.ysi);
        *ierr = 1;
        return;
    }
    const long ngrd = (long) g3.nz*g3.nx*g3.ny;
    hipEventRecord(g3.ev0, g_dev.stream);
    if (!prepare_source3d(isrc, who)) { *ierr = 1; return; }
    hipMemsetAsync(g3.d_err, 0, sizeof(int), g_dev.stream);
    double updates = 0.0;
    SweepLaunch L = launch3d(false);
    L.zsi = s.zsi; L.xsi = s.xsi; L.ysi = s.ysi;
    bool ok = true;
    for (int sw = 1; sw <= 8 && ok; sw++) {
        ok = run_sweep(L, sw, true);
        updates += sweep_updates(sw, true, s);
    }
    int corr = 0;
    if (g3.precision == 0 && chain_ok(false)) {
        if (ok && g3.nsweep > 0) ok = run_chain(L, g3.nsweep);
        for (int k = 0; k < g3.nsweep; k++)
            for (int sw = 1; sw <= 8; sw++)
                updates += sweep_updates(sw, false, s);
    } else if (g3.precision == 0) {
        ok = ok && reset_skipping(L);
        for (int k = 0; k < g3.nsweep && ok; k++)
            for (int sw = 1; sw <= 8 && ok; sw++) {
                ok = run_sweep(L, sw, false);
                updates += sweep_updates(sw, false, s);
            }
    } else if (ok) {
        // mixed precision: FP32 Gauss-Seidel iterations from the FP64
        // initialization, then FP64 correction iterations to convergence.
        const long ncell = (long) (g3.nz - 1)*(g3.nx - 1)*(g3.ny - 1);
        if (!g3.d_tt32) {
            ok = dalloc(&g3.d_tt32, ngrd) && dalloc(&g3.d_slow32, ncell);
            if (!ok) msg("%s: out of GPU memory (mixed precision)", who);
        }
        if (ok) {
            ok = fteik_k_d2f(g3.d_slow32, g3.d_slow, ncell, g_dev.stream) == 0
              && fteik_k_d2f(g3.d_tt32, g3.d_tt, ngrd, g_dev.stream) == 0
              && fteik_k_consts_to_float(g3.d_cst, g3.d_cst32,
                                         g_dev.stream) == 0;
        }
        SweepLaunch L32 = launch3d(true);
        L32.zsi = s.zsi; L32.xsi = s.xsi; L32.ysi = s.ysi;
        if (chain_ok(true)) {
            if (ok && g3.nsweep > 0) ok = run_chain(L32, g3.nsweep);
            for (int k = 0; k < g3.nsweep; k++)
                for (int sw = 1; sw <= 8; sw++)
                    updates += sweep_updates(sw, false, s);
        } else {
            // FP32 iterations skip unchanged tiles as the FP64 ones do
            ok = ok && reset_skipping(L32);
            for (int k = 0; k < g3.nsweep && ok; k++)
                for (int sw = 1; sw <= 8 && ok; sw++) {
                    ok = run_sweep(L32, sw, false);
                    updates += sweep_updates(sw, false, s);
                }
        }
        // merge: the FP64 initialization values are upper bounds of the
        // converged field, so keep the smaller of the two
        if (ok) ok = fteik_k_min_merge_f(g3.d_tt, g3.d_tt32, ngrd,
                                         g3.mixedBias, g_dev.stream) == 0;
        L.maxdelta = g3.d_maxdelta;
        ok = ok && reset_skipping(L);
        const int maxCorr = g3.maxCorr;
        const bool chain = chain_ok(false);
        for (corr = 0; corr < maxCorr && ok; ) {
            hipMemsetAsync(g3.d_maxdelta, 0, sizeof(double), g_dev.stream);
            if (chain) {
                ok = run_chain(L, 1);
                for (int sw = 1; sw <= 8; sw++)
                    updates += sweep_updates(sw, false, s);
            } else {
                for (int sw = 1; sw <= 8 && ok; sw++) {
                    ok = run_sweep(L, sw, false);
                    updates += sweep_updates(sw, false, s);
                }
            }
            corr++;
            double md = 0.0;
            if (ok) ok = d2h(&md, g3.d_maxdelta, sizeof(double));
            if (!ok || md <= g3.convTol) break;
        }
        L.maxdelta = nullptr;
    }
    hipEventRecord(g3.ev1, g_dev.stream);
    int err = 0;
    if (ok) ok = d2h(&err, g3.d_err, sizeof(int));
    if (ok && err) {
        msg("%s: a sweep exceeded its wait limit", who);
        ok = false;
    }
    if (!ok) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) msg("%s: HIP error %s", who, hipGetErrorString(e));
        *ierr = 1;
        return;
    }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, g3.ev0, g3.ev1);
    g3.last_ms = ms;
    g3.last_corr = corr;
    g3.last_updates = updates;
    if (g3.verbose > 2)
        printf("fteik_solver3d_solveSourceLSM: Solver time in seconds=%14.8f\n",
               ms*1.0e-3);
    g3.haveTimes = true;
    g3.srcNumber = isrc;
}

void fteik_solver3d_getTravelTimeField64f(int ngin, int order, double *tt,
                                          int *ierr)
{
    *ierr = 0;
    if (!g3.haveTimes) {
        msg("fteik_solver3d_getTravelTimeField64f: Travel times not yet "
            "computed");
        *ierr = 1;
        return;
    }
    const long ngrd = (long) g3.nz*g3.nx*g3.ny;
    if (ngin != ngrd) {
        msg("fteik_solver3d_getTravelTimeField64f: ngin=%d /= ngrd =%ld",
            ngin, ngrd);
        *ierr = 1;
        return;
    }
    const double *src = g3.d_tt;
    if (order == 1 || order == 2) {
        if (fteik_k_field_reorder(g3.d_scr, g3.d_tt, g3.nz, g3.nx, g3.ny, 1,
                                  order, g_dev.stream)) { *ierr = 1; return; }
        src = g3.d_scr;
    } else if (order != 0) {
        printf("fteik_solver3d_getTravelTimeField64f: Defaulting to natural "
               "order\n");
    }
    if (!d2h(tt, src, ngrd*sizeof(double))) *ierr = 1;
}

void fteik_solver3d_getTravelTimeFieldCell64f(int ncin, int order, double *tt,
                                              int *ierr)
{
    *ierr = 0;
    if (!g3.haveTimes) {
        msg("fteik_solver3d_getTravelTimeFieldCell64f: Travel times not yet "
            "computed");
        *ierr = 1;
        return;
    }
    const long ncell = (long) (g3.nz - 1)*(g3.nx - 1)*(g3.ny - 1);
    if (ncin != ncell) {
        msg("fteik_solver3d_getTravelTimeFieldCell64f: ncin=%d /= ncell =%ld",
            ncin, ncell);
        *ierr = 1;
        return;
    }
    if (order != 0 && order != 1 && order != 2)
        printf("fteik_solver3d_getTravelTimeFieldCell64f: Defaulting to "
               "natural order\n");
    if (fteik_k_field_cells(g3.d_scr, g3.d_tt, g3.nz, g3.nx, g3.ny, 1,
                            (order == 1 || order == 2) ? order : 0,
                            g_dev.stream) ||
        !d2h(tt, g3.d_scr, ncell*sizeof(double))) *ierr = 1;
}

void fteik_solver3d_getTravelTimes64f(int nrec, double *ttr, int *ierr)
{
    *ierr = 0;
    if (nrec <= 0) {
        printf("fteik_solver2d_getTravelTimes64f: No receivers\n");
        return;
    }
    if (!g3.haveTimes) {
        msg("fteik_solver2d_getTravelTimes64f: Travel times not yet computed");
        *ierr = 1;
        return;
    }
    if (!g3.haveRec) {
        msg("fteik_receiver_getTravelTimes64f: Module not initialized");
        msg("fteik_solver3d_getTravelTimes64f: Error getting travel times");
        *ierr = 1;
        return;
    }
    const int n = std::min(nrec, g3.nrec);
    if (nrec != g3.nrec) for (int i = 0; i < nrec; i++) ttr[i] = 99999.0;
    if (fteik_k_receivers(g3.d_ttr, g3.d_tt, g3.d_ri, g3.d_rd, n, g3.nz,
                          g3.nx, g3.dz, g3.dx, g3.dy, 1, g_dev.stream) ||
        !d2h(ttr, g3.d_ttr, (size_t) n*sizeof(double))) *ierr = 1;
}

void fteik_solver3d_traceRaysToReceivers(int *ierr)
{
    msg("fteik_solver3d_traceRaysToReceivers: ray tracing is not part of "
        "this library");
    *ierr = 1;
}

void fteik_solver3d_traceRaysToPoints(int np, const double *zp,
                                      const double *xp, const double *yp,
                                      int *ierr)
{
    (void) np; (void) zp; (void) xp; (void) yp;
    msg("fteik_solver3d_traceRaysToPoints: ray tracing is not part of this "
        "library");
    *ierr = 1;
}

void fteik_amd_evalLocalSolver3d(int n, const double *in, const int *node,
                                 double *outNoInit, double *outInit,
                                 int *ierr)
{
    *ierr = 0;
    if (!g3.cstReady || n < 1) {
        msg("fteik_amd_evalLocalSolver3d: no source prepared");
        *ierr = 1;
        return;
    }
    double *d_in = nullptr, *d_o1 = nullptr, *d_o2 = nullptr;
    int *d_node = nullptr;
    bool ok = dalloc(&d_in, 15*(size_t) n) && dalloc(&d_node, 4*(size_t) n) &&
              dalloc(&d_o1, n) && dalloc(&d_o2, n) &&
              h2d(d_in, in, 15*(size_t) n*sizeof(double)) &&
              h2d(d_node, node, 4*(size_t) n*sizeof(int)) &&
              fteik_k_eval3d(d_in, d_node, g3.d_cst, n, d_o1, d_o2,
                             g_dev.stream) == 0 &&
              d2h(outNoInit, d_o1, (size_t) n*sizeof(double)) &&
              d2h(outInit, d_o2, (size_t) n*sizeof(double));
    dfree(d_in); dfree(d_node); dfree(d_o1); dfree(d_o2);
    if (!ok) *ierr = 1;
}

void fteik_amd_timeSweeps3d(int precision, int nrep, double *msPerLaunch,
                            int *ierr)
{
    // Device time of Gauss-Seidel sweep launches on the field of the last
    // solve: nrep iterations of the 8 sweep directions, hipEvents on the
    // library's stream around them.  The field is left as after the extra
    // sweeps.
    static const char *who = "fteik_amd_timeSweeps3d";
    *ierr = 0;
    *msPerLaunch = 0.0;
    if (!g3.haveTimes || g3.srcNumber < 1 || nrep < 1 ||
        (precision != 0 && precision != 1) || (precision == 1 && !g3.d_tt32)) {
        msg("%s: needs a solved field (a mixed solve for precision 1)", who);
        *ierr = 1;
        return;
    }
    const SrcInfo &s = g3.src[(size_t) (g3.srcNumber - 1)];
    SweepLaunch L = launch3d(precision == 1);
    L.zsi = s.zsi; L.xsi = s.xsi; L.ysi = s.ysi;
    hipMemsetAsync(g3.d_err, 0, sizeof(int), g_dev.stream);
    hipEventRecord(g3.ev0, g_dev.stream);
    bool ok = true;
    for (int r = 0; r < nrep && ok; r++)
        for (int sw = 1; sw <= 8 && ok; sw++) ok = run_sweep(L, sw, false);
    hipEventRecord(g3.ev1, g_dev.stream);
    int err = 0;
    ok = ok && d2h(&err, g3.d_err, sizeof(int)) && err == 0;
    float ms = 0.0f;
    if (ok) hipEventElapsedTime(&ms, g3.ev0, g3.ev1);
    if (!ok) { msg("%s: sweep failed", who); *ierr = 1; return; }
    *msPerLaunch = (double) ms/(8.0*nrep);
}

const char *fteik_amd_version(void)
{
    return "fteik_amd 0.1 (gfx950, HIP)";
}

// ====================================================================
//                         2D solver entry points
// ====================================================================
void fteik_solver2d_free(void) { free2d(); }
void fteik_solver2d_setVerbosity(int verbose) { g2.verbose = verbose; }
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
    if (fteik_k_solve2d(fields, g2.d_slow, g2.nz, g2.nx, g2.dz, g2.dx, g2.eps,
                        0.0, 0.0, 0, 0, g2.nsweep, n, (int) ngrd, g2.d_sa,
                        g2.d_sa + n, g2.d_si, g2.d_si + n, g_dev.stream))
        return false;
    if (g2.nrec > 0) {
        for (int q = 0; q < n; q++) {
            if (fteik_k_receivers(g2.d_ttr, fields + (long) q*ngrd, g2.d_ri,
                                  g2.d_rd, g2.nrec, g2.nz, g2.nx, g2.dz,
                                  g2.dx, 1.0, 0, g_dev.stream) ||
                !d2h(g2.ttimesRec.data() + (size_t) (first + q)*g2.nrec,
                     g2.d_ttr, (size_t) g2.nrec*sizeof(double)))
                return false;
        }
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

void fteik_solver2d_traceRaysToReceivers(int *ierr)
{
    msg("fteik_solver2d_traceRaysToReceivers: ray tracing is not part of "
        "this library");
    *ierr = 1;
}

void fteik_solver2d_traceRaysToPoints(int np, const double *xp,
                                      const double *zp, int *ierr)
{
    (void) np; (void) xp; (void) zp;
    msg("fteik_solver2d_traceRaysToPoints: ray tracing is not part of this "
        "library");
    *ierr = 1;
}

// ====================================================================
//                            Locator
// ====================================================================
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

// float sum in index order (locate.F90:654 SUM over the padded weights)
static // Per-event arrays for the batched grid search: observations and weights
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

float weight_norm(const float *w, int lntf)
{
    float s = 0.0f;
    for (int i = 0; i < lntf; i++) s = s + w[i];
    return 1.0f/s;
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

// ====================================================================
//                           Model helpers
// ====================================================================
int fteik_model_grid2indexF(int i, int j, int k, int nz, int nzx)
{
    return (k - 1)*nzx + (j - 1)*nz + i;
}

int fteik_model_velGrid2indexF(int i, int j, int k, int nzm1, int nzm1_nxm1)
{
    return (k - 1)*nzm1_nxm1 + (j - 1)*nzm1 + i;
}

void fteik_model_index2gridF(int igrd, int *i, int *j, int *k, int *ierr)
{
    const int nz = g3.nz, nx = g3.nx, ny = g3.ny, nzx = nz*nx;
    const int g = igrd - 1;
    if (nzx <= 0) { *i = *j = *k = 0; *ierr = 3; return; }
    *k = g/nzx;
    *j = (g - (*k)*nzx)/nz;
    *i = g - (*k)*nzx - (*j)*nz;
    *k += 1; *j += 1; *i += 1;
    *ierr = 0;
    if (*i < 1 || *i > nz) (*ierr)++;
    if (*j < 1 || *j > nx) (*ierr)++;
    if (*k < 1 || *k > ny) (*ierr)++;
}

int fteik_model_grid2index(const int iz, const int ix, const int iy,
                           const int nz, const int nzx)
{
    return fteik_model_grid2indexF(iz + 1, ix + 1, iy + 1, nz, nzx) - 1;
}

// ====================================================================
//                              Analytic
// ====================================================================
void fteik_analytic_free(void) { free_analytic(); }
void fteik_analytic_setVerbosity(int verbose) { ga.verbose = verbose; }

void fteik_analytic_initialize64f(int nz, int nx, int ny, double z0,
                                  double x0, double y0, double dz, double dx,
                                  double dy, int verbose, int *ierr)
{
    *ierr = 0;
    free_analytic();
    ga.verbose = verbose;
    if (nx < 1 || ny < 1 || nz < 1) {
        printf(" fteik_analytic_initialize64f: Invalid number of grid points "
               "%d %d %d\n", nx, ny, nz);
        *ierr = 1;
        return;
    }
    if (dx <= 0.0 || dy <= 0.0 || dz <= 0.0) {
        printf(" fteik_analytic_initialize64f: Grid spacing must be positive\n");
        *ierr = 1;
        return;
    }
    if (!device_ok("fteik_analytic_initialize64f")) { *ierr = 1; return; }
    ga.nz = nz; ga.nx = nx; ga.ny = ny; ga.dz = dz; ga.dx = dx; ga.dy = dy;
    ga.z0 = z0; ga.x0 = x0; ga.y0 = y0;
    const long ngrd = (long) nz*nx*ny;
    if (!dalloc(&ga.d_tt, ngrd) || !dalloc(&ga.d_scr, ngrd)) {
        *ierr = 1;
        return;
    }
    hipMemsetAsync(ga.d_tt, 0, ngrd*sizeof(double), g_dev.stream);
    sync_ok();
    ga.init = true;
}

void fteik_analytic_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, const double *ysrc,
                                  int *ierr)
{
    // analytic.F90:174-186: the source module checks positions against the
    // model geometry, whose origin the analytic module leaves at zero
    // (analytic.F90:136-140); the solvers use the true positions.
    *ierr = 0;
    ga.zs.clear(); ga.xs.clear(); ga.ys.clear();
    if (nsrc < 1 || !ga.init) {
        msg("fteik_analytic_setSources64f: Failed to set source");
        *ierr = 1;
        return;
    }
    const bool is3d = ga.ny >= 2;
    for (int i = 0; i < nsrc; i++) {
        SrcInfo s;
        if (locate_source(is3d, ga.nz, ga.nx, ga.ny, ga.dz, ga.dx, ga.dy, 0.0,
                          0.0, 0.0, zsrc[i], xsrc[i], is3d ? ysrc[i] : 0.0,
                          i + 1, &s)) {
            msg("fteik_analytic_setSources64f: Failed to set source");
            ga.zs.clear(); ga.xs.clear(); ga.ys.clear();
            *ierr = 1;
            return;
        }
        ga.zs.push_back(zsrc[i]);
        ga.xs.push_back(xsrc[i]);
        ga.ys.push_back(ysrc[i]);
    }
}

void fteik_analytic_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, const double *yrec,
                                    int *ierr)
{
    *ierr = 0;
    if (nrec < 1) {
        printf(" fteik_analytic_setReceivers64f: No receivers to set\n");
        return;
    }
    ga.zr.assign(zrec, zrec + nrec);
    ga.xr.assign(xrec, xrec + nrec);
    ga.yr.assign(yrec, yrec + nrec);
}

void fteik_analytic_setConstantVelocity64f(double v, int *ierr)
{
    *ierr = 0;
    if (v <= 0.0) {
        printf(" fteik_analytic_setConstantVelocity64f: Velocity must be "
               "postiive %f\n", v);
        *ierr = 1;
        return;
    }
    ga.vconst = v;
}

void fteik_analytic_setLinearVelocityGradient64f(double v0, double v1,
                                                 int *ierr)
{
    *ierr = 0;
    ga.vtop = ga.vconst;
    ga.vbot = ga.vconst;
    if (v0 <= 0.0 || v1 <= 0.0) {
        printf(" fteik_analytic_setLinearVelocityGradient64f: velocities must "
               "be positive\n");
        *ierr = 1;
        return;
    }
    if (ga.nz < 2) {
        printf(" fteik_analytic_setLinearVelocityGradient64f: 2 points "
               "required in z\n");
        *ierr = 1;
        return;
    }
    ga.vtop = v0;
    ga.vbot = v1;
}

static void analytic_solve(int isrc, int gradient, int *ierr)
{
    *ierr = 0;
    const int nsrc = (int) ga.zs.size();
    if (!ga.init || isrc < 1 || isrc > nsrc) {
        msg("fteik_analytic: Source = %d must be in range [1,%d]", isrc, nsrc);
        *ierr = 1;
        return;
    }
    if (gradient && ga.nz < 2) { *ierr = 1; return; }
    const size_t s = (size_t) (isrc - 1);
    if (fteik_k_analytic(ga.d_tt, ga.nz, ga.nx, ga.ny, ga.dz, ga.dx, ga.dy,
                         ga.z0, ga.x0, ga.y0, ga.zs[s], ga.xs[s], ga.ys[s],
                         ga.vconst, ga.vtop, ga.vbot, gradient,
                         g_dev.stream) || !sync_ok()) {
        *ierr = 1;
        return;
    }
    ga.haveField = true;
}

void fteik_analytic_solveSourceConstantVelocity64f(int isrc, int *ierr)
{
    analytic_solve(isrc, 0, ierr);
}

void fteik_analytic_solveSourceLinearVelocityGradient64f(int isrc, int *ierr)
{
    analytic_solve(isrc, 1, ierr);
}

void fteik_analytic_getTravelTimeField64f(int ngin, int order, double *tt,
                                          int *ierr)
{
    *ierr = 0;
    const long ngrd = (long) ga.nz*ga.nx*ga.ny;
    if (!ga.init) { *ierr = 1; return; }
    if (ngin < ngrd) {
        printf(" fteik_analytic_getTravelTimeField64f: Insufficient space\n");
        *ierr = 1;
        return;
    }
    if (fteik_k_analytic_reorder(ga.d_scr, ga.d_tt, ga.nz, ga.nx, ga.ny,
                                 (order == 1 || order == 2) ? order : 0,
                                 g_dev.stream) ||
        !d2h(tt, ga.d_scr, ngrd*sizeof(double))) {
        *ierr = 1;
        return;
    }
    for (long i = ngrd; i < ngin; i++) tt[i] = 0.0;
}

static void analytic_receivers(int ldr, double *trec, int gradient, int *ierr)
{
    *ierr = 0;
    const int nsrc = (int) ga.zs.size(), nrec = (int) ga.zr.size();
    if (!ga.init || nsrc < 1 || ldr < nrec) { *ierr = 1; return; }
    if (ldr > nrec) for (long i = 0; i < (long) ldr*nsrc; i++) trec[i] = 0.0;
    if (nrec < 1) return;
    dfree(ga.d_pts);
    if (!dalloc(&ga.d_pts, 4*(size_t) nrec) ||
        !h2d(ga.d_pts, ga.zr.data(), nrec*sizeof(double)) ||
        !h2d(ga.d_pts + nrec, ga.xr.data(), nrec*sizeof(double)) ||
        !h2d(ga.d_pts + 2*nrec, ga.yr.data(), nrec*sizeof(double))) {
        *ierr = 1;
        return;
    }
    for (int s = 0; s < nsrc; s++) {
        if (fteik_k_analytic_points(ga.d_pts + 3*nrec, ga.d_pts,
                                    ga.d_pts + nrec, ga.d_pts + 2*nrec, nrec,
                                    ga.nz, ga.dz, ga.z0, ga.zs[(size_t) s],
                                    ga.xs[(size_t) s], ga.ys[(size_t) s],
                                    ga.vconst, ga.vtop, ga.vbot, gradient,
                                    g_dev.stream) ||
            !d2h(trec + (size_t) s*ldr, ga.d_pts + 3*nrec,
                 nrec*sizeof(double))) {
            *ierr = 1;
            return;
        }
    }
}

void fteik_analytic_getTravelTimesConstantVel64f(int ldr, double *trec,
                                                 int *ierr)
{
    analytic_receivers(ldr, trec, 0, ierr);
}

void fteik_analytic_getTravelTimesGradientVel64f(int ldr, double *trec,
                                                 int *ierr)
{
    analytic_receivers(ldr, trec, 1, ierr);
}

// ====================================================================
//                              Graph
// ====================================================================

void fteik_graph3d_free(void) { gg = Graph(); }

void fteik_graph3d_initialize(int nz, int nx, int ny, int verbose, int *ierr)
{
    *ierr = 0;
    if (nz < 3 || nx < 3 || ny < 3) {
        msg("fteik_graph3d_initialize: Insufficient number of grid points");
        *ierr = -1;
        return;
    }
    gg = Graph();
    gg.nz = nz; gg.nx = nx; gg.ny = ny; gg.verbose = verbose;
    gg.init = true;
}

void fteik_graph3d_makeLevelStructures(int *ierr)
{
    // graph3d.F90:135-177: level sizes are the plane counts of iz+ix+iy
    *ierr = 0;
    if (!gg.init) {
        msg("fteik_graph3d_makeLevelStructures: Graph not initalized");
        *ierr = 1;
        return;
    }
    const int nz = gg.nz, nx = gg.nx, ny = gg.ny;
    gg.nLevels = nx + ny + nz - 2;
    gg.levelPtr.assign((size_t) gg.nLevels + 1, 0);
    gg.levelPtr[0] = 1;
    gg.maxLevelSize = 0;
    for (int L = 0; L < gg.nLevels; L++) {
        long np = 0;
        for (int k = 0; k < ny; k++) {
            int jlo = std::max(0, L - k - (nz - 1)), jhi = std::min(nx - 1, L - k);
            if (jhi >= jlo) np += jhi - jlo + 1;
        }
        gg.levelPtr[(size_t) L + 1] = gg.levelPtr[(size_t) L] + (int) np;
        gg.maxLevelSize = std::max(gg.maxLevelSize, (int) np);
    }
    gg.made = true;
}

void fteik_graph3d_getIJKVF(int nwork, int sweep, int *ijkv, int *ierr)
{
    *ierr = 0;
    if (sweep < 1 || sweep > 8) {
        msg("fteik_graph3d_getIJKVF: Invalid sweep = %d", sweep);
        *ierr = 1;
        return;
    }
    const long ngrd = (long) gg.nz*gg.nx*gg.ny;
    if (nwork < 4*ngrd) {
        msg("fteik_graph3d_getIJKVF: Insufficient space");
        *ierr = 1;
        return;
    }
    if (!gg.made) {
        msg("fteik_graph3d_getIJKVF: ijkv not yet computed");
        *ierr = 1;
        return;
    }
    if (!device_ok("fteik_graph3d_getIJKVF")) { *ierr = 1; return; }
    int *d_ijkv = nullptr, *d_ptr = nullptr;
    bool ok = dalloc(&d_ijkv, 4*(size_t) ngrd) &&
              dalloc(&d_ptr, gg.levelPtr.size()) &&
              h2d(d_ptr, gg.levelPtr.data(), gg.levelPtr.size()*sizeof(int)) &&
              fteik_k_graph_ijkv(d_ijkv, d_ptr, gg.nz, gg.nx, gg.ny, sweep,
                                 g_dev.stream) == 0 &&
              d2h(ijkv, d_ijkv, 4*(size_t) ngrd*sizeof(int));
    dfree(d_ijkv); dfree(d_ptr);
    if (!ok) { *ierr = 1; return; }
    for (long i = 4*ngrd; i < nwork; i++) ijkv[i] = 0;
}

void fteik_graph3d_getLevelPointerF(int nwork, int *levelPtr, int *ierr)
{
    *ierr = 0;
    if (!gg.made) {
        msg("fteik_graph3d_getLevelPointerF: levelPtr not yet made");
        *ierr = 1;
        return;
    }
    if (nwork < gg.nLevels + 1) {
        msg("fteik_graph3d_getLevelPointerF: Insufficient space");
        *ierr = 1;
        return;
    }
    for (int i = 0; i <= gg.nLevels; i++) levelPtr[i] = gg.levelPtr[(size_t) i];
    for (int i = gg.nLevels + 1; i < nwork; i++) levelPtr[i] = 0;
}

int fteik_graph3d_getNumberOfLevels(int *ierr)
{
    *ierr = 0;
    if (gg.nLevels < 1) {
        msg("fteik_graph3d_getNumberOfLevels: nLevels not yet computed");
        *ierr = 1;
    }
    return gg.nLevels;
}

int fteik_graph3d_getMaxLevelSize(int *ierr)
{
    *ierr = 0;
    if (gg.maxLevelSize < 1) {
        msg("fteik_graph3d_getMaxLevelSize: maxLevelSize not yet computed");
        *ierr = 1;
    }
    return gg.maxLevelSize;
}

// ====================================================================
//                         HDF5 archive stubs
// ====================================================================
int fteik_h5io_initializeF(const char *fileName)
{
    (void) fileName;
    msg("fteik_h5io_initializeF: HDF5 archives are not part of this library");
    return -1;
}
int fteik_h5io_finalizeF(void) { return -1; }
int fteik_h5io_writeLevelSchedulesF(void) { return -1; }
int fteik_h5io_writeTravelTimesF(const char *dataName)
{
    (void) dataName;
    return -1;
}
int fteik_h5io_writeVelocityModelF(const char *dataName)
{
    (void) dataName;
    return -1;
}

} // extern "C"

// ====================================================================
//             Batched 3D solver (EXTENSION: no reference counterpart)
// ====================================================================
namespace {

struct BGrid {
    int nz = 0, nx = 0, ny = 0;
    double dz = 0, dx = 0, dy = 0, z0 = 0, x0 = 0, y0 = 0;
    SrcInfo src;
    bool set = false;
    double *d_slow = nullptr, *d_tt = nullptr;
    float *d_slow32 = nullptr, *d_tt32 = nullptr;
    void *d_cst = nullptr, *d_cst32 = nullptr;
};

struct Batch3D {
    bool init = false, haveTimes = false;
    int precision = 0, tile = 6, maxCorr = 16;
    double bias = 1.0/64.0, convTol = 0.0;
    std::vector<BGrid> g;
    fteik::BatchGrid *d_desc = nullptr;
    int *d_map = nullptr, nwg = 0, mapTile = -1;
    unsigned long long *d_flags = nullptr, serial = 0;
    int nflagsCap = 0;
    // sweeps with mailbox tiles (mbTile >= 0): per-grid mailboxes in d_mb,
    // the workgroup orders in d_mbmap (the iteration sweeps' first, then the
    // eight initialization sweeps': entries mbOff[s] .. mbOff[s] + mbN[s] - 1
    // for s = 0 (iterations), 1..8), per-launch sweep arguments d_mbargs
    // (8 x 8 store-wave mailbox tiles with edge-major mailboxes by default:
    // 70 ms against 78 ms for 16 x 16 tiles and 113 ms for the best flag
    // tile on 64 grids of 64..192 per axis)
    int mbTile = 12, mapMbTile = -1;
    int mbOff[9] = {0}, mbN[9] = {0};
    unsigned long long *d_mb = nullptr;
    size_t mbCap = 0;
    // edge-major mailboxes of the grids (mailbox tiles other than 7, 8)
    unsigned long long *d_mbe = nullptr;
    size_t mbeCap = 0;
    int *d_mbmap = nullptr;
    char *d_mbargs = nullptr;
    int *d_err = nullptr;
    double *d_maxdelta = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, last_updates = 0.0;
    int last_corr = 0;
};
Batch3D gb;

void free_batch()
{
    for (BGrid &x : gb.g) {
        dfree(x.d_slow); dfree(x.d_tt); dfree(x.d_slow32); dfree(x.d_tt32);
        dfree(x.d_cst); dfree(x.d_cst32);
    }
    dfree(gb.d_desc); dfree(gb.d_map); dfree(gb.d_flags); dfree(gb.d_err);
    dfree(gb.d_maxdelta); dfree(gb.d_mb); dfree(gb.d_mbmap); dfree(gb.d_mbe);
    dfree(gb.d_mbargs);
    if (gb.ev0) hipEventDestroy(gb.ev0);
    if (gb.ev1) hipEventDestroy(gb.ev1);
    gb = Batch3D();
}

// Device descriptors and the workgroup order for the current tile shape.
bool batch_prepare(const char *who)
{
    const int ng = (int) gb.g.size();
    std::vector<fteik::BatchGrid> desc((size_t) ng);
    std::vector<int> dims(3*(size_t) ng);
    if (gb.mbTile >= 0) {
        // 2 words per node covers both precisions
        size_t words = 0;
        for (const BGrid &x : gb.g) words += 2*(size_t) x.nz*x.nx*x.ny;
        if (words > gb.mbCap) {
            dfree(gb.d_mb);
            gb.mbCap = 0;
            if (!dalloc(&gb.d_mb, words)) {
                msg("%s: out of GPU memory", who);
                return false;
            }
            gb.mbCap = words;
            hipMemsetAsync(gb.d_mb, 0, words*sizeof(unsigned long long),
                           g_dev.stream);
        }
        if (!gb.d_mbargs &&
            !dalloc(&gb.d_mbargs, fteik_k_batch_mb_args_bytes(ng))) {
            msg("%s: out of GPU memory", who);
            return false;
        }
    }
    // edge-major mailboxes (as the single-grid solver: not for the 16-wide
    // tiles without a store wave)
    const bool edge = gb.mbTile >= 0 && gb.mbTile != 7 && gb.mbTile != 8;
    auto edge_words = [](const BGrid &x, bool ydir) -> size_t {
        const int nt = x.nz + x.nx + x.ny;
        return ydir ? 2*(size_t) ((x.ny + 7)/8 + 1)*nt*x.nx
                    : 2*(size_t) ((x.nx + 7)/8 + 1)*nt*x.ny;
    };
    if (edge) {
        size_t words = 0;
        for (const BGrid &x : gb.g) words += edge_words(x, false)
                                           + edge_words(x, true);
        if (words > gb.mbeCap) {
            dfree(gb.d_mbe);
            gb.mbeCap = 0;
            if (!dalloc(&gb.d_mbe, words)) {
                msg("%s: out of GPU memory", who);
                return false;
            }
            gb.mbeCap = words;
            hipMemsetAsync(gb.d_mbe, 0, words*sizeof(unsigned long long),
                           g_dev.stream);
        }
    }
    size_t mboff = 0, mbeoff = 0;
    for (int i = 0; i < ng; i++) {
        const BGrid &x = gb.g[(size_t) i];
        fteik::BatchGrid &d = desc[(size_t) i];
        d.tt = x.d_tt; d.slow = x.d_slow; d.cst = x.d_cst;
        d.tt32 = x.d_tt32; d.slow32 = x.d_slow32; d.cst32 = x.d_cst32;
        d.maxdelta = gb.d_maxdelta + i;
        d.mb = gb.mbTile >= 0 ? gb.d_mb + mboff : nullptr;
        mboff += 2*(size_t) x.nz*x.nx*x.ny;
        d.mbx = d.mby = nullptr;
        d.mbnt = x.nz + x.nx + x.ny;
        if (edge) {
            d.mbx = gb.d_mbe + mbeoff;
            mbeoff += edge_words(x, false);
            d.mby = gb.d_mbe + mbeoff;
            mbeoff += edge_words(x, true);
        }
        d.nz = x.nz; d.nx = x.nx; d.ny = x.ny;
        d.zsi = x.src.zsi; d.xsi = x.src.xsi; d.ysi = x.src.ysi;
        dims[3*(size_t) i] = x.nz; dims[3*(size_t) i + 1] = x.nx;
        dims[3*(size_t) i + 2] = x.ny;
    }
    int nflags = 0;
    fteik_k_batch_map(gb.tile, ng, dims.data(), nullptr, 0, &nflags);
    std::vector<int> map(4*(size_t) std::max(nflags, 1));
    const int nwg = fteik_k_batch_map(gb.tile, ng, dims.data(), map.data(),
                                      nflags, &nflags);
    if (nwg < 1) { msg("%s: empty batch", who); return false; }
    dfree(gb.d_map);
    if (nflags > gb.nflagsCap) {
        dfree(gb.d_flags);
        if (!dalloc(&gb.d_flags, (size_t) nflags)) {
            msg("%s: out of GPU memory", who);
            return false;
        }
        gb.nflagsCap = nflags;
        hipMemsetAsync(gb.d_flags, 0, (size_t) nflags*sizeof(unsigned long long),
                       g_dev.stream);
    }
    if (!gb.d_desc && !dalloc(&gb.d_desc, (size_t) ng)) return false;
    if (!dalloc(&gb.d_map, map.size()) ||
        !h2d(gb.d_desc, desc.data(), desc.size()*sizeof(fteik::BatchGrid)) ||
        !h2d(gb.d_map, map.data(), map.size()*sizeof(int))) {
        msg("%s: GPU error", who);
        return false;
    }
    gb.nwg = nwg;
    gb.mapTile = gb.tile;
    dfree(gb.d_mbmap);
    gb.mapMbTile = -1;
    if (gb.mbTile >= 0) {
        std::vector<int> srcs(3*(size_t) ng);
        for (int i = 0; i < ng; i++) {
            srcs[3*(size_t) i] = desc[(size_t) i].zsi;
            srcs[3*(size_t) i + 1] = desc[(size_t) i].xsi;
            srcs[3*(size_t) i + 2] = desc[(size_t) i].ysi;
        }
        std::vector<int> mbmap;
        for (int sw = 0; sw <= 8; sw++) {
            const int n = fteik_k_batch_mb_map(gb.mbTile, ng, dims.data(),
                                               srcs.data(), sw, nullptr, 0);
            if (n < 0 || (sw == 0 && n < 1)) {
                msg("%s: empty batch", who);
                return false;
            }
            gb.mbOff[sw] = (int) (mbmap.size()/2);
            gb.mbN[sw] = n;
            mbmap.resize(mbmap.size() + 2*(size_t) n);
            if (n > 0 && fteik_k_batch_mb_map(gb.mbTile, ng, dims.data(),
                                              srcs.data(), sw,
                                              mbmap.data() + 2*(size_t)
                                              gb.mbOff[sw], n) != n) {
                msg("%s: workgroup order failed", who);
                return false;
            }
        }
        if (!dalloc(&gb.d_mbmap, mbmap.size()) ||
            !h2d(gb.d_mbmap, mbmap.data(), mbmap.size()*sizeof(int))) {
            msg("%s: GPU error", who);
            return false;
        }
    }
    gb.mapMbTile = gb.mbTile;
    return true;
}

bool batch_sweep(int sweep, bool linit, bool fp32, bool track)
{
    if (gb.mbTile >= 0) {
        const int m = linit ? sweep : 0;
        const int rc = fteik_k_batch_mb_sweep(gb.d_desc, (int) gb.g.size(),
                                              gb.d_mbmap + 2*(size_t)
                                              gb.mbOff[m], gb.mbN[m],
                                              gb.mbTile, sweep, linit ? 1 : 0,
                                              ++gb.serial, gb.d_err,
                                              fp32 ? 1 : 0, track ? 1 : 0,
                                              gb.d_mbargs, g_dev.stream);
        if (rc != 0) msg("fteik_batch3d_solve: sweep launch failed (%d)", rc);
        return rc == 0;
    }
    int rc = fteik_k_batch_sweep(gb.d_desc, gb.d_map, gb.nwg, gb.tile, sweep,
                                 linit ? 1 : 0, ++gb.serial, gb.d_flags,
                                 gb.d_err, fp32 ? 1 : 0, track ? 1 : 0,
                                 g_dev.stream);
    if (rc != 0) msg("fteik_batch3d_solve: sweep launch failed (%d)", rc);
    return rc == 0;
}

double batch_updates_init()
{
    static const int tz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
    static const int tx[8] = {1, -1, 1, -1, 1, -1, 1, -1};
    static const int ty[8] = {1, 1, -1, -1, 1, 1, -1, -1};
    double u = 0.0;
    for (const BGrid &x : gb.g)
        for (int s = 0; s < 8; s++) {
            auto len = [&](int sg, int n, int si) -> double {
                int a0, a1;
                if (sg > 0) { a0 = std::max(2, si); a1 = n; }
                else        { a0 = 1; a1 = si + 1; }
                return (double) std::max(0, a1 - a0 + 1);
            };
            u += len(tz[s], x.nz, x.src.zsi)*len(tx[s], x.nx, x.src.xsi)
               * len(ty[s], x.ny, x.src.ysi);
        }
    return u;
}

double batch_updates_iter()
{
    double u = 0.0;
    for (const BGrid &x : gb.g)
        u += 8.0*(double) (x.nz - 1)*(double) (x.nx - 1)*(double) (x.ny - 1);
    return u;
}

} // namespace

extern "C" {

void fteik_batch3d_free(void) { free_batch(); }

void fteik_batch3d_initialize(int ngrids, int *ierr)
{
    *ierr = 0;
    free_batch();
    if (ngrids < 1) {
        msg("fteik_batch3d_initialize: ngrids = %d must be positive", ngrids);
        *ierr = 1;
        return;
    }
    if (!device_ok("fteik_batch3d_initialize")) { *ierr = 1; return; }
    gb.g.assign((size_t) ngrids, BGrid());
    if (!dalloc(&gb.d_err, 1) || !dalloc(&gb.d_maxdelta, (size_t) ngrids)) {
        msg("fteik_batch3d_initialize: out of GPU memory");
        free_batch();
        *ierr = 1;
        return;
    }
    hipEventCreate(&gb.ev0);
    hipEventCreate(&gb.ev1);
    gb.init = true;
}

void fteik_batch3d_setGrid(int igrid, int nz, int nx, int ny, double z0,
                           double x0, double y0, double dz, double dx,
                           double dy, const double *vel, double zs, double xs,
                           double ys, int *ierr)
{
    static const char *who = "fteik_batch3d_setGrid";
    *ierr = 0;
    gb.haveTimes = false;
    if (!gb.init || igrid < 1 || igrid > (int) gb.g.size()) {
        msg("%s: batch not initialized or grid %d out of range", who, igrid);
        *ierr = 1;
        return;
    }
    if (nz < 3 || nx < 3 || ny < 3 || dz <= 0.0 || dx <= 0.0 || dy <= 0.0) {
        msg("%s: grid %d: need n >= 3 and positive spacing", who, igrid);
        *ierr = 1;
        return;
    }
    BGrid &x = gb.g[(size_t) (igrid - 1)];
    dfree(x.d_slow); dfree(x.d_tt); dfree(x.d_slow32); dfree(x.d_tt32);
    dfree(x.d_cst); dfree(x.d_cst32);
    x = BGrid();
    x.nz = nz; x.nx = nx; x.ny = ny;
    x.z0 = z0; x.x0 = x0; x.y0 = y0; x.dz = dz; x.dx = dx; x.dy = dy;
    if (locate_source(true, nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys,
                      igrid, &x.src)) { *ierr = 1; return; }
    if (x.src.zsi > nz - 2 || x.src.xsi > nx - 2 || x.src.ysi > ny - 2) {
        msg("%s: grid %d: source in the last cell along an axis", who, igrid);
        *ierr = 1;
        return;
    }
    const long ncell = (long) (nz - 1)*(nx - 1)*(ny - 1);
    const long ngrd = (long) nz*nx*ny;
    if (!dalloc(&x.d_slow, (size_t) ncell) || !dalloc(&x.d_tt, (size_t) ngrd) ||
        !dalloc((char **) &x.d_cst, 512) || !dalloc((char **) &x.d_cst32, 512)) {
        msg("%s: out of GPU memory", who);
        *ierr = 1;
        return;
    }
    if (!set_model(who, true, nz, nx, ny, 0, (int) ncell, 0, vel, x.d_slow)) {
        *ierr = 1;
        return;
    }
    x.set = true;
    gb.mapTile = -1;
}

void fteik_batch3d_setMixedParameters(double bias, int maxCorrections,
                                      int *ierr)
{
    *ierr = 0;
    if (!(bias >= 0.0 && bias < 0.1) || maxCorrections < 1 ||
        maxCorrections > 64) {
        msg("fteik_batch3d_setMixedParameters: need 0 <= bias < 0.1 and "
            "1 <= maxCorrections <= 64");
        *ierr = 1;
        return;
    }
    gb.bias = bias;
    gb.maxCorr = maxCorrections;
}

void fteik_batch3d_setPrecision(int mode, int *ierr)
{
    *ierr = (mode == 0 || mode == 1) ? 0 : 1;
    if (*ierr) { msg("fteik_batch3d_setPrecision: mode must be 0 or 1");
                 return; }
    gb.precision = mode;
}

void fteik_batch3d_setTile(int tile, int *ierr)
{
    *ierr = 0;
    if (tile < 0 || tile >= fteik_k_num_tiles()) {
        msg("fteik_batch3d_setTile: tile %d is not available for batches",
            tile);
        *ierr = 1;
        return;
    }
    // a flag tile (mode 0) or a mailbox tile (modes 1, 2) serves every
    // sweep
    if (fteik_k_tile_mode(tile) == 0) {
        gb.tile = tile;
        gb.mbTile = -1;
    } else {
        gb.mbTile = tile;
    }
}

void fteik_batch3d_solve(int nsweep, double eps, int *ierr)
{
    static const char *who = "fteik_batch3d_solve";
    *ierr = 0;
    gb.haveTimes = false;
    if (!gb.init || nsweep < 0) { msg("%s: not initialized", who); *ierr = 1;
                                  return; }
    for (size_t i = 0; i < gb.g.size(); i++) {
        if (!gb.g[i].set) { msg("%s: grid %d not set", who, (int) i + 1);
                            *ierr = 1; return; }
        if (!check_eps(who, eps, gb.g[i].nz, gb.g[i].nx, gb.g[i].ny, true)) {
            *ierr = 1;
            return;
        }
    }
    hipEventRecord(gb.ev0, g_dev.stream);
    for (BGrid &x : gb.g) {
        fteik::SourceGeom sg;
        sg.dz = x.dz; sg.dx = x.dx; sg.dy = x.dy; sg.eps = eps;
        sg.zsa = x.src.zsa; sg.xsa = x.src.xsa; sg.ysa = x.src.ysa;
        sg.zsi = x.src.zsi; sg.xsi = x.src.xsi; sg.ysi = x.src.ysi;
        sg.nz = x.nz; sg.nx = x.nx; sg.ny = x.ny;
        if (fteik_k_source3d_init(x.d_cst, x.d_tt, x.d_slow, &sg,
                                  (long) x.nz*x.nx*x.ny, 0, g_dev.stream)) {
            msg("%s: source set-up failed", who);
            *ierr = 1;
            return;
        }
    }
    if (gb.precision == 1) {
        for (BGrid &x : gb.g) {
            const long ng = (long) x.nz*x.nx*x.ny;
            const long nc = (long) (x.nz - 1)*(x.nx - 1)*(x.ny - 1);
            if ((!x.d_tt32 && !dalloc(&x.d_tt32, (size_t) ng)) ||
                (!x.d_slow32 && !dalloc(&x.d_slow32, (size_t) nc))) {
                msg("%s: out of GPU memory", who);
                *ierr = 1;
                return;
            }
        }
        gb.mapTile = -1;
    }
    if ((gb.mapTile != gb.tile || gb.mapMbTile != gb.mbTile) &&
        !batch_prepare(who)) { *ierr = 1; return; }
    hipMemsetAsync(gb.d_err, 0, sizeof(int), g_dev.stream);
    bool ok = true;
    double updates = batch_updates_init();
    for (int sw = 1; sw <= 8 && ok; sw++) ok = batch_sweep(sw, true, false, false);
    int corr = 0;
    if (gb.precision == 0) {
        for (int k = 0; k < nsweep && ok; k++)
            for (int sw = 1; sw <= 8 && ok; sw++)
                ok = batch_sweep(sw, false, false, false);
        updates += nsweep*batch_updates_iter();
    } else if (ok) {
        for (BGrid &x : gb.g) {
            const long ng = (long) x.nz*x.nx*x.ny;
            const long nc = (long) (x.nz - 1)*(x.nx - 1)*(x.ny - 1);
            ok = ok && fteik_k_d2f(x.d_slow32, x.d_slow, nc, g_dev.stream) == 0
                    && fteik_k_d2f(x.d_tt32, x.d_tt, ng, g_dev.stream) == 0
                    && fteik_k_consts_to_float(x.d_cst, x.d_cst32,
                                               g_dev.stream) == 0;
        }
        for (int k = 0; k < nsweep && ok; k++)
            for (int sw = 1; sw <= 8 && ok; sw++)
                ok = batch_sweep(sw, false, true, false);
        updates += nsweep*batch_updates_iter();
        for (BGrid &x : gb.g)
            ok = ok && fteik_k_min_merge_f(x.d_tt, x.d_tt32,
                                           (long) x.nz*x.nx*x.ny, gb.bias,
                                           g_dev.stream) == 0;
        const size_t ng = gb.g.size();
        std::vector<double> md(ng);
        for (corr = 0; corr < gb.maxCorr && ok; ) {
            hipMemsetAsync(gb.d_maxdelta, 0, ng*sizeof(double), g_dev.stream);
            for (int sw = 1; sw <= 8 && ok; sw++)
                ok = batch_sweep(sw, false, false, true);
            updates += batch_updates_iter();
            corr++;
            ok = ok && d2h(md.data(), gb.d_maxdelta, ng*sizeof(double));
            if (!ok || *std::max_element(md.begin(), md.end()) <= gb.convTol)
                break;
        }
    }
    hipEventRecord(gb.ev1, g_dev.stream);
    int err = 0;
    ok = ok && d2h(&err, gb.d_err, sizeof(int));
    if (ok && err) { msg("%s: a sweep exceeded its wait limit", who);
                     ok = false; }
    if (!ok) { *ierr = 1; return; }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, gb.ev0, gb.ev1);
    gb.last_ms = ms;
    gb.last_updates = updates;
    gb.last_corr = corr;
    gb.haveTimes = true;
}

void fteik_batch3d_getLastSolveInfo(double *ms, int *corrIterations,
                                    double *updates, int *ierr)
{
    *ierr = gb.haveTimes ? 0 : 1;
    *ms = gb.last_ms;
    *corrIterations = gb.last_corr;
    *updates = gb.last_updates;
}

void fteik_batch3d_getTravelTimeField64f(int igrid, int ngrd, int order,
                                         double *tt, int *ierr)
{
    static const char *who = "fteik_batch3d_getTravelTimeField64f";
    *ierr = 0;
    if (!gb.haveTimes || igrid < 1 || igrid > (int) gb.g.size()) {
        msg("%s: no solved grid %d", who, igrid);
        *ierr = 1;
        return;
    }
    const BGrid &x = gb.g[(size_t) (igrid - 1)];
    const long n = (long) x.nz*x.nx*x.ny;
    if (ngrd != n) { msg("%s: ngrd = %d /= %ld", who, ngrd, n); *ierr = 1;
                     return; }
    if (order == 1 || order == 2) {
        double *d = nullptr;
        bool ok = dalloc(&d, (size_t) n) &&
                  fteik_k_field_reorder(d, x.d_tt, x.nz, x.nx, x.ny, 1, order,
                                        g_dev.stream) == 0 &&
                  d2h(tt, d, (size_t) n*sizeof(double));
        dfree(d);
        if (!ok) *ierr = 1;
        return;
    }
    if (!d2h(tt, x.d_tt, (size_t) n*sizeof(double))) *ierr = 1;
}

void fteik_batch3d_getNumberOfGrids(int *ngrids, int *ierr)
{
    *ierr = gb.init ? 0 : 1;
    *ngrids = (int) gb.g.size();
}

// Table itf of the locator from batch grid igrid, on the device: the solved
// field rounded to single precision like locate_setTravelTimeField64f.
void locate_setTravelTimeFieldFromBatch3d(int igrid, int itf, int *ierr)
{
    static const char *who = "locate_setTravelTimeFieldFromBatch3d";
    *ierr = 0;
    if (!gl.init || itf < 1 || itf > gl.ntf) {
        msg("%s: locator not initialized or table %d out of range", who, itf);
        *ierr = 1;
        return;
    }
    if (!gb.haveTimes || igrid < 1 || igrid > (int) gb.g.size()) {
        msg("%s: no solved batch grid %d", who, igrid);
        *ierr = 1;
        return;
    }
    const BGrid &x = gb.g[(size_t) (igrid - 1)];
    if ((long) x.nz*x.nx*x.ny != gl.ngrd) {
        msg("%s: grid %d has %ld points, the locator %d", who, igrid,
            (long) x.nz*x.nx*x.ny, gl.ngrd);
        *ierr = 1;
        return;
    }
    if (fteik_k_tables_d2f(gl.d_tab + (long) (itf - 1)*gl.ldgrd, x.d_tt,
                           gl.ngrd, g_dev.stream) != 0 || !sync_ok())
        *ierr = 1;
}

} // extern "C"

// ====================================================================
//     Learned surrogate for the locator (EXTENSION: no counterpart)
// ====================================================================
namespace {

struct Surrogate {
    bool init = false;
    int ntf = 0, H = 256, step = 0;
    FteikSurrogateGeo geo;
    void *d_geo = nullptr;
    long np = 0;                       // parameter count
    long oW1, ob1, oW[3], ob[3], oW5, ob5;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    unsigned short *d_pb = nullptr;    // BF16 copy of all parameters
    // work buffers (grown on demand)
    long cap = 0;
    float *d_x = nullptr, *d_tgt = nullptr, *d_y = nullptr, *d_dh = nullptr;
    float *d_z[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *d_h[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *d_dy = nullptr, *d_dz = nullptr;
    double *d_loss = nullptr;
    unsigned long long seed = 13;
    // inversion work buffers (grown on demand): event arrays sized by
    // events, the rest by rows R = events * starts
    long invEvents = 0, invRows = 0;
    float *d_io = nullptr, *d_iw = nullptr, *d_ixn = nullptr;
    unsigned char *d_isk = nullptr;
    unsigned long long *d_ibest = nullptr;
    int *d_ifb = nullptr;
    float *d_pos = nullptr, *d_mis = nullptr, *d_zp = nullptr,
          *d_yall = nullptr;
    unsigned short *d_ha = nullptr, *d_hb = nullptr;
};
Surrogate gs;

void free_inv_work()
{
    dfree(gs.d_io); dfree(gs.d_iw); dfree(gs.d_ixn); dfree(gs.d_isk);
    dfree(gs.d_ibest); dfree(gs.d_ifb);
    dfree(gs.d_pos); dfree(gs.d_mis); dfree(gs.d_zp);
    dfree(gs.d_yall); dfree(gs.d_ha); dfree(gs.d_hb);
    gs.invEvents = gs.invRows = 0;
}

// Inversion buffers for n events (lntf observations each) and R rows.
bool inv_work(long n, long lntf, long R)
{
    const int H = gs.H, T = gs.ntf;
    if (n*lntf > gs.invEvents || R > gs.invRows) {
        const long ne = std::max(n*lntf, gs.invEvents);
        const long nr = std::max(R, gs.invRows);
        const long nev = std::max(n, ne/std::max(lntf, 1L));
        free_inv_work();
        if (!(dalloc(&gs.d_io, (size_t) ne) && dalloc(&gs.d_iw, (size_t) ne) &&
              dalloc(&gs.d_isk, (size_t) ne) && dalloc(&gs.d_ixn, (size_t) nev) &&
              dalloc(&gs.d_ibest, (size_t) nev) && dalloc(&gs.d_ifb, (size_t) nev) &&
              dalloc(&gs.d_pos, 3*(size_t) nr) && dalloc(&gs.d_mis, (size_t) nr) &&
              dalloc(&gs.d_zp, (size_t) nr*H) &&
              dalloc(&gs.d_yall, 4*(size_t) nr*T) &&
              dalloc(&gs.d_ha, 4*(size_t) nr*H) &&
              dalloc(&gs.d_hb, 4*(size_t) nr*H))) {
            free_inv_work();
            return false;
        }
        gs.invEvents = ne;
        gs.invRows = nr;
    }
    return true;
}

void free_surrogate()
{
    dfree(gs.d_geo); dfree(gs.d_p); dfree(gs.d_g); dfree(gs.d_m); dfree(gs.d_v);
    dfree(gs.d_pb); dfree(gs.d_x); dfree(gs.d_tgt); dfree(gs.d_y);
    dfree(gs.d_dh); dfree(gs.d_dy); dfree(gs.d_dz); dfree(gs.d_loss);
    for (int l = 0; l < 4; l++) { dfree(gs.d_z[l]); dfree(gs.d_h[l]); }
    free_inv_work();
    gs = Surrogate();
}

bool sur_work(long rows)
{
    if (rows <= gs.cap) return true;
    dfree(gs.d_x); dfree(gs.d_tgt); dfree(gs.d_y); dfree(gs.d_dh);
    dfree(gs.d_dy); dfree(gs.d_dz);
    for (int l = 0; l < 4; l++) { dfree(gs.d_z[l]); dfree(gs.d_h[l]); }
    const size_t R = (size_t) rows, H = (size_t) gs.H, T = (size_t) gs.ntf;
    bool ok = dalloc(&gs.d_x, 3*R) && dalloc(&gs.d_tgt, R*T) &&
              dalloc(&gs.d_y, R*T) && dalloc(&gs.d_dh, R*H) &&
              dalloc(&gs.d_dy, R*T) && dalloc(&gs.d_dz, R*H);
    for (int l = 0; l < 4 && ok; l++)
        ok = dalloc(&gs.d_z[l], R*H) && dalloc(&gs.d_h[l], R*H);
    if (!ok) { gs.cap = 0; return false; }
    gs.cap = rows;
    return true;
}

// Forward pass over `rows` rows of d_x into d_y (network output, before
// the baseline and scale are applied).
bool sur_forward(long rows)
{
    hipStream_t s = g_dev.stream;
    const int H = gs.H;
    const unsigned short *pb = gs.d_pb;
    bool ok = fteik_k_sur_layer1(gs.d_x, gs.d_p + gs.oW1, gs.d_p + gs.ob1,
                                 gs.d_z[0], gs.d_h[0], rows, H, s) == 0;
    for (int l = 0; l < 3 && ok; l++) {
        // z_{l+1} = h_l W^T + b
        ok = fteik_k_gemm_bf16(gs.d_h[l], H, 1, pb + gs.oW[l], 1, H,
                               gs.d_z[l + 1], H, (int) rows, H, H,
                               gs.d_p + gs.ob[l], 0, s) == 0 &&
             fteik_k_sur_silu(gs.d_z[l + 1], gs.d_h[l + 1], rows*H, s) == 0;
    }
    ok = ok && fteik_k_gemm_bf16(gs.d_h[3], H, 1, pb + gs.oW5, 1, H, gs.d_y,
                                 gs.ntf, (int) rows, gs.ntf, H,
                                 gs.d_p + gs.ob5, 0, s) == 0;
    return ok;
}

// Forward + backward of the mean squared error against d_tgt; parameter
// gradients into d_g, loss (mean over rows x ntf) added to d_loss.
bool sur_backward(long rows)
{
    hipStream_t s = g_dev.stream;
    const int H = gs.H, T = gs.ntf;
    const unsigned short *pb = gs.d_pb;
    const float inv = 1.0f/(float) ((double) rows*T);
    bool ok = fteik_k_sur_mse(gs.d_y, gs.d_tgt, gs.d_dy, rows*T, inv,
                              gs.d_loss, s) == 0;
    // output layer: dW5 = dy^T h4, db5 = colsum(dy), dh4 = dy W5
    ok = ok && fteik_k_gemm_bf16(gs.d_dy, 1, T, gs.d_h[3], H, 1,
                                 gs.d_g + gs.oW5, H, T, H, (int) rows,
                                 nullptr, 0, s) == 0 &&
         fteik_k_sur_colsum(gs.d_dy, rows, T, gs.d_g + gs.ob5, s) == 0 &&
         fteik_k_gemm_bf16(gs.d_dy, T, 1, pb + gs.oW5, H, 1, gs.d_dh, H,
                           (int) rows, H, T, nullptr, 0, s) == 0 &&
         fteik_k_sur_silu_bwd(gs.d_dh, gs.d_z[3], gs.d_dz, rows*H, s) == 0;
    for (int l = 2; l >= 0 && ok; l--) {
        // layer l+1 (weights oW[l]) maps h_l -> z_{l+1}; dz holds dz_{l+1}
        ok = fteik_k_gemm_bf16(gs.d_dz, 1, H, gs.d_h[l], H, 1,
                               gs.d_g + gs.oW[l], H, H, H, (int) rows,
                               nullptr, 0, s) == 0 &&
             fteik_k_sur_colsum(gs.d_dz, rows, H, gs.d_g + gs.ob[l], s) == 0 &&
             fteik_k_gemm_bf16(gs.d_dz, H, 1, pb + gs.oW[l], H, 1, gs.d_dh,
                               H, (int) rows, H, H, nullptr, 0, s) == 0 &&
             fteik_k_sur_silu_bwd(gs.d_dh, gs.d_z[l], gs.d_dz, rows*H, s) == 0;
    }
    ok = ok && fteik_k_sur_dw1(gs.d_dz, gs.d_x, rows, H, gs.d_g + gs.oW1, s) == 0
            && fteik_k_sur_colsum(gs.d_dz, rows, H, gs.d_g + gs.ob1, s) == 0;
    return ok;
}

} // namespace

extern "C" {

void fteik_surrogate_free(void) { free_surrogate(); }

void fteik_surrogate_initialize(int ntf, int nz, int nx, int ny, double dz,
                                double dx, double dy, const double *zsta,
                                const double *xsta, const double *ysta,
                                double v0, double scale, int seed, int *ierr)
{
    static const char *who = "fteik_surrogate_initialize";
    *ierr = 0;
    free_surrogate();
    if (ntf < 1 || ntf > 64 || nz < 4 || nx < 4 || ny < 4 || v0 <= 0.0 ||
        scale <= 0.0) {
        msg("%s: need 1 <= ntf <= 64, grids of at least 4 points, v0 > 0, "
            "scale > 0", who);
        *ierr = 1;
        return;
    }
    if (!device_ok(who)) { *ierr = 1; return; }
    gs.ntf = ntf;
    FteikSurrogateGeo &g = gs.geo;
    memset(&g, 0, sizeof(g));
    g.nz = nz; g.nx = nx; g.ny = ny; g.ntf = ntf;
    g.dz = (float) dz; g.dx = (float) dx; g.dy = (float) dy;
    for (int s = 0; s < ntf; s++) {
        g.sz[s] = (float) zsta[s]; g.sx[s] = (float) xsta[s];
        g.sy[s] = (float) ysta[s];
    }
    g.v0 = (float) v0; g.scale = (float) scale;
    if ((int) sizeof(FteikSurrogateGeo) != fteik_k_sur_geo_size()) {
        msg("%s: geometry block size mismatch", who);
        *ierr = 1;
        return;
    }
    const long H = gs.H;
    long o = 0;
    gs.oW1 = o; o += 3*H;
    gs.ob1 = o; o += H;
    for (int l = 0; l < 3; l++) { gs.oW[l] = o; o += H*H; gs.ob[l] = o; o += H; }
    gs.oW5 = o; o += ntf*H;
    gs.ob5 = o; o += ntf;
    gs.np = o;
    // He-uniform initial weights, zero biases (deterministic in `seed`)
    std::vector<float> p((size_t) o, 0.0f);
    unsigned long long st = 0x9e3779b97f4a7c15ull ^ (unsigned long long) seed;
    auto uni = [&]() -> float {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float) ((double) (st >> 11)*(1.0/9007199254740992.0));
    };
    auto fill = [&](long off, long n, int fanin) {
        const float a = sqrtf(6.0f/(float) fanin);
        for (long i = 0; i < n; i++) p[(size_t) (off + i)] = a*(2.0f*uni() - 1.0f);
    };
    fill(gs.oW1, 3*H, 3);
    for (int l = 0; l < 3; l++) fill(gs.oW[l], H*H, (int) H);
    fill(gs.oW5, ntf*H, (int) H);
    for (long i = gs.oW5; i < gs.oW5 + ntf*H; i++) p[(size_t) i] *= 0.1f;
    bool ok = dalloc((char **) &gs.d_geo, sizeof(FteikSurrogateGeo)) &&
              dalloc(&gs.d_p, (size_t) o) && dalloc(&gs.d_g, (size_t) o) &&
              dalloc(&gs.d_m, (size_t) o) && dalloc(&gs.d_v, (size_t) o) &&
              dalloc(&gs.d_pb, (size_t) o) && dalloc(&gs.d_loss, 1) &&
              h2d(gs.d_geo, &g, sizeof(g)) &&
              h2d(gs.d_p, p.data(), (size_t) o*sizeof(float));
    if (ok) {
        hipMemsetAsync(gs.d_m, 0, (size_t) o*sizeof(float), g_dev.stream);
        hipMemsetAsync(gs.d_v, 0, (size_t) o*sizeof(float), g_dev.stream);
        ok = fteik_k_f2bf(gs.d_p, gs.d_pb, o, g_dev.stream) == 0 && sync_ok();
    }
    if (!ok) { msg("%s: GPU error", who); free_surrogate(); *ierr = 1; return; }
    gs.init = true;
}

void fteik_surrogate_getNumberOfParameters(int *np, int *ierr)
{
    *ierr = gs.init ? 0 : 1;
    *np = (int) gs.np;
}

void fteik_surrogate_getParameters(int np, float *p, int *ierr)
{
    *ierr = (gs.init && np == gs.np) ? 0 : 1;
    if (*ierr == 0 && !d2h(p, gs.d_p, (size_t) np*sizeof(float))) *ierr = 1;
}

void fteik_surrogate_setParameters(int np, const float *p, int *ierr)
{
    *ierr = (gs.init && np == gs.np) ? 0 : 1;
    if (*ierr) return;
    if (!h2d(gs.d_p, p, (size_t) np*sizeof(float)) ||
        fteik_k_f2bf(gs.d_p, gs.d_pb, np, g_dev.stream) != 0 || !sync_ok())
        *ierr = 1;
}

// Network output (before baseline and scale) at n normalised positions.
void fteik_surrogate_forward(int n, const float *pos, float *out, int *ierr)
{
    *ierr = 0;
    if (!gs.init || n < 1 || !sur_work(n) ||
        !h2d(gs.d_x, pos, 3*(size_t) n*sizeof(float)) || !sur_forward(n) ||
        !d2h(out, gs.d_y, (size_t) n*gs.ntf*sizeof(float))) {
        msg("fteik_surrogate_forward: failed");
        *ierr = 1;
    }
}

// One forward/backward pass on given data: parameter gradients of the mean
// squared error (over n x ntf outputs) and the loss.
void fteik_surrogate_gradients(int n, const float *pos, const float *tgt,
                               float *grad, double *loss, int *ierr)
{
    *ierr = 0;
    const double zero = 0.0;
    if (!gs.init || n < 1 || !sur_work(n) ||
        !h2d(gs.d_x, pos, 3*(size_t) n*sizeof(float)) ||
        !h2d(gs.d_tgt, tgt, (size_t) n*gs.ntf*sizeof(float)) ||
        !h2d(gs.d_loss, &zero, sizeof(double)) || !sur_forward(n) ||
        !sur_backward(n) ||
        !d2h(grad, gs.d_g, (size_t) gs.np*sizeof(float)) ||
        !d2h(loss, gs.d_loss, sizeof(double))) {
        msg("fteik_surrogate_gradients: failed");
        *ierr = 1;
    }
}

// Trains on samples drawn from the locator's tables (a pool of `pool`
// uniformly drawn cells); Adam with a cosine-decayed learning rate.
void fteik_surrogate_train(int steps, int batch, double lr, int pool,
                           double *lossOut, int *ierr)
{
    static const char *who = "fteik_surrogate_train";
    *ierr = 0;
    *lossOut = 0.0;
    if (!gs.init || !gl.init || gl.ntf != gs.ntf ||
        (long) gs.geo.nz*gs.geo.nx*gs.geo.ny != gl.ngrd || steps < 1 ||
        batch < 64 || pool < batch) {
        msg("%s: surrogate and locator must be initialized for the same "
            "stations and grid", who);
        *ierr = 1;
        return;
    }
    if (!sur_work(batch)) { msg("%s: out of GPU memory", who); *ierr = 1;
                            return; }
    hipStream_t s = g_dev.stream;
    const double zero = 0.0;
    bool ok = true;
    double tail = 0.0;
    const int tailN = std::max(1, std::min(steps, 50));
    for (int it = 0; it < steps && ok; it++) {
        const long first = (long) it*batch;
        ok = fteik_k_sur_samples(gl.d_tab, gl.ldgrd, gs.d_geo, gl.ngrd,
                                 gs.seed, pool, first, batch, gs.d_x,
                                 gs.d_tgt, gs.ntf, s) == 0 &&
             h2d(gs.d_loss, &zero, sizeof(double)) && sur_forward(batch) &&
             sur_backward(batch);
        gs.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f/(1.0f - powf(b1, (float) gs.step));
        const float c2 = 1.0f/(1.0f - powf(b2, (float) gs.step));
        const float lrt = (float) (lr*0.5*(1.0 + cos(3.14159265358979*it/steps)));
        ok = ok && fteik_k_sur_adam(gs.d_p, gs.d_g, gs.d_m, gs.d_v, gs.np, lrt,
                                    b1, b2, 1e-8f, c1, c2, gs.d_pb, s) == 0;
        if (ok && it >= steps - tailN) {
            double l = 0.0;
            ok = d2h(&l, gs.d_loss, sizeof(double));
            tail += l;
        }
    }
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    *lossOut = tail/tailN;
}

// Surrogate-pruned location of events evnmbr[0..n-1] (fresh origin-time
// state): damped Gauss-Newton on the network's misfit from `starts`
// starting points, exact scoring of the 4 x 4 x 4 cells around the best,
// full grid search for events whose best cell lies on the cube's faces.
void locate_locateEventsSurrogate(int n, const int *evnmbr, int starts,
                                  int iterations, int *optIndx, double *t0Opt,
                                  double *objOpt, int *nFallback, int *ierr)
{
    static const char *who = "locate_locateEventsSurrogate";
    *ierr = 0;
    *nFallback = 0;
    if (!gs.init || !gl.init || gl.ntf != gs.ntf || n < 1 || starts < 1 ||
        starts > 9 || iterations < 0 ||
        (long) gs.geo.nz*gs.geo.nx*gs.geo.ny != gl.ngrd) {
        msg("%s: surrogate and locator not set up for the same grid", who);
        *ierr = 1;
        return;
    }
    for (int e = 0; e < n; e++)
        if (evnmbr[e] < 1 || evnmbr[e] > gl.nEvents) {
            msg("%s: invalid event %d", who, evnmbr[e]);
            *ierr = 1;
            return;
        }
    hipStream_t s = g_dev.stream;
    const int H = gs.H, T = gs.ntf;
    const long R = (long) n*starts;
    std::vector<float> o((size_t) n*gl.lntf), w((size_t) n*gl.lntf),
                       xn((size_t) n);
    std::vector<unsigned char> sk((size_t) n*gl.lntf);
    for (int e = 0; e < n; e++) {
        const size_t k = (size_t) (evnmbr[e] - 1)*gl.lntf;
        memcpy(&o[(size_t) e*gl.lntf], &gl.obs[k], gl.lntf*sizeof(float));
        memcpy(&w[(size_t) e*gl.lntf], &gl.wt[k], gl.lntf*sizeof(float));
        memcpy(&sk[(size_t) e*gl.lntf], &gl.skip[k], (size_t) gl.lntf);
        xn[(size_t) e] = weight_norm(&gl.wt[k], gl.lntf);
    }
    if (!inv_work(n, gl.lntf, R)) {
        msg("%s: out of GPU memory", who);
        *ierr = 1;
        return;
    }
    float *d_o = gs.d_io, *d_w = gs.d_iw, *d_xn = gs.d_ixn, *d_pos = gs.d_pos,
          *d_mis = gs.d_mis, *d_zp = gs.d_zp,
          *d_yall = gs.d_yall;
    unsigned short *d_ha = gs.d_ha, *d_hb = gs.d_hb;
    unsigned char *d_sk = gs.d_isk;
    unsigned long long *d_best = gs.d_ibest;
    int *d_fb = gs.d_ifb;
    bool ok = h2d(d_o, o.data(), o.size()*sizeof(float)) &&
              h2d(d_w, w.data(), w.size()*sizeof(float)) &&
              h2d(d_xn, xn.data(), xn.size()*sizeof(float)) &&
              h2d(d_sk, sk.data(), sk.size());
    ok = ok && fteik_k_inv_start(d_pos, R, starts, s) == 0;
    for (int it = 0; it <= iterations && ok; it++) {
        // forward with tangents over 4R rows
        ok = fteik_k_inv_layer1(d_pos, R, gs.d_p + gs.oW1, gs.d_p + gs.ob1,
                                d_zp, d_ha, H, s) == 0;
        unsigned short *hin = d_ha, *hout = d_hb;
        for (int l = 0; l < 3 && ok; l++) {
            // fused GEMM + activation (the GEMM then k_inv_act, bitwise)
            ok = fteik_k_inv_hidden(hin, R, gs.d_pb + gs.oW[l],
                                    gs.d_p + gs.ob[l], hout, H, s) == 0;
            std::swap(hin, hout);
        }
        ok = ok && fteik_k_gemm_bf16(hin, H, 1, gs.d_pb + gs.oW5, 1, H, d_yall,
                                     T, (int) (4*R), T, H, nullptr, 0, s) == 0
                && fteik_k_inv_gn(d_yall, R, T, gs.d_p + gs.ob5, gs.d_geo, d_o,
                                  d_w, gl.lntf, starts, d_pos, d_mis,
                                  it < iterations ? 1 : 0, s) == 0;
    }
    ok = ok && fteik_k_inv_score(d_pos, d_mis, starts, n, gs.d_geo, gl.d_tab,
                                 gl.ldgrd, d_o, d_w, d_sk, d_xn, gl.lntf,
                                 d_best, d_fb, s) == 0;
    std::vector<unsigned long long> best((size_t) n);
    std::vector<int> fb((size_t) n);
    ok = ok && d2h(best.data(), d_best, best.size()*sizeof(unsigned long long))
            && d2h(fb.data(), d_fb, fb.size()*sizeof(int));
    // full grid search for the flagged events
    std::vector<int> redo;
    for (int e = 0; e < n && ok; e++) if (fb[(size_t) e]) redo.push_back(e);
    *nFallback = (int) redo.size();
    if (ok && !redo.empty()) {
        const int m = (int) redo.size();
        std::vector<float> o2((size_t) m*gl.lntf), w2((size_t) m*gl.lntf),
                           xn2((size_t) m);
        std::vector<unsigned char> sk2((size_t) m*gl.lntf);
        std::vector<unsigned long long> b2((size_t) m, ~0ull);
        for (int q = 0; q < m; q++) {
            const size_t e = (size_t) redo[(size_t) q];
            memcpy(&o2[(size_t) q*gl.lntf], &o[e*gl.lntf], gl.lntf*sizeof(float));
            memcpy(&w2[(size_t) q*gl.lntf], &w[e*gl.lntf], gl.lntf*sizeof(float));
            memcpy(&sk2[(size_t) q*gl.lntf], &sk[e*gl.lntf], (size_t) gl.lntf);
            xn2[(size_t) q] = xn[e];
        }
        std::vector<float> h2((size_t) m*gl.lntf);
        mask_skipped(o2.data(), w2.data(), h2.data(), sk2.data(), o2.size());
        float *d_o2 = nullptr, *d_w2 = nullptr, *d_h2 = nullptr,
              *d_xn2 = nullptr;
        unsigned char *d_sk2 = nullptr;
        unsigned long long *d_b2 = nullptr;
        ok = dalloc(&d_o2, o2.size()) && dalloc(&d_w2, w2.size()) &&
             dalloc(&d_h2, h2.size()) &&
             dalloc(&d_xn2, xn2.size()) && dalloc(&d_sk2, sk2.size()) &&
             dalloc(&d_b2, b2.size()) &&
             h2d(d_o2, o2.data(), o2.size()*sizeof(float)) &&
             h2d(d_w2, w2.data(), w2.size()*sizeof(float)) &&
             h2d(d_h2, h2.data(), h2.size()*sizeof(float)) &&
             h2d(d_xn2, xn2.data(), xn2.size()*sizeof(float)) &&
             h2d(d_sk2, sk2.data(), sk2.size()) &&
             h2d(d_b2, b2.data(), b2.size()*sizeof(unsigned long long));
        if (ok) {
            int rc = fteik_k_locate_batch(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                          d_o2, d_w2, d_h2, d_xn2, gl.lntf, m,
                                          d_b2, s);
            if (rc == 2)
                for (int q = 0; q < m && ok; q++)
                    ok = fteik_k_locate(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                        d_o2 + (size_t) q*gl.lntf,
                                        d_w2 + (size_t) q*gl.lntf,
                                        d_sk2 + (size_t) q*gl.lntf, gl.lntf,
                                        xn2[(size_t) q], nullptr, nullptr, 0,
                                        d_b2 + q, s) == 0;
            else
                ok = rc == 0;
        }
        ok = ok && d2h(b2.data(), d_b2, b2.size()*sizeof(unsigned long long));
        for (int q = 0; q < m && ok; q++) best[(size_t) redo[(size_t) q]] = b2[(size_t) q];
        dfree(d_o2); dfree(d_w2); dfree(d_h2); dfree(d_xn2); dfree(d_sk2);
        dfree(d_b2);
    }
    // origin times at the chosen cells (locate.F90:684-688)
    std::vector<int> cells((size_t) n);
    for (int e = 0; e < n; e++) {
        const unsigned long long k = best[(size_t) e];
        unsigned bits = (unsigned) (k >> 32);
        float obj;
        memcpy(&obj, &bits, sizeof(obj));
        cells[(size_t) e] = (int) (k & 0xffffffffull);
        optIndx[e] = cells[(size_t) e] + 1;
        objOpt[e] = (double) obj;
    }
    // the fallback flags and misfits are consumed: their buffers take the
    // chosen cells and their origin times
    std::vector<float> t0v((size_t) n);
    int *d_c = d_fb;
    float *d_t0v = d_mis;
    ok = ok && h2d(d_c, cells.data(), (size_t) n*sizeof(int)) &&
         fteik_k_locate_t0_at(d_t0v, gl.d_tab, gl.ntf, gl.ldgrd, d_o, d_w,
                              d_sk, d_xn, gl.lntf, d_c, n, gl.ngrd, s) == 0 &&
         d2h(t0v.data(), d_t0v, (size_t) n*sizeof(float));
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    for (int e = 0; e < n; e++)
        t0Opt[e] = (double) t0v[(size_t) e] + gl.tepoch[(size_t) (evnmbr[e] - 1)];
}

} // extern "C"

extern "C" {

// Diagnostic: C = A B with A (M x K) and B (K x N) row-major FP32, rounded
// to BF16 on the device and multiplied by the surrogate's MFMA GEMM.
void fteik_amd_gemmBf16(int M, int N, int K, const float *A, const float *B,
                        float *C, int *ierr)
{
    *ierr = 0;
    if (M < 1 || N < 1 || K < 1 || !device_ok("fteik_amd_gemmBf16")) {
        *ierr = 1;
        return;
    }
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    unsigned short *hA = nullptr, *hB = nullptr;
    const size_t a = (size_t) M*K, b = (size_t) K*N, c = (size_t) M*N;
    bool ok = dalloc(&dA, a) && dalloc(&dB, b) && dalloc(&dC, c) &&
              dalloc(&hA, a) && dalloc(&hB, b) &&
              h2d(dA, A, a*sizeof(float)) && h2d(dB, B, b*sizeof(float)) &&
              fteik_k_f2bf(dA, hA, (long) a, g_dev.stream) == 0 &&
              fteik_k_f2bf(dB, hB, (long) b, g_dev.stream) == 0 &&
              fteik_k_gemm_bf16(hA, K, 1, hB, N, 1, dC, N, M, N, K, nullptr,
                                0, g_dev.stream) == 0 &&
              d2h(C, dC, c*sizeof(float));
    dfree(dA); dfree(dB); dfree(dC); dfree(hA); dfree(hB);
    if (!ok) *ierr = 1;
}

// Diagnostic: one hidden layer of the surrogate's inversion over 4 blocks
// of R rows (primal, three tangents), H = 256.  hin (4R x 256), W (256 x
// 256, row n holds output n's weights) and b (256) are rounded to BF16 /
// kept FP32 as the network keeps them; fused receives the fused kernel's
// BF16 output bits, unfused those of the MFMA GEMM followed by the
// separate activation kernel.
void fteik_amd_invHiddenBf16(int R, const float *hin, const float *W,
                             const float *b, unsigned short *fused,
                             unsigned short *unfused, int *ierr)
{
    *ierr = 0;
    const int H = 256;
    if (R < 1 || !device_ok("fteik_amd_invHiddenBf16")) { *ierr = 1; return; }
    const size_t nh = 4*(size_t) R*H, nw = (size_t) H*H;
    float *dh = nullptr, *dw = nullptr, *db = nullptr, *dz = nullptr,
          *dzp = nullptr;
    unsigned short *bh = nullptr, *bw = nullptr, *o1 = nullptr, *o2 = nullptr;
    bool ok = dalloc(&dh, nh) && dalloc(&dw, nw) && dalloc(&db, (size_t) H) &&
              dalloc(&dz, nh) && dalloc(&dzp, (size_t) R*H) &&
              dalloc(&bh, nh) && dalloc(&bw, nw) && dalloc(&o1, nh) &&
              dalloc(&o2, nh) &&
              h2d(dh, hin, nh*sizeof(float)) && h2d(dw, W, nw*sizeof(float)) &&
              h2d(db, b, H*sizeof(float)) &&
              fteik_k_f2bf(dh, bh, (long) nh, g_dev.stream) == 0 &&
              fteik_k_f2bf(dw, bw, (long) nw, g_dev.stream) == 0 &&
              fteik_k_inv_hidden(bh, R, bw, db, o1, H, g_dev.stream) == 0 &&
              fteik_k_gemm_bf16(bh, H, 1, bw, 1, H, dz, H, 4*R, H, H, nullptr,
                                0, g_dev.stream) == 0 &&
              fteik_k_inv_act(dz, R, H, db, dzp, o2, g_dev.stream) == 0 &&
              d2h(fused, o1, nh*sizeof(unsigned short)) &&
              d2h(unfused, o2, nh*sizeof(unsigned short));
    dfree(dh); dfree(dw); dfree(db); dfree(dz); dfree(dzp); dfree(bh);
    dfree(bw); dfree(o1); dfree(o2);
    if (!ok) *ierr = 1;
}

} // extern "C"
