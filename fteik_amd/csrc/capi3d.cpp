// capi3d.cpp -- the 3D level-set solver entry points (solver3D.F90):
// model, sources, receivers, solve, getters, and the sweep tuning
// extensions.
#include "capi_internal.h"

using namespace fteik_capi;

namespace {

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
    std::vector<double> zr, xr, yr;      // receiver positions (model coords)
    int nrec = 0;
    bool haveRec = false;
    double *d_grad = nullptr;            // [x y z] per node, on demand
    bool haveGrad = false;
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
    dfree(g3.d_ri); dfree(g3.d_rd); dfree(g3.d_ttr); dfree(g3.d_grad);
    if (g3.ev0) hipEventDestroy(g3.ev0);
    if (g3.ev1) hipEventDestroy(g3.ev1);
    g3 = Solver3D();
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

} // namespace

bool fteik_capi::solver3d_view(Solver3DView *v)
{
    if (!g3.init || !g3.haveModel) return false;
    v->nz = g3.nz; v->nx = g3.nx; v->ny = g3.ny;
    v->dz = g3.dz; v->dx = g3.dx; v->dy = g3.dy;
    v->z0 = g3.z0; v->x0 = g3.x0; v->y0 = g3.y0;
    v->d_slow = g3.d_slow;
    return true;
}

extern "C" {

void fteik_solver3d_free(void)
{
    free3d();
    model_free();
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
    model_free();
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
    if (!(dz <= kMaxSpacing && dx <= kMaxSpacing && dy <= kMaxSpacing)) {
        msg("fteik_solver3d_initialize: grid spacing must be finite and at "
            "most %g", kMaxSpacing);
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
    // With three or more 8 x 8 tiles per CU the staged-load variant of that
    // tile comes first (257^3: 0.81 against 0.88 ms per sweep); on smaller
    // grids its extra LDS reads per step cost more than its fewer memory
    // requests save (129^2 x 257: 0.88 against 0.84 us per level).
    static const int pref_many[] = {18, 12, 9, 11, 8, 7, 10, 6, 3, 1, 0, 5, 2,
                                    4};
    static const int pref_large[] = {12, 18, 9, 11, 8, 7, 10, 6, 3, 1, 0, 5,
                                     2, 4};
    static const int pref_small[] = {12, 18, 9, 11, 8, 10, 7, 6, 3, 1, 0, 5,
                                     2, 4};
    const long tiles8 = (long) ((nx - 1 + 7)/8)*((ny - 1 + 7)/8);
    const bool many = tiles8 >= 3L*g_dev.ncu;
    const bool large = (long) (nx - 1)*(ny - 1) >= 128L*g_dev.ncu;
    int first = -1;
    for (int t : (many ? pref_many : large ? pref_large : pref_small))
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
    model_set(true, nz, nx, ny);
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
    g3.zr.assign(zrec, zrec + nrec);
    g3.xr.assign(xrec, xrec + nrec);
    g3.yr.assign(yrec, yrec + nrec);
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
    if (!(bias > 0.0 && bias < 0.1) || maxCorrections < 1 ||
        maxCorrections > 64) {
        msg("fteik_solver3d_setMixedParameters: need 0 < bias < 0.1 and "
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

// Times a complete solve of source isrc with every tile shape usable on
// this grid (in the current precision mode, skipping and overlap settings)
// and keeps the fastest: the FP64 sweeps' shape in FP64 mode, both shapes
// in the mixed mode.  Every shape gives the same field bit for bit (any
// order that respects the levels rounds alike), so the solver is left
// holding the solved field of isrc.
void fteik_solver3d_autotune(int isrc, int *tile, int *tile32, int *ierr)
{
    static const char *who = "fteik_solver3d_autotune";
    *ierr = 0;
    *tile = g3.tile;
    *tile32 = g3.tile32;
    if (!g3.init || !g3.haveModel || isrc < 1 || isrc > (int) g3.src.size()) {
        msg("%s: solver, model or source %d not set", who, isrc);
        *ierr = 1;
        return;
    }
    if (g3.schedule != 0) {
        msg("%s: the per-level schedule has no tile shape to tune", who);
        *ierr = 1;
        return;
    }
    const int t0 = g3.tile, t032 = g3.tile32;
    const bool mixed = g3.precision == 1;
    int best = -1;
    double bestMs = 0.0;
    for (int t = 0; t < (int) g3.tileOK.size(); t++) {
        if (!g3.tileOK[(size_t) t]) continue;
        g3.tile = t;
        g3.tile32 = mixed ? t : t032;
        int e = 0;
        fteik_solver3d_solveSourceLSM(isrc, &e);
        if (e == 1) continue;                  // shape failed: not a winner
        if (g3.verbose > 1)
            printf("%s: tile %d: %.3f ms\n", who, t, g3.last_ms);
        if (best < 0 || g3.last_ms < bestMs) { best = t; bestMs = g3.last_ms; }
    }
    if (best < 0) {
        g3.tile = t0;
        g3.tile32 = t032;
        msg("%s: no tile shape solved source %d", who, isrc);
        *ierr = 1;
        return;
    }
    g3.tile = best;
    g3.tile32 = mixed ? best : t032;
    int e = 0;
    fteik_solver3d_solveSourceLSM(isrc, &e);    // leave the winner's solve
    if (e == 1) { *ierr = 1; return; }
    *tile = g3.tile;
    *tile32 = g3.tile32;
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
    g3.haveGrad = false;
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
    bool converged = true;
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
        converged = false;
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
            if (ok && md <= g3.convTol) converged = true;
            if (!ok || converged) break;
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
    if (!converged) {
        // the field is kept, but it need not be the FP64 solution
        msg("%s: mixed precision: %d correction iterations did not converge "
            "(raise maxCorrections with fteik_solver3d_setMixedParameters)",
            who, corr);
        *ierr = 2;
    }
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

// solver3D.F90:1249-1321 on the device, once per solved field.
static bool gradient3d(const char *who)
{
    if (g3.haveGrad) return true;
    const long ngrd = (long) g3.nz*g3.nx*g3.ny;
    if (!g3.d_grad && !dalloc(&g3.d_grad, 3*(size_t) ngrd)) {
        msg("%s: out of GPU memory (gradient)", who);
        return false;
    }
    if (fteik_k_fd_gradient3d(g3.d_grad, g3.d_tt, g3.nz, g3.nx, g3.ny, g3.dz,
                              g3.dx, g3.dy, g_dev.stream) != 0 || !sync_ok()) {
        msg("%s: GPU error (gradient)", who);
        return false;
    }
    g3.haveGrad = true;
    return true;
}

// solver3D.F90:784-822: rays from the points to the solved source, down
// the finite-difference gradient of the field (the reference computes the
// gradient and sets the destinations; its 3D tracer, rays.F90:442-447, is
// an empty stub, so the tracing itself is this library's).
void fteik_solver3d_traceRaysToPoints(int np, const double *zp,
                                      const double *xp, const double *yp,
                                      int *ierr)
{
    static const char *who = "fteik_solver3d_traceRaysToPoints";
    *ierr = 0;
    if (!g3.haveTimes || g3.srcNumber < 1) {
        msg("%s: Travel time field not yet computed", who);
        *ierr = 1;
        return;
    }
    if (!gradient3d(who)) { *ierr = 1; return; }
    RayGeom G;
    G.is3d = true; G.nz = g3.nz; G.nx = g3.nx; G.ny = g3.ny;
    G.dz = g3.dz; G.dx = g3.dx; G.dy = g3.dy;
    G.z0 = g3.z0; G.x0 = g3.x0; G.y0 = g3.y0;
    const SrcInfo &s = g3.src[(size_t) (g3.srcNumber - 1)];
    if (!trace_rays(G, g3.d_slow, g3.d_grad, nullptr, nullptr, s.ztrue,
                    s.xtrue, s.ytrue, np, xp, yp, zp, who))
        *ierr = 1;
}

// solver3D.F90:756-777
void fteik_solver3d_traceRaysToReceivers(int *ierr)
{
    *ierr = 0;
    if (!g3.haveRec || g3.nrec < 1) {
        msg("fteik_solver3d_traceRaysToReceivers: No receivers set");
        *ierr = 1;
        return;
    }
    fteik_solver3d_traceRaysToPoints(g3.nrec, g3.zr.data(), g3.xr.data(),
                                     g3.yr.data(), ierr);
    if (*ierr != 0) {
        msg("fteik_solver3d_traceRaysToReceivers: Failed to compute rays");
        *ierr = 1;
    }
}

// The gradient of the solved field (solver3D.F90:1249-1321): [x y z] per
// node, natural node order, 3*ngrd values.
void fteik_solver3d_getTravelTimeGradient64f(int ngin, double *tgrad,
                                             int *ierr)
{
    static const char *who = "fteik_solver3d_getTravelTimeGradient64f";
    *ierr = 0;
    if (!g3.haveTimes) {
        msg("%s: Travel times not yet computed", who);
        *ierr = 1;
        return;
    }
    const long ngrd = (long) g3.nz*g3.nx*g3.ny;
    if (ngin != ngrd) {
        msg("%s: ngin=%d /= ngrd =%ld", who, ngin, ngrd);
        *ierr = 1;
        return;
    }
    if (!gradient3d(who) ||
        !d2h(tgrad, g3.d_grad, 3*(size_t) ngrd*sizeof(double))) *ierr = 1;
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

void fteik_amd_sqrtStencil64f(int n, const double *x, double *y, int *ierr)
{
    // The iteration stencil's square root of its 2D operators, evaluated on
    // the GPU for n caller-given arguments (exact for normal x >= 2^-767).
    static const char *who = "fteik_amd_sqrtStencil64f";
    *ierr = 0;
    if (n < 1 || !x || !y) { msg("%s: n must be positive", who); *ierr = 1;
                             return; }
    if (!device_ok(who)) { *ierr = 1; return; }
    double *d_x = nullptr, *d_y = nullptr;
    bool ok = dalloc(&d_x, (size_t) n) && dalloc(&d_y, (size_t) n) &&
              h2d(d_x, x, (size_t) n*sizeof(double)) &&
              fteik_k_sqrt_pos(d_x, d_y, n, g_dev.stream) == 0 &&
              d2h(y, d_y, (size_t) n*sizeof(double));
    dfree(d_x);
    dfree(d_y);
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; }
}

void fteik_amd_timeStencilChain3d(int precision, int nsteps,
                                  double *usPerStep, int *ierr)
{
    // Device time of one lane's dependent chain of nsteps iteration
    // stencils, with one wave per SIMD (4 per CU) on the whole GPU: the
    // latency floor of one sweep step.  Uses the mesh constants of the last
    // prepared source.
    static const char *who = "fteik_amd_timeStencilChain3d";
    *ierr = 0;
    *usPerStep = 0.0;
    if (!g3.cstReady || nsteps < 1 || (precision != 0 && precision != 1)) {
        msg("%s: needs a prepared source, nsteps >= 1, precision 0 or 1", who);
        *ierr = 1;
        return;
    }
    const int nwg = 4*std::max(1, g_dev.ncu);
    double *d_out = nullptr;
    bool ok = dalloc(&d_out, 64*(size_t) nwg);
    // warm-up, then the timed launch
    ok = ok && fteik_k_time_stencil_chain(g3.d_cst, 16, nwg, precision,
                                          d_out, g_dev.stream) == 0;
    hipEventRecord(g3.ev0, g_dev.stream);
    ok = ok && fteik_k_time_stencil_chain(g3.d_cst, nsteps, nwg, precision,
                                          d_out, g_dev.stream) == 0;
    hipEventRecord(g3.ev1, g_dev.stream);
    ok = ok && sync_ok();
    float ms = 0.0f;
    if (ok) hipEventElapsedTime(&ms, g3.ev0, g3.ev1);
    dfree(d_out);
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    *usPerStep = 1000.0*(double) ms/(double) nsteps;
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

} // extern "C"
