// capi_model.cpp -- model index helpers (model.F90, fwrapper.c), the
// analytic solutions (analytic.F90), the level structures (graph3d.F90)
// and the HDF5 archive stubs.
#include "capi_internal.h"

using namespace fteik_capi;

namespace {

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

extern "C" {

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

// model.F90:794-813 on the model geometry of the solver initialized last
// (ny = 1 for the 2D solver).  With no geometry set (nz = nx = ny = -1) the
// reference's arithmetic yields out-of-range indices and ierr = 3, as here.
void fteik_model_index2gridF(int igrd, int *i, int *j, int *k, int *ierr)
{
    const int nz = g_model.nz, nx = g_model.nx, ny = g_model.ny;
    const int nzx = (nz > 0 && nx > 0) ? nz*nx : -1;
    const int g = igrd - 1;
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
