// capi_common.cpp -- device set-up and teardown, host-device copies, and the
// host bookkeeping shared by the solvers: source and receiver grid
// positions, velocity-model upload, the model geometry.
#include <cmath>
#include "capi_internal.h"

namespace fteik_capi {

void msg(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
}

Device g_dev;
ModelGeom g_model;
Locator gl;

static void teardown_at_exit(void)
{
    fteik_amd_finalize();
}

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
    g_dev.dev = dev;
    hipDeviceGetAttribute(&g_dev.ncu, hipDeviceAttributeMultiprocessorCount,
                          dev);
    if (hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking)
        != hipSuccess) {
        g_dev.failed = true;
        msg("%s: could not create a HIP stream", who);
        return false;
    }
    g_dev.ready = true;
    // Registered after the HIP runtime is up, so it runs before the
    // runtime's own exit-time teardown: the library's device memory,
    // events and stream are released while the runtime can still do so.
    static bool registered = false;
    if (!registered) {
        registered = true;
        atexit(teardown_at_exit);
    }
    return true;
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

double vmax(const double *v, long n)
{
    double m = v[0];
    for (long i = 0; i < n; i++) {
        if (v[i] != v[i]) return HUGE_VAL;       // NaN
        m = (v[i] > m) ? v[i] : m;
    }
    return m;
}

void model_set(bool is3d, int nz, int nx, int ny)
{
    g_model.is3d = is3d;
    g_model.nz = nz;
    g_model.nx = nx;
    g_model.ny = is3d ? ny : 1;
}

void model_free()
{
    g_model = ModelGeom();
}

// ------------------------------------------------ source / receiver host
// source.F90:136-190 (the grid position of a source; no field arithmetic).
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
    if (!(vmax(vel, n) <= kMaxVelocity)) {
        msg("%s: velocities must be finite and at most %g m/s", who,
            kMaxVelocity);
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

} // namespace fteik_capi

using namespace fteik_capi;

extern "C" {

// Releases every solver's device state, then the library's stream.  Runs
// at process exit (registered when the device is first opened); callers
// may also call it themselves, after which the next entry point that
// needs the GPU opens it again.
void fteik_amd_finalize(void)
{
    if (!g_dev.ready) return;
    hipStreamSynchronize(g_dev.stream);
    fteik_surrogate_free();
    fteik_batch3d_free();
    locate_finalize();
    fteik_analytic_free();
    fteik_solver2d_free();
    fteik_solver3d_free();
    hipStreamSynchronize(g_dev.stream);
    hipStreamDestroy(g_dev.stream);
    g_dev = Device();
}

const char *fteik_amd_version(void)
{
    return "fteik_amd 0.2 (gfx950, HIP)";
}

} // extern "C"
