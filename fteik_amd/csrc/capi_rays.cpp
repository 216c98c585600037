// capi_rays.cpp -- the ray store shared by the 3D and 2D solvers' ray
// tracing (the reference's FTEIK_RAYS64F module state, rays.F90:7-120):
// integrator choice, destination checks, the device trace, and the getters
// of the last trace's paths and times.
#include "capi_internal.h"

namespace fteik_capi {

RayStore gr;

// rays.F90:271-347: destinations must lie in the model.
static bool check_destinations(const RayGeom &G, int np, const double *x,
                               const double *y, const double *z)
{
    const double zmax = G.z0 + (double) (G.nz - 1)*G.dz;
    const double xmax = G.x0 + (double) (G.nx - 1)*G.dx;
    const double ymax = G.is3d ? G.y0 + (double) (G.ny - 1)*G.dy : G.y0;
    bool ok = true;
    for (int i = 0; i < np; i++) {
        if (z[i] < G.z0 || z[i] > zmax) {
            msg("fteik_rays_setRayDestinations: zr(%d)=%12.5E must be in "
                "range [%12.5E%12.5E]", i + 1, z[i], G.z0, zmax);
            ok = false;
        }
        if (x[i] < G.x0 || x[i] > xmax) {
            msg("fteik_rays_setRayDestinations: xr(%d)=%12.5E must be in "
                "range [%12.5E%12.5E]", i + 1, x[i], G.x0, xmax);
            ok = false;
        }
        if (G.is3d && (y[i] < G.y0 || y[i] > ymax)) {
            msg("fteik_rays_setRayDestinations: yr(%d)=%12.5E must be in "
                "range [%12.5E%12.5E]", i + 1, y[i], G.y0, ymax);
            ok = false;
        }
    }
    return ok;
}

bool trace_rays(const RayGeom &G, const double *d_slow, const double *d_g3,
                const double *d_gx, const double *d_gz, double zsrc,
                double xsrc, double ysrc, int np, const double *x,
                const double *y, const double *z, const char *who)
{
    gr.nrays = 0;
    gr.pts.clear(); gr.npts.clear(); gr.time.clear(); gr.status.clear();
    if (np < 1) {
        msg("fteik_rays_setRayDestinations: No origins!");
        msg("%s: Failed to set ray origins", who);
        return false;
    }
    if (!check_destinations(G, np, x, y, z)) {
        msg("%s: Failed to set ray origins", who);
        return false;
    }
    // steps of min(h)/2: a path rarely exceeds 2x the grid's diagonal in
    // grid steps, so 4 (nx + ny + nz) points hold any ray that ends
    const int maxpts = 4*(G.nx + G.ny + G.nz) + 16;
    std::vector<double> start(3*(size_t) np);
    for (int i = 0; i < np; i++) {
        start[3*(size_t) i] = x[i] - G.x0;
        start[3*(size_t) i + 1] = G.is3d ? y[i] - G.y0 : 0.0;
        start[3*(size_t) i + 2] = z[i] - G.z0;
    }
    fteik::RayGrid rg;
    rg.is3d = G.is3d ? 1 : 0; rg.nz = G.nz; rg.nx = G.nx; rg.ny = G.ny;
    rg.dz = G.dz; rg.dx = G.dx; rg.dy = G.dy;
    // the true source position: the solved field's minimum, where the
    // gradient vanishes (the reference's 2D tracer aims at the source
    // cell's first node instead, rays.F90:588-594, which a ray whose source
    // lies inside a cell never reaches)
    const double xs = xsrc - G.x0;
    const double ys = G.is3d ? ysrc - G.y0 : 0.0;
    const double zs = zsrc - G.z0;
    double *d_start = nullptr, *d_pts = nullptr, *d_time = nullptr;
    int *d_n = nullptr, *d_st = nullptr;
    const size_t npt = (size_t) np*maxpts*3;
    bool ok = dalloc(&d_start, start.size()) && dalloc(&d_pts, npt) &&
              dalloc(&d_time, (size_t) np) && dalloc(&d_n, (size_t) np) &&
              dalloc(&d_st, (size_t) np) &&
              h2d(d_start, start.data(), start.size()*sizeof(double)) &&
              fteik_k_trace_rays(&rg, d_slow, d_g3, d_gx, d_gz,
                                 gr.integrator, xs, ys, zs, d_start, np,
                                 maxpts, d_pts, d_n, d_time, d_st,
                                 g_dev.stream) == 0;
    if (ok) {
        gr.pts.resize(npt);
        gr.npts.resize((size_t) np);
        gr.time.resize((size_t) np);
        gr.status.resize((size_t) np);
        ok = d2h(gr.pts.data(), d_pts, npt*sizeof(double)) &&
             d2h(gr.npts.data(), d_n, (size_t) np*sizeof(int)) &&
             d2h(gr.time.data(), d_time, (size_t) np*sizeof(double)) &&
             d2h(gr.status.data(), d_st, (size_t) np*sizeof(int));
    }
    dfree(d_start); dfree(d_pts); dfree(d_time); dfree(d_n); dfree(d_st);
    if (!ok) { msg("%s: GPU error", who); return false; }
    gr.nrays = np;
    gr.maxpts = maxpts;
    gr.is3d = G.is3d;
    gr.x0 = G.x0; gr.y0 = G.y0; gr.z0 = G.z0;
    bool all = true;
    for (int i = 0; i < np; i++)
        if (gr.status[(size_t) i]) {
            msg("fteik_rays_trace: Ray path %d failed to terminate", i + 1);
            all = false;
        }
    if (!all) msg("%s: Failed to compute rays", who);
    return all;
}

void rays_clear()
{
    gr.nrays = 0;
    gr.pts.clear(); gr.npts.clear(); gr.time.clear(); gr.status.clear();
}

} // namespace fteik_capi

using namespace fteik_capi;

extern "C" {

// rays.F90:211-224 (Euler, midpoint or RK4; Euler after an invalid value)
void fteik_rays_setODEIntegrator(int integrator, int *ierr)
{
    *ierr = 0;
    gr.integrator = 1;
    if (integrator < 1 || integrator > 3) {
        msg("fteik_rays_setODEIntegrator: Invalid integrator = %d using Euler",
            integrator);
        *ierr = 1;
        return;
    }
    gr.integrator = integrator;
}

void fteik_rays_getNumberOfRays(int *nrays, int *ierr)
{
    *ierr = 0;
    *nrays = gr.nrays;
}

// Travel times along the last trace's rays (seconds); nrays must match.
void fteik_rays_getTravelTimes64f(int nrays, double *t, int *ierr)
{
    *ierr = 0;
    if (gr.nrays < 1 || nrays != gr.nrays) {
        msg("fteik_rays_getTravelTimes64f: %d rays traced, %d requested",
            gr.nrays, nrays);
        *ierr = 1;
        return;
    }
    for (int i = 0; i < nrays; i++) t[i] = gr.time[(size_t) i];
}

// Points (x, y, z in model coordinates; y = 0 for 2D rays) of ray iray
// (1-based) from the receiver to the source: npts points into xyz[3*nwork].
void fteik_rays_getRayPath(int iray, int nwork, double *xyz, int *npts,
                           double *time, int *ierr)
{
    *ierr = 0;
    *npts = 0;
    *time = 0.0;
    if (iray < 1 || iray > gr.nrays) {
        msg("fteik_rays_getRayPath: ray %d not in [1,%d]", iray, gr.nrays);
        *ierr = 1;
        return;
    }
    const size_t r = (size_t) (iray - 1);
    const int n = gr.npts[r];
    *npts = n;
    *time = gr.time[r];
    if (nwork < n) {
        msg("fteik_rays_getRayPath: nwork = %d < %d points", nwork, n);
        *ierr = 1;
        return;
    }
    const double *p = gr.pts.data() + r*(size_t) gr.maxpts*3;
    for (int k = 0; k < n; k++) {
        xyz[3*k] = p[3*k] + gr.x0;
        xyz[3*k + 1] = gr.is3d ? p[3*k + 1] + gr.y0 : 0.0;
        xyz[3*k + 2] = p[3*k + 2] + gr.z0;
    }
}

} // extern "C"
