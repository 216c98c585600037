// capi_internal.h -- host-side state shared by the C-ABI source files
// (capi_*.cpp).  Not installed.  Each capi file implements one group of the
// entry points of include/fteik/fteik_amd.h; host code only validates
// inputs, computes scalar geometry (grid positions of sources and
// receivers), moves data and launches the kernels declared in kernels.h.
#pragma once
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

namespace fteik_capi {

using fteik::SourceGeom;
using fteik::SweepLaunch;

// Message to stderr with a trailing newline (the reference writes its
// errors to ERROR_UNIT).
void msg(const char *fmt, ...);

// The one GPU of the process and the library's stream.
struct Device {
    bool ready = false;
    bool failed = false;
    hipStream_t stream = nullptr;
    int ncu = 0;
    int dev = 0;
};
extern Device g_dev;

// Opens the device on first use (and registers the exit-time teardown);
// false, with a message naming `who`, when no GPU is usable.
bool device_ok(const char *who);

template <typename T>
inline bool dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **) p, n*sizeof(T)) == hipSuccess;
}

template <typename T>
inline void dfree(T *&p)
{
    if (p) hipFree((void *) p);
    p = nullptr;
}

bool h2d(void *d, const void *h, size_t bytes);
bool d2h(void *h, const void *d, size_t bytes);
bool sync_ok();
double vmin(const double *v, long n);
double vmax(const double *v, long n);     // +inf when an entry is NaN

// Largest accepted velocity and grid spacing.  Together they keep every
// argument of the iteration stencil's 2D square roots, wherever its gate is
// open (at least about 4/(v h)^2), a normal number, the range on which the
// stencil's square root is exact (stencils.h, sqrt_pos).
constexpr double kMaxVelocity = 1e30;
constexpr double kMaxSpacing = 1e30;

// Grid position of a source (source.F90:136-190).
struct SrcInfo {
    double ztrue, xtrue, ytrue;
    double zsa, xsa, ysa;
    int zsi, xsi, ysi;
};

int locate_source(bool is3d, int nz, int nx, int ny, double dz, double dx,
                  double dy, double z0, double x0, double y0, double zs,
                  double xs, double ys, int isrc, SrcInfo *o);
int locate_receivers(bool is3d, int nz, int nx, int ny, double dz, double dx,
                     double dy, double z0, double x0, double y0, int nrec,
                     const double *z, const double *x, const double *y,
                     std::vector<int> &ri, std::vector<double> &rd);
bool set_model(const char *who, bool is3d, int nz, int nx, int ny, int kind,
               int n, int order, const double *vel, double *d_slow);
bool check_eps(const char *who, double eps, int nz, int nx, int ny,
               bool is3d);

// The reference's model module state (model.F90:43, :181-216, :615-636):
// one grid geometry per process, set by whichever solver initialized last
// (ny = 1 for the 2D solver) and reset to -1 by either solver's free.
struct ModelGeom {
    int nz = -1, nx = -1, ny = -1;
    bool is3d = true;
};
extern ModelGeom g_model;
void model_set(bool is3d, int nz, int nx, int ny);
void model_free();

// ---------------------------------------------------------------- locator
// Shared by the locator entry points, the batch-to-table hand-off and the
// surrogate.
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
extern Locator gl;

int padlength32(int alignment, int n);
float weight_norm(const float *w, int lntf);
void mask_skipped(float *o, float *w, float *w2, const unsigned char *sk,
                  size_t n);

// The 3D solver's grid and device slowness (for drivers built on it).
struct Solver3DView {
    int nz = 0, nx = 0, ny = 0;
    double dz = 0, dx = 0, dy = 0, z0 = 0, x0 = 0, y0 = 0;
    const double *d_slow = nullptr;
};
// false when the 3D solver is not initialized or has no model
bool solver3d_view(Solver3DView *v);

// ------------------------------------------------------------------- rays
// Geometry of a ray trace (2D: is3d false, ny = 1).
struct RayGeom {
    bool is3d = true;
    int nz = 0, nx = 0, ny = 1;
    double dz = 0, dx = 0, dy = 1, z0 = 0, x0 = 0, y0 = 0;
};

// The last trace's rays (the reference's FTEIK_RAYS64F state).
struct RayStore {
    int integrator = 1;              // 1 Euler, 2 midpoint, 3 RK4
    int nrays = 0, maxpts = 0;
    bool is3d = true;
    double x0 = 0, y0 = 0, z0 = 0;
    std::vector<double> pts;         // [nrays][maxpts][3], grid-relative
    std::vector<int> npts, status;
    std::vector<double> time;
};
extern RayStore gr;

// Traces rays from the np points (model coordinates) to the source at
// (zsrc, xsrc, ysrc) (model coordinates) through the device gradient (d_g3
// in 3D, d_gx and d_gz in 2D) and slowness; false (with messages naming
// `who`) if a destination lies outside the model or a ray does not reach
// the source.
bool trace_rays(const RayGeom &G, const double *d_slow, const double *d_g3,
                const double *d_gx, const double *d_gz, double zsrc,
                double xsrc, double ysrc, int np, const double *x,
                const double *y, const double *z, const char *who);
void rays_clear();

} // namespace fteik_capi
