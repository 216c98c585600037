/*
 * fteik_oracle.h -- CPU restatement of the reference fteik algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (fteik_amd/)
 * includes, links or calls this code.  It is the checker that the HIP path
 * is compared against (tests/, __graft_entry__.smoke(), bench.py's
 * cpu_baseline leg), and it is itself pinned against the reference build in
 * oracle/_ref and against the golden fixtures in tests/golden.
 *
 * Conventions: arrays are 0-based in memory; grid indices passed to and
 * returned from the functions are 1-based (Fortran numbering), exactly as in
 * the reference.  Travel-time fields are z-fastest [nz x nx x ny] ("natural"
 * ordering), slowness is z-fastest [(nz-1) x (nx-1) x (ny-1)].
 */
#ifndef FTEIK_ORACLE_H
#define FTEIK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_HUGE 99999.0

/* Source information (source.F90:82-236, :429-464). */
typedef struct
{
    double zsa, xsa, ysa;   /* fractional 1-based grid coordinates */
    int zsi, xsi, ysi;      /* 1-based cell indices */
    double szero;           /* slowness of the source cell */
} or_source;

/* Module constants of FTEIK_LOCALSOLVER3D64F (localSolver.F90:363-374). */
typedef struct
{
    double dz, dx, dy;
    double dzi, dxi, dyi, dz2, dx2, dy2, dz2i, dx2i, dy2i;
    double dsum, dsum_nine, dsumi, epsSolver;
    double dz2dx2, dz2dy2, dx2dy2;
    double dz2i_dx2i, dz2i_dy2i, dx2i_dy2i;
    double dz2i_p_dx2i, dz2i_p_dy2i, dx2i_p_dy2i;
    double dz2i_p_dy2i_inv, dz2i_p_dx2i_inv, dx2i_p_dy2i_inv;
    double szero, szero2, zsa, xsa, ysa;
    int zsi, xsi, ysi;
} or_consts3d;

/* ---------------------------------------------------------------- model */
int or_grid2index(int i, int j, int k, int nz, int nzx);
int or_velgrid2index(int i, int j, int k, int nzm1, int nzm1_nxm1);
int or_index2grid(int igrd, int nz, int nx, int ny, int *i, int *j, int *k);
int or_cell_slowness(int is3d, int nz, int nx, int ny, int order, int nc,
                     const double *vel, double *slow);
int or_nodal_slowness(int is3d, int nz, int nx, int ny, int order, int ng,
                      const double *vel, double *slow);
int or_velocity_slowness(int ncell, const double *vel, double *slow);

/* --------------------------------------------------------------- source */
int or_source_locate(int is3d, int nz, int nx, int ny,
                     double dz, double dx, double dy,
                     double z0, double x0, double y0,
                     double zs, double xs, double ys,
                     const double *slow, or_source *src);

/* ---------------------------------------------------------- local solver */
void or_consts3d_init(double dz, double dx, double dy, double eps,
                      const or_source *src, or_consts3d *c);
double or_tana3d(int i, int j, int k, const or_consts3d *c);
void or_tanad3d(int i, int j, int k, const or_consts3d *c,
                double *tanad, double *tzc, double *txc, double *tyc);
void or_seeds3d(const or_consts3d *c, int nz, int nx, int *dest, double *ts8);
double or_noinit3d(const or_consts3d *c, const double *t, const double *s,
                   double tt0);
double or_init3d(const or_consts3d *c, const double *t, const double *s,
                 double tt0, int i, int j, int k, int sweep);

/* --------------------------------------------------------- levels, masks */
void or_sweep_signs(int sweep, int *sgntz, int *sgntx, int *sgnty,
                    int *sgnvz, int *sgnvx, int *sgnvy);
int or_sweep_limits(int sweep, int linit, int nz, int nx, int ny,
                    int zsi, int xsi, int ysi, int *lim);
int or_levels3d(int sweep, int nz, int nx, int ny, int *levelPtr, int *ijkv);
void or_gather_slow3d(int sweep, int i, int j, int k, int nz, int nx, int ny,
                      const double *slow, double *s7);
void or_gather_tt3d(int sweep, int i, int j, int k, int nz, int nx,
                    const double *tt, double *t7, double *tt8);

/* ------------------------------------------------------------ 3D solver */
int or_sweep3d(int sweep, int linit, int schedule, int nz, int nx, int ny,
               const double *slow, const or_consts3d *c, double *tt);
int or_solve3d(int nz, int nx, int ny, double dz, double dx, double dy,
               const double *slow, const or_source *src, int nsweep,
               double eps, int schedule, double *tt);
int or_field_reorder3d(int nz, int nx, int ny, int order, const double *tt,
                       double *out);
int or_field_cell3d(int nz, int nx, int ny, int order, const double *tt,
                    double *out);

/* ------------------------------------------------------------ receivers */
int or_receivers_locate(int is3d, int nz, int nx, int ny, double dz,
                        double dx, double dy, double z0, double x0, double y0,
                        int nrec, const double *z, const double *x,
                        const double *y, int *ri, double *rd);
void or_receivers_interp(int is3d, int nz, int nx, double dz, double dx,
                         double dy, int nrec, const int *ri, const double *rd,
                         const double *tt, double *ttr);

/* ------------------------------------------------------------ 2D solver */
int or_solve2d(int nz, int nx, double dz, double dx, const double *slow,
               const or_source *src, int nsweep, double eps, double *tt);
double or_noinit2d(const double *t4, const double *s3, double dz, double dx);
double or_init2d(int sweep, int nz, int nx, int iz, int ix, double dz,
                 double dx, double eps, const or_source *src,
                 const double *slow, const double *tt);

/* -------------------------------------------------------------- analytic */
int or_analytic_constant(int nz, int nx, int ny, double dz, double dx,
                         double dy, double z0, double x0, double y0,
                         double zs, double xs, double ys, double vconst,
                         double *tt_xfast);
int or_analytic_gradient(int nz, int nx, int ny, double dz, double dx,
                         double dy, double z0, double x0, double y0,
                         double zs, double xs, double ys, double vtop,
                         double vbottom, double *tt_xfast);
double or_analytic_gradient_point(int nz, double dz, double z0, double zs,
                                  double xs, double ys, double vtop,
                                  double vbottom, double z, double x,
                                  double y);

/* --------------------------------------------------------------- locator */
int or_padlength32(int alignment, int n);
void or_locate_observation(int nttimes, const int *obs2tf1,
                           const double *picks, const double *wts, int lntf,
                           double *tepoch, float *obs, float *wt,
                           unsigned char *skip);
void or_locate_objective(int ntf, int ngrd, int ldgrd, const float *tables,
                         const float *obs, const float *wt,
                         const unsigned char *skip, const float *statics,
                         int lntf, float *t0work, float *objwork,
                         int nthreads);
int or_locate_event(int ntf, int ngrd, int ldgrd, const float *tables,
                    const float *obs, const float *wt,
                    const unsigned char *skip, int lntf, double tepoch,
                    int accumulate_t0, int pad_argmin, float *t0work,
                    float *objwork, int nthreads, double *t0opt,
                    double *objopt);

/* ---------------------------------------- gradients and ray tracing
 * (fteik_oracle_rays.c) */
void or_fd_gradient3d(int nz, int nx, int ny, double dz, double dx,
                      double dy, const double *tt, double *tgrad);
void or_fd_gradient2d(int nz, int nx, double dz, double dx, const double *tt,
                      double *tgradx, double *tgradz);
int or_trace_ray3d(int nz, int nx, int ny, double dz, double dx, double dy,
                   const double *slow, const double *tgrad, int integrator,
                   double xs, double ys, double zs, double xr, double yr,
                   double zr, int maxpts, double *pts, int *npts,
                   double *time);
int or_trace_ray2d(int nz, int nx, double dz, double dx, const double *slow,
                   const double *tgradx, const double *tgradz, int integrator,
                   double xs, double zs, double xr, double zr, int maxpts,
                   double *pts, int *npts, double *time);

#ifdef __cplusplus
}
#endif
#endif
