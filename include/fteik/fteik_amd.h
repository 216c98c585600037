/*
 * fteik_amd.h -- C ABI of the MI355X fteik library (libfteik_shared.so).
 *
 * Every entry point below has the name, argument list and error convention
 * of the reference's BIND(C) export it replaces (cited file:line of the
 * reference), so existing C, Fortran and ctypes callers link against this
 * library unchanged:
 *   - scalars by value, arrays by pointer, `int *ierr` last
 *     (0 = success, 1 = failure with a message on stderr);
 *   - sources, receivers, fields and event numbers are 1-based exactly
 *     where the reference's are (isrc, itf, evnmbr, obs2tf, optIndx);
 *   - inputs are copied into library-owned (GPU) state; outputs are copied
 *     into caller buffers; state is one process-global solver of each kind.
 * Entry points marked EXTENSION have no counterpart in the reference.
 *
 * Ordering constants (constants.F90:28-39).
 */
#ifndef FTEIK_AMD_H
#define FTEIK_AMD_H
#include <stdbool.h>

#define FTEIK_NATURAL_ORDERING 0
#define FTEIK_ZXY_ORDERING     0
#define FTEIK_ZX_ORDERING      0
#define FTEIK_XYZ_ORDERING     1
#define FTEIK_XZ_ORDERING      1
#define FTEIK_ZYX_ORDERING     2

#ifdef __cplusplus
extern "C" {
#endif

/* ============================ 3D solver ============================== */
/* solver3D.F90:126-132 */
void fteik_solver3d_initialize(int nz, int nx, int ny,
                               double z0, double x0, double y0,
                               double dz, double dx, double dy,
                               int nsweep, double eps, double convTol,
                               int verbose, int *ierr);
/* solver3D.F90:575 */
void fteik_solver3d_free(void);
/* solver3D.F90:1138 */
void fteik_solver3d_setVelocityModel64f(int ncell, const double *vel,
                                        int *ierr);
/* solver3D.F90:1083 */
void fteik_solver3d_setNodalVelocityModel64f(int ng, int order,
                                             const double *vel, int *ierr);
/* solver3D.F90:1115 */
void fteik_solver3d_setCellVelocityModel64f(int nc, int order,
                                            const double *vel, int *ierr);
/* solver3D.F90:717 */
void fteik_solver3d_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, const double *ysrc,
                                  int *ierr);
/* solver3D.F90:1045 */
void fteik_solver3d_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, const double *yrec,
                                    int *ierr);
/* solver3D.F90:1156 */
void fteik_solver3d_solveSourceLSM(int isrc, int *ierr);
/* solver3D.F90:990 */
void fteik_solver3d_getTravelTimes64f(int nrec, double *ttr, int *ierr);
/* solver3D.F90:844 */
void fteik_solver3d_getTravelTimeField64f(int ngrd, int order, double *tt,
                                          int *ierr);
/* solver3D.F90:915 */
void fteik_solver3d_getTravelTimeFieldCell64f(int ncell, int order,
                                              double *tt, int *ierr);
/* solver3D.F90:737 */
void fteik_solver3d_getNumberOfSources(int *nsrc, int *ierr);
/* solver3D.F90:1024 */
void fteik_solver3d_getNumberOfReceivers(int *nrec, int *ierr);
/* solver3D.F90:683 */
void fteik_solver3d_setNumberOfSweeps(int nsweep, int *ierr);
/* solver3D.F90:640 */
void fteik_solver3d_setSphereToCartEpsilon(double eps, int *ierr);
/* solver3D.F90:191 */
void fteik_solver3d_setVerbosity(int verbose);
/* solver3D.F90:206 */
void fteik_solver3d_setConvergenceTolerance(double convTol);
/* solver3D.F90:756, :784: rays from the receivers / the points (model
 * coordinates) to the solved source, down the finite-difference gradient
 * of the field (dx/ds = -grad T / u, integrator of
 * fteik_rays_setODEIntegrator).  Paths and times are read with the
 * fteik_rays_get* functions.  The reference computes the gradient and
 * checks the destinations but its 3D tracer is an empty stub
 * (rays.F90:442-447); this library traces the rays on the GPU.  ierr = 1
 * if the field is not solved, a point lies outside the model or a ray does
 * not reach the source. */
void fteik_solver3d_traceRaysToReceivers(int *ierr);
void fteik_solver3d_traceRaysToPoints(int np, const double *zp,
                                      const double *xp, const double *yp,
                                      int *ierr);
/* EXTENSION: the finite-difference gradient of the solved field
 * (solver3D.F90:1249-1321: forward differences, backward on the last
 * face), [dT/dx dT/dy dT/dz] per node in natural order: tgrad[3*ngrd]. */
void fteik_solver3d_getTravelTimeGradient64f(int ngrd, double *tgrad,
                                             int *ierr);
/* EXTENSION: 0 = FP64 sweeps (reference semantics, default);
 * 1 = mixed: FP64 initialization sweeps, nsweep FP32 Gauss-Seidel
 *     iterations, promotion to FP64 with a small upward margin, then FP64
 *     iterations until the largest update of an iteration is <= convTol
 *     seconds (see fteik_solver3d_setMixedParameters). */
void fteik_solver3d_setPrecision(int mode, int *ierr);
/* EXTENSION: 0 = pipelined column-block sweeps (default);
 * 1 = one kernel launch per level (diagnostic schedule). */
void fteik_solver3d_setSchedule(int schedule, int *ierr);
/* EXTENSION: mixed-mode parameters: the single-precision field is raised
 * by the relative margin bias (default 2^-6, 0 < bias < 0.1) when it is
 * promoted, so it stays an upper bound of the double-precision solution and
 * the double correction iterations (at most maxCorrections, default 64)
 * converge onto it.  A correction phase that reaches maxCorrections without
 * converging fails the solve with ierr = 1.  Reset to the defaults by
 * fteik_solver3d_initialize. */
void fteik_solver3d_setMixedParameters(double bias, int maxCorrections,
                                       int *ierr);
/* EXTENSION: tile shape of the pipelined FP64 and FP32 sweeps (index into
 * the library's table of workgroup shapes and halo batch sizes; chosen by
 * fteik_solver3d_autotune, or from a measured preference list at
 * initialize).  Flag tiles (mode 0) are accepted only when all their
 * workgroups can be resident; mailbox tiles (modes 1, 2) also launch
 * without co-residency: tiles are dealt to workgroups so that a tile waits
 * only on tiles of lower-numbered workgroups, and workgroups are dispatched
 * in order, so the lowest unfinished tile is always resident.  Every wait has a spin limit; a sweep that exceeds it
 * fails the solve with ierr = 1 rather than hanging. */
void fteik_solver3d_setTile(int tile, int tile32, int *ierr);
void fteik_solver3d_getTile(int *tile, int *tile32, int *ntiles, int *ierr);
/* EXTENSION: the autotuner.  Solves source isrc (1-based) once with every
 * tile shape usable on this grid, in the current precision, tile-skipping
 * and overlap settings, keeps the fastest (the FP64 shape in FP64 mode,
 * both shapes in the mixed mode) and returns it.  All shapes give the same
 * field bit for bit; the solver is left holding the solved field of isrc. */
void fteik_solver3d_autotune(int isrc, int *tile, int *tile32, int *ierr);
/* EXTENSION: skip, in the Gauss-Seidel iterations, every tile of the
 * pipelined sweep whose columns have not changed since its last pass in the
 * same direction and whose upwind tiles are skipped too: such a pass would
 * reproduce the tile's values exactly, so results are unchanged (on by
 * default; mailbox tile shapes only). */
void fteik_solver3d_setTileSkipping(int on, int *ierr);
/* EXTENSION: run the Gauss-Seidel iteration sweeps of a solve as one chain
 * in which each sweep starts while the one before it is still running,
 * wherever the nodes it reads and writes are already final in the one
 * before (results are bitwise those of one sweep after another).  on = 1:
 * chains on grids whose sweep has at most one tile per CU, where a lone
 * sweep leaves CUs idle; 2: on every grid; 0 (default): never.  A chain
 * does every update of every sweep, so it pays only while most tiles still
 * change; tile skipping does not apply inside a chain.  getSweepOverlap
 * reports whether FP64 solves of this grid chain their sweeps and the tile
 * shapes of the FP64 and FP32 chains (-1: none). */
void fteik_solver3d_setSweepOverlap(int on, int *ierr);
/* EXTENSION: tile shapes of the FP64 and FP32 sweep chains (store-wave
 * mailbox shapes whose two sweeps' tiles fit the GPU; chosen by the library
 * at initialize). */
void fteik_solver3d_setChainTile(int tile, int tile32, int *ierr);
void fteik_solver3d_getSweepOverlap(int *on, int *tile, int *tile32,
                                    int *ierr);
/* EXTENSION: device time of the last solve (ms, hipEvent), number of FP64
 * correction iterations run (mixed mode) and node updates performed. */
void fteik_solver3d_getLastSolveInfo(double *ms, int *corrIterations,
                                     double *updates, int *ierr);

/* ============================ 2D solver ============================== */
/* solver2D.F90:124 */
void fteik_solver2d_initialize(int nz, int nx, double z0, double x0,
                               double dz, double dx, int nsweep, double eps,
                               double convTol, int verbose, int *ierr);
/* solver2D.F90:233 */
void fteik_solver2d_free(void);
/* solver2D.F90:1688, :1637, :1665 */
void fteik_solver2d_setVelocityModel64f(int ncell, const double *vel,
                                        int *ierr);
void fteik_solver2d_setNodalVelocityModel64f(int ng, int order,
                                             const double *vel, int *ierr);
void fteik_solver2d_setCellVelocityModel64f(int nc, int order,
                                            const double *vel, int *ierr);
/* solver2D.F90:1554 */
void fteik_solver2d_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, int *ierr);
/* solver2D.F90:1512 */
void fteik_solver2d_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, int *ierr);
/* solver2D.F90:782, :1042, :740 */
void fteik_solver2d_solveSourceLSM(int isrc, int *ierr);
void fteik_solver2d_solveSourceFSM(int isrc, int *ierr);
void fteik_solver2d_solveLSM(int *ierr);
/* solver2D.F90:1295, :1345 */
void fteik_solver2d_getTravelTimeField64f(int ngrd, int order, double *tt,
                                          int *ierr);
void fteik_solver2d_getTravelTimeFieldCell64f(int ncell, int order,
                                              double *tt, int *ierr);
/* solver2D.F90:1398: ttr is [ldr x nsrc] */
void fteik_solver2d_getTravelTimes64f(int ldr, double *ttr, int *ierr);
/* solver2D.F90:268, :1485, :1593, :205, :221 */
void fteik_solver2d_getNumberOfReceivers(int *nrec, int *ierr);
void fteik_solver2d_setNumberOfSweeps(int nsweep, int *ierr);
void fteik_solver2d_setSphereToCartEpsilon(double eps, int *ierr);
void fteik_solver2d_setVerbosity(int verbose);
void fteik_solver2d_setConvergenceTolerance(double convTol);
/* EXTENSION: 0 = pipelined columns, one lane per x column walking down z,
 * no barrier per level (default; grids with nx <= 513, otherwise the
 * level schedule runs); 1 = level by level with a workgroup barrier per
 * level (diagnostic).  Both give the same results bit for bit. */
void fteik_solver2d_setSchedule(int schedule, int *ierr);
/* solver2D.F90:283, :310: rays to the last solved source (as the 3D
 * functions; the reference's 2D tracer, rays.F90:449-669, stops after its
 * first ray). */
void fteik_solver2d_traceRaysToReceivers(int *ierr);
void fteik_solver2d_traceRaysToPoints(int np, const double *xp,
                                      const double *zp, int *ierr);
/* EXTENSION: the finite-difference gradient of the solved 2D field
 * (solver2D.F90:984-1032), dT/dx and dT/dz per node in natural order. */
void fteik_solver2d_getTravelTimeGradient64f(int ngrd, double *tgradx,
                                             double *tgradz, int *ierr);

/* ================================ Rays =============================== */
/* rays.F90:211: 1 = Euler (default), 2 = midpoint, 3 = classical RK4; an
 * invalid value keeps Euler and returns ierr = 1. */
void fteik_rays_setODEIntegrator(int integrator, int *ierr);
/* EXTENSION: the last trace's rays (the reference keeps them in its
 * module, rays.F90:23, without getters).  getRayPath: npts points
 * (x, y, z; model coordinates, y = 0 for 2D) from the receiver to the
 * source into xyz[3*nwork], and the ray's time (seconds: path length times
 * cell slowness, summed over the steps). */
void fteik_rays_getNumberOfRays(int *nrays, int *ierr);
void fteik_rays_getTravelTimes64f(int nrays, double *t, int *ierr);
void fteik_rays_getRayPath(int iray, int nwork, double *xyz, int *npts,
                           double *time, int *ierr);

/* ============================== Locator ============================== */
/* locate.F90:142 */
void locate_initialize(int nEvents, int ntf, int ngrd, int *ierr);
/* locate.F90:233 */
void locate_finalize(void);
/* locate.F90:287 */
void locate_setTravelTimeField64f(int ngrd, int itf, const double *tt,
                                  int *ierr);
/* locate.F90:361 (float, although fteik_fortran.h:208 says double) */
void locate_setTravelTimeField32f(int ngrd, int itf, const float *tt,
                                  int *ierr);
/* locate.F90:431 */
void locate_setObservation64f(int evnmbr, int nttimes, bool lhaveOT,
                              double t0, const int *obs2tf,
                              const double *pickTimes, const double *wts,
                              int *ierr);
/* fwrapper.c:42: 0-based evnmbr and obs2tf */
void locate_setObservation64fC(const int evnmbr, const int nttimes,
                               const bool lhaveOT, const double t0In,
                               const int *obs2tf, const double *pickTimes,
                               const double *wts, int *ierr);
/* locate.F90:511, :541, :573 */
void locate_setNumberOfEvents(int nEvents, int *ierr);
void locate_setNumberOfTravelTimeFields(int ntf, int *ierr);
void locate_setNumberOfGridPointsF(int ngrd, int *ierr);
/* locate.F90:112: optIndx is 1-based */
void locate_locateEvent(int evnmbr, int *optIndx, double *t0Opt,
                        double *objOpt);
/* locate.F90:748 (all arguments by reference) */
void locate_computeStatics32fF(const int *ldo, const int *nobs,
                               const int *nev, const float *tobs,
                               const float *tt, const float *t0,
                               const float *wts, float *tstatic);
/* EXTENSION: 1 (default) keeps the reference's behaviour exactly: the
 * origin-time work array carries over from one locateEvent call to the next
 * (locate.F90:212, :687) and, for fewer than 512 grid points, the zero
 * padding of the objective array takes part in the minimum search
 * (locate.F90:122, :200-203).  0 gives every event a fresh origin-time
 * estimate and searches the grid points only. */
void locate_setReferenceCompatibility(int mode, int *ierr);
/* EXTENSION: locates events evnmbr[0..n-1] (1-based) in one pass over the
 * tables, each with a fresh origin-time estimate. */
void locate_locateEvents(int n, const int *evnmbr, int *optIndx,
                         double *t0Opt, double *objOpt, int *ierr);
/* EXTENSION (no counterpart in the reference): locate_locateEvents without
 * any device allocation, for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Per-call
 * device buffers come from one static device array of about 3.4 MB per GPU,
 * held until the process ends: at most 4096 events per call and at most 64
 * padded observations per event (ntf <= 64), else ierr = 1.  Results equal
 * locate_locateEvents'.  Calls that overlap on one GPU share the array and
 * race. */
void locate_locateEventsNoAlloc(int n, const int *evnmbr, int *optIndx,
                                double *t0Opt, double *objOpt, int *ierr);

/* ========================= Batched 3D solver ========================= */
/* EXTENSION (no counterpart in the reference, which solves one source at a
 * time, solver3D.F90:1156): a batch of ngrids independent single-source
 * solves, each on its own grid (sizes may differ), advanced together: one
 * launch per sweep direction covers every grid.  Each grid follows
 * fteik_solver3d_solveSourceLSM exactly (FP64 fields are bitwise those of
 * the single-grid solver), or the mixed mode of fteik_solver3d_setPrecision.
 * vel holds (nz-1)(nx-1)(ny-1) cell velocities in natural order. */
void fteik_batch3d_initialize(int ngrids, int *ierr);
void fteik_batch3d_free(void);
void fteik_batch3d_setGrid(int igrid, int nz, int nx, int ny, double z0,
                           double x0, double y0, double dz, double dx,
                           double dy, const double *vel, double zs, double xs,
                           double ys, int *ierr);
void fteik_batch3d_setPrecision(int mode, int *ierr);
/* promotion margin and correction cap of the batch's mixed mode (see
 * fteik_solver3d_setMixedParameters) */
void fteik_batch3d_setMixedParameters(double bias, int maxCorrections,
                                      int *ierr);
/* tile shape of the batch (see fteik_solver3d_setTile), flag (mode 0) or
 * mailbox (modes 1, 2); it serves every sweep (default: mailbox tile 12) */
void fteik_batch3d_setTile(int tile, int *ierr);
/* EXTENSION: the batch autotuner: times fteik_batch3d_solve(nsweep, eps)
 * with every tile shape (flag and mailbox tiles), keeps the fastest and
 * returns it; the batch is left solved with it. */
void fteik_batch3d_autotune(int nsweep, double eps, int *tile, int *ierr);
void fteik_batch3d_solve(int nsweep, double eps, int *ierr);
void fteik_batch3d_getLastSolveInfo(double *ms, int *corrIterations,
                                    double *updates, int *ierr);
void fteik_batch3d_getTravelTimeField64f(int igrid, int ngrd, int order,
                                         double *tt, int *ierr);
void fteik_batch3d_getNumberOfGrids(int *ngrids, int *ierr);
/* EXTENSION: locator table itf (1-based) from batch grid igrid without a
 * host round trip; same rounding as locate_setTravelTimeField64f. */
void locate_setTravelTimeFieldFromBatch3d(int igrid, int itf, int *ierr);
/* EXTENSION: the reciprocal station-as-source driver of fteik.c:186-235
 * (with lreciprocal, receivers become sources), on the device: the nsta
 * stations (model coordinates) are solved as sources on the 3D solver's
 * grid and model in one batch (fteik_batch3d_*, precision 0 = FP64, 1 =
 * mixed; nsweep, eps as for the solver) and station s's field becomes
 * locator table itf0 + s (1-based), rounded as
 * locate_setTravelTimeField64f rounds.  Needs fteik_solver3d_initialize +
 * a velocity model, and locate_initialize with ngrd = nz*nx*ny and
 * ntf >= itf0 + nsta - 1.  The batch is left solved (its fields can be
 * read with fteik_batch3d_getTravelTimeField64f).  ierr = 2: the mixed
 * mode's corrections did not converge (tables still set). */
void locate_setTravelTimeFieldsFromStations3d(int nsta, const double *zsta,
                                              const double *xsta,
                                              const double *ysta, int nsweep,
                                              double eps, int precision,
                                              int itf0, int *ierr);

/* ===================== Learned locator surrogate ===================== */
/* EXTENSION (no counterpart in the reference): an MLP 3 -> 256 -> 256 ->
 * 256 -> 256 -> ntf (SiLU) from a normalised hypocentre to the stations'
 * travel times, trained on the GPU (BF16 MFMA, FP32 accumulation, Adam) on
 * cells sampled from the locator's tables.  It predicts
 * (T - |x - station|/v0)/scale.  Stations are given in metres relative to
 * the grid origin.  Positions are normalised to [-1, 1]^3 as (x, y, z). */
void fteik_surrogate_initialize(int ntf, int nz, int nx, int ny, double dz,
                                double dx, double dy, const double *zsta,
                                const double *xsta, const double *ysta,
                                double v0, double scale, int seed, int *ierr);
void fteik_surrogate_free(void);
void fteik_surrogate_getNumberOfParameters(int *np, int *ierr);
void fteik_surrogate_getParameters(int np, float *p, int *ierr);
void fteik_surrogate_setParameters(int np, const float *p, int *ierr);
/* network outputs (before baseline and scale) at n positions */
void fteik_surrogate_forward(int n, const float *pos, float *out, int *ierr);
/* one forward/backward pass: gradients of the mean squared error */
void fteik_surrogate_gradients(int n, const float *pos, const float *tgt,
                               float *grad, double *loss, int *ierr);
/* training on the locator's tables; lossOut = mean loss of the last steps */
void fteik_surrogate_train(int steps, int batch, double lr, int pool,
                           double *lossOut, int *ierr);
/* surrogate-pruned location (fresh origin-time state): Gauss-Newton on the
 * surrogate's misfit from `starts` (<= 9) points, exact scoring of the
 * 4 x 4 x 4 cells around the best (re-centred on the best cell up to 4
 * times while it lies on a face of the cube), full grid search for events
 * whose best cell still lies on a face (counted in nFallback). */
void locate_locateEventsSurrogate(int n, const int *evnmbr, int starts,
                                  int iterations, int *optIndx, double *t0Opt,
                                  double *objOpt, int *nFallback, int *ierr);
/* EXTENSION: on = 1 (default): proposals from the network; on = 0: from
 * the straight-ray baseline |x - station| / v0 alone (the network's output
 * taken as zero), with the same Gauss-Newton steps, cube scoring and
 * full-search fallback -- the method without the learned part. */
void fteik_surrogate_setNetworkProposals(int on, int *ierr);
/* diagnostic: C = A B (row-major FP32 in/out) through the BF16 MFMA GEMM */
void fteik_amd_gemmBf16(int M, int N, int K, const float *A, const float *B,
                        float *C, int *ierr);
/* EXTENSION (diagnostic; no counterpart in the reference): one hidden layer
 * of the surrogate's inversion (4 blocks of R rows: primal and three
 * tangents, 256 wide) by the fused MFMA kernel and by the GEMM followed by
 * the activation kernel; BF16 output bits of both (4R x 256 each). */
void fteik_amd_invHiddenBf16(int R, const float *hin, const float *W,
                             const float *b, unsigned short *fused,
                             unsigned short *unfused, int *ierr);

/* =========================== Model helpers =========================== */
/* model.F90:771 */
int fteik_model_grid2indexF(int i, int j, int k, int nz, int nzx);
/* model.F90:827 */
int fteik_model_velGrid2indexF(int i, int j, int k, int nzm1, int nzm1_nxm1);
/* model.F90:794 (on the geometry of the solver initialized last, the
 * reference's shared model module; ny = 1 after fteik_solver2d_initialize) */
void fteik_model_index2gridF(int igrd, int *i, int *j, int *k, int *ierr);
/* fwrapper.c:8: 0-based */
int fteik_model_grid2index(const int iz, const int ix, const int iy,
                           const int nz, const int nzx);

/* ============================== Analytic ============================= */
/* analytic.F90:100, :208, :194, :346, :313, :408, :174 */
void fteik_analytic_initialize64f(int nz, int nx, int ny, double z0,
                                  double x0, double y0, double dz, double dx,
                                  double dy, int verbose, int *ierr);
void fteik_analytic_free(void);
void fteik_analytic_setVerbosity(int verbose);
void fteik_analytic_setReceivers64f(int nrec, const double *zrec,
                                    const double *xrec, const double *yrec,
                                    int *ierr);
void fteik_analytic_setConstantVelocity64f(double vconst, int *ierr);
void fteik_analytic_setLinearVelocityGradient64f(double v0, double v1,
                                                 int *ierr);
void fteik_analytic_setSources64f(int nsrc, const double *zsrc,
                                  const double *xsrc, const double *ysrc,
                                  int *ierr);
/* analytic.F90:538, :574, :686, :747, :795 */
void fteik_analytic_solveSourceConstantVelocity64f(int isrc, int *ierr);
void fteik_analytic_solveSourceLinearVelocityGradient64f(int isrc, int *ierr);
void fteik_analytic_getTravelTimeField64f(int ngrd, int order, double *tt,
                                          int *ierr);
void fteik_analytic_getTravelTimesConstantVel64f(int ldr, double *trec,
                                                 int *ierr);
void fteik_analytic_getTravelTimesGradientVel64f(int ldr, double *trec,
                                                 int *ierr);

/* =============================== Graph =============================== */
/* graph3d.F90:82, :122, :392, :471, :448, :432, :359.  Levels are computed
 * in closed form on the GPU (level = flipped iz+ix+iy, each level sorted by
 * linear index); the solver itself does not need them. */
void fteik_graph3d_initialize(int nz, int nx, int ny, int verbose, int *ierr);
void fteik_graph3d_makeLevelStructures(int *ierr);
void fteik_graph3d_getIJKVF(int nwork, int sweep, int *ijkv, int *ierr);
void fteik_graph3d_getLevelPointerF(int nwork, int *levelPtr, int *ierr);
int fteik_graph3d_getNumberOfLevels(int *ierr);
int fteik_graph3d_getMaxLevelSize(int *ierr);
void fteik_graph3d_free(void);

/* ========================== HDF5 archive I/O ========================= */
/* modules.F90:845-977 / h5io.F90: HDF5 is not part of this library; the
 * symbols exist so python/fteik3d.py's binding table loads.  All return
 * -1 (failure). */
int fteik_h5io_initializeF(const char *fileName);
int fteik_h5io_finalizeF(void);
int fteik_h5io_writeLevelSchedulesF(void);
int fteik_h5io_writeTravelTimesF(const char *dataName);
int fteik_h5io_writeVelocityModelF(const char *dataName);

/* ============================ Diagnostics ============================ */
/* EXTENSION: evaluates the device 3D local solvers on n nodes.  in holds
 * n rows [tv te tn tev ten tnv tnve tt0 s1..s7], node n rows
 * [sweep i j k] (1-based).  Mesh constants come from the source of the
 * last solve (or fteik_amd_prepareSource3d). */
void fteik_amd_evalLocalSolver3d(int n, const double *in, const int *node,
                                 double *outNoInit, double *outInit,
                                 int *ierr);
/* EXTENSION: computes the solver constants for source isrc without
 * solving (so the stencil diagnostic can run). */
void fteik_amd_prepareSource3d(int isrc, int *ierr);
/* EXTENSION: device time per sweep launch (ms) of nrep Gauss-Seidel
 * iterations (8 sweeps each) on the field of the last solve, in FP64
 * (precision 0) or FP32 (precision 1, after a mixed solve), timed with
 * hipEvents on the library's stream.  Leaves the field as after the extra
 * sweeps. */
void fteik_amd_timeSweeps3d(int precision, int nrep, double *msPerLaunch,
                            int *ierr);
/* EXTENSION: device time (microseconds per step) of one dependent chain of
 * nsteps iteration stencils per lane, one wave per SIMD on the whole GPU,
 * FP64 (precision 0) or FP32 (1): the latency floor of a sweep step,
 * which bounds a sweep at (number of levels) x this.  Needs the mesh
 * constants of a solved or prepared source. */
void fteik_amd_timeStencilChain3d(int precision, int nsteps,
                                  double *usPerStep, int *ierr);
/* EXTENSION (diagnostic): y = the square root the FP64 iteration stencil
 * uses for its 2D operators, evaluated on the GPU for n arguments x.  It
 * equals the correctly rounded sqrt for normal x >= 2^-767, the range of
 * every such argument wherever its operator's gate is open (velocities and
 * grid spacings are bounded at 1e30 to keep it so). */
void fteik_amd_sqrtStencil64f(int n, const double *x, double *y, int *ierr);
/* EXTENSION: library build string. */
const char *fteik_amd_version(void);
/* EXTENSION: releases every solver's device state, then the library's HIP
 * stream.  Registered to run at process exit when the GPU is first opened;
 * a caller may also call it, after which the next call that needs the GPU
 * opens it again. */
void fteik_amd_finalize(void);

#ifdef __cplusplus
}
#endif
#endif /* FTEIK_AMD_H */
