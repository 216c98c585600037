// capi_batch.cpp -- the batched multi-grid 3D solver and its hand-off
// of solved fields to the locator's tables (extensions).
#include "capi_internal.h"

using namespace fteik_capi;

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

// Grid igrid of the batch: geometry, source and the slowness, either from
// host cell velocities (vel) or copied from a device slowness (d_slow).
static bool batch_set_grid(const char *who, int igrid, int nz, int nx,
                           int ny, double z0, double x0, double y0, double dz,
                           double dx, double dy, const double *vel,
                           const double *d_slow, double zs, double xs,
                           double ys)
{
    gb.haveTimes = false;
    if (!gb.init || igrid < 1 || igrid > (int) gb.g.size()) {
        msg("%s: batch not initialized or grid %d out of range", who, igrid);
        return false;
    }
    if (nz < 3 || nx < 3 || ny < 3 || dz <= 0.0 || dx <= 0.0 || dy <= 0.0) {
        msg("%s: grid %d: need n >= 3 and positive spacing", who, igrid);
        return false;
    }
    if (!(dz <= kMaxSpacing && dx <= kMaxSpacing && dy <= kMaxSpacing)) {
        msg("%s: grid %d: spacing must be finite and at most %g", who, igrid,
            kMaxSpacing);
        return false;
    }
    BGrid &x = gb.g[(size_t) (igrid - 1)];
    dfree(x.d_slow); dfree(x.d_tt); dfree(x.d_slow32); dfree(x.d_tt32);
    dfree(x.d_cst); dfree(x.d_cst32);
    x = BGrid();
    x.nz = nz; x.nx = nx; x.ny = ny;
    x.z0 = z0; x.x0 = x0; x.y0 = y0; x.dz = dz; x.dx = dx; x.dy = dy;
    if (locate_source(true, nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys,
                      igrid, &x.src)) return false;
    if (x.src.zsi > nz - 2 || x.src.xsi > nx - 2 || x.src.ysi > ny - 2) {
        msg("%s: grid %d: source in the last cell along an axis", who, igrid);
        return false;
    }
    const long ncell = (long) (nz - 1)*(nx - 1)*(ny - 1);
    const long ngrd = (long) nz*nx*ny;
    if (!dalloc(&x.d_slow, (size_t) ncell) || !dalloc(&x.d_tt, (size_t) ngrd) ||
        !dalloc((char **) &x.d_cst, 512) || !dalloc((char **) &x.d_cst32, 512)) {
        msg("%s: out of GPU memory", who);
        return false;
    }
    if (d_slow) {
        if (hipMemcpyAsync(x.d_slow, d_slow, (size_t) ncell*sizeof(double),
                           hipMemcpyDeviceToDevice, g_dev.stream)
            != hipSuccess || !sync_ok()) {
            msg("%s: GPU error", who);
            return false;
        }
    } else if (!set_model(who, true, nz, nx, ny, 0, (int) ncell, 0, vel,
                          x.d_slow)) {
        return false;
    }
    x.set = true;
    gb.mapTile = -1;
    return true;
}

void fteik_batch3d_setGrid(int igrid, int nz, int nx, int ny, double z0,
                           double x0, double y0, double dz, double dx,
                           double dy, const double *vel, double zs, double xs,
                           double ys, int *ierr)
{
    *ierr = batch_set_grid("fteik_batch3d_setGrid", igrid, nz, nx, ny, z0,
                           x0, y0, dz, dx, dy, vel, nullptr, zs, xs, ys)
            ? 0 : 1;
}

void fteik_batch3d_setMixedParameters(double bias, int maxCorrections,
                                      int *ierr)
{
    *ierr = 0;
    if (!(bias > 0.0 && bias < 0.1) || maxCorrections < 1 ||
        maxCorrections > 64) {
        msg("fteik_batch3d_setMixedParameters: need 0 < bias < 0.1 and "
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
    bool converged = true;
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
        converged = false;
        for (corr = 0; corr < gb.maxCorr && ok; ) {
            hipMemsetAsync(gb.d_maxdelta, 0, ng*sizeof(double), g_dev.stream);
            for (int sw = 1; sw <= 8 && ok; sw++)
                ok = batch_sweep(sw, false, false, true);
            updates += batch_updates_iter();
            corr++;
            ok = ok && d2h(md.data(), gb.d_maxdelta, ng*sizeof(double));
            if (ok && *std::max_element(md.begin(), md.end()) <= gb.convTol)
                converged = true;
            if (!ok || converged) break;
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
    if (!converged) {
        msg("%s: mixed precision: %d correction iterations did not converge",
            who, corr);
        *ierr = 2;
    }
}

// Times fteik_batch3d_solve(nsweep, eps) with every tile shape the batch
// can use (flag and mailbox tiles) and keeps the fastest; the batch is left
// solved with it.  The fields do not depend on the shape.
void fteik_batch3d_autotune(int nsweep, double eps, int *tile, int *ierr)
{
    static const char *who = "fteik_batch3d_autotune";
    *ierr = 0;
    *tile = gb.mbTile >= 0 ? gb.mbTile : gb.tile;
    if (!gb.init) { msg("%s: batch not initialized", who); *ierr = 1;
                    return; }
    const int keepTile = gb.tile, keepMb = gb.mbTile;
    int best = -1;
    double bestMs = 0.0;
    for (int t = 0; t < fteik_k_num_tiles(); t++) {
        int e = 0;
        fteik_batch3d_setTile(t, &e);
        if (e) continue;
        fteik_batch3d_solve(nsweep, eps, &e);
        if (e == 1) continue;
        if (best < 0 || gb.last_ms < bestMs) { best = t; bestMs = gb.last_ms; }
    }
    gb.tile = keepTile;
    gb.mbTile = keepMb;
    if (best < 0) { msg("%s: no tile shape solved the batch", who);
                    *ierr = 1; return; }
    int e = 0;
    fteik_batch3d_setTile(best, &e);
    fteik_batch3d_solve(nsweep, eps, &e);
    if (e == 1) { *ierr = 1; return; }
    *tile = best;
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

// The reciprocal station-as-source driver (fteik.c:186-235, :264-326 with
// lreciprocal: receivers become sources), batched and on the device: every
// station of the 3D solver's grid and model is solved as a source in one
// batch, and its field becomes locator table itf0 + s, rounded to single
// precision as locate_setTravelTimeField64f rounds (locate.F90:336-339).
// The batch stays solved, so the fields can still be read with
// fteik_batch3d_getTravelTimeField64f.
void locate_setTravelTimeFieldsFromStations3d(int nsta, const double *zsta,
                                              const double *xsta,
                                              const double *ysta, int nsweep,
                                              double eps, int precision,
                                              int itf0, int *ierr)
{
    static const char *who = "locate_setTravelTimeFieldsFromStations3d";
    *ierr = 0;
    Solver3DView v;
    if (!solver3d_view(&v)) {
        msg("%s: the 3D solver's grid and model are not set", who);
        *ierr = 1;
        return;
    }
    if (!gl.init || nsta < 1 || itf0 < 1 || itf0 + nsta - 1 > gl.ntf ||
        (long) v.nz*v.nx*v.ny != gl.ngrd) {
        msg("%s: locator not initialized for %d tables from %d on this grid",
            who, nsta, itf0);
        *ierr = 1;
        return;
    }
    if (precision != 0 && precision != 1) {
        msg("%s: precision must be 0 (FP64) or 1 (mixed)", who);
        *ierr = 1;
        return;
    }
    fteik_batch3d_initialize(nsta, ierr);
    if (*ierr) return;
    for (int s = 0; s < nsta; s++)
        if (!batch_set_grid(who, s + 1, v.nz, v.nx, v.ny, v.z0, v.x0, v.y0,
                            v.dz, v.dx, v.dy, nullptr, v.d_slow, zsta[s],
                            xsta[s], ysta[s])) {
            msg("%s: station %d", who, s + 1);
            *ierr = 1;
            return;
        }
    fteik_batch3d_setPrecision(precision, ierr);
    fteik_batch3d_solve(nsweep, eps, ierr);
    if (*ierr == 1) return;
    const int solved = *ierr;          // 2: mixed corrections unconverged
    for (int s = 0; s < nsta && *ierr != 1; s++)
        locate_setTravelTimeFieldFromBatch3d(s + 1, itf0 + s, ierr);
    if (*ierr == 0) *ierr = solved;
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
