// batch3d.hip -- many single-source 3D solves, each on its own grid, swept
// together: one launch per sweep direction covers every tile of every grid.
//
// The reference solves one source at a time (solver3D.F90:1156-1243); here a
// batch of (grid, source) pairs of uneven sizes is advanced sweep by sweep
// with the same pipelined column-tile body as the single-grid solver
// (sweep_pencil.h).  Workgroup ids are ordered by tile diagonal first
// (a + b of the tile's flipped position), then grid, then a: every tile's
// upwind producers in its own grid have lower ids, and the first tiles of
// all grids start together.  Grids do not interact.
//
// The launch is not cooperative (the batch is larger than the GPU), so
// forward progress rests on in-order dispatch: a waiting workgroup only
// waits on lower ids, which were dispatched before it on their own XCD and
// are running or done.  Every wait keeps the spin limit of the body.
#include <hip/hip_runtime.h>
#include "sweep_pencil.h"

namespace fteik {

struct BatchTile { int grid, tile, flag, pad; };

template <typename R, bool INIT, int BX, int BY, int HB>
__global__ __launch_bounds__(BX*BY)
void batch_sweep_kernel(const BatchGrid* __restrict__ grids,
                        const BatchTile* __restrict__ map, int sweep,
                        unsigned long long serial,
                        unsigned long long* __restrict__ flags,
                        int* __restrict__ err, int fp32, int track)
{
    const BatchTile m = map[blockIdx.x];
    const BatchGrid& g = grids[m.grid];
    SweepArgs<R> a;
    a.tt = (R*) (fp32 ? g.tt32 : g.tt);
    a.slow = (const R*) (fp32 ? g.slow32 : g.slow);
    a.cst = (const Consts<R>*) (fp32 ? g.cst32 : g.cst);
    a.nz = g.nz; a.nx = g.nx; a.ny = g.ny;
    sweep_signs(sweep, a.sz, a.sx, a.sy);
    int b[6];
    flipped_box(sweep, INIT, g.nz, g.nx, g.ny, g.zsi, g.xsi, g.ysi, b);
    a.fz0 = b[0]; a.fz1 = b[1]; a.fx0 = b[2]; a.fx1 = b[3];
    a.fy0 = b[4]; a.fy1 = b[5];
    a.nbx = (g.nx + BX - 1)/BX;
    a.nby = (g.ny + BY - 1)/BY;
    // this tile's flag is flags[m.flag]; the body addresses flags by the
    // grid-local tile number, so hand it the grid's base
    a.flags = flags + (m.flag - m.tile);
    a.err = err;
    a.serial = serial;
    a.maxdelta = track ? g.maxdelta : nullptr;
    a.mb = nullptr;
    a.mbx = a.mby = nullptr; a.mbnt = 0;
    a.colmod = nullptr; a.plan = nullptr;
    if (a.fz0 > a.fz1 || a.fx0 > a.fx1 || a.fy0 > a.fy1) {
        if (threadIdx.x == 0)
            __hip_atomic_store(flags + m.flag, (serial << 32) | 0xffffffffull,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    pencil_body<R, INIT, BX, BY, HB>(a, m.tile);
}

// Sweep of a batch with mailbox tiles (modes 1, 2; INIT: initialization
// sweep): workgroup w sweeps map entry (grid, tile); entries are ordered by
// tile diagonal, then grid, then x, so every workgroup waits only on
// lower-numbered ones and in-order dispatch keeps the lowest unfinished one
// resident (the launch is larger than the GPU).  args[g]: grid g's sweep
// arguments for this launch; each grid has its own mailbox.
template <typename R, bool INIT, int BX, int BY, int HB, int MODE, int KB,
          bool ST>
__global__ __launch_bounds__(BX*BY + 64*MODE)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(HB))))
void batch_mb_kernel(const SweepArgs<R>* __restrict__ args,
                     const int* __restrict__ map)
{
    // read through the constant address space: the loads are invariant, so
    // the register allocator may reload a field where it is used (as it
    // does kernel arguments) instead of keeping it live
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) SweepArgs<R> CArgs;
    const SweepArgs<R> a = ((CArgs*) args)[map[2*blockIdx.x]];
#else   // the host pass only parses the kernel
    const SweepArgs<R> a = args[map[2*blockIdx.x]];
#endif
    pencil_mb_body<R, INIT, BX, BY, HB, MODE, false, false, KB, ST>(
        a, dyn_of(a), map[2*blockIdx.x + 1]);
}

// Sweep arguments of every grid for one mailbox sweep (linit: an
// initialization sweep, whose box depends on the grid's source).
template <typename R>
__global__ void batch_mb_args_kernel(const BatchGrid* __restrict__ grids,
                                     int ngrids, int sweep, int linit,
                                     int bx, int by,
                                     unsigned long long serial,
                                     int* __restrict__ err, int fp32,
                                     int track,
                                     SweepArgs<R>* __restrict__ args)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= ngrids) return;
    const BatchGrid g = grids[i];
    SweepArgs<R> a;
    a.tt = (R*) (fp32 ? g.tt32 : g.tt);
    a.slow = (const R*) (fp32 ? g.slow32 : g.slow);
    a.cst = (const Consts<R>*) (fp32 ? g.cst32 : g.cst);
    a.flags = nullptr;
    a.err = err;
    a.serial = serial;
    a.nz = g.nz; a.nx = g.nx; a.ny = g.ny;
    sweep_signs(sweep, a.sz, a.sx, a.sy);
    int b[6];
    flipped_box(sweep, linit != 0, g.nz, g.nx, g.ny, g.zsi, g.xsi, g.ysi, b);
    a.fz0 = b[0]; a.fz1 = b[1]; a.fx0 = b[2]; a.fx1 = b[3];
    a.fy0 = b[4]; a.fy1 = b[5];
    a.nbx = max((a.fx1 - a.fx0 + bx)/bx, 0);
    a.nby = max((a.fy1 - a.fy0 + by)/by, 0);
    a.maxdelta = track ? g.maxdelta : nullptr;
    a.mb = g.mb;
    a.mbx = g.mbx; a.mby = g.mby; a.mbnt = g.mbnt;
    a.colmod = nullptr; a.plan = nullptr; a.prog = nullptr;
    a.pserial = 0; a.par = 0; a.psz = a.psx = a.psy = 0;
    args[i] = a;
}

template <typename R, bool INIT>
static const void* batch_mb_fn(int tile)
{
#define BM(n) case n: return (const void*) batch_mb_kernel<R, INIT, \
        kTiles[n].bx, kTiles[n].by, kTiles[n].hb, kTiles[n].mode, \
        kTiles[n].kb, kTiles[n].st != 0>;
    switch (tile) {
        BM(7) BM(8) BM(9) BM(10) BM(11) BM(12) BM(13)
        BM(14) BM(15) BM(16) BM(17) BM(18)
        default: return nullptr;
    }
#undef BM
}
static_assert(kTiles[7].mode == 1 && kTiles[13].mode == 2 && kNumTiles == 19,
              "batch_mb_fn lists the mailbox tiles");

template <typename R, bool INIT>
static const void* batch_fn_t(int tile)
{
#define BF(n) case n: return (const void*) batch_sweep_kernel<R, INIT, \
        kTiles[n].bx, kTiles[n].by, kTiles[n].hb>;
    switch (tile) {
        BF(0) BF(1) BF(2) BF(3) BF(4) BF(5) BF(6)
        default: return nullptr;
    }
#undef BF
}
// the batch uses the flag-exchange tiles only (kTiles[0..6])

} // namespace fteik

using namespace fteik;

// Workgroup order of a batch for tile shape `tile`: tiles sorted by
// diagonal, then grid, then x; flag slots are grid-major.  map receives
// 4 ints per workgroup; returns the workgroup count (or -1 if map_cap is too
// small).  Host-side bookkeeping only.
extern "C" int fteik_k_batch_map(int tile, int ngrids, const int* dims,
                                 int* map, int map_cap, int* nflags)
{
    if (tile < 0 || tile >= kNumTiles) return -1;
    const TileShape ts = kTiles[tile];
    int total = 0, maxdiag = 0;
    for (int g = 0; g < ngrids; g++) {
        const int nbx = (dims[3*g + 1] + ts.bx - 1)/ts.bx;
        const int nby = (dims[3*g + 2] + ts.by - 1)/ts.by;
        total += nbx*nby;
        maxdiag = (nbx + nby - 2 > maxdiag) ? nbx + nby - 2 : maxdiag;
    }
    *nflags = total;
    if (total > map_cap) return -1;
    int w = 0;
    for (int d = 0; d <= maxdiag; d++) {
        int base = 0;
        for (int g = 0; g < ngrids; g++) {
            const int nbx = (dims[3*g + 1] + ts.bx - 1)/ts.bx;
            const int nby = (dims[3*g + 2] + ts.by - 1)/ts.by;
            for (int bx = 0; bx < nbx; bx++) {
                const int by = d - bx;
                if (by < 0 || by >= nby) continue;
                map[4*w] = g;
                map[4*w + 1] = bx*nby + by;
                map[4*w + 2] = base + bx*nby + by;
                map[4*w + 3] = 0;
                w++;
            }
            base += nbx*nby;
        }
    }
    return w;
}

extern "C" int fteik_k_batch_sweep(const BatchGrid* grids, const int* map,
                                   int nwg, int tile, int sweep, int linit,
                                   unsigned long long serial,
                                   unsigned long long* flags, int* err,
                                   int fp32, int track, hipStream_t stream)
{
    if (tile < 0 || tile >= kNumTiles || kTiles[tile].mode != 0 || nwg < 1)
        return 2;
    const TileShape ts = kTiles[tile];
    const void* fn = fp32 ? (linit ? batch_fn_t<float, true>(tile)
                                   : batch_fn_t<float, false>(tile))
                          : (linit ? batch_fn_t<double, true>(tile)
                                   : batch_fn_t<double, false>(tile));
    const BatchTile* m = (const BatchTile*) map;
    void* args[] = {(void*) &grids, (void*) &m, (void*) &sweep,
                    (void*) &serial, (void*) &flags, (void*) &err,
                    (void*) &fp32, (void*) &track};
    hipError_t e = hipLaunchKernel(fn, dim3(nwg), dim3(ts.bx*ts.by), args, 0,
                                   stream);
    return e == hipSuccess ? 0 : 1;
}

// Workgroup order of a mailbox sweep of a batch for tile shape `tile`:
// entries (grid, tile) by diagonal, then grid, then x; 2 ints each.  sweep
// 0: an iteration sweep (box: flipped columns 1 .. n-1, the same for every
// direction); sweep 1..8: that initialization sweep, whose box runs from
// each grid's source (srcs: zsi, xsi, ysi per grid, 1-based) and may be
// empty.  Returns the entry count (with map == nullptr nothing is
// written), or -1 if map_cap is too small.  Host-side bookkeeping only.
extern "C" int fteik_k_batch_mb_map(int tile, int ngrids, const int* dims,
                                    const int* srcs, int sweep, int* map,
                                    int map_cap)
{
    if (tile < 0 || tile >= kNumTiles || kTiles[tile].mode == 0 ||
        sweep < 0 || sweep > 8 || (sweep > 0 && !srcs)) return -1;
    const TileShape ts = kTiles[tile];
    auto counts = [&](int g, int& nbx, int& nby) {
        const int nz = dims[3*g], nx = dims[3*g + 1], ny = dims[3*g + 2];
        if (sweep == 0) {
            nbx = (nx - 1 + ts.bx - 1)/ts.bx;
            nby = (ny - 1 + ts.by - 1)/ts.by;
            return;
        }
        int b[6];
        flipped_box(sweep, true, nz, nx, ny, srcs[3*g], srcs[3*g + 1],
                    srcs[3*g + 2], b);
        const bool empty = b[0] > b[1] || b[2] > b[3] || b[4] > b[5];
        nbx = empty ? 0 : (b[3] - b[2] + ts.bx)/ts.bx;
        nby = empty ? 0 : (b[5] - b[4] + ts.by)/ts.by;
    };
    int total = 0, maxdiag = 0;
    for (int g = 0; g < ngrids; g++) {
        int nbx, nby;
        counts(g, nbx, nby);
        total += nbx*nby;
        if (nbx > 0 && nby > 0)
            maxdiag = (nbx + nby - 2 > maxdiag) ? nbx + nby - 2 : maxdiag;
    }
    if (!map) return total;
    if (total > map_cap) return -1;
    int w = 0;
    for (int d = 0; d <= maxdiag; d++)
        for (int g = 0; g < ngrids; g++) {
            int nbx, nby;
            counts(g, nbx, nby);
            for (int bx = 0; bx < nbx; bx++) {
                const int by = d - bx;
                if (by < 0 || by >= nby) continue;
                map[2*w] = g;
                map[2*w + 1] = bx*nby + by;
                w++;
            }
        }
    return w;
}

extern "C" size_t fteik_k_batch_mb_args_bytes(int ngrids)
{
    return (size_t) ngrids*sizeof(SweepArgs<double>);
}

// One mailbox iteration sweep of a batch.  argbuf: device buffer of
// fteik_k_batch_mb_args_bytes(ngrids) bytes; map: nwg entries from
// fteik_k_batch_mb_map.
extern "C" int fteik_k_batch_mb_sweep(const BatchGrid* grids, int ngrids,
                                      const int* map, int nwg, int tile,
                                      int sweep, int linit,
                                      unsigned long long serial,
                                      int* err, int fp32, int track,
                                      void* argbuf, hipStream_t stream)
{
    if (tile < 0 || tile >= kNumTiles || kTiles[tile].mode == 0 || nwg < 0 ||
        ngrids < 1 || sweep < 1 || sweep > 8 || !argbuf) return 2;
    if (nwg == 0) return 0;                 // every grid's box is empty
    const TileShape ts = kTiles[tile];
    const void* fn = linit
        ? (fp32 ? batch_mb_fn<float, true>(tile)
                : batch_mb_fn<double, true>(tile))
        : (fp32 ? batch_mb_fn<float, false>(tile)
                : batch_mb_fn<double, false>(tile));
    if (fp32)
        hipLaunchKernelGGL(batch_mb_args_kernel<float>, dim3((ngrids + 63)/64),
                           dim3(64), 0, stream, grids, ngrids, sweep, linit,
                           ts.bx, ts.by, serial, err, fp32, track,
                           (SweepArgs<float>*) argbuf);
    else
        hipLaunchKernelGGL(batch_mb_args_kernel<double>,
                           dim3((ngrids + 63)/64), dim3(64), 0, stream, grids,
                           ngrids, sweep, linit, ts.bx, ts.by, serial, err,
                           fp32, track, (SweepArgs<double>*) argbuf);
    void* args[] = {&argbuf, (void*) &map};
    hipError_t e = hipLaunchKernel(fn, dim3(nwg),
                                   dim3(ts.bx*ts.by + 64*ts.mode), args, 0,
                                   stream);
    return e == hipSuccess ? 0 : 1;
}
