// sweep3d.hip -- 3D level-set fast sweeping on gfx950: the single-grid
// solver's kernels and launchers.  One sweep direction is one launch of the
// pipelined column-tile kernel (sweep_pencil.h); kernel boundaries between
// sweeps give the coherence the next sweep needs.  A launch is cooperative
// when all its tiles fit on the GPU at once; otherwise (the default 8 x 8
// tiles at 257^3) it relies on in-order workgroup dispatch: a tile waits
// only on tiles of lower-numbered workgroups (tiles are dealt in diagonal
// order, tile_of_rank), so the lowest unfinished one is always resident and
// every wait ends.  Each wait also has a spin limit that turns a hang into
// a reported error.  Schedule 1 is a one-launch-per-level cross-check and
// fallback.
#include <hip/hip_runtime.h>
#include <cstring>
#include <stdint.h>
#include "sweep_pencil.h"

namespace fteik {

template <typename R>
__global__ void fill_kernel(R* __restrict__ a, long n, R v)
{
    long i = (long) blockIdx.x*blockDim.x + threadIdx.x;
    long stride = (long) gridDim.x*blockDim.x;
    for (; i < n; i += stride) a[i] = v;
}

// localSolver.F90:878-956 + solver3D.F90:1201-1202: mesh constants, the
// source-cell slowness and the 8 analytic seed times.  Single thread.
template <typename R>
__global__ void source3d_init_kernel(Consts<R>* __restrict__ cst,
                                     R* __restrict__ tt,
                                     const R* __restrict__ slow,
                                     SourceGeom g)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Consts<R> c;
    const R one = R(1);
    c.dz = R(g.dz); c.dx = R(g.dx); c.dy = R(g.dy);
    c.zsa = R(g.zsa); c.xsa = R(g.xsa); c.ysa = R(g.ysa);
    c.zsi = g.zsi; c.xsi = g.xsi; c.ysi = g.ysi; c.pad = 0;
    long icell = (long) (g.ysi - 1)*(g.nz - 1)*(g.nx - 1)
               + (long) (g.xsi - 1)*(g.nz - 1) + g.zsi - 1;
    c.szero = slow[icell];
    c.szero2 = c.szero*c.szero;
    R hmin = rmin(rmin(c.dz, c.dx), c.dy);
    c.epsSolver = R(g.eps)*hmin*c.szero;
    R dz2 = c.dz*c.dz, dx2 = c.dx*c.dx, dy2 = c.dy*c.dy;
    c.dzi = one/c.dz; c.dxi = one/c.dx; c.dyi = one/c.dy;
    c.dz2i = one/dz2; c.dx2i = one/dx2; c.dy2i = one/dy2;
    R dsum = c.dz2i + c.dx2i + c.dy2i;
    c.dsum_nine = dsum*R(9);
    c.dsumi = one/dsum;
    c.dz2dx2 = one/(dz2*dx2);
    c.dz2dy2 = one/(dz2*dy2);
    c.dx2dy2 = one/(dx2*dy2);
    c.dz2i_dx2i = c.dz2i*c.dx2i;
    c.dz2i_dy2i = c.dz2i*c.dy2i;
    c.dx2i_dy2i = c.dx2i*c.dy2i;
    c.dz2i_p_dy2i_inv = one/(c.dz2i + c.dy2i);
    c.dz2i_p_dx2i_inv = one/(c.dz2i + c.dx2i);
    c.dx2i_p_dy2i_inv = one/(c.dx2i + c.dy2i);
    c.dz2i_p_dx2i = c.dz2i + c.dx2i;
    c.dz2i_p_dy2i = c.dz2i + c.dy2i;
    c.dx2i_p_dy2i = c.dx2i + c.dy2i;
    *cst = c;
    if (tt) {
        for (int n = 0; n < 8; n++) {
            int i = c.zsi + (n & 1), j = c.xsi + ((n >> 1) & 1),
                k = c.ysi + ((n >> 2) & 1);
            long g0 = (long) (k - 1)*g.nz*g.nx + (long) (j - 1)*g.nz + i - 1;
            tt[g0] = tana3d(i, j, k, c);
        }
    }
}

// Derive single-precision constants from the double ones (mixed mode).
__global__ void consts_to_float_kernel(const Consts<double>* __restrict__ cd,
                                       Consts<float>* __restrict__ cf)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Consts<double>& a = *cd;
    Consts<float> b;
#define CV(f) b.f = (float) a.f
    CV(dz); CV(dx); CV(dy); CV(dzi); CV(dxi); CV(dyi); CV(dz2i); CV(dx2i);
    CV(dy2i); CV(dsum_nine); CV(dsumi); CV(epsSolver); CV(dz2dx2); CV(dz2dy2);
    CV(dx2dy2); CV(dz2i_dx2i); CV(dz2i_dy2i); CV(dx2i_dy2i); CV(dz2i_p_dx2i);
    CV(dz2i_p_dy2i); CV(dx2i_p_dy2i); CV(dz2i_p_dy2i_inv); CV(dz2i_p_dx2i_inv);
    CV(dx2i_p_dy2i_inv); CV(szero); CV(szero2); CV(zsa); CV(xsa); CV(ysa);
#undef CV
    b.zsi = a.zsi; b.xsi = a.xsi; b.ysi = a.ysi; b.pad = 0;
    *cf = b;
}

template <typename R, bool INIT, int BX, int BY, int HB>
__global__ __launch_bounds__(BX*BY)
void sweep3d_pencil_kernel(SweepArgs<R> a)
{
    pencil_body<R, INIT, BX, BY, HB>(a, blockIdx.x);
}

// compute waves + halo wave (+ store wave in mode 2); at least 3 waves per
// SIMD so two workgroups share a CU.  SKIP: iteration sweep with tile
// skipping (a separate instance, so the dense sweep keeps its registers)
template <typename R, bool INIT, int BX, int BY, int HB, int MODE, bool SKIP,
          int KB, bool ST>
__global__ __launch_bounds__(BX*BY + 64*MODE)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(HB))))
void sweep3d_pencil_mb_kernel(SweepArgs<R> a)
{
    if (SKIP) {
        // the plan lists the processed tiles; the other workgroups exit
        if ((int) blockIdx.x >= a.plan[0]) return;
        pencil_mb_body<R, INIT, BX, BY, HB, MODE, SKIP, false, KB, ST>(
            a, dyn_of(a), a.plan[1 + blockIdx.x]);
    } else {
        const int wg = tile_of_rank((int) blockIdx.x, a.nbx, a.nby);
        if (wg < 0) return;
        pencil_mb_body<R, INIT, BX, BY, HB, MODE, SKIP, false, KB, ST>(
            a, dyn_of(a), wg);
    }
}

// A chain of nq consecutive iteration sweeps (directions dir0 + q, q = 0..
// nq-1, cyclically; serials a.serial + q) in one launch: workgroup
// q*ntile + tile sweeps that tile in sweep q, overlapping sweep q - 1 as
// far as the data dependences between them allow (see the CHAIN body).
// Tiles stay in row-major order here: a sweep's tiles become ready in the
// order the preceding sweep, in another direction, releases them, and the
// diagonal order of single sweeps (tile_of_rank) measured 35 % slower at
// 257^3.
// The iteration-sweep box is the same in every direction in flipped
// coordinates, so the tiles are too.  The launch need not be resident at
// once: a workgroup waits only on lower-numbered ones (its sweep's upwind
// tiles, the preceding sweep), and workgroups are dispatched in order, so
// the lowest unfinished workgroup is always resident and never blocked.
// Consecutive sweeps alternate between two progress-word sets.
template <typename R, int BX, int BY, int HB>
__global__ __launch_bounds__(BX*BY + 128)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(HB))))
void sweep3d_chain_kernel(SweepArgs<R> a, int dir0)
{
    const int ntile = a.nbx*a.nby;
    const int q = (int) blockIdx.x / ntile;
    const int wg = (int) blockIdx.x - q*ntile;
    pencil_mb_body<R, false, BX, BY, HB, 2, false, true>(
        a, chain_dyn(dir0, q, a.serial), wg);
}

// Tile-skipping plan of one iteration sweep (one workgroup).  A tile whose
// columns have not changed since its last pass in this sweep direction
// (colmod <= lastproc) and whose upwind tiles are all skipped would
// reproduce its current values exactly: every node's update reads the same
// upwind values as in that pass, and a sweep is idempotent on its own
// output.  Unrolling "and whose upwind tiles are skipped" gives: a tile is
// processed iff some tile of its upwind quadrant (fbx' <= fbx and
// fby' <= fby) has a changed column -- a 2D prefix-OR over the tiles.
// Writes the plan (count, processed tiles in dispatch order, skip flags) and
// stamps every tile's lastproc with this launch's serial.
constexpr int kPlanMaxTiles = 4096;

// Tiles cover the box [fx0, fx1] x [fy0, fy1] (flipped columns) from its
// first column; the column row before the box is read by the first tiles
// and counts as theirs.
// Dirty marks for the plan below: a tile is marked when one of its columns
// (or one of the column row before the box, for the first tiles) changed
// after the tile's last pass in this direction.  One thread per column, so
// the column stamps are read by the whole GPU rather than one workgroup;
// concurrent marks of a tile all write 1.
__global__ __launch_bounds__(256)
void skip_mark_kernel(const unsigned long long* __restrict__ colmod,
                      const unsigned long long* __restrict__ lastproc,
                      int* __restrict__ mark, int nx, int ny, int sx, int sy,
                      int fx0, int fx1, int fy0, int fy1, int bx, int by,
                      int nby)
{
    const int lx = fx1 - fx0 + 2, ly = fy1 - fy0 + 2;
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= lx*ly) return;
    const int fk = fy0 - 1 + c / lx, fj = fx0 - 1 + c % lx;
    const int t = (max(fj - fx0, 0)/bx)*nby + max(fk - fy0, 0)/by;
    const int j = unflip(fj, nx, sx), k = unflip(fk, ny, sy);
    if (colmod[(long) (k - 1)*nx + (j - 1)] > lastproc[t]) mark[t] = 1;
}

__global__ __launch_bounds__(1024)
void skip_plan_kernel(unsigned long long* __restrict__ lastproc,
                      int* __restrict__ plan, int nbx, int nby,
                      unsigned long long serial)
{
    __shared__ unsigned char s_dirty[kPlanMaxTiles];
    __shared__ int s_off[kPlanMaxTiles];     // per diagonal: count, offset
    const int nt = nbx*nby;
    // dirty marks of the tiles, set by skip_mark_kernel; cleared here for
    // the next sweep
    int* const mark = plan + 1 + 2*nt;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        s_dirty[t] = mark[t] ? 1 : 0;
        mark[t] = 0;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < nbx; p += blockDim.x) {
        unsigned char acc = 0;
        for (int q = 0; q < nby; q++) {
            acc |= s_dirty[p*nby + q];
            s_dirty[p*nby + q] = acc;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nby; q += blockDim.x) {
        unsigned char acc = 0;
        for (int p = 0; p < nbx; p++) {
            acc |= s_dirty[p*nby + q];
            s_dirty[p*nby + q] = acc;
        }
    }
    __syncthreads();
    // compaction of the processed tiles in the dispatch order of
    // tile_of_rank (by diagonal fbx + fby, then fbx): each thread counts the
    // processed tiles of its diagonals, one wave turns the counts into
    // offsets, then each thread writes its diagonals' tiles
    const int nd = nbx + nby - 1;            // <= kPlanMaxTiles
    for (int dg = threadIdx.x; dg < nd; dg += blockDim.x) {
        const int lo = max(0, dg - (nby - 1)), hi = min(dg, nbx - 1);
        int n = 0;
        for (int p = lo; p <= hi; p++) n += s_dirty[p*nby + dg - p];
        s_off[dg] = n;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int carry = 0;
        for (int d0 = 0; d0 < nd; d0 += 64) {
            const int dg = d0 + lane;
            const int v = (dg < nd) ? s_off[dg] : 0;
            int inc = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(inc, o, 64);
                if (lane >= o) inc += u;
            }
            if (dg < nd) s_off[dg] = carry + inc - v;
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) plan[0] = carry;
    }
    __syncthreads();
    for (int dg = threadIdx.x; dg < nd; dg += blockDim.x) {
        const int lo = max(0, dg - (nby - 1)), hi = min(dg, nbx - 1);
        int w = 1 + s_off[dg];
        for (int p = lo; p <= hi; p++) {
            const int t = p*nby + dg - p;
            if (s_dirty[t]) plan[w++] = t;
        }
    }
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        plan[1 + nt + t] = s_dirty[t] ? 0 : 1;
        lastproc[t] = serial;
    }
}

// Simple reference schedule: one launch per level (used for cross-checks
// and as a fallback for grids whose column blocks cannot be co-resident).
template <typename R, bool INIT>
__global__ __launch_bounds__(256)
void sweep3d_level_kernel(SweepArgs<R> a, int L)
{
    const int lx = a.fx1 - a.fx0 + 1, ly = a.fy1 - a.fy0 + 1;
    long idx = (long) blockIdx.x*blockDim.x + threadIdx.x;
    if (idx >= (long) lx*ly) return;
    const int dk = (int) (idx / lx), dj = (int) (idx - (long) dk*lx);
    const int fj = a.fx0 + dj, fk = a.fy0 + dk;
    const int fi = L - fj - fk;
    if (fi < a.fz0 || fi > a.fz1) return;
    const int nz = a.nz, nx = a.nx, ny = a.ny;
    const int i = unflip(fi, nz, a.sz), j = unflip(fj, nx, a.sx),
              k = unflip(fk, ny, a.sy);
    const long nzx = (long) nz*nx;
    auto G = [&](int ii, int jj, int kk) -> long {
        return (long) (kk - 1)*nzx + (long) (jj - 1)*nz + ii - 1;
    };
    const Consts<R> c = *a.cst;
    Nbr<R> nb;
    nb.tv   = a.tt[G(i - a.sz, j, k)];
    nb.te   = a.tt[G(i, j - a.sx, k)];
    nb.tn   = a.tt[G(i, j, k - a.sy)];
    nb.tev  = a.tt[G(i - a.sz, j - a.sx, k)];
    nb.ten  = a.tt[G(i, j - a.sx, k - a.sy)];
    nb.tnv  = a.tt[G(i - a.sz, j, k - a.sy)];
    nb.tnve = a.tt[G(i - a.sz, j - a.sx, k - a.sy)];
    Slw<R> sl;
    slowness7(a.slow, nz, nx, ny, i, j, k, a.sz, a.sx, a.sy, sl);
    R v = a.tt[G(i, j, k)];
    R vn = INIT ? init3d(c, nb, sl, v, i, j, k, a.sz, a.sx, a.sy)
                : noinit3d(c, nb, sl, v);
    a.tt[G(i, j, k)] = vn;
    if (a.maxdelta && v - vn > R(0))
        atomicMax((unsigned long long*) a.maxdelta,
                  (unsigned long long) __double_as_longlong((double) (v - vn)));
}

// ------------------------------------------------------------ launchers
template <typename R, bool INIT, bool SKIP>
static const void* pencil_fn_t(int tile)
{
#define PF(n) case n: return SKIP ? nullptr \
        : (const void*) sweep3d_pencil_kernel<R, INIT, \
        kTiles[n].bx, kTiles[n].by, kTiles[n].hb>;
#define PM(n) case n: return (const void*) sweep3d_pencil_mb_kernel<R, INIT, \
        kTiles[n].bx, kTiles[n].by, kTiles[n].hb, kTiles[n].mode, SKIP, \
        kTiles[n].kb, kTiles[n].st != 0>;
    switch (tile) {
        PF(0) PF(1) PF(2) PF(3) PF(4) PF(5) PF(6)
        PM(7) PM(8) PM(9) PM(10) PM(11) PM(12) PM(13)
        PM(14) PM(15) PM(16) PM(17) PM(18)
        default: return nullptr;
    }
#undef PF
#undef PM
}
static_assert(kNumTiles == 19 && kTiles[6].mode == 0 && kTiles[7].mode == 1,
              "pencil_fn_t lists every tile");

static int tile_threads(int tile)
{
    return kTiles[tile].bx*kTiles[tile].by + 64*kTiles[tile].mode;
}

// variant 0: iteration sweep, 1: init sweep, 2: iteration sweep with tile
// skipping (mailbox tiles only)
template <typename R>
static const void* pencil_fn(int variant, int tile)
{
    switch (variant) {
        case 0: return pencil_fn_t<R, false, false>(tile);
        case 1: return pencil_fn_t<R, true, false>(tile);
        case 2: return pencil_fn_t<R, false, true>(tile);
        default: return nullptr;
    }
}

template <typename R>
static int launch_sweep_t(const SweepLaunch& L, int sweep, bool linit,
                          hipStream_t stream)
{
    SweepArgs<R> a;
    a.tt = (R*) L.tt; a.slow = (const R*) L.slow;
    a.cst = (const Consts<R>*) L.cst;
    a.flags = L.flags; a.err = L.err; a.serial = L.serial;
    a.nz = L.nz; a.nx = L.nx; a.ny = L.ny;
    sweep_signs(sweep, a.sz, a.sx, a.sy);
    int b[6];
    flipped_box(sweep, linit, L.nz, L.nx, L.ny, L.zsi, L.xsi, L.ysi, b);
    a.fz0 = b[0]; a.fz1 = b[1]; a.fx0 = b[2]; a.fx1 = b[3];
    a.fy0 = b[4]; a.fy1 = b[5];
    a.maxdelta = L.maxdelta;
    a.mb = L.mb;
    a.mbx = L.mbx; a.mby = L.mby; a.mbnt = L.mbnt;
    a.colmod = nullptr; a.plan = nullptr;
    if (a.fz0 > a.fz1 || a.fx0 > a.fx1 || a.fy0 > a.fy1) return 0;
    if (L.schedule == 1) {
        const int nlev = (a.fz1 + a.fx1 + a.fy1) - (a.fz0 + a.fx0 + a.fy0) + 1;
        const long ncol = (long) (a.fx1 - a.fx0 + 1)*(a.fy1 - a.fy0 + 1);
        const int nblk = (int) ((ncol + 255)/256);
        for (int l = 0; l < nlev; l++) {
            int Lv = a.fz0 + a.fx0 + a.fy0 + l;
            if (linit)
                hipLaunchKernelGGL((sweep3d_level_kernel<R, true>), dim3(nblk),
                                   dim3(256), 0, stream, a, Lv);
            else
                hipLaunchKernelGGL((sweep3d_level_kernel<R, false>), dim3(nblk),
                                   dim3(256), 0, stream, a, Lv);
        }
        return hipGetLastError() == hipSuccess ? 0 : 1;
    }
    const int tile = (L.tile >= 0 && L.tile < kNumTiles) ? L.tile : 0;
    const TileShape ts = kTiles[tile];
    if (ts.mode == 0) {
        // flag-exchange tiles cover the columns from 0
        a.nbx = (L.nx + ts.bx - 1)/ts.bx;
        a.nby = (L.ny + ts.by - 1)/ts.by;
    } else {
        // mailbox tiles cover the box's columns
        a.nbx = (a.fx1 - a.fx0 + ts.bx)/ts.bx;
        a.nby = (a.fy1 - a.fy0 + ts.by)/ts.by;
    }
    const int nblk = a.nbx*a.nby;
    if (nblk > L.max_flags) return 2;
    void* args[] = {&a};
    hipError_t e;
    if (ts.mode >= 1 && !L.mb) return 2;
    const bool skip = ts.mode >= 1 && !linit && L.colmod && L.plan &&
                      nblk <= kPlanMaxTiles;
    if (skip) {
        a.colmod = L.colmod; a.plan = L.plan;
        const int ncol = (a.fx1 - a.fx0 + 2)*(a.fy1 - a.fy0 + 2);
        hipLaunchKernelGGL(skip_mark_kernel, dim3((ncol + 255)/256),
                           dim3(256), 0, stream, L.colmod,
                           L.lastproc + (long) (sweep - 1)*nblk,
                           L.plan + 1 + 2*nblk, L.nx, L.ny, a.sx, a.sy,
                           a.fx0, a.fx1, a.fy0, a.fy1, ts.bx, ts.by, a.nby);
        hipLaunchKernelGGL(skip_plan_kernel, dim3(1), dim3(1024), 0, stream,
                           L.lastproc + (long) (sweep - 1)*nblk, L.plan,
                           a.nbx, a.nby, L.serial);
    }
    const void* fn = pencil_fn<R>(linit ? 1 : (skip ? 2 : 0), tile);
    if (L.cooperative)
        e = hipLaunchCooperativeKernel(fn, dim3(nblk), dim3(tile_threads(tile)),
                                       args, 0, stream);
    else
        e = hipLaunchKernel(fn, dim3(nblk), dim3(tile_threads(tile)), args, 0,
                            stream);
    return e == hipSuccess ? 0 : 1;
}

} // namespace fteik

using namespace fteik;

extern "C" int fteik_k_sweep3d(const SweepLaunch* L, int sweep, int linit,
                               hipStream_t stream)
{
    if (L->fp32) return launch_sweep_t<float>(*L, sweep, linit != 0, stream);
    return launch_sweep_t<double>(*L, sweep, linit != 0, stream);
}

// Workgroups of kernel fn that the GPU keeps resident per CU.  The
// occupancy query counts registers per CU, but a workgroup's waves are
// placed on the SIMDs starting from the first one: a workgroup of more than
// four waves puts two on the first SIMD, and a second one fits only if that
// SIMD has room for four such waves (measured: 5- and 6-wave workgroups at
// 152 registers run one per CU although the query reports two).
static int resident_per_cu(const void* fn, int threads)
{
    int api = 0;
    hipFuncAttributes fa;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, fn, threads, 0)
            != hipSuccess ||
        hipFuncGetAttributes(&fa, fn) != hipSuccess) return 0;
    const int waves = (threads + 63)/64;
    if (waves <= 4) return api;
    const int regs = ((fa.numRegs + 7)/8)*8;
    const int per_simd = regs > 0 ? min(8, 512/regs) : 8;
    return min(api, per_simd/((waves + 3)/4));
}

template <typename R>
static const void* chain_fn(int tile)
{
#define PC(n) case n: return (const void*) sweep3d_chain_kernel<R, \
        kTiles[n].bx, kTiles[n].by, kTiles[n].hb>;
    switch (tile) {
        PC(10) PC(11) PC(12) PC(13)
        default: return nullptr;
    }
#undef PC
}
static_assert(kTiles[10].mode == 2 && kTiles[11].mode == 2 &&
              kTiles[12].mode == 2 && kTiles[13].mode == 2 &&
              kTiles[13].kb == 1 && kNumTiles == 19,
              "chain_fn lists the mode-2 tiles with per-step mailbox reads");

// Tiles of the iteration-sweep box (flipped columns 1 .. n - 1).
static void chain_tiles(int nx, int ny, const TileShape& ts, int& nbx,
                        int& nby)
{
    nbx = (nx - 1 + ts.bx - 1)/ts.bx;
    nby = (ny - 1 + ts.by - 1)/ts.by;
}

extern "C" int fteik_k_chain_fits(int nx, int ny, int fp32, int tile)
{
    // a sweep's tiles must be resident together with at least one more
    // sweep's for the chain to overlap anything
    const void* fn = fp32 ? chain_fn<float>(tile) : chain_fn<double>(tile);
    if (!fn || nx < 2 || ny < 2) return 0;
    const TileShape ts = kTiles[tile];
    int nbx, nby, ncu = 0, dev = 0;
    chain_tiles(nx, ny, ts, nbx, nby);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess) return 0;
    const int per_cu = resident_per_cu(fn, tile_threads(tile));
    return (long) per_cu*ncu >= 2L*nbx*nby ? 1 : 0;
}

template <typename R>
static int launch_chain_t(const SweepLaunch& L, int dir0, int nq,
                          hipStream_t stream)
{
    const void* fn = chain_fn<R>(L.tile);
    if (!fn || !L.mb || !L.prog || nq < 1 || dir0 < 0 || dir0 > 7) return 2;
    const TileShape ts = kTiles[L.tile];
    SweepArgs<R> a;
    memset(&a, 0, sizeof(a));
    a.tt = (R*) L.tt; a.slow = (const R*) L.slow;
    a.cst = (const Consts<R>*) L.cst;
    a.flags = L.flags; a.err = L.err; a.serial = L.serial;
    a.nz = L.nz; a.nx = L.nx; a.ny = L.ny;
    int b[6];
    flipped_box(dir0 + 1, false, L.nz, L.nx, L.ny, 1, 1, 1, b);
    a.fz0 = b[0]; a.fz1 = b[1]; a.fx0 = b[2]; a.fx1 = b[3];
    a.fy0 = b[4]; a.fy1 = b[5];
    if (a.fz0 > a.fz1 || a.fx0 > a.fx1 || a.fy0 > a.fy1) return 0;
    chain_tiles(L.nx, L.ny, ts, a.nbx, a.nby);
    if (a.nbx*a.nby > L.max_flags) return 2;
    a.maxdelta = L.maxdelta;
    a.mb = L.mb;
    a.mbx = a.mby = nullptr;   // chains keep the per-node mailbox
    a.mbnt = 0;
    a.prog = L.prog;
    void* args[] = {&a, &dir0};
    hipError_t e = hipLaunchKernel(fn, dim3(nq*a.nbx*a.nby),
                                   dim3(tile_threads(L.tile)), args, 0,
                                   stream);
    return e == hipSuccess ? 0 : 1;
}

extern "C" int fteik_k_sweep3d_chain(const SweepLaunch* L, int dir0, int nq,
                                     hipStream_t stream)
{
    if (L->fp32) return launch_chain_t<float>(*L, dir0, nq, stream);
    return launch_chain_t<double>(*L, dir0, nq, stream);
}

extern "C" int fteik_k_pencil_fits(int nx, int ny, int fp32, int tile)
{
    // cooperative launch needs every block of the grid resident at once
    if (tile < 0 || tile >= kNumTiles) return 0;
    const TileShape ts = kTiles[tile];
    const int nblk = ((nx + ts.bx - 1)/ts.bx)*((ny + ts.by - 1)/ts.by);
    int ncu = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess) return 0;
    for (int v = 0; v < (ts.mode >= 1 ? 3 : 2); v++) {
        const void* fn = fp32 ? pencil_fn<float>(v, tile)
                              : pencil_fn<double>(v, tile);
        const int per_cu = resident_per_cu(fn, tile_threads(tile));
        if ((long) per_cu*ncu < nblk) return 0;
    }
    return 1;
}

extern "C" int fteik_k_num_tiles(void) { return kNumTiles; }

extern "C" void fteik_k_tile_shape(int tile, int* bx, int* by, int* hb,
                                   int* mode)
{
    const TileShape ts = kTiles[(tile >= 0 && tile < kNumTiles) ? tile : 0];
    *bx = ts.bx; *by = ts.by; *hb = ts.hb; *mode = ts.mode;
}

extern "C" int fteik_k_tile_mode(int tile)
{
    return (tile >= 0 && tile < kNumTiles) ? kTiles[tile].mode : -1;
}

extern "C" int fteik_k_max_flags(int nx, int ny)
{
    // flag slots for the smallest workgroup shape
    int m = 0;
    for (int t = 0; t < kNumTiles; t++) {
        const int n = ((nx + kTiles[t].bx - 1)/kTiles[t].bx)
                    * ((ny + kTiles[t].by - 1)/kTiles[t].by);
        m = (n > m) ? n : m;
    }
    return m;
}

extern "C" int fteik_k_source3d_init(void* cst, void* tt, const void* slow,
                                     const SourceGeom* g, long ngrd, int fp32,
                                     hipStream_t stream)
{
    const int nb = (int) std::min<long>((ngrd + 255)/256, 4096);
    if (fp32) {
        if (tt) hipLaunchKernelGGL(fill_kernel<float>, dim3(nb), dim3(256), 0,
                                   stream, (float*) tt, ngrd, (float) kHuge);
        hipLaunchKernelGGL(source3d_init_kernel<float>, dim3(1), dim3(64), 0,
                           stream, (Consts<float>*) cst, (float*) tt,
                           (const float*) slow, *g);
    } else {
        if (tt) hipLaunchKernelGGL(fill_kernel<double>, dim3(nb), dim3(256), 0,
                                   stream, (double*) tt, ngrd, kHuge);
        hipLaunchKernelGGL(source3d_init_kernel<double>, dim3(1), dim3(64), 0,
                           stream, (Consts<double>*) cst, (double*) tt,
                           (const double*) slow, *g);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

extern "C" int fteik_k_consts_to_float(const void* cd, void* cf,
                                       hipStream_t stream)
{
    hipLaunchKernelGGL(consts_to_float_kernel, dim3(1), dim3(64), 0, stream,
                       (const Consts<double>*) cd, (Consts<float>*) cf);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
