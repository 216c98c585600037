// sweep_pencil.h -- the pipelined column-tile sweep shared by the
// single-grid solver (sweep3d.hip) and the batched solver (batch3d.hip).
//
// One sweep of the reference (autoCode.f90:1001-1316 and siblings: every
// masked node updated once, level by level, levels = flipped iz+ix+iy)
// runs as a pipelined column wavefront:
//
//   * the (x, y) plane of the sweep's flipped frame is cut into tiles of
//     BX x BY columns; one workgroup owns a tile, one thread owns a column
//     and walks down it, one node per step;
//   * at step t every column processes the node on global level t, so all
//     nodes a node depends on were produced at steps t-1..t-3; values of
//     the last steps live in an LDS ring, one barrier per step;
//   * the upwind tiles (x-1, y-1, corner) publish a progress counter every
//     HB steps; a tile pulls its halo columns for the next HB steps with
//     coherent loads once its producers are far enough ahead;
//   * own old times and cell slownesses of the next HB nodes of a column
//     are prefetched into registers one batch ahead;
//   * tiles are numbered upwind-first, so every dependency points to a
//     lower workgroup id; every wait has a spin limit that turns a hang
//     into a reported error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "stencils.h"
#include "kernels.h"

namespace fteik {

// Tile shapes of the pipelined sweep: BX x BY columns per workgroup; HB is
// the halo batch (mode 0) or the prefetch depth of the compute waves (modes
// 1, 2).  The library picks one per grid at initialize; the autotuner
// (fteik_solver3d_autotune) times them.
// mode 0: upwind tiles publish a progress flag every hb steps and the
//         downwind tile pulls its halo after the flag (pencil_body);
// mode 1: boundary values go to a tagged mailbox and a dedicated halo wave
//         polls it ahead of the compute waves (pencil_mb_body);
// mode 2: as mode 1, plus a store wave that takes the global stores off the
//         compute waves.
// kb: mailbox steps a halo lane fetches per round trip (edge-major
//     mailbox, modes 1 and 2): it waits for the last step of the batch to
//     be posted, then loads the batch's words together, so a tile trails
//     its upwind neighbour by about kb steps but pays one mailbox round
//     trip per kb steps instead of one per step.
// st: staged loads (8 x 8 store-wave tiles with prefetch depth 4): the
//     compute wave loads a batch's cell layers and own old times
//     cooperatively, several lanes per column along z, into LDS, instead of
//     one cache line per lane and load (see pencil_mb_body).  Fewer memory
//     requests per step: faster where every CU holds several tiles (a 257^2
//     sweep plane), slower on small grids, where the extra LDS reads of a
//     step are not hidden.
struct TileShape { int bx, by, hb, mode, kb = 1, st = 0; };
constexpr TileShape kTiles[] = {
    {16, 16, 8, 0}, {16, 16, 4, 0}, {16, 16, 2, 0}, {8, 8, 4, 0},
    {8, 8, 2, 0}, {32, 8, 4, 0}, {16, 8, 4, 0},
    {16, 16, 4, 1}, {16, 8, 4, 1}, {8, 8, 4, 1},
    {16, 16, 4, 2}, {16, 8, 4, 2}, {8, 8, 4, 2},
    // 4 waves in at most 128 registers: four workgroups per CU, so two
    // sweeps' tiles of a 257^2 chain are resident together
    {16, 8, 2, 2},
    // batched mailbox reads
    {8, 8, 4, 2, 4}, {8, 8, 4, 2, 2}, {16, 8, 4, 2, 4}, {8, 8, 4, 2, 8},
    // staged loads
    {8, 8, 4, 2, 1, 1},
};
// Waves per SIMD the kernels of a tile shape are compiled for.  The
// hardware places a workgroup's waves on the SIMDs from the first one, so a
// 5- or 6-wave workgroup at 3 waves per SIMD leaves no room for a second.
constexpr int waves_per_simd(int hb) { return hb <= 2 ? 4 : 3; }
constexpr int kNumTiles = (int) (sizeof(kTiles)/sizeof(kTiles[0]));
constexpr unsigned long long kSpinLimit = 1ull << 22;

// ---------------------------------------------------------------- utils
__device__ __forceinline__ void store_coherent(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_coherent(float* p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_coherent(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float load_coherent(const float* p)
{
    return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Flipped 0-based coordinate -> actual 1-based index.
__device__ __forceinline__ int unflip(int f, int n, int sgn)
{
    return (sgn > 0) ? f + 1 : n - f;
}

// Tile (fbx*nby + fby) swept by the k-th workgroup of a launch over an
// nbx x nby plane of tiles: tiles in order of their diagonal fbx + fby, then
// of fbx.  A sweep starts at tile (0, 0) and tile (fbx, fby) can start about
// one tile width of levels after its upwind neighbours (fbx - 1, fby) and
// (fbx, fby - 1), which come earlier in this order.  When not every
// workgroup of a launch is resident, workgroups are dispatched in order, so
// the tiles are dispatched about when their inputs arrive; in row-major
// order the last rows of tiles would wait for slots freed by the first and
// lengthen the sweep.  The loop is uniform (scalar) and runs once per
// workgroup.  -1 when k is not a tile rank (k >= nbx*nby).
__device__ __forceinline__ int tile_of_rank(int k, int nbx, int nby)
{
    for (int d = 0; d < nbx + nby - 1; d++) {
        const int lo = max(0, d - (nby - 1)), hi = min(d, nbx - 1);
        const int n = hi - lo + 1;
        if (k < n) return (lo + k)*nby + (d - lo - k);
        k -= n;
    }
    return -1;
}

// ------------------------------------------------------- pencil sweep
template <typename R>
struct SweepArgs {
    R* tt;
    const R* slow;
    const Consts<R>* cst;
    unsigned long long* flags;   // one per block, (serial << 32) | step+1
    int* err;
    unsigned long long serial;
    int nz, nx, ny;
    int sz, sx, sy;              // time signs of the sweep
    int fz0, fz1, fx0, fx1, fy0, fy1;   // update box, flipped 0-based
    int nbx, nby;                // blocks along flipped x, y
    double* maxdelta;            // optional: max |t_new - t_old| (as bits)
    unsigned long long* mb;      // mailbox (mode 1): 2 words/node FP64, 1 FP32
    // Edge-major mailbox (sweeps outside chains; null: the per-node one
    // above).  The last column of x-tile ex posts its node of row fk at
    // step t as entry (ex*mbnt + t)*ny + fk of mbx, the last row of y-tile
    // ey its node of column fj as entry (ey*mbnt + t)*nx + fj of mby: the
    // values one halo edge reads at a step are contiguous, a few cache
    // lines instead of one per column.  Entries have the per-node layout's
    // size (2 words FP64, 1 FP32).
    unsigned long long* mbx;
    unsigned long long* mby;
    int mbnt;                    // steps per tile edge (> the last step)
    // Tile skipping (mailbox tiles, iteration sweeps; null = off):
    //   colmod[k*nx + j]  launch serial of the last change of column (j, k)
    //   plan[0]           number of tiles processed in this launch
    //   plan[1 + w]       tile of workgroup w (w < plan[0])
    //   plan[1 + ntiles + t]  1 if tile t is skipped in this launch
    //   plan[1 + 2*ntiles + t]  dirty mark of tile t (plan-building scratch)
    unsigned long long* colmod;
    const int* plan;
    // Sweep chains (consecutive iteration sweeps in one launch; sweep q
    // uses progress set par = q & 1):
    //   prog[par*ntiles + tile] = (serial << 32) | last step stored, or
    //                     (serial << 32) | kDone once the tile is finished
    //   pserial           serial of the preceding sweep (0: none)
    //   psz, psx, psy     its signs
    unsigned long long* prog;
    unsigned long long pserial;
    int par;
    int psz, psx, psy;
};

// The per-sweep values of a mailbox sweep, kept apart from SweepArgs so a
// kernel that runs several sweeps (chains) need not copy SweepArgs.
struct SweepDyn {
    int sz, sx, sy;                  // time signs of the sweep
    int psz, psx, psy;               // those of the preceding sweep (chains)
    int par;                         // progress-word set (chains)
    unsigned long long serial;       // launch serial / mailbox tag
    unsigned long long pserial;      // preceding sweep's serial (0: none)
};

template <typename R>
__device__ __forceinline__ SweepDyn dyn_of(const SweepArgs<R>& a)
{
    SweepDyn d;
    d.sz = a.sz; d.sx = a.sx; d.sy = a.sy;
    d.psz = a.psz; d.psx = a.psx; d.psy = a.psy;
    d.par = a.par; d.serial = a.serial; d.pserial = a.pserial;
    return d;
}

constexpr unsigned kDone = 0x7fffffffu;
// s_waitcnt operand: vmcnt(0), expcnt and lgkmcnt not waited for
constexpr int kWaitVmcnt0 = 0x0f70;

// Device-coherent field access for sweep chains: the preceding sweep of a
// chain runs in the same launch on other CUs (other L2s).
__device__ __forceinline__ double ld_coh(const double* p)
{
    return __longlong_as_double((long long) __hip_atomic_load(
        (unsigned long long*) p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ float ld_coh(const float* p)
{
    return __uint_as_float(__hip_atomic_load(
        (unsigned*) p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_coh(double* p, double v)
{
    __hip_atomic_store((unsigned long long*) p,
                       (unsigned long long) __double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(float* p, float v)
{
    __hip_atomic_store((unsigned*) p, __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <typename R>
__device__ __forceinline__ void slowness7(const R* __restrict__ slow, int nz,
                                          int nx, int ny, int i, int j, int k,
                                          int sz, int sx, int sy, Slw<R>& s)
{
    // autoCode.f90:407-445: all 19 cells come from the clamped 2x2x2 cell
    // neighbourhood; i1 = i - sgnv is im for + sweeps and ip for - sweeps.
    const int im = max(i - 1, 1), ip = min(i, nz - 1);
    const int jm = max(j - 1, 1), jp = min(j, nx - 1);
    const int km = max(k - 1, 1), kp = min(k, ny - 1);
    const long sz1 = nz - 1, szx = (long) (nz - 1)*(nx - 1);
#define SC(a, b, c) slow[(long) ((c) - 1)*szx + (long) ((b) - 1)*sz1 + (a) - 1]
    const R c000 = SC(im, jm, km), c100 = SC(ip, jm, km);
    const R c010 = SC(im, jp, km), c110 = SC(ip, jp, km);
    const R c001 = SC(im, jm, kp), c101 = SC(ip, jm, kp);
    const R c011 = SC(im, jp, kp), c111 = SC(ip, jp, kp);
#undef SC
    // select by the sweep's sgnv: index 0 -> (i-1 side), 1 -> (i side)
    const bool zi = (sz < 0), xj = (sx < 0), yk = (sy < 0);
    // cells in plane i1 over (j, k)
    const R a00 = zi ? c100 : c000, a10 = zi ? c110 : c010;
    const R a01 = zi ? c101 : c001, a11 = zi ? c111 : c011;
    s.s1 = rmin(rmin(rmin(a00, a01), a10), a11);
    // cells in plane j1 over (i, k)
    const R b00 = xj ? c010 : c000, b10 = xj ? c110 : c100;
    const R b01 = xj ? c011 : c001, b11 = xj ? c111 : c101;
    s.s2 = rmin(rmin(rmin(b00, b10), b01), b11);
    // cells in plane k1 over (i, j)
    const R d00 = yk ? c001 : c000, d10 = yk ? c101 : c100;
    const R d01 = yk ? c011 : c010, d11 = yk ? c111 : c110;
    s.s3 = rmin(rmin(rmin(d00, d01), d10), d11);
    // line (i1, j1) over k
    const R e0 = zi ? (xj ? c110 : c100) : (xj ? c010 : c000);
    const R e1 = zi ? (xj ? c111 : c101) : (xj ? c011 : c001);
    s.s4 = rmin(e0, e1);
    // line (i1, k1) over j
    const R f0 = zi ? (yk ? c101 : c100) : (yk ? c001 : c000);
    const R f1 = zi ? (yk ? c111 : c110) : (yk ? c011 : c010);
    s.s5 = rmin(f0, f1);
    // line (j1, k1) over i
    const R g0 = xj ? (yk ? c011 : c010) : (yk ? c001 : c000);
    const R g1 = xj ? (yk ? c111 : c110) : (yk ? c101 : c100);
    s.s6 = rmin(g0, g1);
    // cell (i1, j1, k1)
    s.s7 = zi ? (xj ? (yk ? c111 : c110) : (yk ? c101 : c100))
              : (xj ? (yk ? c011 : c010) : (yk ? c001 : c000));
}

// The 7 min-reduced slownesses from the node's 2x2x2 cell neighbourhood
// given as u (upwind, the i1/j1/k1 side of autoCode.f90:407-409) and d
// (other side) along each axis: c[z][x][y], index 0 = u, 1 = d.
template <typename R>
__device__ __forceinline__ void slowness7_ud(const R (&zu)[4], const R (&zd)[4],
                                             Slw<R>& s)
{
    // zu/zd hold (x,y) = (u,u), (u,d), (d,u), (d,d); lmin: the cell values
    // come from memory and are never NaN (a minimum is exact, so the order
    // of the reductions does not matter)
    s.s4 = lmin(zu[0], zu[1]);                                 // line i1,j1
    s.s5 = lmin(zu[0], zu[2]);                                 // line i1,k1
    s.s6 = lmin(zu[0], zd[0]);                                 // line j1,k1
    s.s1 = lmin(lmin(s.s4, zu[2]), zu[3]);                     // plane i1
    s.s2 = lmin(lmin(s.s6, zu[1]), zd[1]);                     // plane j1
    s.s3 = lmin(lmin(s.s5, zd[0]), zd[2]);                     // plane k1
    s.s7 = zu[0];                                              // cell
}

template <typename R, bool INIT, int BX, int BY, int HB>
__device__ __forceinline__ void pencil_body(const SweepArgs<R>& a, const int wg)
{
    constexpr int NT = BX*BY;
    constexpr int RING = (HB + 3 <= 8) ? 8 : 16;     // LDS ring depth
    constexpr int HCOLS = BX + BY + 1;
    static_assert(HB + 3 <= RING, "ring too shallow");
    __shared__ R ring[RING][(BX + 1)*(BY + 1)];
    __shared__ int s_abort;
    const int fbx = wg / a.nby, fby = wg - (wg / a.nby)*a.nby;
    const int fj0 = fbx*BX, fk0 = fby*BY;
    const int lj = threadIdx.x % BX, lk = threadIdx.x / BX;
    const int fj = fj0 + lj, fk = fk0 + lk;
    const int nz = a.nz, nx = a.nx, ny = a.ny;
    // extended box: update box plus the one frozen layer its nodes read
    const int ez0 = a.fz0 - 1, ex0 = a.fx0 - 1, ey0 = a.fy0 - 1;
    const int jlo = max(fj0, ex0), jhi = min(min(fj0 + BX - 1, a.fx1), nx - 1);
    const int klo = max(fk0, ey0), khi = min(min(fk0 + BY - 1, a.fy1), ny - 1);
    unsigned long long* myflag = a.flags + wg;
    const unsigned long long hi = a.serial << 32;
    if (jlo > jhi || klo > khi) {
        if (threadIdx.x == 0)
            __hip_atomic_store(myflag, hi | 0xffffffffull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const int tlo = ez0 + jlo + klo;
    const int thi = a.fz1 + jhi + khi;
    const bool colE = (fj >= jlo && fj <= jhi && fk >= klo && fk <= khi);
    const bool colB = colE && fj >= a.fx0 && fk >= a.fy0;
    // Loads of threads outside the extended box use a clamped in-grid
    // column; their values are never used.
    const int cj = min(max(fj, jlo), jhi), ck = min(max(fk, klo), khi);
    const int j = unflip(cj, nx, a.sx), k = unflip(ck, ny, a.sy);
    R* const col = a.tt + (long) (k - 1)*nz*nx + (long) (j - 1)*nz;
    // the column's four cell columns (autoCode.f90:407-445 clamps)
    const int jm = max(j - 1, 1), jp = min(j, nx - 1);
    const int km = max(k - 1, 1), kp = min(k, ny - 1);
    const int xu = (a.sx > 0) ? jm : jp, xd = (a.sx > 0) ? jp : jm;
    const int yu = (a.sy > 0) ? km : kp, yd = (a.sy > 0) ? kp : km;
    const long szx = (long) (nz - 1)*(nx - 1);
    const R* const c_uu = a.slow + (long) (yu - 1)*szx + (long) (xu - 1)*(nz - 1);
    const R* const c_ud = a.slow + (long) (yd - 1)*szx + (long) (xu - 1)*(nz - 1);
    const R* const c_du = a.slow + (long) (yu - 1)*szx + (long) (xd - 1)*(nz - 1);
    const R* const c_dd = a.slow + (long) (yd - 1)*szx + (long) (xd - 1)*(nz - 1);
    const int fdep = cj + ck;           // node of step t is fi = t - fdep
    const Consts<R> c = *a.cst;
    const int me = (lk + 1)*(BX + 1) + (lj + 1);
    const int cW = me - 1, cS = me - (BX + 1), cSW = me - (BX + 2);
    R dmax = R(0);

    // 0-based offset of flipped node fi in its column, clamped to the box
    auto node_off = [&](int fi) -> int {
        fi = min(max(fi, ez0), a.fz1);
        return (a.sz > 0) ? fi : nz - 1 - fi;
    };
    // 0-based z offset of the cell layer on the upwind side of flipped
    // node f (its i1 plane); the downwind layer of f is the upwind one of
    // f + 1.  Clamped as max(i-1,1) / min(i,nz-1) clamp the cell index.
    auto layer_off = [&](int f) -> int {
        const int zc = (a.sz > 0) ? f : nz - f;          // 1-based cell
        return min(max(zc, 1), nz - 1) - 1;
    };

    // Register double buffer: own old times and cell layers of the next
    // HB steps are loaded one batch ahead, so no step waits on memory.
    R tn[HB], cn[HB + 1][4];
    auto prefetch = [&](int tb) {
        const int f0 = tb - fdep;
        #pragma unroll
        for (int q = 0; q < HB; q++) tn[q] = col[node_off(f0 + q)];
        #pragma unroll
        for (int q = 0; q <= HB; q++) {
            const int z = layer_off(f0 + q);
            cn[q][0] = c_uu[z]; cn[q][1] = c_ud[z];
            cn[q][2] = c_du[z]; cn[q][3] = c_dd[z];
        }
    };

    if (threadIdx.x == 0) {
        s_abort = 0;
        __hip_atomic_store(myflag, hi | (unsigned long long) tlo,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // The initialization sweeps (heavier stencil, once per solve) load each
    // batch at its start instead, which keeps their register count low
    // enough for two workgroups per CU.
    if (!INIT) prefetch(tlo);
    __syncthreads();

    for (int tb = tlo; tb <= thi; tb += HB) {
        if (INIT) prefetch(tb);
        R tc[HB], cc[HB + 1][4];
        #pragma unroll
        for (int q = 0; q < HB; q++) tc[q] = tn[q];
        #pragma unroll
        for (int q = 0; q <= HB; q++) {
            cc[q][0] = cn[q][0]; cc[q][1] = cn[q][1];
            cc[q][2] = cn[q][2]; cc[q][3] = cn[q][3];
        }
        if (!INIT && tb + HB <= thi) prefetch(tb + HB);

        // halo batch: steps tb-3 .. tb+HB-2 of the upwind columns (the three
        // oldest are needed by the first batch only)
        {
            const int need = tb + HB - 2;
            if (threadIdx.x < 3) {
                int pbx = fbx - (threadIdx.x != 1 ? 1 : 0);
                int pby = fby - (threadIdx.x != 0 ? 1 : 0);
                if (pbx >= 0 && pby >= 0) {
                    const unsigned long long* pf = a.flags + pbx*a.nby + pby;
                    const unsigned long long want =
                        hi | (unsigned long long) (need + 1);
                    unsigned long long spins = 0;
                    while (__hip_atomic_load(pf, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) < want) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > kSpinLimit) {
                            atomicExch(a.err, 1);
                            s_abort = 1;
                            break;
                        }
                    }
                }
            }
            __syncthreads();
            if (s_abort) {
                // release the blocks that wait on this one; the error flag
                // marks the whole sweep as failed
                if (threadIdx.x == 0)
                    __hip_atomic_store(myflag, hi | 0xffffffffull,
                                       __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            for (int h = threadIdx.x; h < HCOLS*(HB + 2); h += NT) {
                int hc = h / (HB + 2), d = h - (h / (HB + 2))*(HB + 2);
                int st = tb - 3 + d;
                int hj, hk, slot;
                if (hc < BY)            { hj = fj0 - 1; hk = fk0 + hc;
                                          slot = (hc + 1)*(BX + 1); }
                else if (hc < BY + BX)  { hj = fj0 + (hc - BY); hk = fk0 - 1;
                                          slot = hc - BY + 1; }
                else                    { hj = fj0 - 1; hk = fk0 - 1; slot = 0; }
                int fi = st - hj - hk;
                if (hj >= 0 && hk >= 0 && hj < nx && hk < ny && fi >= 0 &&
                    fi < nz) {
                    int i = unflip(fi, nz, a.sz);
                    int jj = unflip(hj, nx, a.sx), kk = unflip(hk, ny, a.sy);
                    long g = (long) (kk - 1)*nz*nx + (long) (jj - 1)*nz + i - 1;
                    ring[st & (RING - 1)][slot] = load_coherent(a.tt + g);
                }
            }
            __syncthreads();
        }

        #pragma unroll
        for (int q = 0; q < HB; q++) {
            const int t = tb + q;
            if (t > thi) break;                       // block-uniform
            const int fi = t - fj - fk;
            if (colE && fi >= ez0 && fi <= a.fz1) {
                R v = tc[q];
                if (colB && fi >= a.fz0) {
                    const int s1 = (t - 1) & (RING - 1),
                              s2 = (t - 2) & (RING - 1),
                              s3 = (t - 3) & (RING - 1);
                    Nbr<R> nb;
                    nb.tv   = ring[s1][me];
                    nb.te   = ring[s1][cW];
                    nb.tn   = ring[s1][cS];
                    nb.tev  = ring[s2][cW];
                    nb.ten  = ring[s2][cSW];
                    nb.tnv  = ring[s2][cS];
                    nb.tnve = ring[s3][cSW];
                    Slw<R> sl;
                    slowness7_ud<R>(cc[q], cc[q + 1], sl);
                    R vn;
                    if (INIT) {
                        const int i = unflip(fi, nz, a.sz);
                        vn = init3d(c, nb, sl, v, i, j, k, a.sz, a.sx, a.sy);
                    } else {
                        vn = noinit3d(c, nb, sl, v);
                    }
                    dmax = rmax(dmax, v - vn);
                    v = vn;
                    store_coherent(col + node_off(fi), v);
                }
                ring[t & (RING - 1)][me] = v;
            }
            if (q == HB - 1 || t == thi) {
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                if (threadIdx.x == 0)
                    __hip_atomic_store(myflag,
                                       hi | (unsigned long long) (t + 1),
                                       __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __syncthreads();
            }
        }
    }
    if (threadIdx.x == 0)
        __hip_atomic_store(myflag, hi | 0xffffffffull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (a.maxdelta && dmax > R(0)) {
        // positive doubles order like their bit patterns
        double dd = (double) dmax;
        atomicMax((unsigned long long*) a.maxdelta,
                  (unsigned long long) __double_as_longlong(dd));
    }
}

// ------------------------------------------------------------ mailbox
// A node value tagged with the launch serial in each 64-bit word, so a
// reader can tell this sweep's value from an older one without any
// ordering between separate stores: FP64 uses two words (low and high
// halves), FP32 one.
// FP64: both words go out in one 16-byte device-coherent (sc1) store and
// come back in one such load: half the memory requests of two 8-byte
// accesses.  Each word still carries the tag, so a reader that sees the two
// halves from different stores rejects the pair.
typedef unsigned long long MbPair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void mb_put(unsigned long long* mb, long g,
                                       double v, unsigned tag)
{
    const unsigned long long b = (unsigned long long) __double_as_longlong(v);
    const unsigned long long t = (unsigned long long) tag << 32;
    MbPair w;
    w.x = t | (b & 0xffffffffull);
    w.y = t | (b >> 32);
    asm volatile("global_store_dwordx4 %0, %1, off sc1"
                 : : "v"(mb + 2*g), "v"(w) : "memory");
}
__device__ __forceinline__ void mb_put(unsigned long long* mb, long g,
                                       float v, unsigned tag)
{
    __hip_atomic_store(mb + g, ((unsigned long long) tag << 32)
                               | (unsigned long long) __float_as_uint(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool mb_get(unsigned long long* mb, long g,
                                       unsigned tag, double& v)
{
    MbPair w;
    asm volatile("global_load_dwordx4 %0, %1, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=v"(w) : "v"(mb + 2*g) : "memory");
    const unsigned long long lo = w.x, hi = w.y;
    if ((unsigned) (lo >> 32) != tag || (unsigned) (hi >> 32) != tag)
        return false;
    v = __longlong_as_double((long long) ((hi << 32) | (lo & 0xffffffffull)));
    return true;
}
__device__ __forceinline__ bool mb_get(unsigned long long* mb, long g,
                                       unsigned tag, float& v)
{
    const unsigned long long w = __hip_atomic_load(mb + g, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned) (w >> 32) != tag) return false;
    v = __uint_as_float((unsigned) (w & 0xffffffffull));
    return true;
}

// KB mailbox entries of this launch in one round trip: the loads are
// issued together and waited for once (in one asm block, so no result
// register is live while a load is in flight outside it).  ok[q] tells
// whether entry q carried the tag.
template <int KB>
__device__ __forceinline__ void mb_get_batch(unsigned long long* mb,
                                             const long (&g)[KB],
                                             unsigned tag, double (&v)[KB],
                                             bool (&ok)[KB])
{
    MbPair w[KB];
    if constexpr (KB == 2) {
        asm volatile("global_load_dwordx4 %0, %2, off sc1\n\t"
                     "global_load_dwordx4 %1, %3, off sc1\n\t"
                     "s_waitcnt vmcnt(0)"
                     : "=&v"(w[0]), "=&v"(w[1])
                     : "v"(mb + 2*g[0]), "v"(mb + 2*g[1]) : "memory");
    } else if constexpr (KB == 4) {
        asm volatile("global_load_dwordx4 %0, %4, off sc1\n\t"
                     "global_load_dwordx4 %1, %5, off sc1\n\t"
                     "global_load_dwordx4 %2, %6, off sc1\n\t"
                     "global_load_dwordx4 %3, %7, off sc1\n\t"
                     "s_waitcnt vmcnt(0)"
                     : "=&v"(w[0]), "=&v"(w[1]), "=&v"(w[2]), "=&v"(w[3])
                     : "v"(mb + 2*g[0]), "v"(mb + 2*g[1]),
                       "v"(mb + 2*g[2]), "v"(mb + 2*g[3]) : "memory");
    } else {
        static_assert(KB == 8, "batches of 2, 4 or 8 entries");
        asm volatile("global_load_dwordx4 %0, %8, off sc1\n\t"
                     "global_load_dwordx4 %1, %9, off sc1\n\t"
                     "global_load_dwordx4 %2, %10, off sc1\n\t"
                     "global_load_dwordx4 %3, %11, off sc1\n\t"
                     "global_load_dwordx4 %4, %12, off sc1\n\t"
                     "global_load_dwordx4 %5, %13, off sc1\n\t"
                     "global_load_dwordx4 %6, %14, off sc1\n\t"
                     "global_load_dwordx4 %7, %15, off sc1\n\t"
                     "s_waitcnt vmcnt(0)"
                     : "=&v"(w[0]), "=&v"(w[1]), "=&v"(w[2]), "=&v"(w[3]),
                       "=&v"(w[4]), "=&v"(w[5]), "=&v"(w[6]), "=&v"(w[7])
                     : "v"(mb + 2*g[0]), "v"(mb + 2*g[1]),
                       "v"(mb + 2*g[2]), "v"(mb + 2*g[3]),
                       "v"(mb + 2*g[4]), "v"(mb + 2*g[5]),
                       "v"(mb + 2*g[6]), "v"(mb + 2*g[7]) : "memory");
    }
    #pragma unroll
    for (int q = 0; q < KB; q++) {
        const unsigned long long lo = w[q].x, hi = w[q].y;
        ok[q] = (unsigned) (lo >> 32) == tag && (unsigned) (hi >> 32) == tag;
        v[q] = __longlong_as_double((long long) ((hi << 32) |
                                                 (lo & 0xffffffffull)));
    }
}
template <int KB>
__device__ __forceinline__ void mb_get_batch(unsigned long long* mb,
                                             const long (&g)[KB],
                                             unsigned tag, float (&v)[KB],
                                             bool (&ok)[KB])
{
    unsigned long long w[KB];
    #pragma unroll
    for (int q = 0; q < KB; q++)
        w[q] = __hip_atomic_load(mb + g[q], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    #pragma unroll
    for (int q = 0; q < KB; q++) {
        ok[q] = (unsigned) (w[q] >> 32) == tag;
        v[q] = __uint_as_float((unsigned) (w[q] & 0xffffffffull));
    }
}

__device__ __forceinline__ int lds_load(int* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Relaxed read of a step counter; the caller branches on the value before
// touching the data it guards, and LDS operations of one wave execute in
// order, so a compiler fence after the wait is all the ordering needed.
__device__ __forceinline__ int lds_load_rx(int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Mailbox sweep of one tile: BX*BY compute threads (one column each), one
// halo wave and, in mode 2, one store wave.
//   * The halo wave polls the mailbox for the upwind columns' values of the
//     coming steps and writes them into the LDS ring.
//   * Compute waves wait on LDS step counters instead of workgroup
//     barriers: wave w needs the halo and wave w - 1 (the rows before its
//     own), so the waves of a tile pipeline.
//   * Mode 1: compute waves store their values and post tile-boundary
//     values themselves.  Mode 2: the store wave copies each step's values
//     from the ring to the field and the mailbox, so compute waves issue no
//     global stores and waiting for their prefetched loads never waits for
//     stores (one counter covers both on this GPU).
// No progress flags and no wait for store completion are needed: every
// mailbox word carries the launch serial.
//
// SKIP (iteration sweeps): only the tiles of a.plan run; a processed tile
// stamps the columns it changed in a.colmod.
//
// CHAIN (iteration sweeps, mode 2): the preceding sweep of the chain may
// still be running.  Node m may be read or overwritten once the preceding
// sweep A has finalized it and read it for the last time, i.e. once A has
// passed level L_A(m) + 3 in the A tiles holding m's column and its
// neighbours.  Over the nodes a tile touches at step t (its columns, the
// halo row before them and the one after, node and z-upwind node) that
// bound is
//     z as in A:  t + 3 + sum over x, y flipped of (n + 1 - 2 f0)
//     z flipped:  n_z - t_first + 3 + sum over x, y of (2 (f0 + B) if as in
//                 A, else n - 1)
// (f0 = the tile's first flipped column, B = its width).  The halo wave
// polls the progress words of those A tiles (at most 3 x 3) ahead of the
// steps it releases.  The store wave publishes the tile's own progress with
// a release after its stores; field loads and stores are device-coherent.
template <typename R, bool INIT, int BX, int BY, int HB, int MODE,
          bool SKIP, bool CHAIN = false, int KB = 1, bool ST = false>
__device__ __forceinline__ void pencil_mb_body(const SweepArgs<R>& a,
                                               const SweepDyn& d,
                                               const int wg)
{
    static_assert(KB == 1 || (!CHAIN && 8 % KB == 0),
                  "batched mailbox reads: edge-major mailbox, KB | 8");
    static_assert(!(INIT && SKIP), "tile skipping is for iteration sweeps");
    static_assert(!CHAIN || (!INIT && !SKIP && MODE == 2),
                  "sweep chains: iteration sweeps of mode-2 tiles");
    // MODE 1: compute waves store their own values (two waves of support
    // fewer per workgroup); MODE 2: a store wave does (see above)
    constexpr bool SW = (MODE == 2);
    constexpr int NC = BX*BY;
    constexpr int NW = NC/64;
    constexpr int RING = 16;
    constexpr int HCOLS = BX + BY + 1;
    static_assert(NC % 64 == 0 && HCOLS <= 64 && HB + 3 <= RING, "tile");
    __shared__ R ring[RING][(BX + 1)*(BY + 1)];
    // STAGE (staged-load tiles, iteration sweeps outside chains): the cell
    // layers of a batch of steps are loaded by the compute wave
    // cooperatively, 8 lanes per cell column along z, and kept in
    // cbuf[cell column][layer], the own old times likewise (HB lanes per
    // column) in tbuf: the 81 cell columns and 64 node columns of the tile
    // cost a few cache lines per load instruction instead of one per lane.
    static_assert(!ST || (MODE == 2 && NC == 64 && HB == 4 && KB == 1),
                  "staged loads: 8 x 8 store-wave tiles, prefetch depth 4");
    constexpr bool STAGE = ST && !INIT && !CHAIN;
    constexpr int NSL = (BX + 1)*(BY + 1);        // cell columns of the tile
    constexpr int NSE = (NSL*8 + 63)/64;          // load rounds of 64 lanes
    __shared__ R cbuf[STAGE ? NSL*8 : 1];
    __shared__ R tbuf[STAGE ? NC*HB : 1];         // own old times of a batch
    // s_wdone[w]: last step compute wave w completed (wave w reads only
    // wave w-1's cells and the halo, so waves pipeline instead of stepping
    // together); s_hready: halo released through step s_hready - 1;
    // s_sdone: last step the store wave wrote out
    __shared__ int s_wdone[NW], s_hready, s_sdone, s_abort;
    // Tiles cover the box's columns from (fx0, fy0): the column row before
    // the box (fx0 - 1 or fy0 - 1) is not updated in this launch and is
    // read from the field.  At n = 257 the 256 updated columns per axis
    // make exactly 256/BX by 256/BY tiles.
    const int fbx = wg / a.nby, fby = wg - (wg / a.nby)*a.nby;
    const int fj0 = a.fx0 + fbx*BX, fk0 = a.fy0 + fby*BY;
    const int nz = a.nz, nx = a.nx, ny = a.ny;
    const int ez0 = a.fz0 - 1, ex0 = a.fx0 - 1, ey0 = a.fy0 - 1;
    const int jlo = fj0, jhi = min(fj0 + BX - 1, a.fx1);
    const int klo = fk0, khi = min(fk0 + BY - 1, a.fy1);
    if (jlo > jhi || klo > khi) return;
    const int tlo = ez0 + jlo + klo;
    const int thi = a.fz1 + jhi + khi;
    const unsigned tag = (unsigned) d.serial;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NW; w++) s_wdone[w] = tlo - 1;
        s_hready = tlo - 3; s_sdone = tlo - 1; s_abort = 0;
    }
    __syncthreads();
    auto wdone_min = [&]() -> int {
        int m = lds_load(&s_wdone[0]);
        #pragma unroll
        for (int w = 1; w < NW; w++) m = min(m, lds_load(&s_wdone[w]));
        return m;
    };

    if (SW && threadIdx.x >= NC + 64) {
        // --------------------------------------------------- store wave
        // NW columns per lane; their field offsets, diagonal and ring cell
        // are computed once
        const int sl = threadIdx.x - NC - 64;
        long sbase[NW];
        int sdep[NW], sring[NW], sflag[NW], sfj[NW], sfk[NW];
        #pragma unroll
        for (int c = 0; c < NW; c++) {
            const int cix = sl + 64*c;
            const int lj = cix % BX, lk = cix / BX;
            const int fj = fj0 + lj, fk = fk0 + lk;
            const bool colE = (fj >= jlo && fj <= jhi && fk >= klo &&
                               fk <= khi);
            const bool post = (lj == BX - 1 || lk == BY - 1);
            sflag[c] = colE ? (1 | (post ? 2 : 0) | (lj == BX - 1 ? 4 : 0)
                               | (lk == BY - 1 ? 8 : 0)) : 0;
            sdep[c] = fj + fk;
            sfj[c] = fj;
            sfk[c] = fk;
            sring[c] = (lk + 1)*(BX + 1) + lj + 1;
            sbase[c] = colE ? (long) (unflip(fk, ny, d.sy) - 1)*nz*nx
                              + (long) (unflip(fj, nx, d.sx) - 1)*nz : 0;
        }
        for (int t = tlo; t <= thi; t++) {
            unsigned long long spins = 0;
            while (wdone_min() < t) {
                if (lds_load(&s_abort) || ++spins > kSpinLimit) return;
                __builtin_amdgcn_s_sleep(1);
            }
            #pragma unroll
            for (int c = 0; c < NW; c++) {
                const int fi = t - sdep[c];
                if (!(sflag[c] & 1) || fi < ez0 || fi > a.fz1) continue;
                const bool upd = fi >= a.fz0, post = (sflag[c] & 2) != 0;
                if (!upd && !post) continue;
                const long g = sbase[c] + ((d.sz > 0) ? fi : nz - 1 - fi);
                const R v = ring[t & (RING - 1)][sring[c]];
                if (upd) {
                    if (CHAIN) st_coh(a.tt + g, v);
                    else a.tt[g] = v;
                }
                if (post) {
                    if (!CHAIN && a.mbx) {
                        if (sflag[c] & 4)
                            mb_put(a.mbx, ((long) fbx*a.mbnt + t)*ny + sfk[c],
                                   v, tag);
                        if (sflag[c] & 8)
                            mb_put(a.mby, ((long) fby*a.mbnt + t)*nx + sfj[c],
                                   v, tag);
                    } else {
                        mb_put(a.mb, g, v, tag);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            if (sl == 0)
                __hip_atomic_store(&s_sdone, t, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            if (CHAIN && (((t - tlo) & 3) == 3 || t == thi)) {
                // this wave's field and mailbox stores are device-coherent
                // (written through the L2): once they have completed, the
                // progress word may announce them.  No cache write-back or
                // invalidation is involved on either side.
                __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                if (sl == 0)
                    __hip_atomic_store(a.prog + (long) d.par*a.nbx*a.nby + wg,
                                       (d.serial << 32) |
                                       (t == thi ? kDone : (unsigned) t),
                                       __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        return;
    }

    if (threadIdx.x >= NC) {
        // ---------------------------------------------------- halo wave
        const int hl = threadIdx.x - NC;
        int hj = 0, hk = 0, slot = 0;
        if (hl < BY)            { hj = fj0 - 1; hk = fk0 + hl;
                                  slot = (hl + 1)*(BX + 1); }
        else if (hl < BY + BX)  { hj = fj0 + (hl - BY); hk = fk0 - 1;
                                  slot = hl - BY + 1; }
        else                    { hj = fj0 - 1; hk = fk0 - 1; slot = 0; }
        // halo columns inside the extended box (one column row before it)
        const bool hcol = hl < HCOLS && hj >= ex0 && hk >= ey0 &&
                          hj <= a.fx1 && hk <= a.fy1;
        const long gcol = hcol ? (long) (unflip(hk, ny, d.sy) - 1)*nz*nx
                               + (long) (unflip(hj, nx, d.sx) - 1)*nz : 0;
        // columns outside the box, and those of a skipped upwind tile, are
        // not written in this launch: they are read from the field instead
        // of the mailbox
        bool direct = hj < a.fx0 || hk < a.fy0;
        if (SKIP && hcol && !direct) {
            const int ut = ((hj - a.fx0)/BX)*a.nby + (hk - a.fy0)/BY;
            direct = a.plan[1 + a.nbx*a.nby + ut] != 0;
        }
        // Values read from the field do not change during the launch: they
        // are loaded HD steps ahead (a double-buffered register window) so
        // their latency is not paid once per step.  Mailbox words are
        // polled one step at a time.
        constexpr int HD = 8;
        const int s_beg = tlo - 3, s_end = thi - 1;
        // Sweep chains: lane l < 9 watches one tile of the preceding sweep
        // (see above).  Invariant: once step s is released, the preceding
        // sweep is done with every node this tile touches through step
        // s + 2*HD (the compute waves load their columns up to 2*HB steps
        // ahead, the direct halo values up to 2*HD).
        int ptile = -1;
        unsigned long long seen = 0;
        int verified = s_beg - 1;
        if (CHAIN && d.pserial) {
            auto trange = [&](int f0, int bw, bool same, int n, int b0,
                              int nb, int& t0, int& t1) {
                int lo = same ? f0 - 1 : n - 1 - f0 - bw;
                int hi = same ? f0 + bw : n - f0;
                lo = min(max(lo, b0), n - 1);
                hi = min(max(hi, b0), n - 1);
                t0 = min((lo - b0)/bw, nb - 1);
                t1 = min((hi - b0)/bw, nb - 1);
            };
            int tx0, tx1, ty0, ty1;
            trange(fj0, BX, d.psx == d.sx, nx, a.fx0, a.nbx, tx0, tx1);
            trange(fk0, BY, d.psy == d.sy, ny, a.fy0, a.nby, ty0, ty1);
            const int ntx = tx1 - tx0 + 1;
            if (hl < ntx*(ty1 - ty0 + 1))
                ptile = (tx0 + hl % ntx)*a.nby + ty0 + hl / ntx;
        }
        auto verify = [&](int T) {
            if (!CHAIN || !d.pserial || T <= verified) return;
            int need;
            if (d.psz == d.sz) {
                need = T + 3;
                if (d.psx != d.sx) need += nx + 1 - 2*fj0;
                if (d.psy != d.sy) need += ny + 1 - 2*fk0;
            } else {
                need = nz - s_beg + 3;
                need += (d.psx == d.sx) ? 2*(fj0 + BX) : nx - 1;
                need += (d.psy == d.sy) ? 2*(fk0 + BY) : ny - 1;
            }
            need = min(max(need, 0), (int) kDone);
            const unsigned long long req = (d.pserial << 32) | (unsigned) need;
            if (ptile >= 0) {
                const unsigned long long* pw =
                    a.prog + (long) (1 - d.par)*a.nbx*a.nby + ptile;
                unsigned long long spins = 0;
                while (seen < req) {
                    seen = __hip_atomic_load(pw, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                    if (seen >= req) break;
                    __builtin_amdgcn_s_sleep(2);
                    // another workgroup gave up: stop waiting for it
                    if ((++spins & 1023) == 0 &&
                        __hip_atomic_load(a.err, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                    if (spins > kSpinLimit) {
                        atomicExch(a.err, 1);
                        __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                }
            }
            // the loads this releases are device-coherent and issued after
            // the progress words were read
            verified = T;
        };
        R cur[HD], nxt[HD];
        auto fetch = [&](int s0, R* buf) {
            #pragma unroll
            for (int q = 0; q < HD; q++) {
                const int fi = s0 + q - hj - hk;
                buf[q] = R(0);
                if (direct && hcol && fi >= ez0 && fi <= a.fz1) {
                    const R* p = a.tt + gcol + ((d.sz > 0) ? fi : nz - 1 - fi);
                    buf[q] = CHAIN ? ld_coh(p) : *p;
                }
            }
        };
        // after a wait-limit failure elsewhere the tile gives up at once
        if (CHAIN && __hip_atomic_load(a.err, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
        verify(s_beg + 3*HD - 1);
        if (CHAIN && lds_load(&s_abort)) return;
        fetch(s_beg, cur);
        for (int s0 = s_beg; s0 <= s_end; s0 += HD) {
            verify(s0 + 3*HD - 1);
            if (CHAIN && lds_load(&s_abort)) return;
            if (s0 + HD <= s_end) fetch(s0 + HD, nxt);
            R mbv[KB > 1 ? KB : 1];
            #pragma unroll
            for (int q = 0; q < HD; q++) {
                const int s = s0 + q;
                if (s > s_end) break;                 // wave-uniform
                if constexpr (KB > 1) {
                    if (q % KB == 0 && !direct && hcol && a.mbx) {
                        // this lane's steps of the batch s .. s + KB - 1
                        const int lo = max(s, ez0 + hj + hk);
                        const int hi = min(min(s + KB - 1, s_end),
                                           a.fz1 + hj + hk);
                        const bool xs = hl < BY || hl == BX + BY;
                        auto addr = [&](int st) -> long {
                            return xs ? ((long) (fbx - 1)*a.mbnt + st)*ny + hk
                                      : ((long) (fby - 1)*a.mbnt + st)*nx + hj;
                        };
                        unsigned long long* mbp = xs ? a.mbx : a.mby;
                        if (lo <= hi) {
                            // the batch's last word first (the upwind tile
                            // posts its steps in order), then all at once
                            R vl;
                            unsigned long long sp = 0;
                            bool fail = false;
                            while (!mb_get(mbp, addr(hi), tag, vl)) {
                                __builtin_amdgcn_s_sleep(1);
                                if (++sp > kSpinLimit) { fail = true; break; }
                            }
                            long gq[KB];
                            bool okq[KB];
                            #pragma unroll
                            for (int qq = 0; qq < KB; qq++)
                                gq[qq] = addr(min(max(s + qq, lo), hi));
                            if (!fail) mb_get_batch<KB>(mbp, gq, tag, mbv, okq);
                            #pragma unroll
                            for (int qq = 0; qq < KB; qq++) {
                                const int st = s + qq;
                                if (fail || st < lo || st > hi || okq[qq])
                                    continue;
                                // not visible yet although a later step is
                                sp = 0;
                                while (!mb_get(mbp, gq[qq], tag, mbv[qq])) {
                                    __builtin_amdgcn_s_sleep(1);
                                    if (++sp > kSpinLimit) { fail = true;
                                                             break; }
                                }
                            }
                            if (fail) {
                                atomicExch(a.err, 1);
                                __hip_atomic_store(&s_abort, 1,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
                // the ring slot of step s was last read at step s - RING + 3
                const int need = s - RING + 3;
                unsigned long long spins = 0;
                while (wdone_min() < need) {
                    if (++spins > kSpinLimit || lds_load(&s_abort)) return;
                    __builtin_amdgcn_s_sleep(1);
                }
                const int fi = s - hj - hk;
                if (hcol && fi >= ez0 && fi <= a.fz1) {
                    R v = cur[q];
                    if (KB > 1 && !direct && a.mbx) {
                        v = mbv[q % (KB > 1 ? KB : 1)];
                    } else if (!direct) {
                        unsigned long long* mbp = a.mb;
                        long g = gcol + ((d.sz > 0) ? fi : nz - 1 - fi);
                        if (!CHAIN && a.mbx) {
                            // x-side lanes and the corner read the last
                            // column of x-tile fbx - 1, the others the last
                            // row of y-tile fby - 1
                            if (hl < BY || hl == BX + BY) {
                                mbp = a.mbx;
                                g = ((long) (fbx - 1)*a.mbnt + s)*ny + hk;
                            } else {
                                mbp = a.mby;
                                g = ((long) (fby - 1)*a.mbnt + s)*nx + hj;
                            }
                        }
                        spins = 0;
                        while (!mb_get(mbp, g, tag, v)) {
                            __builtin_amdgcn_s_sleep(1);
                            if (++spins > kSpinLimit) {
                                atomicExch(a.err, 1);
                                __hip_atomic_store(&s_abort, 1,
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                                break;
                            }
                        }
                    }
                    ring[s & (RING - 1)][slot] = v;
                }
                if (lds_load(&s_abort)) return;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                if (hl == 0)
                    __hip_atomic_store(&s_hready, s + 1, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            #pragma unroll
            for (int q = 0; q < HD; q++) cur[q] = nxt[q];
        }
        return;
    }

    // -------------------------------------------------------- compute waves
    const int wv = threadIdx.x / 64;
    const int lj = threadIdx.x % BX, lk = threadIdx.x / BX;
    const int fj = fj0 + lj, fk = fk0 + lk;
    const bool colE = (fj >= jlo && fj <= jhi && fk >= klo && fk <= khi);
    const bool colB = colE && fj >= a.fx0 && fk >= a.fy0;
    const bool post = colE && (lj == BX - 1 || lk == BY - 1);
    const int cj = min(max(fj, jlo), jhi), ck = min(max(fk, klo), khi);
    const int j = unflip(cj, nx, d.sx), k = unflip(ck, ny, d.sy);
    const long colbase = (long) (k - 1)*nz*nx + (long) (j - 1)*nz;
    const R* const col = a.tt + colbase;
    const int jm = max(j - 1, 1), jp = min(j, nx - 1);
    const int km = max(k - 1, 1), kp = min(k, ny - 1);
    const int xu = (d.sx > 0) ? jm : jp, xd = (d.sx > 0) ? jp : jm;
    const int yu = (d.sy > 0) ? km : kp, yd = (d.sy > 0) ? kp : km;
    const long szx = (long) (nz - 1)*(nx - 1);
    const R* const c_uu = a.slow + (long) (yu - 1)*szx + (long) (xu - 1)*(nz - 1);
    const R* const c_ud = a.slow + (long) (yd - 1)*szx + (long) (xu - 1)*(nz - 1);
    const R* const c_du = a.slow + (long) (yu - 1)*szx + (long) (xd - 1)*(nz - 1);
    const R* const c_dd = a.slow + (long) (yd - 1)*szx + (long) (xd - 1)*(nz - 1);
    const int fdep = cj + ck;
    const Consts<R> c = *a.cst;
    const int me = (lk + 1)*(BX + 1) + (lj + 1);
    const int cW = me - 1, cS = me - (BX + 1), cSW = me - (BX + 2);
    R dmax = R(0);
    auto node_off = [&](int fi) -> int {
        fi = min(max(fi, ez0), a.fz1);
        return (d.sz > 0) ? fi : nz - 1 - fi;
    };
    auto layer_off = [&](int f) -> int {
        const int zc = (d.sz > 0) ? f : nz - f;
        return min(max(zc, 1), nz - 1) - 1;
    };
    R tn[HB], cn[HB + 1][4];
    auto prefetch = [&](int tb) {
        const int f0 = tb - fdep;
        if (!STAGE) {
            #pragma unroll
            for (int q = 0; q < HB; q++)
                tn[q] = CHAIN ? ld_coh(col + node_off(f0 + q))
                              : col[node_off(f0 + q)];
        }
        if (!STAGE) {
            #pragma unroll
            for (int q = 0; q <= HB; q++) {
                const int z = layer_off(f0 + q);
                cn[q][0] = c_uu[z]; cn[q][1] = c_ud[z];
                cn[q][2] = c_du[z]; cn[q][3] = c_dd[z];
            }
        }
    };
    // STAGE: element e = lane + 64 i of a batch's cell loads is layer
    // m = e % 8 of cell column e / 8 = (xa, yb) (x slot 0..BX, y slot
    // 0..BY; slot a lies between node columns fj0 + a - 1 and fj0 + a).  A
    // node column (lj, lk) reads slots (lj, lk), (lj, lk + 1), (lj + 1, lk),
    // (lj + 1, lk + 1) as cells uu, ud, du, dd (the clamps of
    // autoCode.f90:407-445 are those of the slot's cell column).  For the
    // batch of steps tb .. tb + 3, slot (a, b) holds the layers of flipped
    // nodes f = tb - (fj0 + fk0) - a - b + m, m = 0 .. 7: its four readers
    // need m = q + (a - lj) + (b - lk) and the next one at step tb + q.
    const R* sptr[STAGE ? NSE : 1];
    int sab[STAGE ? NSE : 1];
    R sv[STAGE ? NSE : 1];
    int cbase[4] = {0, 0, 0, 0};
    const int sm = threadIdx.x & 7;
    const int sD = fj0 + fk0;
    if constexpr (STAGE) {
        const long szx = (long) (nz - 1)*(nx - 1);
        #pragma unroll
        for (int i = 0; i < NSE; i++) {
            const int sl = min((int) (threadIdx.x + 64*i) >> 3, NSL - 1);
            const int sa = sl % (BX + 1), sb = sl / (BX + 1);
            const int cx = (d.sx > 0) ? min(max(fj0 + sa, 1), nx - 1)
                                      : min(max(nx - fj0 - sa, 1), nx - 1);
            const int cy = (d.sy > 0) ? min(max(fk0 + sb, 1), ny - 1)
                                      : min(max(ny - fk0 - sb, 1), ny - 1);
            sptr[i] = a.slow + (long) (cy - 1)*szx + (long) (cx - 1)*(nz - 1);
            sab[i] = sa + sb;
        }
        // LDS element of cell uu (slot (lj, lk), delta 0), ud (lj, lk + 1),
        // du (lj + 1, lk) (delta 1 each) and dd (lj + 1, lk + 1) (delta 2)
        const int slj = threadIdx.x % BX, slk = threadIdx.x / BX;
        cbase[0] = (slk*(BX + 1) + slj)*8;
        cbase[1] = ((slk + 1)*(BX + 1) + slj)*8 + 1;
        cbase[2] = (slk*(BX + 1) + slj + 1)*8 + 1;
        cbase[3] = ((slk + 1)*(BX + 1) + slj + 1)*8 + 2;
    }
    // STAGE: the batch's own old times, element e = lane + 64 i being step
    // e % HB of node column e / HB (HB lanes per column along z)
    constexpr int NTE = STAGE ? NC*HB/64 : 1;
    const R* tptr[NTE];
    int tdep[NTE];
    if constexpr (STAGE) {
        #pragma unroll
        for (int i = 0; i < NTE; i++) {
            const int cix = (int) (threadIdx.x + 64*i) / HB;
            const int tj = min(max(fj0 + cix % BX, jlo), jhi);
            const int tk = min(max(fk0 + cix / BX, klo), khi);
            tptr[i] = a.tt + (long) (unflip(tk, ny, d.sy) - 1)*nz*nx
                      + (long) (unflip(tj, nx, d.sx) - 1)*nz;
            tdep[i] = tj + tk;
        }
    }
    R tv_[NTE];
    auto stage_load = [&](int tb) {
        #pragma unroll
        for (int i = 0; i < NSE; i++)
            sv[i] = sptr[i][layer_off(tb - sD - sab[i] + sm)];
        #pragma unroll
        for (int i = 0; i < NTE; i++)
            tv_[i] = tptr[i][node_off(tb + (((int) threadIdx.x + 64*i) % HB)
                                      - tdep[i])];
    };
    if (CHAIN) {
        // the halo wave's release of step tlo - 3 covers the loads below
        unsigned long long spins = 0;
        while (lds_load(&s_hready) < tlo - 2) {
            if (lds_load(&s_abort) || ++spins > kSpinLimit) {
                atomicExch(a.err, 1);
                __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                return;
            }
            // waits for the preceding sweep: leave the SIMD to others
            __builtin_amdgcn_s_sleep(8);
        }
    }
    // cached step counters (s_hready, s_wdone[wv -+ 1], s_sdone)
    int c_h = tlo - 3, c_wm = (wv > 0) ? tlo - 1 : 1 << 30,
        c_wp = (wv < NW - 1) ? tlo - 1 : 1 << 30, c_s = SW ? tlo - 1 : 1 << 30;
    if (!INIT) prefetch(tlo);
    if constexpr (STAGE) stage_load(tlo);
    for (int tb = tlo; tb <= thi; tb += HB) {
        if (INIT) prefetch(tb);
        R tc[HB], cc[HB + 1][4];
        #pragma unroll
        for (int q = 0; q < HB; q++) tc[q] = tn[q];
        if constexpr (STAGE) {
            // this batch's cell layers into the buffer (the wave's own LDS
            // accesses stay in order: the last batch's reads are done)
            #pragma unroll
            for (int i = 0; i < NSE; i++) {
                const int e = (int) threadIdx.x + 64*i;
                if (e < NSL*8) cbuf[e] = sv[i];
            }
            #pragma unroll
            for (int i = 0; i < NTE; i++) tbuf[threadIdx.x + 64*i] = tv_[i];
            if (tb + HB <= thi) stage_load(tb + HB);
        } else {
            #pragma unroll
            for (int q = 0; q <= HB; q++) {
                cc[q][0] = cn[q][0]; cc[q][1] = cn[q][1];
                cc[q][2] = cn[q][2]; cc[q][3] = cn[q][3];
            }
        }
        if (!INIT && tb + HB <= thi) prefetch(tb + HB);
        #pragma unroll
        for (int q = 0; q < HB; q++) {
            const int t = tb + q;
            if (t > thi) break;                       // workgroup-uniform
            // STAGE: this step's slowness terms need no other wave: they are
            // formed before the wait below
            Slw<R> sls;
            if constexpr (STAGE) {
                R zu[4], zd[4];
                #pragma unroll
                for (int c4 = 0; c4 < 4; c4++) {
                    zu[c4] = cbuf[cbase[c4] + q];
                    zd[c4] = cbuf[cbase[c4] + q + 1];
                }
                slowness7_ud<R>(zu, zd, sls);
            }
            // halo of steps < t, the wave above (w - 1) past step t - 1, and
            // the readers of the ring slot this step overwrites (wave w + 1,
            // the store wave) past step t - RING + 3.  The counters only
            // grow: values read earlier that already allow this step are not
            // read again.
            if (c_h < t || c_wm < t - 1 || c_wp < t - RING + 3 ||
                c_s < t - RING + 3) {
                unsigned long long spins = 0;
                for (;;) {
                    c_h = lds_load_rx(&s_hready);
                    if (wv > 0) c_wm = lds_load_rx(&s_wdone[wv - 1]);
                    if (wv < NW - 1) c_wp = lds_load_rx(&s_wdone[wv + 1]);
                    if (SW) c_s = lds_load_rx(&s_sdone);
                    if (c_h >= t && c_wm >= t - 1 && c_wp >= t - RING + 3 &&
                        c_s >= t - RING + 3) break;
                    if (lds_load(&s_abort) || ++spins > kSpinLimit) {
                        atomicExch(a.err, 1);
                        __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        return;
                    }
                    // a waiting wave yields its issue slots to the waves of
                    // other tiles sharing the SIMD
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const int fi = t - fj - fk;
            if constexpr (SW && !INIT) {
                // Iteration sweeps with a store wave: every lane runs the
                // stencil at every step, and the lanes whose node is not
                // updated keep their old value, so the step has no branch.
                // The ring values of lanes outside the box rows [ez0, fz1]
                // or columns are read only by lanes that are outside too
                // (the readers of slot t are the node's own column and its
                // x, y and xy successors at steps t + 1 .. t + 3, whose
                // flipped row is the same or one more), and the store wave
                // reads in-box nodes only.
                R v = tc[q];
                if constexpr (STAGE) v = tbuf[threadIdx.x*HB + q];
                const int s1 = (t - 1) & (RING - 1), s2 = (t - 2) & (RING - 1),
                          s3 = (t - 3) & (RING - 1);
                Nbr<R> nb;
                nb.tv   = ring[s1][me];
                nb.te   = ring[s1][cW];
                nb.tn   = ring[s1][cS];
                nb.tev  = ring[s2][cW];
                nb.ten  = ring[s2][cSW];
                nb.tnv  = ring[s2][cS];
                nb.tnve = ring[s3][cSW];
                Slw<R> sl;
                if constexpr (STAGE) {
                    sl = sls;
                } else {
                    slowness7_ud<R>(cc[q], cc[q + 1], sl);
                }
                const R vn = noinit3d(c, nb, sl, v);
                const bool upd = colB && fi >= a.fz0 && fi <= a.fz1;
                dmax = upd ? rmax(dmax, v - vn) : dmax;
                v = upd ? vn : v;
                ring[t & (RING - 1)][me] = v;
            } else if (colE && fi >= ez0 && fi <= a.fz1) {
                R v = tc[q];
                if constexpr (STAGE) v = tbuf[threadIdx.x*HB + q];
                if (colB && fi >= a.fz0) {
                    const int s1 = (t - 1) & (RING - 1),
                              s2 = (t - 2) & (RING - 1),
                              s3 = (t - 3) & (RING - 1);
                    Nbr<R> nb;
                    nb.tv   = ring[s1][me];
                    nb.te   = ring[s1][cW];
                    nb.tn   = ring[s1][cS];
                    nb.tev  = ring[s2][cW];
                    nb.ten  = ring[s2][cSW];
                    nb.tnv  = ring[s2][cS];
                    nb.tnve = ring[s3][cSW];
                    Slw<R> sl;
                    if constexpr (STAGE) {
                        sl = sls;
                    } else {
                        slowness7_ud<R>(cc[q], cc[q + 1], sl);
                    }
                    R vn;
                    if (INIT) {
                        const int i = unflip(fi, nz, d.sz);
                        vn = init3d(c, nb, sl, v, i, j, k, d.sz, d.sx, d.sy);
                    } else {
                        vn = noinit3d(c, nb, sl, v);
                    }
                    dmax = rmax(dmax, v - vn);
                    v = vn;
                    if (!SW) a.tt[colbase + node_off(fi)] = v;
                }
                if (!SW && post) {
                    if (!CHAIN && a.mbx) {
                        if (lj == BX - 1)
                            mb_put(a.mbx, ((long) fbx*a.mbnt + t)*ny + fk, v,
                                   tag);
                        if (lk == BY - 1)
                            mb_put(a.mby, ((long) fby*a.mbnt + t)*nx + fj, v,
                                   tag);
                    } else {
                        mb_put(a.mb, colbase + node_off(fi), v, tag);
                    }
                }
                ring[t & (RING - 1)][me] = v;
            }
            // LDS operations of one wave execute in order: the counter store
            // cannot overtake the ring store, so only the compiler needs to
            // be kept from reordering them
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if ((threadIdx.x & 63) == 0)
                __hip_atomic_store(&s_wdone[wv], t, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
    }
    if (a.maxdelta && dmax > R(0)) {
        double dd = (double) dmax;
        atomicMax((unsigned long long*) a.maxdelta,
                  (unsigned long long) __double_as_longlong(dd));
    }
    if (SKIP && colB && dmax > R(0))
        a.colmod[(long) (k - 1)*nx + (j - 1)] = d.serial;
}

__host__ __device__ inline void sweep_signs(int sweep, int& sz, int& sx,
                                            int& sy)
{
    // prefetch.py:4-8
    static const int tz[8] = { 1,  1,  1,  1, -1, -1, -1, -1};
    static const int tx[8] = { 1, -1,  1, -1,  1, -1,  1, -1};
    static const int ty[8] = { 1,  1, -1, -1,  1,  1, -1, -1};
    sz = tz[sweep - 1]; sx = tx[sweep - 1]; sy = ty[sweep - 1];
}

// solver3D.F90:493-565 box, converted to the flipped 0-based frame.
// Per-sweep values of sweep q of a chain starting in direction dir0 + 1
// with serial serial0.
__device__ __forceinline__ SweepDyn chain_dyn(int dir0, int q,
                                              unsigned long long serial0)
{
    SweepDyn d;
    sweep_signs((dir0 + q) % 8 + 1, d.sz, d.sx, d.sy);
    d.serial = serial0 + q;
    d.par = q & 1;
    d.pserial = 0;
    d.psz = d.psx = d.psy = 0;
    if (q > 0) {
        sweep_signs((dir0 + q - 1) % 8 + 1, d.psz, d.psx, d.psy);
        d.pserial = serial0 + q - 1;
    }
    return d;
}

__host__ __device__ inline void flipped_box(int sweep, bool linit, int nz,
                                            int nx, int ny, int zsi, int xsi,
                                            int ysi, int* b)
{
    int sz, sx, sy;
    sweep_signs(sweep, sz, sx, sy);
    auto axis = [&](int sgn, int n, int si, int* lo, int* hi) {
        // actual 1-based limits, then flip
        int a0, a1;
        if (sgn > 0) { a0 = linit ? max(2, si) : 2; a1 = n; }
        else         { a0 = 1; a1 = linit ? si + 1 : n - 1; }
        if (sgn > 0) { *lo = a0 - 1; *hi = a1 - 1; }
        else         { *lo = n - a1; *hi = n - a0; }
    };
    axis(sz, nz, zsi, &b[0], &b[1]);
    axis(sx, nx, xsi, &b[2], &b[3]);
    axis(sy, ny, ysi, &b[4], &b[5]);
}

} // namespace fteik
