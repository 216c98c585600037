// sweep2d.hip -- 2D level-set fast sweeping on gfx950.
//
// solver2D.F90:782-978: 4 initialization sweeps (lexicographic in the
// reference, :825-867) and nsweep x 4 Gauss-Seidel sweeps over the levels
// ix+iz (:897-925).  Every sweep is evaluated here level by level -- a
// topological order of the same dependency graph, so the results equal the
// reference's bit for bit -- by one workgroup per source.  Two schedules:
// solve2d_pipe_kernel (default) pipelines one column per lane without
// barriers; solve2d_kernel (level by level, one barrier per level) is the
// cross-check and serves grids wider than 513 nodes.  Several sources
// (solveLSM) run as independent workgroups of one launch.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "stencils.h"
#include "kernels.h"

namespace fteik {

constexpr int T2 = 512;
constexpr unsigned long long kSpinLimit2D = 1ull << 24;

// localSolver.F90:160-193
__device__ __forceinline__ void signs2d(int sweep, int& sgntz, int& sgntx)
{
    sgntz = (sweep <= 2) ? 1 : -1;
    sgntx = (sweep == 1 || sweep == 3) ? 1 : -1;
}

__global__ __launch_bounds__(T2)
void solve2d_kernel(double* __restrict__ ttall, const double* __restrict__ slow,
                    int nz, int nx, double dz, double dx, double eps,
                    int nsweep, long ld_tt, const double* __restrict__ zsa_v,
                    const double* __restrict__ xsa_v,
                    const int* __restrict__ zsi_v, const int* __restrict__ xsi_v)
{
    const int src = blockIdx.x;
    double* tt = ttall + (long) src*ld_tt;
    const long ngrd = (long) nz*nx;
    __shared__ Consts<double> cs;
    if (threadIdx.x == 0) {
        // localSolver.F90:299-345
        Consts<double> c;
        c.zsa = zsa_v[src]; c.xsa = xsa_v[src]; c.ysa = 1.0;
        c.zsi = zsi_v[src]; c.xsi = xsi_v[src]; c.ysi = 1;
        c.szero = slow[(long) (c.xsi - 1)*(nz - 1) + c.zsi - 1];
        c.szero2 = c.szero*c.szero;
        c.epsSolver = eps;
        c.dz = dz; c.dx = dx; c.dy = 1.0;
        const double dz2 = dz*dz, dx2 = dx*dx;
        c.dzi = 1.0/dz; c.dxi = 1.0/dx; c.dyi = 1.0;
        c.dz2i = 1.0/dz2; c.dx2i = 1.0/dx2; c.dy2i = 1.0;
        c.dz2i_dx2i = c.dz2i*c.dx2i;
        c.dz2i_p_dx2i = c.dz2i + c.dx2i;
        c.dz2i_p_dx2i_inv = 1.0/(c.dz2i + c.dx2i);
        cs = c;
    }
    __syncthreads();
    const Consts<double> c = cs;
    for (long g = threadIdx.x; g < ngrd; g += blockDim.x) tt[g] = kHuge;
    __syncthreads();
    if (threadIdx.x < 4) {
        const int i = c.zsi + (threadIdx.x & 1), j = c.xsi + (threadIdx.x >> 1);
        tt[(long) (j - 1)*nz + i - 1] = tana2d(i, j, c);
    }
    __syncthreads();
    const int zsi = c.zsi, xsi = c.xsi;
    const int nsw = 4 + 4*nsweep;
    for (int q = 0; q < nsw; q++) {
        const bool linit = q < 4;
        const int sweep = (q & 3) + 1;
        int sgntz, sgntx;
        signs2d(sweep, sgntz, sgntx);
        // actual 1-based box (solver2D.F90:826-866 init, :377-454 iterate)
        int z0, z1, x0, x1;
        if (sgntz > 0) { z0 = linit ? max(2, zsi) : 2; z1 = nz; }
        else           { z0 = 1; z1 = linit ? zsi + 1 : nz - 1; }
        if (sgntx > 0) { x0 = linit ? max(2, xsi) : 2; x1 = nx; }
        else           { x0 = 1; x1 = linit ? xsi + 1 : nx - 1; }
        // flipped 0-based box
        const int fz0 = (sgntz > 0) ? z0 - 1 : nz - z1;
        const int fz1 = (sgntz > 0) ? z1 - 1 : nz - z0;
        const int fx0 = (sgntx > 0) ? x0 - 1 : nx - x1;
        const int fx1 = (sgntx > 0) ? x1 - 1 : nx - x0;
        if (fz0 > fz1 || fx0 > fx1) continue;
        for (int L = fz0 + fx0; L <= fz1 + fx1; L++) {
            for (int fj = fx0 + (int) threadIdx.x; fj <= fx1; fj += blockDim.x) {
                const int fi = L - fj;
                if (fi < fz0 || fi > fz1) continue;
                const int iz = (sgntz > 0) ? fi + 1 : nz - fi;
                const int ix = (sgntx > 0) ? fj + 1 : nx - fj;
                const long me = (long) (ix - 1)*nz + iz - 1;
                const double tv  = tt[(long) (ix - 1)*nz + iz - sgntz - 1];
                const double te  = tt[(long) (ix - sgntx - 1)*nz + iz - 1];
                const double tev = tt[(long) (ix - sgntx - 1)*nz + iz - sgntz - 1];
                const double t   = tt[me];
                // solver2D.F90:491-502 / localSolver.F90:222-231
                const int izc = iz - (sgntz > 0 ? 1 : 0);
                const int ixc = ix - (sgntx > 0 ? 1 : 0);
                const double s1 = rmin(slow[(long) (max(ix - 1, 1) - 1)*(nz - 1) + izc - 1],
                                       slow[(long) (min(ix, nx - 1) - 1)*(nz - 1) + izc - 1]);
                const double s2 = rmin(slow[(long) (ixc - 1)*(nz - 1) + max(iz - 1, 1) - 1],
                                       slow[(long) (ixc - 1)*(nz - 1) + min(iz, nz - 1) - 1]);
                const double s3 = slow[(long) (ixc - 1)*(nz - 1) + izc - 1];
                double tupd;
                if (linit)
                    tupd = init2d(c, sgntz, sgntx, iz, ix, tv, te, tev, t, s1,
                                  s2, s3);
                else
                    tupd = noinit2d(c, tv, te, tev, s1, s2, s3);
                tt[me] = rmin(t, tupd);
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ pipelined
// The same sweeps without a barrier per level: one workgroup per source, one
// lane per flipped x column (nx - 1 <= 512 columns, up to 8 waves), each
// lane walking down its column one node per step (step L: node fi = L - fj
// of column fj), so a step's nodes are independent (one level).  A lane
// keeps its column's node above (fi - 1) in a register; the left column's
// value of the step before comes from lane - 1 by a wave shift, and the one
// of two steps before is the shifted value the lane received at the last
// step.  Lane 0 of wave w > 0 takes lane 63 of wave w - 1 from a ring in LDS
// after that wave's step counter has passed the step; wave 0's lane 0 reads
// the frozen column left of the box.  Own old times and slownesses are
// staged through LDS a batch of S2 steps at a time, loaded cooperatively
// along the columns (a load instruction touches a few cache lines instead
// of one per lane), and the new times leave the same way.  The order of
// the node updates is a topological order of each sweep's dependency
// graph, as in the level kernel, so the results are the same bit for bit.
constexpr int R2D = 16;      // ring slots: wave w may lead wave w + 1 by that
constexpr int T2P = 512;     // largest workgroup (8 waves)
constexpr int S2 = 8;        // steps per staged batch
constexpr int CW2 = S2 + 1;  // cell rows of a staged cell column
constexpr int NRT = (65*S2 + 63)/64;    // old-time pieces per lane
constexpr int NRC = (65*CW2 + 63)/64;   // cell pieces per lane

__device__ __forceinline__ double wave_shr1(double x)
{
    // lane i gets lane i - 1's value (lane 0 keeps its own)
    const long long b = __double_as_longlong(x);
    int lo = (int) b, hi = (int) (b >> 32);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, false);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}

// TMAX threads at most (TMAX/64 waves)
template <int TMAX>
__global__ __launch_bounds__(TMAX)
void solve2d_pipe_kernel(double* __restrict__ ttall,
                         const double* __restrict__ slow, int nz, int nx,
                         double dz, double dx, double eps, int nsweep,
                         long ld_tt, const double* __restrict__ zsa_v,
                         const double* __restrict__ xsa_v,
                         const int* __restrict__ zsi_v,
                         const int* __restrict__ xsi_v, int* err,
                         double* __restrict__ scratch)
{
    const int src = blockIdx.x;
    double* tt = ttall + (long) src*ld_tt;
    const long ngrd = (long) nz*nx;
    __shared__ Consts<double> cs;
    __shared__ double ring[TMAX/64][R2D];
    __shared__ int s_done[TMAX/64];
    // per wave: old times of a batch (64 node columns + the frozen left
    // column), its cell window (65 columns x CW2 rows), its new times;
    // rows padded against LDS bank conflicts
    __shared__ double s_t[TMAX/64][65][S2 + 1];
    __shared__ double s_c[TMAX/64][65][CW2 + 2];
    __shared__ double s_o[TMAX/64][64][S2 + 1];
    __shared__ int s_fail;
    if (threadIdx.x == 0) {
        // localSolver.F90:299-345
        Consts<double> c;
        c.zsa = zsa_v[src]; c.xsa = xsa_v[src]; c.ysa = 1.0;
        c.zsi = zsi_v[src]; c.xsi = xsi_v[src]; c.ysi = 1;
        c.szero = slow[(long) (c.xsi - 1)*(nz - 1) + c.zsi - 1];
        c.szero2 = c.szero*c.szero;
        c.epsSolver = eps;
        c.dz = dz; c.dx = dx; c.dy = 1.0;
        const double dz2 = dz*dz, dx2 = dx*dx;
        c.dzi = 1.0/dz; c.dxi = 1.0/dx; c.dyi = 1.0;
        c.dz2i = 1.0/dz2; c.dx2i = 1.0/dx2; c.dy2i = 1.0;
        c.dz2i_dx2i = c.dz2i*c.dx2i;
        c.dz2i_p_dx2i = c.dz2i + c.dx2i;
        c.dz2i_p_dx2i_inv = 1.0/(c.dz2i + c.dx2i);
        cs = c;
    }
    __syncthreads();
    const Consts<double> c = cs;
    for (long g = threadIdx.x; g < ngrd; g += blockDim.x) tt[g] = kHuge;
    __syncthreads();
    if (threadIdx.x < 4) {
        const int i = c.zsi + (threadIdx.x & 1), j = c.xsi + (threadIdx.x >> 1);
        tt[(long) (j - 1)*nz + i - 1] = tana2d(i, j, c);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    const int zsi = c.zsi, xsi = c.xsi;
    const int nsw = 4 + 4*nsweep;
    bool fail = false;
    if (threadIdx.x == 0) s_fail = 0;
    for (int q = 0; q < nsw; q++) {
        const bool linit = q < 4;
        const int sweep = (q & 3) + 1;
        int sgntz, sgntx;
        signs2d(sweep, sgntz, sgntx);
        int z0, z1, x0, x1;
        if (sgntz > 0) { z0 = linit ? max(2, zsi) : 2; z1 = nz; }
        else           { z0 = 1; z1 = linit ? zsi + 1 : nz - 1; }
        if (sgntx > 0) { x0 = linit ? max(2, xsi) : 2; x1 = nx; }
        else           { x0 = 1; x1 = linit ? xsi + 1 : nx - 1; }
        const int fz0 = (sgntz > 0) ? z0 - 1 : nz - z1;
        const int fz1 = (sgntz > 0) ? z1 - 1 : nz - z0;
        const int fx0 = (sgntx > 0) ? x0 - 1 : nx - x1;
        const int fx1 = (sgntx > 0) ? x1 - 1 : nx - x0;
        if (fz0 > fz1 || fx0 > fx1) continue;            // uniform
        // waves with columns in the box, their first and last steps (one
        // step early, for the left column's value of two steps before)
        const int wbox = (fx1 - fx0)/64 + 1;
        const int Lf = fz0 + fx0 + 64*wv - 1;
        const int Ll = fz1 + min(fx0 + 64*wv + 63, fx1);
        if (lane == 0 && wv < nwv) s_done[wv] = Lf - 1;
        __syncthreads();
        if (s_fail) break;                                // uniform
        auto sweep_columns = [&](auto kind) {
        constexpr bool LINIT = decltype(kind)::value;
        if (wv < wbox) {
            // This wave's columns are F .. F + 63 (flipped), lane j has
            // column fj = F + j and at step L the node on flipped row
            // fi = L - fj.  In flipped cells (cell column fj - 1 lies
            // upwind of node column fj, cell row fi - 1 upwind of node row
            // fi) a node reads the cells (fj-1, fi-1), (min(fj, nx-2),
            // fi-1) and (fj-1, min(fi, nz-2)) (solver2D.F90:491-502: slow
            // at izc over xa, xb and at ixc over max(iz-1,1), min(iz,nz-1)).
            // Per batch of S2 steps the wave loads, S2 lanes along each
            // column: the S2 old times of each of its 64 node columns (and,
            // wave 0, of the frozen column left of the box) and S2 + 1 rows
            // of each of its 65 cell columns F-1 .. F+63 (cell column c
            // from row L0 - (F-1+c) - 2), into LDS one batch ahead through
            // registers; the new times of a batch go to LDS and out to the
            // field S2 lanes per column after it.  No lane reads the field
            // values another lane of this sweep writes (old times are the
            // lane's own, ahead of its writes), so deferring the stores to
            // the end of a batch changes nothing the sweep reads.
            const int F = fx0 + 64*wv;
            const int fj = F + lane;
            const bool colA = fj <= fx1;
            const int cfj = min(fj, fx1);
            const int ix = (sgntx > 0) ? cfj + 1 : nx - cfj;
            auto nodecol = [&](int f) -> const double* {   // column f, clamped
                f = min(f, fx1);
                return tt + (long) (((sgntx > 0) ? f + 1 : nx - f) - 1)*nz;
            };
            const double* const lcol = tt + (long) (((sgntx > 0) ? fx0 + 1
                                                     : nx - fx0) - sgntx - 1)*nz;
            auto zoff = [&](int fi) -> int {        // 0-based node in column
                fi = min(max(fi, fz0 - 1), fz1);
                return (sgntz > 0) ? fi : nz - 1 - fi;
            };
            auto cell = [&](int fc, int fr) -> const double* {  // flipped cell
                fc = min(fc, nx - 2);
                fr = min(max(fr, 0), nz - 2);
                const int c0 = (sgntx > 0) ? fc : nx - 2 - fc;
                const int r0 = (sgntz > 0) ? fr : nz - 2 - fr;
                return slow + (long) c0*(nz - 1) + r0;
            };
            double (*const st)[S2 + 1] = s_t[wv];
            double (*const sc)[CW2 + 2] = s_c[wv];
            double (*const so)[S2 + 1] = s_o[wv];
            double rt[NRT], rc[NRC];
            // batch from step L0 into registers: old-time piece i is
            // column lane/S2 + (64/S2) i, element lane%S2 (column 64:
            // wave 0's frozen left column; pieces past it are harmless
            // in-range reads); cell piece i is slot lane + 64 i of the
            // 65 x CW2 cell window
            auto fetch = [&](int L0) {
                #pragma unroll
                for (int i = 0; i < NRT; i++) {
                    const int c = lane/S2 + (64/S2)*i, e = lane % S2;
                    const int cc = min(c, 63);
                    const double* p = nodecol(F + cc) + zoff(L0 - F - cc + e);
                    if (i == NRT - 1 && wv == 0 && c == 64)
                        p = lcol + zoff(L0 - fx0 + e);
                    rt[i] = *p;
                }
                #pragma unroll
                for (int i = 0; i < NRC; i++) {
                    const int q = min(lane + 64*i, 65*CW2 - 1);
                    const int c = q / CW2, e = q - c*CW2;
                    rc[i] = *cell(F - 1 + c, L0 - (F - 1 + c) - 2 + e);
                }
            };
            auto park = [&]() {                     // registers to LDS
                #pragma unroll
                for (int i = 0; i < NRT; i++) {
                    const int c = lane/S2 + (64/S2)*i, e = lane % S2;
                    if (c <= 64) st[c][e] = rt[i];
                }
                #pragma unroll
                for (int i = 0; i < NRC; i++) {
                    const int q = lane + 64*i;
                    const int c = q / CW2, e = q - c*CW2;
                    if (q < 65*CW2) sc[c][e] = rc[i];
                }
            };
            // the new times of the batch from step L0 to the field; lanes
            // with no node to write store to their scratch slot, so every
            // lane issues the same stores
            double* const scr = scratch + (long) src*T2P + threadIdx.x;
            auto flush = [&](int L0) {
                #pragma unroll
                for (int i = 0; i < S2; i++) {
                    const int c = lane/S2 + (64/S2)*i, e = lane % S2;
                    const int fi = L0 - F - c + e;
                    const bool act = F + c <= fx1 && fi >= fz0 && fi <= fz1;
                    double* const dst = act ? (double*) nodecol(F + c)
                        + ((sgntz > 0) ? fi : nz - 1 - fi) : scr;
                    *dst = so[c][e];
                }
            };
            double P = *(nodecol(fj) + zoff(fz0 - 1));  // node above the box
            double tep = 0.0;
            int cl = Lf - 2, cr = -(1 << 30);       // cached counters
            fetch(Lf);
            for (int L0 = Lf; L0 <= Ll; L0 += S2) {
                park();
                if (L0 > Lf) flush(L0 - S2);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_wave_barrier();
                if (L0 + S2 <= Ll) fetch(L0 + S2);
                #pragma unroll
                for (int k = 0; k < S2; k++) {
                    const int L = L0 + k;
                    if (L > Ll) break;                   // wave-uniform
                    const int fi = L - fj;
                    double te = wave_shr1(P);
                    if (wv == 0) {
                        te = (lane == 0) ? st[64][k] : te;
                    } else if (L - 1 <= fz1 + fx0 + 64*wv - 1) {
                        // lane 63 of the wave before, after its step L - 1
                        // (counters are read as wave-uniform values)
                        unsigned spins = 0;
                        while (cl < L - 1) {
                            cl = __builtin_amdgcn_readfirstlane(
                                __hip_atomic_load(&s_done[wv - 1],
                                                  __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP));
                            if (cl >= L - 1) break;
                            if (++spins > kSpinLimit2D) { fail = true; break; }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        const double rv = ring[wv - 1][L % R2D];
                        te = (lane == 0) ? rv : te;
                    }
                    const double tev = tep;
                    if (colA && fi >= fz0 && fi <= fz1) {
                        const int iz = (sgntz > 0) ? fi + 1 : nz - fi;
                        const double cA = sc[lane][k];          // (fj-1, fi-1)
                        const double cB = sc[lane][k + 1];      // (fj-1, min(fi, nz-2))
                        const double cC = (fj <= nx - 2) ? sc[lane + 1][k + 1]
                                                         : cA;  // (min(fj, nx-2), fi-1)
                        const double s1 = rmin(cA, cC);
                        const double s2 = rmin(cA, cB);
                        const double s3 = cA;
                        const double t = st[lane][k];
                        double tupd;
                        if constexpr (LINIT)
                            tupd = init2d(c, sgntz, sgntx, iz, ix, P, te, tev,
                                          t, s1, s2, s3);
                        else
                            tupd = noinit2d(c, P, te, tev, s1, s2, s3);
                        P = rmin(t, tupd);
                    }
                    so[lane][k] = P;
                    tep = te;
                    // lane 63's value for the next wave's step L + 1, once
                    // that wave is past step L + 1 - R2D (the slot's reader)
                    if (wv + 1 < wbox) {
                        unsigned spins = 0;
                        while (cr < L + 1 - R2D) {
                            cr = __builtin_amdgcn_readfirstlane(
                                __hip_atomic_load(&s_done[wv + 1],
                                                  __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP));
                            if (cr >= L + 1 - R2D) break;
                            if (++spins > kSpinLimit2D) { fail = true; break; }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        if (lane == 63) ring[wv][(L + 1) % R2D] = P;
                    }
                    // LDS operations of one wave execute in order
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane == 0)
                        __hip_atomic_store(&s_done[wv], L, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (fail) break;
                }
                if (fail) break;
            }
            // the last batch's new times
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_wave_barrier();
            if (!fail) {
                const int nb = (Ll - Lf)/S2;
                flush(Lf + nb*S2);
            }
        }
        };
        if (linit) sweep_columns(std::true_type{});
        else sweep_columns(std::false_type{});
        if (fail && lane == 0) {
            atomicExch(err, 1);
            s_fail = 1;
        }
        __syncthreads();                  // the sweep's field is complete
    }
}

} // namespace fteik

using namespace fteik;

extern "C" int fteik_k_solve2d(double* tt, const double* slow, int nz, int nx,
                               double dz, double dx, double eps, double zsa,
                               double xsa, int zsi, int xsi, int nsweep,
                               int nsrc_batch, int ld_tt, const double* zsa_v,
                               const double* xsa_v, const int* zsi_v,
                               const int* xsi_v, int schedule, int* err,
                               double* scratch, hipStream_t stream)
{
    (void) zsa; (void) xsa; (void) zsi; (void) xsi;
    // pipelined columns when every flipped column of a sweep box has a lane
    if (schedule == 0 && nx - 1 <= T2P && err && scratch) {
        const int threads = 64*((nx - 1 + 63)/64);
        hipLaunchKernelGGL((solve2d_pipe_kernel<T2P>), dim3(nsrc_batch),
                           dim3(threads), 0, stream, tt, slow, nz, nx, dz, dx,
                           eps, nsweep, (long) ld_tt, zsa_v, xsa_v, zsi_v,
                           xsi_v, err, scratch);
    } else {
        hipLaunchKernelGGL(solve2d_kernel, dim3(nsrc_batch), dim3(T2), 0,
                           stream, tt, slow, nz, nx, dz, dx, eps, nsweep,
                           (long) ld_tt, zsa_v, xsa_v, zsi_v, xsi_v);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
