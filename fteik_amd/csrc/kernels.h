// kernels.h -- internal launcher interface between the host C-ABI layer
// (capi_*.cpp) and the HIP kernels.  Not installed; plain C structs only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fteik {

struct SourceGeom {
    double dz, dx, dy, eps;
    double zsa, xsa, ysa;
    int zsi, xsi, ysi;
    int nz, nx, ny;
};

struct SweepLaunch {
    void* tt;                    // double* or float* field (natural order)
    const void* slow;            // double* or float* slowness
    const void* cst;             // Consts<double|float>*
    unsigned long long* flags;   // per-block progress flags
    int* err;                    // device error flag
    double* maxdelta;            // optional convergence monitor (bits)
    unsigned long long serial;   // launch serial for the flags
    int max_flags;
    int nz, nx, ny;
    int zsi, xsi, ysi;
    int fp32;                    // 0: double, 1: float
    int schedule;                // 0: pipelined column blocks, 1: per level
    int cooperative;
    int tile;                    // tile shape index of the pipelined sweep
    unsigned long long* mb;      // mailbox of the mode-1 tiles
    // edge-major mailbox of sweeps outside chains (null: per-node mailbox):
    // x edges (nx + 7)/8 + 1 tiles by mbnt steps by ny, y edges likewise
    unsigned long long* mbx;
    unsigned long long* mby;
    int mbnt;
    // tile skipping state (null: off): per-column serial of the last change
    // (nx*ny), per-direction serial of each tile's last pass (8*max_flags),
    // and the launch plan (1 + 2*max_flags ints)
    unsigned long long* colmod;
    unsigned long long* lastproc;
    int* plan;
    // sweep chains: progress words of the two workgroup sets
    // (2*max_flags, zeroed once; serials only grow)
    unsigned long long* prog;
};

// One grid of a batched solve (batch3d.hip); device pointers.
struct BatchGrid {
    void* tt; const void* slow; const void* cst;        // FP64
    void* tt32; const void* slow32; const void* cst32;  // FP32 (mixed)
    double* maxdelta;                                   // per-grid monitor
    unsigned long long* mb;        // per-grid mailbox (mailbox-tile sweeps)
    unsigned long long* mbx;       // per-grid edge-major mailbox (null: the
    unsigned long long* mby;       // per-node one), as SweepLaunch::mbx/mby
    int mbnt;
    int nz, nx, ny, zsi, xsi, ysi;
};

// Grid of a ray trace (rays.hip): 2D grids have ny = 1 and is3d = 0.
struct RayGrid {
    int is3d, nz, nx, ny;
    double dz, dx, dy;
};

} // namespace fteik

extern "C" {
// rays.hip
int fteik_k_fd_gradient3d(double* g, const double* tt, int nz, int nx,
                          int ny, double dz, double dx, double dy,
                          hipStream_t stream);
int fteik_k_fd_gradient2d(double* gx, double* gz, const double* tt, int nz,
                          int nx, double dz, double dx, hipStream_t stream);
int fteik_k_trace_rays(const fteik::RayGrid* rg, const double* slow,
                       const double* g3, const double* gx, const double* gz,
                       int integrator, double xs, double ys, double zs,
                       const double* start, int nray, int maxpts,
                       double* pts, int* npts, double* time, int* status,
                       hipStream_t stream);
// batch3d.hip
int fteik_k_batch_map(int tile, int ngrids, const int* dims, int* map,
                      int map_cap, int* nflags);
int fteik_k_batch_sweep(const fteik::BatchGrid* grids, const int* map,
                        int nwg, int tile, int sweep, int linit,
                        unsigned long long serial, unsigned long long* flags,
                        int* err, int fp32, int track, hipStream_t stream);
int fteik_k_batch_mb_map(int tile, int ngrids, const int* dims,
                         const int* srcs, int sweep, int* map, int map_cap);
size_t fteik_k_batch_mb_args_bytes(int ngrids);
int fteik_k_batch_mb_sweep(const fteik::BatchGrid* grids, int ngrids,
                           const int* map, int nwg, int tile, int sweep,
                           int linit, unsigned long long serial, int* err,
                           int fp32, int track, void* argbuf,
                           hipStream_t stream);
// sweep3d.hip
int fteik_k_sweep3d(const fteik::SweepLaunch* L, int sweep, int linit,
                    hipStream_t stream);
int fteik_k_pencil_fits(int nx, int ny, int fp32, int tile);
int fteik_k_chain_fits(int nx, int ny, int fp32, int tile);
void fteik_k_tile_shape(int tile, int* bx, int* by, int* hb, int* mode);
int fteik_k_sweep3d_chain(const fteik::SweepLaunch* L, int dir0, int nq,
                          hipStream_t stream);
int fteik_k_num_tiles(void);
int fteik_k_tile_mode(int tile);
int fteik_k_max_flags(int nx, int ny);
int fteik_k_source3d_init(void* cst, void* tt, const void* slow,
                          const fteik::SourceGeom* g, long ngrd, int fp32,
                          hipStream_t stream);
int fteik_k_consts_to_float(const void* cd, void* cf, hipStream_t stream);

// util.hip
int fteik_k_slowness_from_cells(double* slow, const double* vel, int nz,
                                int nx, int ny, int is3d, int order,
                                hipStream_t stream);
int fteik_k_slowness_from_nodes(double* slow, const double* vel, int nz,
                                int nx, int ny, int is3d, int order,
                                hipStream_t stream);
int fteik_k_slowness_from_velocity32(double* slow, const float* vel, long n,
                                     hipStream_t stream);
int fteik_k_field_reorder(double* out, const double* tt, int nz, int nx,
                          int ny, int is3d, int order, hipStream_t stream);
int fteik_k_field_cells(double* out, const double* tt, int nz, int nx,
                        int ny, int is3d, int order, hipStream_t stream);
int fteik_k_receivers(double* ttr, const double* tt, const int* ri,
                      const double* rd, int nrec, int nz, int nx, double dz,
                      double dx, double dy, int is3d, hipStream_t stream);
// receiver times of nfield fields fstride doubles apart, field q's to
// ttr + q*nrec (nfield <= 65535)
int fteik_k_receivers_many(double* ttr, const double* tt, int nfield,
                           long fstride, const int* ri, const double* rd,
                           int nrec, int nz, int nx, double dz, double dx,
                           double dy, int is3d, hipStream_t stream);
int fteik_k_d2f(float* out, const double* in, long n, hipStream_t stream);
int fteik_k_fill_u64(unsigned long long* a, long n, unsigned long long v,
                     hipStream_t stream);
int fteik_k_f2d_up(double* out, const float* in, long n, hipStream_t stream);
int fteik_k_analytic(double* tt_xfast, int nz, int nx, int ny, double dz,
                     double dx, double dy, double z0, double x0, double y0,
                     double zs, double xs, double ys, double vconst,
                     double vtop, double vbot, int gradient,
                     hipStream_t stream);
int fteik_k_analytic_points(double* out, const double* z, const double* x,
                            const double* y, int n, int nz, double dz,
                            double z0, double zs, double xs, double ys,
                            double vconst, double vtop, double vbot,
                            int gradient, hipStream_t stream);
int fteik_k_analytic_reorder(double* out, const double* tt_xfast, int nz,
                             int nx, int ny, int order, hipStream_t stream);

// sweep2d.hip
int fteik_k_solve2d(double* tt, const double* slow, int nz, int nx,
                    double dz, double dx, double eps, double zsa, double xsa,
                    int zsi, int xsi, int nsweep, int nsrc_batch,
                    int ld_tt, const double* zsa_v, const double* xsa_v,
                    const int* zsi_v, const int* xsi_v, int schedule,
                    int* err, double* scratch, hipStream_t stream);
// doubles of scratch per source the pipelined 2D sweeps need
constexpr int kSolve2dScratch = 512;

// locate.hip
int fteik_k_tables_d2f(float* tables, const double* tt, int ngrd,
                       hipStream_t stream);
int fteik_k_locate(const float* tables, int ntf, int ngrd, long ldgrd,
                   const float* obs, const float* wt, const unsigned char* skip,
                   int lntf, float xnorm, float* t0work, float* objwork,
                   int accumulate, unsigned long long* best,
                   hipStream_t stream);
int fteik_k_locate_batch(const float* tables, int ntf, int ngrd, long ldgrd,
                         const float* obs, const float* wt, const float* wt2,
                         const float* xnorm, int lntf, int nev,
                         unsigned long long* best, hipStream_t s);

// stencil_eval.hip (diagnostics: device stencils on given node inputs,
// and the latency of the stencil's dependency chain)
int fteik_k_time_stencil_chain(const void* cst, int nsteps, int nwg,
                               int fp32, void* out, hipStream_t stream);
int fteik_k_eval3d(const double* in, const int* node, const void* cst,
                   int n, double* out_noinit, double* out_init,
                   hipStream_t stream);
int fteik_k_sqrt_pos(const double* x, double* y, long n, hipStream_t stream);
}

extern "C" {
// util.hip (continued)
int fteik_k_min_merge_f(double* tt, const float* tt32, long n, double bias,
                        hipStream_t stream);
int fteik_k_graph_ijkv(int* ijkv, const int* levelPtr, int nz, int nx, int ny,
                       int sweep, hipStream_t stream);
int fteik_k_statics(float* tstatic, const float* tobs, const float* tt,
                    const float* wts, const float* t0, int ldo, int nobs,
                    int nev, hipStream_t stream);
// locate.hip (continued)
int fteik_k_locate_t0_at(float* t0, const float* tables, int ntf, long ldgrd,
                         const float* obs, const float* wt,
                         const unsigned char* skip, const float* xnorm,
                         int lntf, const int* cells, int nev, int ngrd,
                         hipStream_t stream);
}

extern "C" {
// surrogate.hip
int fteik_k_gemm_bf16(const void* A, long sam, long sak, const void* B,
                      long sbk, long sbn, float* C, long ldc, int M, int N,
                      int K, const float* bias, int accumulate,
                      hipStream_t stream);
// GEMMs with fused epilogues (k <= 256 except the partial form)
int fteik_k_gemm_parts(int M, int N, int K);
int fteik_k_gemm_partial(const void* A, long sam, long sak, const void* B,
                         long sbk, long sbn, int M, int N, int K, float* P,
                         hipStream_t stream);
// split-K partial products of up to 4 GEMMs in one launch (each into its
// own P, fteik_k_gemm_parts(M, N, K) parts of M x N)
struct FteikGemmPart {
    const void* A; long sam, sak;
    const void* B; long sbk, sbn;
    int M, N, K;
    float* P;
};
int fteik_k_gemm_partials(int n, const FteikGemmPart* parts,
                          hipStream_t stream);
int fteik_k_gemm_silu(const void* A, long sam, long sak, const void* B,
                      long sbk, long sbn, int M, int N, int K,
                      const float* bias, float* z, void* h, long ld,
                      hipStream_t stream);
int fteik_k_f2bf(const float* a, void* b, long n, hipStream_t stream);
void fteik_k_sur_release(void);
int fteik_k_sur_layer1(const float* x, const float* W1, const float* b1,
                       float* z, void* h, long rows, int H, hipStream_t stream);
// A range of parameters whose gradient is the in-order sum of nparts parts
// src[z*pstride + i - off].
struct FteikGradSeg {
    long off, len, pstride;
    const float* src;
    int nparts;
    // weights: the segment is a row-major matrix with ncol columns whose
    // BF16 transpose is also kept, at tcopy with row stride ldt (null:
    // none)
    unsigned short* tcopy;
    int ncol, ldt;
};
// One fused training pass of the surrogate over 32-row blocks (forward,
// squared error, backward through the activations); see surrogate.hip.
struct FteikMlpArgs {
    const float* x;              // [pool][3]
    const float* tgt;            // [pool][T]
    long pool, first;            // row r of the pass is row (first + r) % pool
    long rows;
    int T, TP;                   // outputs, and T rounded up to 16
    const float* p;              // FP32 parameters (W1, biases)
    const unsigned short* pb;    // BF16 parameters, weights [out][in]
    const unsigned short* pbT;   // BF16 transposed weights [in][out]
    long oW1, ob1, oW[3], ob[3], oW5, ob5;   // offsets in p and pb
    long oWT[3], oW5T;           // offsets in pbT (W5T rows padded to TP)
    float* z[4];                 // pre-activations [rows][256]
    unsigned short* h[4];        // activations [rows][256]
    unsigned short* dz[4];       // pre-activation gradients (dz[0] unused)
    unsigned short* dy;          // output gradient [rows][T]
    float* colp;                 // bias-gradient parts: 5 slots of
                                 // [nblk][256] (slot 3: [nblk][T])
    float* dw1p;                 // first-layer weight-gradient parts [nblk][768]
    float inv;                   // 1 / (rows T)
    double* loss;
};
constexpr int kMlpRows = 32;     // rows per workgroup of the fused pass
int fteik_k_mlp_train(const FteikMlpArgs* a, hipStream_t stream);
int fteik_k_sur_transpose(const void* src, int rows, int cols, void* dst,
                          int ldt, hipStream_t stream);
constexpr int kSurMaxSeg = 12;
int fteik_k_sur_grad_adam(const FteikGradSeg* segs, int nseg, float* g,
                          float* p, float* m, float* v, void* pb, long n,
                          int adam, float lr, float b1, float b2, float eps,
                          float c1, float c2, double* lclear,
                          hipStream_t stream);
int fteik_k_sur_samples(const float* tab, long ldgrd, const void* geo,
                        long ngrd, unsigned long long seed, long pool,
                        long first, long rows, float* x, float* tgt, int ldt,
                        double* loss, hipStream_t stream);
int fteik_k_sur_geo_size(void);
int fteik_k_inv_start(float* pos, long R, int starts, hipStream_t stream);
int fteik_k_inv_hidden(const void* hin, long R, const void* W,
                       const float* b, void* hout, int H, hipStream_t s);
// Output layer of the inversion network, station-major: y[s*ldy + q] =
// sum_k W5[s][k] h[q][k] for rows q < NR of h (NR x 256 BF16) and stations
// s < T <= 64.  y must hold ((T + 31)/32)*32 rows of ldy >= NR floats, ldy
// a multiple of 32: whole 32 x 32 tiles are stored.
int fteik_k_inv_out(const void* hin, long NR, const void* W5, int T,
                    float* y, long ldy, hipStream_t s);
int fteik_k_inv_layer1(const float* pos, long R, const float* W1,
                       const float* b1, float* zp, void* h, int H,
                       hipStream_t stream);
int fteik_k_inv_act(const float* zall, long R, int H, const float* b,
                    float* zp, void* h, hipStream_t stream);
int fteik_k_inv_gn(const float* y, long R, int ldy, const float* b5,
                   const void* geo, const float* obs, const float* wt,
                   int lntf, int starts, float* pos, float* misfit, int apply,
                   hipStream_t stream);
int fteik_k_inv_score(const float* pos, const float* misfit, int starts,
                      int nev, const void* geo, const float* tab, long ldgrd,
                      const float* obs, const float* wt,
                      const unsigned char* skip, const float* xnorm, int lntf,
                      unsigned long long* best, int* fallback,
                      hipStream_t stream);
}

// Host mirror of the surrogate's geometry block (surrogate.hip GeoArgs).
struct FteikSurrogateGeo {
    int nz, nx, ny, ntf;
    float dz, dx, dy;
    float sx[64], sy[64], sz[64];
    float v0, scale;
};
