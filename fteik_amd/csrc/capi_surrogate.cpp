// capi_surrogate.cpp -- the learned surrogate of the locator: training,
// inference and surrogate-pruned location (extensions), plus the MFMA
// diagnostics.
#include "capi_internal.h"

using namespace fteik_capi;

// ====================================================================
//     Learned surrogate for the locator (EXTENSION: no counterpart)
// ====================================================================
namespace {

struct Surrogate {
    bool init = false;
    int ntf = 0, H = 256, step = 0;
    FteikSurrogateGeo geo;
    void *d_geo = nullptr;
    long np = 0;                       // parameter count
    long oW1, ob1, oW[3], ob[3], oW5, ob5;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    unsigned short *d_pb = nullptr;    // BF16 copy of all parameters
    // BF16 transposes of the weights ([in][out]; W5's rows padded to TP
    // outputs with zeros) for the backward products
    unsigned short *d_pbT = nullptr;
    long oWT[3] = {0, 0, 0}, oW5T = 0;
    int TP = 16;
    // work buffers (grown on demand)
    long cap = 0;
    float *d_x = nullptr, *d_tgt = nullptr, *d_y = nullptr;
    float *d_z[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *d_h[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *d_dy = nullptr;
    unsigned short *d_dz[3] = {nullptr, nullptr, nullptr};   // dz_1..dz_3
    // gradient parts: split-K partial products of the weight gradients
    // (grown on demand), column sums per 32-row block of the bias
    // gradients (5 slots), first-layer sums per 32-row block
    float *d_P = nullptr, *d_colp = nullptr, *d_dw1p = nullptr;
    long capP = 0;
    // training samples of the whole pool (made once per training call)
    float *d_poolx = nullptr, *d_pooltgt = nullptr;
    long poolCap = 0;
    double *d_loss = nullptr;          // two slots, alternating by step
    unsigned long long seed = 13;
    // inversion work buffers (grown on demand): event arrays sized by
    // events, the rest by rows R = events * starts
    long invEvents = 0, invRows = 0;
    float *d_io = nullptr, *d_iw = nullptr, *d_ixn = nullptr;
    unsigned char *d_isk = nullptr;
    unsigned long long *d_ibest = nullptr;
    int *d_ifb = nullptr;
    float *d_pos = nullptr, *d_mis = nullptr,
          *d_yall = nullptr;
    unsigned short *d_ha = nullptr, *d_hb = nullptr;
    // proposals from the network (default) or from the straight-ray
    // baseline alone (network output taken as zero: the ablation)
    bool useNet = true;
    float *d_zero = nullptr;           // ntf zero biases for the baseline
};
Surrogate gs;

void free_inv_work()
{
    dfree(gs.d_io); dfree(gs.d_iw); dfree(gs.d_ixn); dfree(gs.d_isk);
    dfree(gs.d_ibest); dfree(gs.d_ifb);
    dfree(gs.d_pos); dfree(gs.d_mis);
    dfree(gs.d_yall); dfree(gs.d_ha); dfree(gs.d_hb);
    gs.invEvents = gs.invRows = 0;
}

// Inversion buffers for n events (lntf observations each) and R rows.
bool inv_work(long n, long lntf, long R)
{
    const int H = gs.H, T = gs.ntf;
    if (n*lntf > gs.invEvents || R > gs.invRows) {
        const long ne = std::max(n*lntf, gs.invEvents);
        const long nr = std::max(R, gs.invRows);
        const long nev = std::max(n, ne/std::max(lntf, 1L));
        free_inv_work();
        if (!(dalloc(&gs.d_io, (size_t) ne) && dalloc(&gs.d_iw, (size_t) ne) &&
              dalloc(&gs.d_isk, (size_t) ne) && dalloc(&gs.d_ixn, (size_t) nev) &&
              dalloc(&gs.d_ibest, (size_t) nev) && dalloc(&gs.d_ifb, (size_t) nev) &&
              dalloc(&gs.d_pos, 3*(size_t) nr) && dalloc(&gs.d_mis, (size_t) nr) &&
              // station-major outputs: whole 32-station tiles of rows
              // padded to a multiple of 32 (fteik_k_inv_out)
              dalloc(&gs.d_yall, (size_t) ((T + 31)/32*32)*
                                 (size_t) ((4*nr + 31)/32*32)) &&
              dalloc(&gs.d_ha, 4*(size_t) nr*H) &&
              dalloc(&gs.d_hb, 4*(size_t) nr*H))) {
            free_inv_work();
            return false;
        }
        gs.invEvents = ne;
        gs.invRows = nr;
    }
    return true;
}

void free_surrogate()
{
    dfree(gs.d_geo); dfree(gs.d_p); dfree(gs.d_g); dfree(gs.d_m); dfree(gs.d_v);
    dfree(gs.d_pb); dfree(gs.d_x); dfree(gs.d_tgt); dfree(gs.d_y);
    dfree(gs.d_dy); dfree(gs.d_P); dfree(gs.d_pbT);
    for (int l = 0; l < 3; l++) dfree(gs.d_dz[l]);
    dfree(gs.d_colp); dfree(gs.d_dw1p); dfree(gs.d_loss);
    dfree(gs.d_poolx); dfree(gs.d_pooltgt);
    dfree(gs.d_zero);
    for (int l = 0; l < 4; l++) { dfree(gs.d_z[l]); dfree(gs.d_h[l]); }
    free_inv_work();
    fteik_k_sur_release();
    gs = Surrogate();
}

long row_blocks(long rows) { return (rows + kMlpRows - 1)/kMlpRows; }

bool sur_work(long rows)
{
    if (rows <= gs.cap) return true;
    dfree(gs.d_x); dfree(gs.d_tgt); dfree(gs.d_y); dfree(gs.d_dy);
    dfree(gs.d_colp); dfree(gs.d_dw1p);
    for (int l = 0; l < 3; l++) dfree(gs.d_dz[l]);
    for (int l = 0; l < 4; l++) { dfree(gs.d_z[l]); dfree(gs.d_h[l]); }
    const size_t R = (size_t) rows, H = (size_t) gs.H, T = (size_t) gs.ntf;
    const size_t nb = (size_t) row_blocks(rows);
    bool ok = dalloc(&gs.d_x, 3*R) && dalloc(&gs.d_tgt, R*T) &&
              dalloc(&gs.d_y, R*T) && dalloc(&gs.d_dy, R*T) &&
              dalloc(&gs.d_colp, 5*nb*H) && dalloc(&gs.d_dw1p, nb*3*H);
    for (int l = 0; l < 3 && ok; l++) ok = dalloc(&gs.d_dz[l], R*H);
    for (int l = 0; l < 4 && ok; l++)
        ok = dalloc(&gs.d_z[l], R*H) && dalloc(&gs.d_h[l], R*H);
    if (!ok) { gs.cap = 0; return false; }
    gs.cap = rows;
    return true;
}

// BF16 transposes of the weights from the BF16 copy (after it changed
// other than by the Adam kernel, which keeps them current itself).
bool sur_transposes()
{
    hipStream_t s = g_dev.stream;
    const int H = gs.H;
    bool ok = true;
    for (int l = 0; l < 3 && ok; l++)
        ok = fteik_k_sur_transpose(gs.d_pb + gs.oW[l], H, H,
                                   gs.d_pbT + gs.oWT[l], H, s) == 0;
    return ok && fteik_k_sur_transpose(gs.d_pb + gs.oW5, gs.ntf, H,
                                       gs.d_pbT + gs.oW5T, gs.TP, s) == 0;
}

// Forward pass over `rows` rows of d_x: network output (before the
// baseline and scale) into d_y.
bool sur_forward(long rows)
{
    hipStream_t s = g_dev.stream;
    const int H = gs.H, T = gs.ntf;
    const unsigned short *pb = gs.d_pb;
    bool ok = fteik_k_sur_layer1(gs.d_x, gs.d_p + gs.oW1, gs.d_p + gs.ob1,
                                 gs.d_z[0], gs.d_h[0], rows, H, s) == 0;
    for (int l = 0; l < 3 && ok; l++) {
        // z_{l+1} = h_l W^T + b, h_{l+1} = silu(z_{l+1})
        ok = fteik_k_gemm_silu(gs.d_h[l], H, 1, pb + gs.oW[l], 1, H,
                               (int) rows, H, H, gs.d_p + gs.ob[l],
                               gs.d_z[l + 1], gs.d_h[l + 1], H, s) == 0;
    }
    return ok && fteik_k_gemm_bf16(gs.d_h[3], H, 1, pb + gs.oW5, 1, H, gs.d_y,
                                   T, (int) rows, T, H, gs.d_p + gs.ob5, 0,
                                   s) == 0;
}

// One training pass over `rows` rows, row r being row (first + r) % pool
// of x and tgt (the mean squared error over rows x ntf added to *loss):
// the fused pass of 32-row blocks through the network and back, then the
// weight-gradient products as split-K parts; the segment table for
// k_grad_adam lists every parameter gradient's parts in parameter order.
bool sur_train_pass(long rows, const float *x, const float *tgt, long pool,
                    long first, double *loss, FteikGradSeg *seg, int *nseg)
{
    hipStream_t s = g_dev.stream;
    const int H = gs.H, T = gs.ntf, M = (int) rows;
    const long nrb = row_blocks(rows);
    const int nsH = fteik_k_gemm_parts(H, H, M);
    const int ns5 = fteik_k_gemm_parts(T, H, M);
    const long needP = 3L*nsH*H*H + (long) ns5*T*H;
    if (needP > gs.capP) {
        dfree(gs.d_P);
        gs.capP = 0;
        if (!dalloc(&gs.d_P, (size_t) needP)) return false;
        gs.capP = needP;
    }
    float *P5 = gs.d_P + 3L*nsH*H*H;
    FteikMlpArgs a;
    memset(&a, 0, sizeof(a));
    a.x = x; a.tgt = tgt; a.pool = pool; a.first = first;
    a.rows = rows; a.T = T; a.TP = gs.TP;
    a.p = gs.d_p; a.pb = gs.d_pb; a.pbT = gs.d_pbT;
    a.oW1 = gs.oW1; a.ob1 = gs.ob1; a.oW5 = gs.oW5; a.ob5 = gs.ob5;
    a.oW5T = gs.oW5T;
    for (int l = 0; l < 3; l++) {
        a.oW[l] = gs.oW[l]; a.ob[l] = gs.ob[l]; a.oWT[l] = gs.oWT[l];
        a.dz[l + 1] = gs.d_dz[l];
    }
    for (int l = 0; l < 4; l++) { a.z[l] = gs.d_z[l]; a.h[l] = gs.d_h[l]; }
    a.dy = gs.d_dy; a.colp = gs.d_colp; a.dw1p = gs.d_dw1p;
    a.inv = 1.0f/(float) ((double) rows*T);
    a.loss = loss;
    // weight gradients in one launch: dW_l = dz_{l+1}^T h_l, dW5 = dy^T h3
    FteikGemmPart parts[4];
    for (int l = 0; l < 3; l++)
        parts[l] = FteikGemmPart{gs.d_dz[l], 1, H, gs.d_h[l], H, 1, H, H, M,
                                 gs.d_P + (long) l*nsH*H*H};
    parts[3] = FteikGemmPart{gs.d_dy, 1, T, gs.d_h[3], H, 1, T, H, M, P5};
    bool ok = fteik_k_mlp_train(&a, s) == 0 &&
              fteik_k_gemm_partials(4, parts, s) == 0;
    auto colp = [&](int slot) { return gs.d_colp + slot*nrb*H; };
    int q = 0;
    auto add = [&](long off, long len, const float *src, int np, long ps,
                   unsigned short *tc, int ncol, int ldt) {
        seg[q].off = off; seg[q].len = len; seg[q].src = src;
        seg[q].nparts = np; seg[q].pstride = ps;
        seg[q].tcopy = tc; seg[q].ncol = ncol; seg[q].ldt = ldt;
        q++;
    };
    add(gs.oW1, 3L*H, gs.d_dw1p, (int) nrb, 3L*H, nullptr, 0, 0);
    add(gs.ob1, H, colp(4), (int) nrb, H, nullptr, 0, 0);
    for (int l = 0; l < 3; l++) {
        add(gs.oW[l], (long) H*H, gs.d_P + (long) l*nsH*H*H, nsH, (long) H*H,
            gs.d_pbT + gs.oWT[l], H, H);
        add(gs.ob[l], H, colp(l), (int) nrb, H, nullptr, 0, 0);
    }
    add(gs.oW5, (long) T*H, P5, ns5, (long) T*H, gs.d_pbT + gs.oW5T, H,
        gs.TP);
    add(gs.ob5, T, colp(3), (int) nrb, T, nullptr, 0, 0);
    *nseg = q;
    return ok;
}

} // namespace

extern "C" {

void fteik_surrogate_free(void) { free_surrogate(); }

void fteik_surrogate_initialize(int ntf, int nz, int nx, int ny, double dz,
                                double dx, double dy, const double *zsta,
                                const double *xsta, const double *ysta,
                                double v0, double scale, int seed, int *ierr)
{
    static const char *who = "fteik_surrogate_initialize";
    *ierr = 0;
    free_surrogate();
    if (ntf < 1 || ntf > 64 || nz < 4 || nx < 4 || ny < 4 || v0 <= 0.0 ||
        scale <= 0.0) {
        msg("%s: need 1 <= ntf <= 64, grids of at least 4 points, v0 > 0, "
            "scale > 0", who);
        *ierr = 1;
        return;
    }
    if (!device_ok(who)) { *ierr = 1; return; }
    gs.ntf = ntf;
    FteikSurrogateGeo &g = gs.geo;
    memset(&g, 0, sizeof(g));
    g.nz = nz; g.nx = nx; g.ny = ny; g.ntf = ntf;
    g.dz = (float) dz; g.dx = (float) dx; g.dy = (float) dy;
    for (int s = 0; s < ntf; s++) {
        g.sz[s] = (float) zsta[s]; g.sx[s] = (float) xsta[s];
        g.sy[s] = (float) ysta[s];
    }
    g.v0 = (float) v0; g.scale = (float) scale;
    if ((int) sizeof(FteikSurrogateGeo) != fteik_k_sur_geo_size()) {
        msg("%s: geometry block size mismatch", who);
        *ierr = 1;
        return;
    }
    const long H = gs.H;
    long o = 0;
    gs.oW1 = o; o += 3*H;
    gs.ob1 = o; o += H;
    for (int l = 0; l < 3; l++) { gs.oW[l] = o; o += H*H; gs.ob[l] = o; o += H; }
    gs.oW5 = o; o += ntf*H;
    gs.ob5 = o; o += ntf;
    gs.np = o;
    // He-uniform initial weights, zero biases (deterministic in `seed`)
    std::vector<float> p((size_t) o, 0.0f);
    unsigned long long st = 0x9e3779b97f4a7c15ull ^ (unsigned long long) seed;
    auto uni = [&]() -> float {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float) ((double) (st >> 11)*(1.0/9007199254740992.0));
    };
    auto fill = [&](long off, long n, int fanin) {
        const float a = sqrtf(6.0f/(float) fanin);
        for (long i = 0; i < n; i++) p[(size_t) (off + i)] = a*(2.0f*uni() - 1.0f);
    };
    fill(gs.oW1, 3*H, 3);
    for (int l = 0; l < 3; l++) fill(gs.oW[l], H*H, (int) H);
    fill(gs.oW5, ntf*H, (int) H);
    for (long i = gs.oW5; i < gs.oW5 + ntf*H; i++) p[(size_t) i] *= 0.1f;
    gs.TP = (ntf + 15)/16*16;
    for (int l = 0; l < 3; l++) gs.oWT[l] = (long) l*H*H;
    gs.oW5T = 3*H*H;
    const size_t nT = (size_t) (3*H*H + H*gs.TP);
    bool ok = dalloc((char **) &gs.d_geo, sizeof(FteikSurrogateGeo)) &&
              dalloc(&gs.d_p, (size_t) o) && dalloc(&gs.d_g, (size_t) o) &&
              dalloc(&gs.d_m, (size_t) o) && dalloc(&gs.d_v, (size_t) o) &&
              dalloc(&gs.d_pb, (size_t) o) && dalloc(&gs.d_pbT, nT) &&
              dalloc(&gs.d_loss, 2) &&
              dalloc(&gs.d_zero, 64) &&
              h2d(gs.d_geo, &g, sizeof(g)) &&
              h2d(gs.d_p, p.data(), (size_t) o*sizeof(float));
    if (ok) {
        hipMemsetAsync(gs.d_m, 0, (size_t) o*sizeof(float), g_dev.stream);
        hipMemsetAsync(gs.d_v, 0, (size_t) o*sizeof(float), g_dev.stream);
        hipMemsetAsync(gs.d_zero, 0, 64*sizeof(float), g_dev.stream);
        hipMemsetAsync(gs.d_pbT, 0, nT*sizeof(unsigned short), g_dev.stream);
        ok = fteik_k_f2bf(gs.d_p, gs.d_pb, o, g_dev.stream) == 0 &&
             sur_transposes() && sync_ok();
    }
    if (!ok) { msg("%s: GPU error", who); free_surrogate(); *ierr = 1; return; }
    gs.init = true;
}

void fteik_surrogate_getNumberOfParameters(int *np, int *ierr)
{
    *ierr = gs.init ? 0 : 1;
    *np = (int) gs.np;
}

void fteik_surrogate_getParameters(int np, float *p, int *ierr)
{
    *ierr = (gs.init && np == gs.np) ? 0 : 1;
    if (*ierr == 0 && !d2h(p, gs.d_p, (size_t) np*sizeof(float))) *ierr = 1;
}

void fteik_surrogate_setParameters(int np, const float *p, int *ierr)
{
    *ierr = (gs.init && np == gs.np) ? 0 : 1;
    if (*ierr) return;
    if (!h2d(gs.d_p, p, (size_t) np*sizeof(float)) ||
        fteik_k_f2bf(gs.d_p, gs.d_pb, np, g_dev.stream) != 0 ||
        !sur_transposes() || !sync_ok())
        *ierr = 1;
}

// Network output (before baseline and scale) at n normalised positions.
void fteik_surrogate_forward(int n, const float *pos, float *out, int *ierr)
{
    *ierr = 0;
    if (!gs.init || n < 1 || !sur_work(n) ||
        !h2d(gs.d_x, pos, 3*(size_t) n*sizeof(float)) || !sur_forward(n) ||
        !d2h(out, gs.d_y, (size_t) n*gs.ntf*sizeof(float))) {
        msg("fteik_surrogate_forward: failed");
        *ierr = 1;
    }
}

// One forward/backward pass on given data: parameter gradients of the mean
// squared error (over n x ntf outputs) and the loss.
void fteik_surrogate_gradients(int n, const float *pos, const float *tgt,
                               float *grad, double *loss, int *ierr)
{
    *ierr = 0;
    const double zero = 0.0;
    FteikGradSeg seg[kSurMaxSeg];
    int nseg = 0;
    if (!gs.init || n < 1 || !sur_work(n) ||
        !h2d(gs.d_x, pos, 3*(size_t) n*sizeof(float)) ||
        !h2d(gs.d_tgt, tgt, (size_t) n*gs.ntf*sizeof(float)) ||
        !h2d(gs.d_loss, &zero, sizeof(double)) ||
        !sur_train_pass(n, gs.d_x, gs.d_tgt, n, 0, gs.d_loss, seg, &nseg) ||
        fteik_k_sur_grad_adam(seg, nseg, gs.d_g, gs.d_p, gs.d_m, gs.d_v,
                              gs.d_pb, gs.np, 0, 0.0f, 0.0f, 0.0f, 0.0f,
                              0.0f, 0.0f, nullptr, g_dev.stream) != 0 ||
        !d2h(grad, gs.d_g, (size_t) gs.np*sizeof(float)) ||
        !d2h(loss, gs.d_loss, sizeof(double))) {
        msg("fteik_surrogate_gradients: failed");
        *ierr = 1;
    }
}

// Trains on samples drawn from the locator's tables (a pool of `pool`
// uniformly drawn cells); Adam with a cosine-decayed learning rate.
void fteik_surrogate_train(int steps, int batch, double lr, int pool,
                           double *lossOut, int *ierr)
{
    static const char *who = "fteik_surrogate_train";
    *ierr = 0;
    *lossOut = 0.0;
    if (!gs.init || !gl.init || gl.ntf != gs.ntf ||
        (long) gs.geo.nz*gs.geo.nx*gs.geo.ny != gl.ngrd || steps < 1 ||
        batch < 64 || pool < batch) {
        msg("%s: surrogate and locator must be initialized for the same "
            "stations and grid", who);
        *ierr = 1;
        return;
    }
    if (!sur_work(batch)) { msg("%s: out of GPU memory", who); *ierr = 1;
                            return; }
    hipStream_t s = g_dev.stream;
    // the pool's samples, once: a step's rows are `batch` consecutive
    // entries of the pool (wrapping), the same cells and values the
    // per-step draw would give
    if (gs.poolCap < pool) {
        dfree(gs.d_poolx); dfree(gs.d_pooltgt);
        gs.poolCap = 0;
        if (!dalloc(&gs.d_poolx, 3*(size_t) pool) ||
            !dalloc(&gs.d_pooltgt, (size_t) pool*gs.ntf)) {
            dfree(gs.d_poolx); dfree(gs.d_pooltgt);
            msg("%s: out of GPU memory", who);
            *ierr = 1;
            return;
        }
        gs.poolCap = pool;
    }
    bool ok = fteik_k_sur_samples(gl.d_tab, gl.ldgrd, gs.d_geo, gl.ngrd,
                                  gs.seed, pool, 0, pool, gs.d_poolx,
                                  gs.d_pooltgt, gs.ntf, nullptr, s) == 0 &&
              hipMemsetAsync(gs.d_loss, 0, 2*sizeof(double), s) == hipSuccess;
    double tail = 0.0;
    const int tailN = std::max(1, std::min(steps, 50));
    FteikGradSeg seg[kSurMaxSeg];
    int nseg = 0;
    for (int it = 0; it < steps && ok; it++) {
        // step it accumulates its loss in slot it & 1; its Adam launch
        // clears the other slot for the next step (no host synchronization
        // per step, so the host queues the next steps while the GPU works)
        double *const loss = gs.d_loss + (it & 1);
        ok = sur_train_pass(batch, gs.d_poolx, gs.d_pooltgt, pool,
                            ((long) it*batch) % pool, loss, seg, &nseg);
        gs.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f/(1.0f - powf(b1, (float) gs.step));
        const float c2 = 1.0f/(1.0f - powf(b2, (float) gs.step));
        const float lrt = (float) (lr*0.5*(1.0 + cos(3.14159265358979*it/steps)));
        ok = ok && fteik_k_sur_grad_adam(seg, nseg, gs.d_g, gs.d_p, gs.d_m,
                                         gs.d_v, gs.d_pb, gs.np, 1, lrt, b1,
                                         b2, 1e-8f, c1, c2,
                                         gs.d_loss + ((it + 1) & 1), s) == 0;
        if (ok && it >= steps - tailN) {
            double l = 0.0;
            ok = d2h(&l, loss, sizeof(double));
            tail += l;
        }
    }
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    *lossOut = tail/tailN;
}

// Proposals from the network (on = 1, default) or from the straight-ray
// baseline |x - station| / v0 alone (on = 0), with the same Gauss-Newton
// steps, cube scoring and full-search fallback: what the network adds.
void fteik_surrogate_setNetworkProposals(int on, int *ierr)
{
    *ierr = 0;
    if (on != 0 && on != 1) {
        msg("fteik_surrogate_setNetworkProposals: on must be 0 or 1");
        *ierr = 1;
        return;
    }
    gs.useNet = on == 1;
}

// Surrogate-pruned location of events evnmbr[0..n-1] (fresh origin-time
// state): damped Gauss-Newton on the network's misfit from `starts`
// starting points, exact scoring of the 4 x 4 x 4 cells around the best,
// full grid search for events whose best cell lies on the cube's faces.
void locate_locateEventsSurrogate(int n, const int *evnmbr, int starts,
                                  int iterations, int *optIndx, double *t0Opt,
                                  double *objOpt, int *nFallback, int *ierr)
{
    static const char *who = "locate_locateEventsSurrogate";
    *ierr = 0;
    *nFallback = 0;
    if (!gs.init || !gl.init || gl.ntf != gs.ntf || n < 1 || starts < 1 ||
        starts > 9 || iterations < 0 ||
        (long) gs.geo.nz*gs.geo.nx*gs.geo.ny != gl.ngrd) {
        msg("%s: surrogate and locator not set up for the same grid", who);
        *ierr = 1;
        return;
    }
    for (int e = 0; e < n; e++)
        if (evnmbr[e] < 1 || evnmbr[e] > gl.nEvents) {
            msg("%s: invalid event %d", who, evnmbr[e]);
            *ierr = 1;
            return;
        }
    hipStream_t s = g_dev.stream;
    const int H = gs.H, T = gs.ntf;
    const long R = (long) n*starts;
    std::vector<float> o((size_t) n*gl.lntf), w((size_t) n*gl.lntf),
                       xn((size_t) n);
    std::vector<unsigned char> sk((size_t) n*gl.lntf);
    for (int e = 0; e < n; e++) {
        const size_t k = (size_t) (evnmbr[e] - 1)*gl.lntf;
        memcpy(&o[(size_t) e*gl.lntf], &gl.obs[k], gl.lntf*sizeof(float));
        memcpy(&w[(size_t) e*gl.lntf], &gl.wt[k], gl.lntf*sizeof(float));
        memcpy(&sk[(size_t) e*gl.lntf], &gl.skip[k], (size_t) gl.lntf);
        xn[(size_t) e] = weight_norm(&gl.wt[k], gl.lntf);
    }
    if (!inv_work(n, gl.lntf, R)) {
        msg("%s: out of GPU memory", who);
        *ierr = 1;
        return;
    }
    float *d_o = gs.d_io, *d_w = gs.d_iw, *d_xn = gs.d_ixn, *d_pos = gs.d_pos,
          *d_mis = gs.d_mis,
          *d_yall = gs.d_yall;
    unsigned short *d_ha = gs.d_ha, *d_hb = gs.d_hb;
    unsigned char *d_sk = gs.d_isk;
    unsigned long long *d_best = gs.d_ibest;
    int *d_fb = gs.d_ifb;
    bool ok = h2d(d_o, o.data(), o.size()*sizeof(float)) &&
              h2d(d_w, w.data(), w.size()*sizeof(float)) &&
              h2d(d_xn, xn.data(), xn.size()*sizeof(float)) &&
              h2d(d_sk, sk.data(), sk.size());
    ok = ok && fteik_k_inv_start(d_pos, R, starts, s) == 0;
    // network outputs station-major: output s of row q at d_yall[s*ldy + q]
    const long ldy = (4*R + 31)/32*32;
    if (ok && !gs.useNet)
        // baseline proposals: network outputs and tangents all zero, so the
        // Gauss-Newton steps fit the straight-ray times alone
        ok = fteik_k_fill_u64((unsigned long long *) d_yall, T*ldy/2, 0ull,
                              s) == 0;
    for (int it = 0; it <= iterations && ok; it++) {
        if (gs.useNet) {
            // forward with tangents over 4R rows
            ok = fteik_k_inv_layer1(d_pos, R, gs.d_p + gs.oW1,
                                    gs.d_p + gs.ob1, nullptr, d_ha, H, s) == 0;
            unsigned short *hin = d_ha, *hout = d_hb;
            for (int l = 0; l < 3 && ok; l++) {
                // fused GEMM + activation (the GEMM then k_inv_act, bitwise)
                ok = fteik_k_inv_hidden(hin, R, gs.d_pb + gs.oW[l],
                                        gs.d_p + gs.ob[l], hout, H, s) == 0;
                std::swap(hin, hout);
            }
            ok = ok && fteik_k_inv_out(hin, 4*R, gs.d_pb + gs.oW5, T, d_yall,
                                       ldy, s) == 0;
        }
        ok = ok && fteik_k_inv_gn(d_yall, R, (int) ldy,
                                  gs.useNet ? gs.d_p + gs.ob5 : gs.d_zero,
                                  gs.d_geo, d_o, d_w, gl.lntf, starts, d_pos,
                                  d_mis, it < iterations ? 1 : 0, s) == 0;
    }
    ok = ok && fteik_k_inv_score(d_pos, d_mis, starts, n, gs.d_geo, gl.d_tab,
                                 gl.ldgrd, d_o, d_w, d_sk, d_xn, gl.lntf,
                                 d_best, d_fb, s) == 0;
    std::vector<unsigned long long> best((size_t) n);
    std::vector<int> fb((size_t) n);
    ok = ok && d2h(best.data(), d_best, best.size()*sizeof(unsigned long long))
            && d2h(fb.data(), d_fb, fb.size()*sizeof(int));
    // full grid search for the flagged events
    std::vector<int> redo;
    for (int e = 0; e < n && ok; e++) if (fb[(size_t) e]) redo.push_back(e);
    *nFallback = (int) redo.size();
    if (ok && !redo.empty()) {
        const int m = (int) redo.size();
        std::vector<float> o2((size_t) m*gl.lntf), w2((size_t) m*gl.lntf),
                           xn2((size_t) m);
        std::vector<unsigned char> sk2((size_t) m*gl.lntf);
        std::vector<unsigned long long> b2((size_t) m, ~0ull);
        for (int q = 0; q < m; q++) {
            const size_t e = (size_t) redo[(size_t) q];
            memcpy(&o2[(size_t) q*gl.lntf], &o[e*gl.lntf], gl.lntf*sizeof(float));
            memcpy(&w2[(size_t) q*gl.lntf], &w[e*gl.lntf], gl.lntf*sizeof(float));
            memcpy(&sk2[(size_t) q*gl.lntf], &sk[e*gl.lntf], (size_t) gl.lntf);
            xn2[(size_t) q] = xn[e];
        }
        std::vector<float> h2((size_t) m*gl.lntf);
        mask_skipped(o2.data(), w2.data(), h2.data(), sk2.data(), o2.size());
        float *d_o2 = nullptr, *d_w2 = nullptr, *d_h2 = nullptr,
              *d_xn2 = nullptr;
        unsigned char *d_sk2 = nullptr;
        unsigned long long *d_b2 = nullptr;
        ok = dalloc(&d_o2, o2.size()) && dalloc(&d_w2, w2.size()) &&
             dalloc(&d_h2, h2.size()) &&
             dalloc(&d_xn2, xn2.size()) && dalloc(&d_sk2, sk2.size()) &&
             dalloc(&d_b2, b2.size()) &&
             h2d(d_o2, o2.data(), o2.size()*sizeof(float)) &&
             h2d(d_w2, w2.data(), w2.size()*sizeof(float)) &&
             h2d(d_h2, h2.data(), h2.size()*sizeof(float)) &&
             h2d(d_xn2, xn2.data(), xn2.size()*sizeof(float)) &&
             h2d(d_sk2, sk2.data(), sk2.size()) &&
             h2d(d_b2, b2.data(), b2.size()*sizeof(unsigned long long));
        if (ok) {
            int rc = fteik_k_locate_batch(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                          d_o2, d_w2, d_h2, d_xn2, gl.lntf, m,
                                          d_b2, s);
            if (rc == 2)
                for (int q = 0; q < m && ok; q++)
                    ok = fteik_k_locate(gl.d_tab, gl.ntf, gl.ngrd, gl.ldgrd,
                                        d_o2 + (size_t) q*gl.lntf,
                                        d_w2 + (size_t) q*gl.lntf,
                                        d_sk2 + (size_t) q*gl.lntf, gl.lntf,
                                        xn2[(size_t) q], nullptr, nullptr, 0,
                                        d_b2 + q, s) == 0;
            else
                ok = rc == 0;
        }
        ok = ok && d2h(b2.data(), d_b2, b2.size()*sizeof(unsigned long long));
        for (int q = 0; q < m && ok; q++) best[(size_t) redo[(size_t) q]] = b2[(size_t) q];
        dfree(d_o2); dfree(d_w2); dfree(d_h2); dfree(d_xn2); dfree(d_sk2);
        dfree(d_b2);
    }
    // origin times at the chosen cells (locate.F90:684-688)
    std::vector<int> cells((size_t) n);
    for (int e = 0; e < n; e++) {
        const unsigned long long k = best[(size_t) e];
        unsigned bits = (unsigned) (k >> 32);
        float obj;
        memcpy(&obj, &bits, sizeof(obj));
        cells[(size_t) e] = (int) (k & 0xffffffffull);
        optIndx[e] = cells[(size_t) e] + 1;
        objOpt[e] = (double) obj;
    }
    // the fallback flags and misfits are consumed: their buffers take the
    // chosen cells and their origin times
    std::vector<float> t0v((size_t) n);
    int *d_c = d_fb;
    float *d_t0v = d_mis;
    ok = ok && h2d(d_c, cells.data(), (size_t) n*sizeof(int)) &&
         fteik_k_locate_t0_at(d_t0v, gl.d_tab, gl.ntf, gl.ldgrd, d_o, d_w,
                              d_sk, d_xn, gl.lntf, d_c, n, gl.ngrd, s) == 0 &&
         d2h(t0v.data(), d_t0v, (size_t) n*sizeof(float));
    if (!ok) { msg("%s: GPU error", who); *ierr = 1; return; }
    for (int e = 0; e < n; e++)
        t0Opt[e] = (double) t0v[(size_t) e] + gl.tepoch[(size_t) (evnmbr[e] - 1)];
}

} // extern "C"

extern "C" {

// Diagnostic: C = A B with A (M x K) and B (K x N) row-major FP32, rounded
// to BF16 on the device and multiplied by the surrogate's MFMA GEMM.
void fteik_amd_gemmBf16(int M, int N, int K, const float *A, const float *B,
                        float *C, int *ierr)
{
    *ierr = 0;
    if (M < 1 || N < 1 || K < 1 || !device_ok("fteik_amd_gemmBf16")) {
        *ierr = 1;
        return;
    }
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    unsigned short *hA = nullptr, *hB = nullptr;
    const size_t a = (size_t) M*K, b = (size_t) K*N, c = (size_t) M*N;
    bool ok = dalloc(&dA, a) && dalloc(&dB, b) && dalloc(&dC, c) &&
              dalloc(&hA, a) && dalloc(&hB, b) &&
              h2d(dA, A, a*sizeof(float)) && h2d(dB, B, b*sizeof(float)) &&
              fteik_k_f2bf(dA, hA, (long) a, g_dev.stream) == 0 &&
              fteik_k_f2bf(dB, hB, (long) b, g_dev.stream) == 0 &&
              fteik_k_gemm_bf16(hA, K, 1, hB, N, 1, dC, N, M, N, K, nullptr,
                                0, g_dev.stream) == 0 &&
              d2h(C, dC, c*sizeof(float));
    dfree(dA); dfree(dB); dfree(dC); dfree(hA); dfree(hB);
    if (!ok) *ierr = 1;
}

// Diagnostic: one hidden layer of the surrogate's inversion over 4 blocks
// of R rows (primal, three tangents), H = 256.  hin (4R x 256), W (256 x
// 256, row n holds output n's weights) and b (256) are rounded to BF16 /
// kept FP32 as the network keeps them; fused receives the fused kernel's
// BF16 output bits, unfused those of the MFMA GEMM followed by the
// separate activation kernel.
void fteik_amd_invHiddenBf16(int R, const float *hin, const float *W,
                             const float *b, unsigned short *fused,
                             unsigned short *unfused, int *ierr)
{
    *ierr = 0;
    const int H = 256;
    if (R < 1 || !device_ok("fteik_amd_invHiddenBf16")) { *ierr = 1; return; }
    const size_t nh = 4*(size_t) R*H, nw = (size_t) H*H;
    float *dh = nullptr, *dw = nullptr, *db = nullptr, *dz = nullptr,
          *dzp = nullptr;
    unsigned short *bh = nullptr, *bw = nullptr, *o1 = nullptr, *o2 = nullptr;
    bool ok = dalloc(&dh, nh) && dalloc(&dw, nw) && dalloc(&db, (size_t) H) &&
              dalloc(&dz, nh) && dalloc(&dzp, (size_t) R*H) &&
              dalloc(&bh, nh) && dalloc(&bw, nw) && dalloc(&o1, nh) &&
              dalloc(&o2, nh) &&
              h2d(dh, hin, nh*sizeof(float)) && h2d(dw, W, nw*sizeof(float)) &&
              h2d(db, b, H*sizeof(float)) &&
              fteik_k_f2bf(dh, bh, (long) nh, g_dev.stream) == 0 &&
              fteik_k_f2bf(dw, bw, (long) nw, g_dev.stream) == 0 &&
              fteik_k_inv_hidden(bh, R, bw, db, o1, H, g_dev.stream) == 0 &&
              fteik_k_gemm_bf16(bh, H, 1, bw, 1, H, dz, H, 4*R, H, H, nullptr,
                                0, g_dev.stream) == 0 &&
              fteik_k_inv_act(dz, R, H, db, dzp, o2, g_dev.stream) == 0 &&
              d2h(fused, o1, nh*sizeof(unsigned short)) &&
              d2h(unfused, o2, nh*sizeof(unsigned short));
    dfree(dh); dfree(dw); dfree(db); dfree(dz); dfree(dzp); dfree(bh);
    dfree(bw); dfree(o1); dfree(o2);
    if (!ok) *ierr = 1;
}

} // extern "C"
