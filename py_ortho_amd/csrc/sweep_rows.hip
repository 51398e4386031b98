// The reference's unfused plane sweep on gfx950: _dense.ncc_warp for every source followed by
// _dense.combine (orthomosaic/_dense.pyx:13-144), the path densify._sweep takes for more than 16
// sources (densify.py:343-353).  Its window sums are running sums along whole rows and down whole
// columns in the reference's order, so every float32 result equals the reference's bit for bit for
// any input -- including images whose FP64 partial sums round (the fused kernel forms the sums over
// short runs, equal to these only when no partial sum rounds).  Any number of sources, any window
// radius and any top-k.
//
// Per pass of P hypotheses (the workspace holds their intermediate maps):
//   rows  - one thread per (hypothesis, source, row): walks the row left to right, warping pixel
//           x + r (FP64 projection and bilinear, as the fused kernel), keeping the last 2r + 2
//           samples in an LDS ring, and writes the four float32 horizontal sums of every pixel
//           (ncc_warp phases 1-2, _dense.pyx:36-85);
//   cols  - one thread per (hypothesis, source, column): walks the column top to bottom with the
//           four FP64 vertical running sums and writes the NCC (phase 3, :86-110);
//   merge - one thread per pixel: per hypothesis in order, the mean of the top_k source scores
//           (descending insertion order) and the streaming winner update (combine, :113-144).
// Sources go to the rows / cols kernels in groups of 16 as kernel arguments.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "om_common.h"

namespace om {
namespace rows {

constexpr int kGroup = 16;          // sources per rows / cols launch (kernel-argument table)
constexpr int kMaxSrc = 256;        // sources of one sweep
constexpr int kWave = 64;

struct Src {
    const float *img;
    int h, w;
    double g[18];
};

struct Pass {
    const float *ref, *mu, *sd, *ray;
    int H, W, r;
    int mode;
    const float *ha, *hb;
    int hoff, i0, ia, np;          // hypotheses [ia, ia + np) of this pass; tables indexed from i0
    int pinhole;
    int s0, ns, n;                 // sources [s0, s0 + ns) of this launch, n in all
    float *hs;                     // (np, n, 4, H, W)
    float *ncc;                    // (np, n, H, W)
    Src src[kGroup];
};

__device__ __forceinline__ float hyp_value(const Pass &P, int i, int64_t p)
{
    switch (P.mode) {
    case OM_HYP_LINEAR: {
        const float t = (float)i * P.hb[p];  // _fast.pyx:638-640, no contraction
        return P.ha[p] + t;
    }
    case OM_HYP_REFINE: {
        float v = (float)(i - P.hoff) * P.hb[p];  // _fast.pyx:654-658
        v = P.ha[p] + v;
        return (v >= 1e-6f || v != v) ? v : 1e-6f;
    }
    case OM_HYP_CONST:
        return P.ha[i - P.i0];
    default:
        return P.ha[(int64_t)(i - P.i0) * ((int64_t)P.H * P.W) + p];
    }
}

// the warped sample of reference pixel p for source S at inverse depth inv (NaN = invalid),
// _dense.pyx:37-65 (_sweep.h:49-80)
__device__ __forceinline__ float warp_px(const Pass &P, const Src &S, int64_t p, float inv)
{
    const int64_t HW = (int64_t)P.H * P.W;
    const double d = 1.0 / (double)inv;
    const double rx = P.ray[p], ry = P.ray[HW + p], rz = P.ray[2 * HW + p];
    const double *g = S.g;
    const double X = g[9] + d * ((g[0] * rx + g[1] * ry) + g[2] * rz);
    const double Y = g[10] + d * ((g[3] * rx + g[4] * ry) + g[5] * rz);
    const double Z = g[11] + d * ((g[6] * rx + g[7] * ry) + g[8] * rz);
    if (Z <= 1e-6)
        return NAN;
    double nx, ny, u, v;
    div2_shared(X, Y, Z, nx, ny);
    if (P.pinhole) {
        u = g[12] * nx + g[16];  // f * (1 + r2 * 0) == f for finite r2
        v = g[12] * ny + g[17];
    } else {
        const double rr = nx * nx + ny * ny;
        const double dd = g[12] * (1.0 + rr * (g[13] + rr * (g[14] + rr * g[15])));
        u = dd * nx + g[16];
        v = dd * ny + g[17];
    }
    const double wl = (double)S.w - 1.001, hl = (double)S.h - 1.001;
    if (!(u >= 0.0 && v >= 0.0 && u <= wl && v <= hl))
        return NAN;
    const double val = bil_f64(S.img, S.w, u, v);
    return isfinite(val) ? (float)val : NAN;
}

// one thread per (hypothesis, source of the group, row): horizontal running sums (_dense.pyx:67-85)
__global__ void __launch_bounds__(kWave) rows_kernel(const Pass P)
{
    extern __shared__ float ring[];      // [2r + 2][64] pairs (J, I)
    const int R2 = 2 * P.r + 2;
    const int64_t t = (int64_t)blockIdx.x * kWave + threadIdx.x;
    const int64_t total = (int64_t)P.np * P.ns * P.H;
    const bool live = t < total;
    const int y = live ? (int)(t % P.H) : 0;
    const int sl = live ? (int)((t / P.H) % P.ns) : 0;
    const int q = live ? (int)(t / ((int64_t)P.H * P.ns)) : 0;
    if (!live)
        return;   // no barriers below: a lane's ring slots are its own
    const int i = P.ia + q;
    const Src &S = P.src[sl];
    const int W = P.W, r = P.r;
    const int64_t HW = (int64_t)P.H * W, row = (int64_t)y * W;
    float *rj = ring, *ri = ring + R2 * kWave;
    const int lane = threadIdx.x;
    double sv = 0.0, sj = 0.0, sjj = 0.0, sij = 0.0;
    auto sample = [&](int x, float &J, float &I) {
        const int64_t p = row + x;
        J = warp_px(P, S, p, hyp_value(P, i, p));
        I = P.ref[p];
    };
    const int lim = r < W ? r : W;
    for (int x = 0; x < lim; ++x) {
        float J, I;
        sample(x, J, I);
        const bool ok = J == J;
        const float jv = ok ? J : 0.f;
        sv = sv + (double)(ok ? 1.0f : 0.0f);
        sj = sj + (double)jv;
        sjj = sjj + (double)(jv * jv);
        sij = sij + (double)(jv * I);
        rj[(x % R2) * kWave + lane] = J;
        ri[(x % R2) * kWave + lane] = I;
    }
    float *out = P.hs + ((int64_t)q * P.n + (P.s0 + sl)) * 4 * HW + row;
    for (int x = 0; x < W; ++x) {
        const int xa = x + r;
        if (xa < W) {
            float J, I;
            sample(xa, J, I);
            const bool ok = J == J;
            const float jv = ok ? J : 0.f;
            sv = sv + (double)(ok ? 1.0f : 0.0f);
            sj = sj + (double)jv;
            sjj = sjj + (double)(jv * jv);
            sij = sij + (double)(jv * I);
            rj[(xa % R2) * kWave + lane] = J;
            ri[(xa % R2) * kWave + lane] = I;
        }
        const int xo = x - r - 1;
        if (xo >= 0) {
            const float J = rj[(xo % R2) * kWave + lane], I = ri[(xo % R2) * kWave + lane];
            const bool ok = J == J;
            const float jv = ok ? J : 0.f;
            sv = sv - (double)(ok ? 1.0f : 0.0f);
            sj = sj - (double)jv;
            sjj = sjj - (double)(jv * jv);
            sij = sij - (double)(jv * I);
        }
        out[x] = (float)sv;
        out[HW + x] = (float)sj;
        out[2 * HW + x] = (float)sjj;
        out[3 * HW + x] = (float)sij;
    }
}

// one thread per (hypothesis, source of the group, column): vertical running sums and NCC
// (_dense.pyx:86-110)
__global__ void __launch_bounds__(256) cols_kernel(const Pass P)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)P.np * P.ns * P.W;
    if (t >= total)
        return;
    const int x = (int)(t % P.W);
    const int sl = (int)((t / P.W) % P.ns);
    const int q = (int)(t / ((int64_t)P.W * P.ns));
    const int H = P.H, W = P.W, r = P.r;
    const int64_t HW = (int64_t)H * W;
    const float *hs = P.hs + ((int64_t)q * P.n + (P.s0 + sl)) * 4 * HW + x;
    float *out = P.ncc + ((int64_t)q * P.n + (P.s0 + sl)) * HW + x;
    const int k = 2 * r + 1;
    const double kk = 1.0 / (double)(k * k);
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    const int lim = r < H ? r : H;
    for (int y = 0; y < lim; ++y) {
        const int64_t o = (int64_t)y * W;
        c0 = c0 + (double)hs[o];
        c1 = c1 + (double)hs[HW + o];
        c2 = c2 + (double)hs[2 * HW + o];
        c3 = c3 + (double)hs[3 * HW + o];
    }
    for (int y = 0; y < H; ++y) {
        if (y + r < H) {
            const int64_t o = (int64_t)(y + r) * W;
            c0 = c0 + (double)hs[o];
            c1 = c1 + (double)hs[HW + o];
            c2 = c2 + (double)hs[2 * HW + o];
            c3 = c3 + (double)hs[3 * HW + o];
        }
        if (y - r - 1 >= 0) {
            const int64_t o = (int64_t)(y - r - 1) * W;
            c0 = c0 - (double)hs[o];
            c1 = c1 - (double)hs[HW + o];
            c2 = c2 - (double)hs[2 * HW + o];
            c3 = c3 - (double)hs[3 * HW + o];
        }
        const int64_t p = (int64_t)y * W + x;
        float res;
        const double cov = c0 * kk;
        if (cov <= 0.9) {
            res = -1.0f;
        } else {
            const double num[3] = {c1 * kk, c2 * kk, c3 * kk};
            double mq[3];
            quotients_shared<3>(num, cov, mq);  // cov > 0.9: IEEE quotients
            const double mj = mq[0], m2 = mq[1], mij = mq[2];
            double sdj = m2 - mj * mj;
            sdj = sdj > 1e-6 ? sqrt(sdj) : 1e-3;
            const double ncc = (mij - (double)P.mu[p] * mj) / ((double)P.sd[p] * sdj);
            res = ncc > 1.0 ? 1.0f : (ncc < -1.0 ? -1.0f : (float)ncc);
        }
        out[(int64_t)y * W] = res;
    }
}

struct State {
    float *best, *prev, *spb, *snb;
    int *idx;
    const float *ncc;
    int64_t HW;
    int n, kk, ia, np;
};

// combine (_dense.pyx:124-144): the n source scores sorted descending by insertion, the mean of
// the first kk in that order, the streaming winner update; hypotheses of the pass in order
__global__ void __launch_bounds__(256) merge_kernel(const State St)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= St.HW)
        return;
    float best = St.best[p], prev = St.prev[p], spb = St.spb[p], snb = St.snb[p];
    int idx = St.idx[p];
    float buf[kMaxSrc];
    for (int q = 0; q < St.np; ++q) {
        const int i = St.ia + q;
        const float *s = St.ncc + (int64_t)q * St.n * St.HW + p;
        for (int a = 0; a < St.n; ++a)
            buf[a] = s[(int64_t)a * St.HW];
        for (int a = 1; a < St.n; ++a) {
            const float t = buf[a];
            int c = a - 1;
            while (c >= 0 && buf[c] < t) {
                buf[c + 1] = buf[c];
                c -= 1;
            }
            buf[c + 1] = t;
        }
        float m = 0.0f;
        for (int a = 0; a < St.kk; ++a)
            m = m + buf[a];
        m = m / (float)St.kk;
        if (idx == i - 1)
            snb = m;
        if (m > best) {
            best = m;
            spb = prev;
            snb = -2.0f;
            idx = i;
        }
        prev = m;
    }
    St.best[p] = best;
    St.prev[p] = prev;
    St.spb[p] = spb;
    St.snb[p] = snb;
    St.idx[p] = idx;
}

struct Plan {
    int n, kk, P;        // sources, top-k used, hypotheses per pass
    size_t bytes;        // workspace
};

static int plan(const om_sweep_desc *d, Plan &pl)
{
    if (!d)
        return fail(OM_EINVAL, "om_sweep_rows: null descriptor");
    if (d->n_src < 1 || d->n_src > kMaxSrc)
        return fail(OM_ENSRC, "om_sweep_rows takes 1..%d sources (got %d)", kMaxSrc, d->n_src);
    if (d->r < 0)
        return fail(OM_EINVAL, "negative window radius");
    if (d->H < 0 || d->W < 0 || d->nh < 0 || d->i0 < 0)
        return fail(OM_ESHAPE, "negative size");
    if (!tap_size_ok(d->H, d->W))
        return fail(OM_ESHAPE, "sweep: reference images of 2^30 pixels or more, or with a side of 2^24 or more, "
                               "are not supported");
    if (d->top_k < 1)
        return fail(OM_EINVAL, "top_k must be >= 1");
    if (d->hyp_mode < OM_HYP_MAPS || d->hyp_mode > OM_HYP_CONST)
        return fail(OM_EINVAL, "unknown hypothesis generator %d", d->hyp_mode);
    if (!d->src)
        return fail(OM_EINVAL, "null source list");
    for (int s = 0; s < d->n_src; ++s)
        if (!tap_size_ok(d->src[s].h, d->src[s].w))
            return fail(OM_ESHAPE, "source %d: images of 2^30 pixels or more, or with a side of 2^24 or more, are "
                                   "not supported", s);
    pl.n = d->n_src;
    pl.kk = d->top_k < d->n_src ? d->top_k : d->n_src;
    const int64_t HW = (int64_t)d->H * d->W;
    int P = d->chunks;
    if (P <= 0) {
        // enough (hypothesis, source, row) walks to fill the chip, within 4 GiB of intermediate maps
        const int64_t walks = (int64_t)std::min(pl.n, kGroup) * std::max(d->H, 1);
        P = (int)std::max<int64_t>(1, cdiv(256 * 1024, walks));
        const int64_t per = (int64_t)pl.n * HW * 5 * 4;
        if (per > 0)
            P = (int)std::min<int64_t>(P, std::max<int64_t>(1, (int64_t(4) << 30) / per));
    }
    P = std::max(1, std::min(P, std::max(d->nh, 1)));
    pl.P = P;
    pl.bytes = (size_t)P * pl.n * (size_t)HW * 5 * sizeof(float);
    return OM_OK;
}

}  // namespace rows
}  // namespace om

using namespace om;

extern "C" size_t om_sweep_rows_workspace_bytes(const om_sweep_desc *d)
{
    rows::Plan pl;
    if (rows::plan(d, pl) != OM_OK)
        return 0;
    return pl.bytes;
}

extern "C" int om_sweep_rows(const om_sweep_desc *d, void *stream)
{
    rows::Plan pl;
    int rc = rows::plan(d, pl);
    if (rc)
        return rc;
    if (d->H == 0 || d->W == 0 || d->nh == 0)
        return OM_OK;
    if (!d->workspace || d->workspace_bytes < pl.bytes)
        return fail(OM_EINVAL, "om_sweep_rows: workspace of %zu bytes needed", pl.bytes);
    hipStream_t st = as_stream(stream);
    const int64_t HW = (int64_t)d->H * d->W;
    rows::Pass P;
    memset(&P, 0, sizeof P);
    P.ref = d->ref;
    P.mu = d->mu;
    P.sd = d->sd;
    P.ray = d->ray;
    P.H = d->H;
    P.W = d->W;
    P.r = d->r;
    P.mode = d->hyp_mode;
    P.ha = d->hyp_a;
    P.hb = d->hyp_b;
    P.hoff = d->hyp_offset;
    P.i0 = d->i0;
    P.n = pl.n;
    P.hs = (float *)d->workspace;
    P.ncc = P.hs + (size_t)pl.P * pl.n * 4 * HW;
    P.pinhole = 1;
    for (int s = 0; s < pl.n; ++s)
        if (d->src[s].geo[13] != 0.0 || d->src[s].geo[14] != 0.0 || d->src[s].geo[15] != 0.0)
            P.pinhole = 0;
    rows::State S{d->best, d->prev, d->s_prev_best, d->s_next_best, d->idx, P.ncc, HW, pl.n, pl.kk, 0, 0};
    const size_t ring = (size_t)(2 * d->r + 2) * rows::kWave * 2 * sizeof(float);
    ProfScope ps("sweep_rows", st);
    for (int ia = d->i0; ia < d->i0 + d->nh; ia += pl.P) {
        P.ia = ia;
        P.np = std::min(pl.P, d->i0 + d->nh - ia);
        for (int s0 = 0; s0 < pl.n; s0 += rows::kGroup) {
            P.s0 = s0;
            P.ns = std::min(rows::kGroup, pl.n - s0);
            for (int s = 0; s < P.ns; ++s) {
                const om_src &o = d->src[s0 + s];
                P.src[s].img = o.img;
                P.src[s].h = o.h;
                P.src[s].w = o.w;
                memcpy(P.src[s].g, o.geo, sizeof(double) * 18);
            }
            const int64_t nr = (int64_t)P.np * P.ns * d->H, nc = (int64_t)P.np * P.ns * d->W;
            hipLaunchKernelGGL(rows::rows_kernel, dim3((unsigned)cdiv(nr, rows::kWave)), dim3(rows::kWave), ring, st,
                               P);
            if ((rc = launch_check("sweep rows_kernel")))
                return rc;
            hipLaunchKernelGGL(rows::cols_kernel, dim3((unsigned)cdiv(nc, 256)), dim3(256), 0, st, P);
            if ((rc = launch_check("sweep cols_kernel")))
                return rc;
        }
        S.ia = ia;
        S.np = P.np;
        hipLaunchKernelGGL(rows::merge_kernel, dim3((unsigned)cdiv(HW, 256)), dim3(256), 0, st, S);
        if ((rc = launch_check("sweep merge_kernel")))
            return rc;
    }
    return OM_OK;
}
