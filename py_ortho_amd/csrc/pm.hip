// Red-black slanted-plane PatchMatch repair for gfx950.
//
// Reference semantics: _dense.pm_refine and its helpers (orthomosaic/_dense.pyx:405-666):
// PMCtx :408, _proj :429-445, _bil :448-458, _pm_cost :461-556, _rnd :559-563, the pass loop
// :606-664.
//
// Parallel form of a sequential algorithm.  Within one colour pass every pixel reads only pixels
// of the other colour (all propagation offsets are odd, _dense.pyx:583-586) and its own state, so
// the pixels of a pass are independent except for the single xorshift32 stream, which the
// reference consumes in row-major order: 8 draws per active pixel whose depth is finite after
// propagation, nothing for the others (:647-655).  A pass therefore runs as
//   A  pm_propagate  one thread per pixel of the colour: current cost + 8 propagated planes,
//   B  pm_rank_rows / pm_scan_rows  each drawing pixel's rank in row-major order, the pass total,
//      and the stream state after the pass,
//   C  pm_random     the stream jumped ahead to the pixel's offset (8 * rank) with precomputed
//                    GF(2) powers of the xorshift32 step matrix, then the 2 random tries.
// The draws each pixel sees are exactly the reference's, so the result equals the sequential one.
//
// Arithmetic.  Costs are evaluated in FP64 with the reference's operation order (the library is
// compiled with -ffp-contract=off), accumulating the window sums per source in sample order, so
// every accept/reject decision is the reference's.  One documented difference: when a cost
// evaluation returns early (plane facing away from the ray, or d <= 0) the reference leaves its
// photometric score variable holding the previous evaluation's value, possibly another pixel's;
// here it holds -1 at the start of each pixel.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include <algorithm>

#include "om_common.h"

namespace om {

// ---------------------------------------------------------------- xorshift32 jump-ahead
struct XsJump {
    uint32_t m[64][32];  // m[j][c] = M^(2^j) e_c, M = the xorshift32 step (13, 17, 5) over GF(2)
};

constexpr uint32_t xs_step(uint32_t s)
{
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

constexpr uint32_t xs_apply_cols(const uint32_t (&cols)[32], uint32_t s)
{
    uint32_t v = 0;
    for (int c = 0; c < 32; ++c)
        if ((s >> c) & 1u)
            v ^= cols[c];
    return v;
}

constexpr XsJump make_xs_jump()
{
    XsJump t{};
    for (int c = 0; c < 32; ++c)
        t.m[0][c] = xs_step(1u << c);
    for (int j = 1; j < 64; ++j)
        for (int c = 0; c < 32; ++c)
            t.m[j][c] = xs_apply_cols(t.m[j - 1], t.m[j - 1][c]);
    return t;
}

__constant__ XsJump kXsJump = make_xs_jump();

__device__ inline uint32_t xs_jump(uint32_t s, uint64_t n)
{
    for (int j = 0; j < 64 && n; ++j, n >>= 1) {
        if (n & 1u) {
            uint32_t v = 0;
#pragma unroll
            for (int c = 0; c < 32; ++c)
                v ^= kXsJump.m[j][c] & (0u - ((s >> c) & 1u));
            s = v;
        }
    }
    return s;
}

__device__ inline double xs_rnd(uint32_t &st)
{
    st ^= st << 13;
    st ^= st >> 17;
    st ^= st << 5;
    return (double)(st & 0xFFFFFFu) / 16777216.0;
}

// ---------------------------------------------------------------- kernel arguments
struct PMK {
    const float *ref, *ray;
    const uint8_t *active;
    float *depth, *normal, *score;
    int H, W, S, r, step, kk, pinhole, seed;
    double cr0, cr1, cr2, geo_w, geo_tol, dmin, dmax;
    double *best, *sc;         // (H, W) per-pixel cost and photometric score of the pass
    int *rank;                 // (H, W) draw rank of the pixel within its row, -1 = no draws
    int *row_cnt, *row_off;    // (H,)
    uint32_t *st;              // 2 stream states: before / after the current pass (by parity)
    unsigned long long *changed;
    const struct PMCam *cam;   // (S,) per-source tables in device memory
    const float *depth_from, *score_from;  // batched runs: start maps copied in first (or NULL)
    // cost memo of the single-view path: per pixel the (cost, photometric score) of its 8 neighbour
    // candidates and its own plane, whether its plane changed at its last visit, and the list of
    // (pixel, slot) evaluations a pass needs
    double *cc;
    uint8_t *chg;
    int *items, *nitems;
    unsigned long long *nevals;  // cost evaluations the memo path performed (all passes of a call)
    // first-iteration item order: items bucketed by the tile of source 0 their plane's centre
    // falls in (tiles of 2^sort_shift pixels, sort_tx x sort_ty of them), or by their own tile of
    // the reference image (sort_rx x sort_ry tiles of 2^sort_rshift) when it falls outside or the
    // item has no plane
    int *sorted, *bins;
    int sort_tx, sort_ty, sort_shift, sort_rx, sort_ry, sort_rshift;
    const int *sort_use;  // set for an evaluation after a sort: nonzero = read the sorted list
    // random refinement with the geometric bound: the margin by which a photometric score can
    // exceed 1 (+inf: no bound), the largest finite |value| of the reference and of the sources
    // (float bits), each pixel's stream state before its current try, and the list of tries to
    // evaluate with its two counters (one per try)
    double *eps;
    uint32_t *absmax;
    uint32_t *pst;
    int *ritems, *nritems;
};

// One source view as the cost kernels read it (device memory, uniform loads): the reference's
// par row (R 9, C 3, f, k1, k2, k3, cx, cy), image, optional depth map, size.
struct PMCam {
    double par[18];
    const float *img;
    const float *dep;
    int iw, ih;
    int pad[2];
    double wl, hl;  // iw - 1.001, ih - 1.001: the bilinear's bounds (_dense.pyx:451)
    // precision variant only: R (9) and Cr - C (3) rounded to float, f, cx, cy, bounds, k1..k3
    float rf[9], bf[3], ff, cxf, cyf, wlf, hlf;
    float k1f, k2f, k3f;
    int pad2;
};

// Host-side staging of the source tables, passed by value to pm_setup_kernel.
struct PMCamArgs {
    PMCam cam[16];
    int S;
};

// world point -> source pixel (_dense.pyx:429-445)
__device__ __forceinline__ bool pm_proj(const PMK &P, int s, double X, double Y, double Z, double &u, double &v,
                                        double &z)
{
    const OM_CONST_AS double *p = as_const(P.cam)[s].par;
    const double dx = X - p[9], dy = Y - p[10], dz = Z - p[11];
    const double xc = p[0] * dx + p[1] * dy + p[2] * dz;
    const double yc = p[3] * dx + p[4] * dy + p[5] * dz;
    const double zc = p[6] * dx + p[7] * dy + p[8] * dz;
    if (zc <= 1e-6)
        return false;
    double nx, ny;
    div2_shared(xc, yc, zc, nx, ny);
    if (P.pinhole) {
        u = p[12] * nx + p[16];  // f * (1 + r2 * 0) == f for finite r2
        v = p[12] * ny + p[17];
    } else {
        const double q = nx * nx + ny * ny;
        const double dd = p[12] * (1.0 + q * (p[13] + q * (p[14] + q * p[15])));
        u = dd * nx + p[16];
        v = dd * ny + p[17];
    }
    z = zc;
    return true;
}

// bilinear sample, NaN outside [0, w-1.001] x [0, h-1.001] (_dense.pyx:448-458)
__device__ __forceinline__ double pm_bil(const float *im, int w, double wl, double hl, double u, double v)
{
    if (!(u >= 0.0 && v >= 0.0 && u <= wl && v <= hl))
        return NAN;
    return bil_f64(im, w, u, v);
}

// waves per SIMD the PatchMatch cost kernels are compiled for (0: the compiler's choice).  4 waves
// (<= 128 VGPRs) hide the gather latency better than the 3 that 137 VGPRs allow: propagation of a
// 2000x1500 view with 8 sources 16.7 -> 15.2 ms per launch on MI355X (5 waves need scratch spills
// and are slower).
#ifndef OM_PM_WAVES
#define OM_PM_WAVES 4
#endif
#if OM_PM_WAVES > 0
#define OM_PM_ATTR __attribute__((amdgpu_waves_per_eu(OM_PM_WAVES, OM_PM_WAVES)))
#else
#define OM_PM_ATTR
#endif
// The reduced-precision cost kernels (FP32 arithmetic, fewer live registers) take 5 waves: at most
// 96 VGPRs, 20-28 bytes of spills; C4's FP16 propagation 390.5 -> 365.8 ms a step, same results.
#ifndef OM_PM_LP_WAVES
#define OM_PM_LP_WAVES 5
#endif
// (The FP64 evaluation at 5 waves, 88 bytes of spills: C2 194.4 -> 195.7 ms a step, the 48-view
// block's propagation 2082 -> 2089 ms; it stays at OM_PM_WAVES.)
#ifndef OM_PM_EVAL64_WAVES
#define OM_PM_EVAL64_WAVES OM_PM_WAVES
#endif
#define OM_PM_EVAL_WAVES (FMT ? OM_PM_LP_WAVES : OM_PM_EVAL64_WAVES)
#define OM_PM_EVAL_ATTR __attribute__((amdgpu_waves_per_eu(OM_PM_EVAL_WAVES, OM_PM_EVAL_WAVES)))

// waves per SIMD of the random-refinement kernels (each lane a different pixel, two cost
// evaluations with the pixel's state live around them)
#ifndef OM_PM_RANDOM_WAVES
#define OM_PM_RANDOM_WAVES OM_PM_WAVES
#endif
#if OM_PM_RANDOM_WAVES > 0
#define OM_PM_RANDOM_ATTR __attribute__((amdgpu_waves_per_eu(OM_PM_RANDOM_WAVES, OM_PM_RANDOM_WAVES)))
#else
#define OM_PM_RANDOM_ATTR
#endif

// Streams the batched memoised propagation spreads its views over (1: the launch stream alone).
// Views alternate between them, so one view's short need / decide kernels and its evaluation's
// tail overlap another view's evaluation.
#ifndef OM_PM_SIDE_STREAMS
#define OM_PM_SIDE_STREAMS 2
#endif

// waves per SIMD of the kernels that evaluate the random tries past the geometric bound (one lane
// per try, each a different pixel with a random plane: their gathers share nothing, and waves wait
// 62 % of their time).  5 waves (96 VGPRs with spills) measured slower than 4: the 48-view block's
// random refinement 526 -> 557 ms a step.
#ifndef OM_PM_TRY_WAVES
#define OM_PM_TRY_WAVES OM_PM_WAVES
#endif
#if OM_PM_TRY_WAVES > 0
#define OM_PM_TRY_ATTR __attribute__((amdgpu_waves_per_eu(OM_PM_TRY_WAVES, OM_PM_TRY_WAVES)))
#else
#define OM_PM_TRY_ATTR
#endif

#ifndef OM_PM_GROUP
#define OM_PM_GROUP 4
#endif
constexpr int kGroup = OM_PM_GROUP;  // sources accumulated together in the cost's sample loop
constexpr int kPMThreads = 128;
constexpr int kPassThreads = 1024;  // need / decide workgroups: one shared-counter update each
constexpr int kSortBins = 65536;  // most tiles of source 0 the first-iteration item order uses

// per-thread LDS slots for the S source scores of one cost evaluation
__device__ __forceinline__ double *pm_ncc_scratch()
{
    __shared__ double ncc_lds[kPMThreads * 16];
    return ncc_lds + threadIdx.x * 16;
}

// slanted-window cost (_dense.pyx:461-556); photo is written only on the full path
// element i of a source image stored in format FMT (precision variant), as float
template <int FMT>
__device__ __forceinline__ float pm_ld_img(const void *base, size_t i)
{
    if (FMT == OM_FMT_F32)
        return ((const float *)base)[i];
    const uint16_t v = ((const uint16_t *)base)[i];
    if (FMT == OM_FMT_BF16)
        return __uint_as_float((uint32_t)v << 16);
    return (float)__builtin_bit_cast(_Float16, v);
}

// The window samples of pixel (y, x) that lie inside the image (the reference's loops skip the
// others, _dense.pyx:480-491): rows a = -r + ka st for ka in [ka0, ka1], columns likewise, walked
// row by row as one sequence of ns samples.  Each step hands out the current sample's reference
// value and ray and loads the next sample's, so their latency overlaps the current sample's work;
// loads use 32-bit byte offsets from the image and ray-plane bases (images of fewer than 2^30
// pixels).
struct PMWindow {
    int ns, ka0, kb0, kb1, ka, kb;
    uint32_t off;
    float In, qxn, qyn, qzn;

    __device__ __forceinline__ uint32_t offset(const PMK &P, int y, int x) const
    {
        return (__umul24((uint32_t)(y - P.r + ka * P.step), (uint32_t)P.W) + (uint32_t)(x - P.r + kb * P.step)) << 2;
    }
    __device__ __forceinline__ void load(const PMK &P)
    {
        const size_t HW = (size_t)P.H * P.W;
        In = *(const float *)((const char *)P.ref + off);
        qxn = *(const float *)((const char *)P.ray + off);
        qyn = *(const float *)((const char *)(P.ray + HW) + off);
        qzn = *(const float *)((const char *)(P.ray + 2 * HW) + off);
    }
    __device__ __forceinline__ void start(const PMK &P, int y, int x)
    {
        const int r = P.r, st = P.step, nw = 2 * r / st;
        ka0 = y < r ? (r - y + st - 1) / st : 0;
        const int ka1 = min(nw, (P.H - 1 - y + r) / st);
        kb0 = x < r ? (r - x + st - 1) / st : 0;
        kb1 = min(nw, (P.W - 1 - x + r) / st);
        ns = (ka1 >= ka0 && kb1 >= kb0) ? (ka1 - ka0 + 1) * (kb1 - kb0 + 1) : 0;
    }
    // back to the first sample (each source group walks the window again)
    __device__ __forceinline__ void rewind(const PMK &P, int y, int x)
    {
        ka = ka0;
        kb = kb0;
        off = ns > 0 ? offset(P, y, x) : 0u;
        In = qxn = qyn = qzn = 0.f;
        if (ns > 0)
            load(P);
    }
    // sample j's values; then the load of sample j + 1 (sample j again after the last)
    __device__ __forceinline__ void step(const PMK &P, int y, int x, int j, float &I, float &qx, float &qy, float &qz)
    {
        I = In;
        qx = qxn;
        qy = qyn;
        qz = qzn;
        if (j + 1 < ns) {
            if (++kb > kb1) {
                kb = kb0;
                ++ka;
            }
            off = offset(P, y, x);
        }
        load(P);
    }
};

constexpr int kGroupLP = 8;  // sources accumulated together by the precision variant

// Precision variant of the photometric part of pm_cost: the plane (num = d n.ray_p) in FP64,
// rounded once; per sample the ray-plane intersection, the projection into each source and the
// bilinear tap of its FP16 / BF16 image in FP32; window sums in FP32; the variances and their
// 1e-8 guard in FP64.  -> the S NCC values in ncc_l.
template <int FMT>
__device__ __forceinline__ void pm_ncc_lp(const PMK &P, int y, int x, double num, double n0, double n1, double n2,
                                          double *ncc_l)
{
    const int S = P.S;
    const float numf = (float)num, nf0 = (float)n0, nf1 = (float)n1, nf2 = (float)n2;
    PMWindow win;
    win.start(P, y, x);
#pragma unroll 1
    for (int g0 = 0; g0 < S; g0 += kGroupLP) {
        float sI[kGroupLP], sJ[kGroupLP], sII[kGroupLP], sJJ[kGroupLP], sIJ[kGroupLP];
        int cnt[kGroupLP];
#pragma unroll
        for (int u = 0; u < kGroupLP; ++u) {
            sI[u] = sJ[u] = sII[u] = sJJ[u] = sIJ[u] = 0.f;
            cnt[u] = 0;
        }
        int nsamp = 0;
        win.rewind(P, y, x);
        for (int j = 0; j < win.ns; ++j) {
            float I, qx, qy, qz;
            win.step(P, y, x, j, I, qx, qy, qz);
            if (!isfinite(I))
                continue;
            const float den = fmaf(nf0, qx, fmaf(nf1, qy, nf2 * qz));
            if (den >= -1e-9f)
                continue;
            const float t = numf * __builtin_amdgcn_rcpf(den);
            nsamp += 1;
#pragma unroll
            for (int u = 0; u < kGroupLP; ++u) {
                const int s = g0 + u;
                if (s < S) {
                    const OM_CONST_AS PMCam &cs = as_const(P.cam)[s];
                    const float dx = fmaf(t, qx, cs.bf[0]), dy = fmaf(t, qy, cs.bf[1]), dz = fmaf(t, qz, cs.bf[2]);
                    const float zc = fmaf(cs.rf[6], dx, fmaf(cs.rf[7], dy, cs.rf[8] * dz));
                    if (zc > 1e-6f) {
                        const float xc = fmaf(cs.rf[0], dx, fmaf(cs.rf[1], dy, cs.rf[2] * dz));
                        const float yc = fmaf(cs.rf[3], dx, fmaf(cs.rf[4], dy, cs.rf[5] * dz));
                        const float iz = __builtin_amdgcn_rcpf(zc);
                        const float nx = xc * iz, ny = yc * iz;
                        float dd = cs.ff;   // radial model of _dense.pyx:438-442 in FP32
                        if (!P.pinhole) {
                            const float r2 = fmaf(nx, nx, ny * ny);
                            dd = cs.ff * fmaf(r2, fmaf(r2, fmaf(r2, cs.k3f, cs.k2f), cs.k1f), 1.0f);
                        }
                        const float pu = fmaf(dd, nx, cs.cxf), pv = fmaf(dd, ny, cs.cyf);
                        if (pu >= 0.f && pv >= 0.f && pu <= cs.wlf && pv <= cs.hlf) {
                            const int x0 = (int)pu, y0 = (int)pv;
                            const float fx = pu - (float)x0, fy = pv - (float)y0;
                            const size_t o = (size_t)y0 * cs.iw + x0;
                            const float a00 = pm_ld_img<FMT>(cs.img, o), a01 = pm_ld_img<FMT>(cs.img, o + 1);
                            const float a10 = pm_ld_img<FMT>(cs.img, o + cs.iw),
                                        a11 = pm_ld_img<FMT>(cs.img, o + cs.iw + 1);
                            const float t0 = fmaf(fx, a01 - a00, a00), t1 = fmaf(fx, a11 - a10, a10);
                            const float J = fmaf(fy, t1 - t0, t0);
                            if (isfinite(J)) {
                                sI[u] += I;
                                sJ[u] += J;
                                sII[u] = fmaf(I, I, sII[u]);
                                sJJ[u] = fmaf(J, J, sJJ[u]);
                                sIJ[u] = fmaf(I, J, sIJ[u]);
                                cnt[u] += 1;
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kGroupLP; ++u) {
            const int s = g0 + u;
            if (s < S) {
                double v;
                if (nsamp < 4 || cnt[u] < 0.7 * nsamp) {
                    v = -1.0;
                } else {
                    const float im = __builtin_amdgcn_rcpf((float)cnt[u]);
                    const float mi = sI[u] * im, mj = sJ[u] * im;
                    const double vi = (double)(sII[u] * im) - (double)mi * mi;
                    const double vj = (double)(sJJ[u] * im) - (double)mj * mj;
                    const float cv = fmaf(-mi, mj, sIJ[u] * im);
                    v = (vi < 1e-8 || vj < 1e-8) ? 0.0 : (double)(cv * __builtin_amdgcn_rsqf((float)(vi * vj)));
                }
                ncc_l[s] = v;
            }
        }
    }
}

// the geometric term of the cost (_dense.pyx:544-556): over the sources with a depth map, the
// mean relative depth disagreement at the plane's point on the pixel's ray, each clipped to 1
template <int SM>
__device__ __forceinline__ double pm_geo(const PMK &P, double rpx, double rpy, double rpz, double d)
{
    const int S = P.S;
    const double X = P.cr0 + d * rpx, Y = P.cr1 + d * rpy, Z = P.cr2 + d * rpz;
    double geo = 0.0;
    int ng = 0;
#pragma unroll
    for (int s = 0; s < SM; ++s) {
        const OM_CONST_AS PMCam &cs = as_const(P.cam)[s < S ? s : 0];
        if (s < S && cs.dep) {
            ng += 1;
            double e = 1.0, u, v, z;
            if (pm_proj(P, s, X, Y, Z, u, v, z)) {
                const int ui = (int)(u + 0.5), vi = (int)(v + 0.5);
                if (ui >= 0 && vi >= 0 && ui < cs.iw && vi < cs.ih) {
                    const double dj = cs.dep[(size_t)vi * cs.iw + ui];
                    if (isfinite(dj) && dj > 0) {
                        e = fabs(z - dj) / dj / P.geo_tol;
                        if (e > 1)
                            e = 1;
                    }
                }
            }
            geo += e;
        }
    }
    if (ng)
        geo /= ng;
    return geo;
}

template <int SM, int FMT>
__device__ double pm_cost(const PMK &P, int y, int x, double d, double n0, double n1, double n2, double &photo)
{
    const int H = P.H, W = P.W;
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    const double rpx = P.ray[p], rpy = P.ray[HW + p], rpz = P.ray[2 * HW + p];
    const double ndr = n0 * rpx + n1 * rpy + n2 * rpz;
    if (ndr >= -1e-6 || d <= 0)
        return 10.0;
    const double num = d * ndr;
    const int S = P.S;
    double *ncc_l = pm_ncc_scratch();  // this thread's S NCC values (LDS)
    if (FMT != 0) {
        pm_ncc_lp<FMT>(P, y, x, num, n0, n1, n2, ncc_l);
    } else {
    int nsamp = 0;
    PMWindow win;
    win.start(P, y, x);
    // Sources in groups of kGroup, the group loop kept rolled: per-sample geometry is recomputed
    // once per group, only kGroup sets of accumulators and one group's camera parameters are live.
    // Each source still sums its samples in window order, so every sum is the reference's.
#pragma unroll 1
    for (int g0 = 0; g0 < S; g0 += kGroup) {
        double sI[kGroup], sJ[kGroup], sII[kGroup], sJJ[kGroup], sIJ[kGroup];
        int cnt[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            sI[u] = sJ[u] = sII[u] = sJJ[u] = sIJ[u] = 0.0;
            cnt[u] = 0;
        }
        nsamp = 0;
        win.rewind(P, y, x);
        for (int j = 0; j < win.ns; ++j) {
            float If, qxf, qyf, qzf;
            win.step(P, y, x, j, If, qxf, qyf, qzf);
            const double I = If, qx = qxf, qy = qyf, qz = qzf;
            if (!isfinite(I))
                continue;
            const double den = n0 * qx + n1 * qy + n2 * qz;
            if (den >= -1e-9)
                continue;
            const double t = num / den;
            const double X = P.cr0 + t * qx, Y = P.cr1 + t * qy, Z = P.cr2 + t * qz;
            nsamp += 1;
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int s = g0 + u;
                if (s < S) {
                    double pu, pv, pz;
                    if (pm_proj(P, s, X, Y, Z, pu, pv, pz)) {
                        const OM_CONST_AS PMCam &cs = as_const(P.cam)[s];
                        const double J = pm_bil(cs.img, cs.iw, cs.wl, cs.hl, pu, pv);
                        if (isfinite(J)) {
                            sI[u] += I;
                            sJ[u] += J;
                            sII[u] += I * I;
                            sJJ[u] += J * J;
                            sIJ[u] += I * J;
                            cnt[u] += 1;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            const int s = g0 + u;
            if (s < S) {
                double v;
                if (nsamp < 4 || cnt[u] < 0.7 * nsamp) {
                    v = -1.0;
                } else {
                    const double m = (double)cnt[u];
                    const double mi = sI[u] / m, mj = sJ[u] / m;
                    const double vi = sII[u] / m - mi * mi;
                    const double vj = sJJ[u] / m - mj * mj;
                    const double cv = sIJ[u] / m - mi * mj;
                    v = (vi < 1e-8 || vj < 1e-8) ? 0.0 : cv / sqrt(vi * vj);
                }
                ncc_l[s] = v;
            }
        }
    }
    }
    double ncc[SM];
#pragma unroll
    for (int s = 0; s < SM; ++s)
        ncc[s] = s < S ? ncc_l[s] : 0.0;
    // descending insertion sort of the S values, the reference's exact procedure (:520-526)
#pragma unroll
    for (int a = 1; a < SM; ++a) {
        if (a < S) {
            const double t = ncc[a];
            bool moving = true;
#pragma unroll
            for (int c = a - 1; c >= 0; --c) {
                if (moving) {
                    if (ncc[c] < t) {
                        ncc[c + 1] = ncc[c];
                    } else {
                        ncc[c + 1] = t;
                        moving = false;
                    }
                }
            }
            if (moving)
                ncc[0] = t;
        }
    }
    double ph = 0.0;
#pragma unroll
    for (int a = 0; a < SM; ++a)
        if (a < P.kk)
            ph += ncc[a];
    ph /= P.kk;
    photo = ph;
    const double geo = P.geo_w > 0 ? pm_geo<SM>(P, rpx, rpy, rpz, d) : 0.0;
    return (1.0 - ph) + P.geo_w * geo;
}

// what changes from one colour pass to the next
struct PassK {
    int colour, parity;
    double dsc, nsc;  // 0.05 / (it+1), 0.3 / (it+1) (_dense.pyx:609-610)
    int first;        // first iteration: no memoised costs yet
};

__device__ __forceinline__ bool colour_pixel(const PMK &P, const PassK &Q, int64_t t, int &y, int &x)
{
    const int W2 = (P.W + 1) / 2;
    y = (int)(t / W2);
    const int j = (int)(t - (int64_t)y * W2);
    x = 2 * j + ((Q.colour + y) & 1);
    return y < P.H && x < P.W;
}

// Pixels of a colour pass in tiles of kNeedTY rows x kNeedTX colour columns, one wave each, tiles in
// row-major order: thread t of the need kernel (whose order the item list and so the evaluation
// kernel's lanes follow) and of the random refinement.  A wave's windows then cover a compact
// patch of the reference and of each source instead of a 128-pixel strip of one row (8 iterations
// on a 2000x1500 view with 8 sources: 229 -> 214 ms with 4 x 16 tiles; 8 x 8 is as fast, 2 x 32 and
// 16 x 4 slower).
#ifndef OM_NEED_TY
#define OM_NEED_TY 4
#endif
constexpr int kNeedTY = OM_NEED_TY, kNeedTX = 64 / OM_NEED_TY;

__host__ __device__ inline int64_t need_threads(int H, int W)
{
    const int64_t W2 = (W + 1) / 2;
    return cdiv(H, kNeedTY) * cdiv(W2, kNeedTX) * 64;
}

__device__ __forceinline__ bool tile_pixel(const PMK &P, const PassK &Q, int64_t t, int &y, int &x)
{
    const int W2 = (P.W + 1) / 2, ntx = (W2 + kNeedTX - 1) / kNeedTX;
    const int64_t tile = t / 64;
    const int k = (int)(t - tile * 64);
    const int ty = (int)(tile / ntx), tx = (int)(tile - (int64_t)ty * ntx);
    y = ty * kNeedTY + k / kNeedTX;
    const int j = tx * kNeedTX + k % kNeedTX;
    x = 2 * j + ((Q.colour + y) & 1);
    return y < P.H && j < W2 && x < P.W;
}

__device__ __forceinline__ void count_changes(const PMK &P, int n)
{
    if (n)
        atomicAdd(P.changed, (unsigned long long)n);
}

// A: current plane + propagation from the 8 neighbours (:614-645), pixel t of the colour
template <int SM, int FMT>
__device__ __forceinline__ void pm_propagate_px(const PMK &P, const PassK &Q, int64_t t)
{
    static const int oy[8] = {-1, 1, 0, 0, -5, 5, 0, 0};
    static const int ox[8] = {0, 0, -1, 1, 0, 0, -5, 5};
    int y, x;
    if (!colour_pixel(P, Q, t, y, x))
        return;
    const size_t HW = (size_t)P.H * P.W, p = (size_t)y * P.W + x;
    if (!P.active[p])
        return;
    double photo = -1.0;
    double n0 = P.normal[p], n1 = P.normal[HW + p], n2 = P.normal[2 * HW + p];
    double best, sc;
    const float d0 = P.depth[p];
    if (isfinite(d0)) {
        best = pm_cost<SM, FMT>(P, y, x, d0, n0, n1, n2, photo);
        sc = photo;
    } else {
        best = 1e9;
        sc = -1;
    }
    int changed = 0;
    for (int k = 0; k < 8; ++k) {
        const int yn = y + oy[k], xn = x + ox[k];
        if (yn < 0 || xn < 0 || yn >= P.H || xn >= P.W)
            continue;
        const size_t q = (size_t)yn * P.W + xn;
        const float dq = P.depth[q];
        if (!isfinite(dq))
            continue;
        const double b0 = P.normal[q], b1 = P.normal[HW + q], b2 = P.normal[2 * HW + q];
        const double dn = b0 * P.ray[q] + b1 * P.ray[HW + q] + b2 * P.ray[2 * HW + q];
        const double ndr = b0 * P.ray[p] + b1 * P.ray[HW + p] + b2 * P.ray[2 * HW + p];
        if (ndr >= -1e-6)
            continue;
        const double dp = (double)dq * dn / ndr;
        if (dp < P.dmin || dp > P.dmax)
            continue;
        const double cst = pm_cost<SM, FMT>(P, y, x, dp, b0, b1, b2, photo);
        if (cst < best) {
            best = cst;
            sc = photo;
            P.depth[p] = (float)dp;
            P.normal[p] = (float)b0;
            P.normal[HW + p] = (float)b1;
            P.normal[2 * HW + p] = (float)b2;
            changed += 1;
        }
    }
    P.best[p] = best;
    P.sc[p] = sc;
    count_changes(P, changed);
}

// B1: for row y, the rank of every drawing pixel (active, this colour, finite depth) in x order
// (a whole workgroup of 256 threads per row)
__device__ __forceinline__ void pm_rank_row(const PMK &P, const PassK &Q, int y)
{
    __shared__ int wsum[4];
    __shared__ int carry_s;
    const int W2 = (P.W + 1) / 2;
    const int x0 = (Q.colour + y) & 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0)
        carry_s = 0;
    __syncthreads();
    for (int j0 = 0; j0 < W2; j0 += 256) {
        const int j = j0 + threadIdx.x;
        const int x = 2 * j + x0;
        const size_t p = (size_t)y * P.W + x;
        const bool flag = j < W2 && x < P.W && P.active[p] && isfinite(P.depth[p]);
        const unsigned long long m = __ballot(flag);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0)
            wsum[wv] = __popcll(m);
        __syncthreads();
        int off = carry_s;
        for (int w = 0; w < wv; ++w)
            off += wsum[w];
        if (j < W2 && x < P.W)
            P.rank[p] = flag ? off + pre : -1;
        __syncthreads();
        if (threadIdx.x == 0)
            carry_s += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        P.row_cnt[y] = carry_s;
}

// B2: exclusive scan of the row counts; stream state after the pass = jump(state, 8 * total)
// (one workgroup of 256 threads per view)
__device__ __forceinline__ void pm_scan_view(const PMK &P, const PassK &Q)
{
    __shared__ long long part[256];
    const int H = P.H;
    const int per = (H + 255) / 256;
    const int a = threadIdx.x * per, b = min(H, a + per);
    long long s = 0;
    for (int y = a; y < b; ++y)
        s += P.row_cnt[y];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long acc = 0;
        for (int t = 0; t < 256; ++t) {
            const long long v = part[t];
            part[t] = acc;
            acc += v;
        }
        P.st[Q.parity ^ 1] = xs_jump(P.st[Q.parity], 8ull * (unsigned long long)acc);
    }
    __syncthreads();
    long long off = part[threadIdx.x];
    for (int y = a; y < b; ++y) {
        P.row_off[y] = (int)off;
        off += P.row_cnt[y];
    }
}

// a random try's plane from the pixel's plane and the stream state st, which advances past the
// try's 4 draws; false when the reference skips the try (:648-656)
__device__ __forceinline__ bool pm_try_plane(const PMK &P, const PassK &Q, size_t p, uint32_t &st, double &dp,
                                             double &b0, double &b1, double &b2)
{
    const size_t HW = (size_t)P.H * P.W;
    dp = (double)P.depth[p] * (1 + Q.dsc * (2 * xs_rnd(st) - 1));
    b0 = (double)P.normal[p] + Q.nsc * (2 * xs_rnd(st) - 1);
    b1 = (double)P.normal[HW + p] + Q.nsc * (2 * xs_rnd(st) - 1);
    b2 = (double)P.normal[2 * HW + p] + Q.nsc * (2 * xs_rnd(st) - 1);
    const double nn = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    if (nn < 1e-9 || dp < P.dmin || dp > P.dmax)
        return false;
    b0 /= nn;
    b1 /= nn;
    b2 /= nn;
    return true;
}

// both random tries of pixel p in the reference's order from stream state st (:646-664); best and
// sc in and out -> the accepted count
template <int SM, int FMT>
__device__ __forceinline__ int pm_random_seq(const PMK &P, const PassK &Q, size_t p, int y, int x, uint32_t st,
                                             double &best, double &sc)
{
    const size_t HW = (size_t)P.H * P.W;
    int changed = 0;
    double photo = -1.0;
    for (int tr = 0; tr < 2; ++tr) {
        double dp, b0, b1, b2;
        if (!pm_try_plane(P, Q, p, st, dp, b0, b1, b2))
            continue;
        const double cst = pm_cost<SM, FMT>(P, y, x, dp, b0, b1, b2, photo);
        if (cst < best) {
            best = cst;
            sc = photo;
            P.depth[p] = (float)dp;
            P.normal[p] = (float)b0;
            P.normal[HW + p] = (float)b1;
            P.normal[2 * HW + p] = (float)b2;
            changed += 1;
        }
    }
    return changed;
}

// C: the 2 random refinements with the pixel's own slice of the stream (:646-664); score write
template <int SM, int FMT>
__device__ __forceinline__ int pm_random_px(const PMK &P, const PassK &Q, int64_t t)
{
    int y, x;
    if (!tile_pixel(P, Q, t, y, x))
        return 0;
    const size_t p = (size_t)y * P.W + x;
    if (!P.active[p])
        return 0;
    double best = P.best[p], sc = P.sc[p];
    const int rk = P.rank[p];
    int changed = 0;
    if (rk >= 0)
        changed = pm_random_seq<SM, FMT>(P, Q, p, y, x,
                                         xs_jump(P.st[Q.parity], 8ull * (unsigned long long)(P.row_off[y] + rk)),
                                         best, sc);
    P.score[p] = (float)sc;
    if (P.chg && changed)
        P.chg[p] = 1;
    return changed;
}

// ---------------------------------------------------------------- random refinement, bounded
// A try's cost is (1 - photo) + geo_w geo (:556).  The photometric score photo is a mean of top-k
// NCC values cv / sqrt(vi vj) with both variances >= 1e-8 (:538-541), so it exceeds 1 only by the
// rounding of the window sums: for windows of at most m samples and finite values of magnitude
// <= M_I (reference) and M_J (sources), by less than 1.5 gamma_{m+3} 1e8 (M_I + M_J)^2 plus the
// rounding of the mean, with gamma_n = n u / (1 - n u) (Higham's bound for recursive sums and
// Cauchy-Schwarz on the exact moments).  A try whose geometric part alone, less twice that margin,
// is not below the pixel's current cost therefore loses the reference's comparison `cst < best`
// whatever its photometric score: it is rejected after the S projections of the geometric term.
// The other tries are listed and evaluated one lane per try, so the lanes stay busy however few
// remain.  Every decision, and so every output, is the reference's.
//
// A degenerate plane costs 10 (:474-475) and is accepted only by a pixel whose cost exceeds 10;
// such a pixel would then take the photometric score its previous evaluation left behind, which a
// rejected try no longer computes.  Costs above 10 need geo_w > 8 or wildly rounded NCC values, so
// those pixels, if any, run both tries in sequence as before; every other pixel rejects degenerate
// tries at once.  Used when geo_w > 0 (with no geometric term the bound cannot reject).

// largest finite |a[i]|, merged into *out as float bits (non-negative floats order like their bits)
__device__ __forceinline__ void pm_absmax_part(const float *a, int64_t n, uint32_t *out, int64_t i0, int64_t stride)
{
    float m = 0.f;
    for (int64_t i = i0; i < n; i += stride) {
        const float v = fabsf(a[i]);
        if (v <= 3.402823466e38f)  // finite: NaN and inf fail
            m = fmaxf(m, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.f)
        atomicMax(out, __float_as_uint(m));
}

// block b of nb: the reference image into absmax[0], every source image into absmax[1]
__device__ __forceinline__ void pm_absmax_view(const PMK &P, int b, int nb)
{
    const int64_t i0 = (int64_t)b * blockDim.x + threadIdx.x, stride = (int64_t)nb * blockDim.x;
    pm_absmax_part(P.ref, (int64_t)P.H * P.W, P.absmax, i0, stride);
    for (int s = 0; s < P.S; ++s) {
        const OM_CONST_AS PMCam &cs = as_const(P.cam)[s];
        pm_absmax_part(cs.img, (int64_t)cs.iw * cs.ih, P.absmax + 1, i0, stride);
    }
}

// the margin: twice the most a photometric score can exceed 1, after the mean and the cost's
// subtraction; +inf (no try rejected early) when it is too large to help
__device__ __forceinline__ void pm_eps_view(const PMK &P)
{
    const double u = 0x1p-53;
    const int nw = 2 * P.r / P.step + 1;        // window samples per side at most
    const double m = (double)nw * nw + 4.0;
    const double gam = m * u / (1.0 - m * u);
    const double s = (double)__uint_as_float(P.absmax[0]) + (double)__uint_as_float(P.absmax[1]);
    double d = 2.0 * gam * 1.0000001e8 * s * s;  // one NCC value over 1 (bound 1.5 gamma_{m+3} ...)
    d = d + (double)(P.kk + 16) * u * (1.0 + d);  // the top-k mean, 1 - photo and the final sum
    const double e = 2.0 * d + 1e-12;
    *P.eps = e < 0.5 ? e : INFINITY;
}

// the workgroup's entries of a list, reserved with one atomic; -> the first entry of this lane
// (lanes in thread order), n its entry count (every thread of the workgroup calls this)
__device__ __forceinline__ int pm_block_reserve(int *count, int n)
{
    __shared__ int wbase[kPassThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += v;
    }
    if (lane == 63)
        wbase[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            const int c = wbase[w];
            wbase[w] = acc;
            acc += c;
        }
        const int b0 = acc ? atomicAdd(count, acc) : 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            wbase[w] += b0;
    }
    __syncthreads();
    return wbase[wave] + incl - n;
}

// both tries of a pixel whose cost exceeds 10, in sequence; the pixel's second-try pass then only
// writes its score
template <int SM>
__device__ __forceinline__ int pm_random_fallback(const PMK &P, const PassK &Q, size_t p)
{
    const int y = (int)(p / P.W), x = (int)(p - (size_t)y * P.W);
    const uint32_t st = P.pst[p];
    double best = P.best[p], sc = P.sc[p];
    const int changed = pm_random_seq<SM, 0>(P, Q, p, y, x, st, best, sc);
    P.best[p] = best;
    P.sc[p] = sc;
    if (P.chg && changed)
        P.chg[p] = 1;
    P.rank[p] = -1;
    return changed;
}

// try tr of pixel t: its plane, the degenerate-plane test, the geometric term and the bound;
// the try is listed unless the bound rejects it.  Pixels without a listed second try get their
// score here; a pixel whose cost exceeds 10 is listed for both tries in sequence.
template <int SM>
__device__ __forceinline__ void pm_try_list_px(const PMK &P, const PassK &Q, int64_t t, int tr)
{
    int y, x;
    size_t p = 0;
    bool listed = false, fallback = false;
    if (tile_pixel(P, Q, t, y, x)) {
        p = (size_t)y * P.W + x;
        if (P.active[p]) {
            const int rk = P.rank[p];
            if (rk >= 0) {
                uint32_t st;
                if (tr == 0) {
                    st = xs_jump(P.st[Q.parity], 8ull * (unsigned long long)(P.row_off[y] + rk));
                } else {
                    st = P.pst[p];
                    for (int k = 0; k < 4; ++k)
                        (void)xs_rnd(st);
                }
                const double best = P.best[p];
                if (tr == 0 && best > 10.0) {
                    P.pst[p] = st;
                    fallback = true;
                } else {
                    P.pst[p] = st;
                    double dp, b0, b1, b2;
                    if (pm_try_plane(P, Q, p, st, dp, b0, b1, b2)) {
                        const size_t HW = (size_t)P.H * P.W;
                        const double rpx = P.ray[p], rpy = P.ray[HW + p], rpz = P.ray[2 * HW + p];
                        const double ndr = b0 * rpx + b1 * rpy + b2 * rpz;
                        if (!(ndr >= -1e-6 || dp <= 0)) {
                            const double g = P.geo_w * pm_geo<SM>(P, rpx, rpy, rpz, dp);
                            listed = !(g - *P.eps >= best);
                        }
                    }
                }
            }
            if (tr == 1 && !listed)
                P.score[p] = (float)P.sc[p];
        }
    }
    const int at = pm_block_reserve(P.nritems + tr, listed ? 1 : 0);
    if (listed)
        P.ritems[at] = (int)p;
    if (tr == 0) {  // pixels with costs above 10 (in practice none), from the list's far end
        const int af = pm_block_reserve(P.nritems + 2, fallback ? 1 : 0);
        if (fallback)
            P.ritems[(int64_t)P.H * ((P.W + 1) / 2) - 1 - af] = (int)p;
    }
}

// listed try i: the full cost and the reference's comparison; -> 1 when accepted
template <int SM>
__device__ __forceinline__ int pm_try_eval(const PMK &P, const PassK &Q, int64_t i, int tr)
{
    const size_t p = (size_t)P.ritems[i];
    const int y = (int)(p / P.W), x = (int)(p - (size_t)y * P.W);
    const size_t HW = (size_t)P.H * P.W;
    uint32_t st = P.pst[p];
    double dp, b0, b1, b2;
    pm_try_plane(P, Q, p, st, dp, b0, b1, b2);  // true: the try was listed
    double photo = -1.0;
    const double cst = pm_cost<SM, 0>(P, y, x, dp, b0, b1, b2, photo);
    int changed = 0;
    if (cst < P.best[p]) {
        P.best[p] = cst;
        P.sc[p] = photo;
        P.depth[p] = (float)dp;
        P.normal[p] = (float)b0;
        P.normal[HW + p] = (float)b1;
        P.normal[2 * HW + p] = (float)b2;
        if (P.chg)
            P.chg[p] = 1;
        changed = 1;
    }
    if (tr == 1)
        P.score[p] = (float)P.sc[p];
    return changed;
}

// ---------------------------------------------------------------- memoised propagation
// The cost of a plane at a pixel depends only on the pixel, the plane and inputs that a call never
// changes, so a candidate whose plane is the same as at the pixel's previous visit has the same
// cost.  A neighbour's candidate plane is a function of the neighbour's depth and normal only, and
// those change only at the neighbour's own visit; the pixel's own plane changes only at its own.
// Per colour pass: (1) list the (pixel, slot) costs that must be computed -- slot 8 is the pixel's
// own plane, slots 0..7 its neighbours' -- (2) compute them, one lane per item, (3) replay the
// reference's accept sequence (_dense.pyx:614-645) from the memo.  Bit-identical to evaluating
// every candidate, with dense lanes however few candidates changed.
__device__ const int kOffY[8] = {-1, 1, 0, 0, -5, 5, 0, 0};
__device__ const int kOffX[8] = {0, 0, -1, 1, 0, 0, -5, 5};

// the candidate plane that neighbour k proposes to pixel p (_dense.pyx:621-635); false if none
__device__ __forceinline__ bool pm_candidate(const PMK &P, int y, int x, int k, double &dp, double &b0, double &b1,
                                             double &b2)
{
    const int yn = y + kOffY[k], xn = x + kOffX[k];
    if (yn < 0 || xn < 0 || yn >= P.H || xn >= P.W)
        return false;
    const size_t HW = (size_t)P.H * P.W, p = (size_t)y * P.W + x, q = (size_t)yn * P.W + xn;
    const float dq = P.depth[q];
    if (!isfinite(dq))
        return false;
    b0 = P.normal[q];
    b1 = P.normal[HW + q];
    b2 = P.normal[2 * HW + q];
    const double dn = b0 * P.ray[q] + b1 * P.ray[HW + q] + b2 * P.ray[2 * HW + q];
    const double ndr = b0 * P.ray[p] + b1 * P.ray[HW + p] + b2 * P.ray[2 * HW + p];
    if (ndr >= -1e-6)
        return false;
    dp = (double)dq * dn / ndr;
    return !(dp < P.dmin || dp > P.dmax);
}

__global__ void pm_memo_reset_kernel(const PMK P, int64_t HW, int clear_chg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *P.nitems = 0;
        if (clear_chg)
            *P.nevals = 0ull;
    }
    if (clear_chg && i < HW)
        P.chg[i] = 0;
}

// (1) which costs the pass needs, appended to the item list a wave at a time (every lane of a
// wave belongs to one view)
__device__ __forceinline__ void pm_need_px(const PMK &P, const PassK &Q, int64_t t)
{
    int y, x;
    unsigned mask = 0;
    size_t p = 0;
    if (tile_pixel(P, Q, t, y, x)) {
        p = (size_t)y * P.W + x;
        if (P.active[p]) {
            if (Q.first || P.chg[p])
                mask |= 1u << 8;
            for (int k = 0; k < 8; ++k) {
                const int yn = y + kOffY[k], xn = x + kOffX[k];
                if (yn < 0 || xn < 0 || yn >= P.H || xn >= P.W)
                    continue;
                if (Q.first || P.chg[(size_t)yn * P.W + xn])
                    mask |= 1u << k;
            }
        }
    }
    // Pixel-major order within the wave: a pixel's items sit together, so the lanes of an
    // evaluation wave that share a pixel share its window's reference values and rays (8 iterations
    // on a 2000x1500 view with 8 sources: 212 -> 207 ms against slot-major order).  The waves of the workgroup take consecutive ranges of one
    // reservation (one pair of atomics per workgroup: the shared counters are the contended part of
    // this kernel).
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int wbase[kPassThreads / 64];
    {
        int total = 0;
        for (int slot = 0; slot < 9; ++slot)
            total += __popcll(__ballot((mask >> slot) & 1u));
        if (lane == 0)
            wbase[wave] = total;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                const int n = wbase[w];
                wbase[w] = acc;
                acc += n;
            }
            const int b0 = acc ? atomicAdd(P.nitems, acc) : 0;
            if (acc)
                atomicAdd(P.nevals, (unsigned long long)acc);
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
                wbase[w] += b0;
        }
        __syncthreads();
    }
    // this lane's first slot: the wave's base plus the exclusive prefix of the lanes' item counts
    const int mine = __popc(mask);
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += v;
    }
    int at = wbase[wave] + incl - mine;
    for (int slot = 0; slot < 9; ++slot)
        if ((mask >> slot) & 1u)
            P.items[at++] = (int)(p * 16 + slot);
}

__global__ void __launch_bounds__(kPassThreads) pm_need_kernel(const PMK P, const PassK Q)
{
    pm_need_px(P, Q, (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
}

// First iteration: the planes are still the random start, so a wave of neighbouring pixels samples
// the sources far apart.  The items are put in the order of the tile of source 0 that each plane's
// centre projects to (a counting sort; the order inside a bin is whatever the atomics give).
// Each item writes its own memo slot, so the order changes no result.
__device__ __forceinline__ int pm_item_bin(const PMK &P, int it, bool *has_plane = nullptr)
{
    const size_t p = (size_t)(it >> 4);
    const int slot = it & 15;
    const int y = (int)(p / P.W), x = (int)(p - (size_t)y * P.W);
    const size_t HW = (size_t)P.H * P.W;
    const int n0 = P.sort_tx * P.sort_ty;
    const int own = n0 + (y >> P.sort_rshift) * P.sort_rx + (x >> P.sort_rshift);   // reference tile
    double dp;
    if (has_plane)
        *has_plane = false;
    if (slot == 8) {
        const float d0 = P.depth[p];
        if (!isfinite(d0))
            return own;
        dp = d0;
    } else {
        double b0, b1, b2;
        if (!pm_candidate(P, y, x, slot, dp, b0, b1, b2))
            return own;
    }
    if (has_plane)
        *has_plane = true;
    const double X = P.cr0 + dp * P.ray[p], Y = P.cr1 + dp * P.ray[HW + p], Z = P.cr2 + dp * P.ray[2 * HW + p];
    double u, v, z;
    const double ts = (double)(1 << P.sort_shift);
    if (pm_proj(P, 0, X, Y, Z, u, v, z) && u >= 0.0 && v >= 0.0 && u < ts * P.sort_tx && v < ts * P.sort_ty)
        return (int)(v / ts) * P.sort_tx + (int)(u / ts);
    return own;   // outside source 0
}

__global__ void pm_sort_zero_kernel(int *bins, int nb)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x)
        bins[i] = 0;
}

// Whether to sort: a PatchMatch started from a sweep already has coherent planes (neighbouring
// items share a bin and sorting gains nothing); the random start scatters them.  One workgroup
// looks at 64 runs of 64 consecutive items spread over the list, counting only items that have
// a plane: bins[nb + 1] = 1 when on average fewer than 4 of them share a bin within a run.
__global__ void __launch_bounds__(256) pm_sort_probe_kernel(const PMK P, int nb)
{
    __shared__ int s_distinct, s_seen;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t n = *P.nitems;
    if (threadIdx.x == 0) {
        s_distinct = 0;
        s_seen = 0;
    }
    __syncthreads();
    for (int run = wave; run < 64; run += 4) {
        const int64_t i = (n / 64 / 64) * 64 * run + lane;   // runs spread evenly over the list
        bool plane = false;
        const int bin = i < n ? pm_item_bin(P, P.items[i], &plane) : 0;
        const bool on = plane;
        unsigned long long todo = __ballot(on);
        const int seen = __popcll(todo);
        int distinct = 0;
        while (todo) {
            const int b = __shfl(bin, __ffsll((long long)todo) - 1, 64);
            todo &= ~__ballot(on && bin == b);
            ++distinct;
        }
        if (lane == 0) {
            atomicAdd(&s_distinct, distinct);
            atomicAdd(&s_seen, seen);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        P.bins[nb + 1] = s_seen > 0 && 4 * s_distinct >= s_seen;
}

__global__ void __launch_bounds__(256) pm_sort_hist_kernel(const PMK P, int nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P.bins[nb + 1] && i < *P.nitems)
        atomicAdd(&P.bins[pm_item_bin(P, P.items[i])], 1);
}

// exclusive scan of the bin counts in place (one workgroup)
__global__ void __launch_bounds__(1024) pm_sort_scan_kernel(int *bins, int nb)
{
    __shared__ int wtot[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (!bins[nb + 1])
        return;
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + tid;
        const int x = i < nb ? bins[i] : 0;
        int y = x;
        for (int o = 1; o < 64; o <<= 1) {
            const int z = __shfl_up(y, o, 64);
            if (lane >= o)
                y += z;
        }
        if (lane == 63)
            wtot[wave] = y;
        __syncthreads();
        int before = 0, all = 0;
        for (int w = 0; w < 16; ++w) {
            before += w < wave ? wtot[w] : 0;
            all += wtot[w];
        }
        if (i < nb)
            bins[i] = carry + before + y - x;
        carry += all;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) pm_sort_scatter_kernel(const PMK P, int nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P.bins[nb + 1] && i < *P.nitems) {
        const int it = P.items[i];
        P.sorted[atomicAdd(&P.bins[pm_item_bin(P, it)], 1)] = it;
    }
}

// enqueue the sort; returns the view whose evaluation reads the sorted list when the sort ran
static PMK pm_sort_items(const PMK &P, int64_t ncol, hipStream_t st)
{
    const int nb = P.sort_tx * P.sort_ty + P.sort_rx * P.sort_ry;
    const unsigned g = (unsigned)cdiv(9 * ncol, 256);
    hipLaunchKernelGGL(pm_sort_zero_kernel, dim3((unsigned)cdiv(nb + 2, 256)), dim3(256), 0, st, P.bins, nb + 2);
    hipLaunchKernelGGL(pm_sort_probe_kernel, dim3(1), dim3(256), 0, st, P, nb);
    hipLaunchKernelGGL(pm_sort_hist_kernel, dim3(g), dim3(256), 0, st, P, nb);
    hipLaunchKernelGGL(pm_sort_scan_kernel, dim3(1), dim3(1024), 0, st, P.bins, nb);
    hipLaunchKernelGGL(pm_sort_scatter_kernel, dim3(g), dim3(256), 0, st, P, nb);
    PMK Q = P;
    Q.sort_use = P.bins + nb + 1;
    return Q;
}

// (2) one cost per lane.  The grid covers the largest possible list, the item count is known only
// on the device: the blocks that have items are mapped so that each XCD (hardware dispatch is
// round-robin over the 8 XCDs) gets an equal, contiguous share -- equal work, and neighbouring
// items in one L2 -- and the surplus blocks of every XCD exit at once.
template <int SM, int FMT>
__device__ __forceinline__ void pm_eval_item(const PMK &P, const int *items, int64_t i);

template <int SM, int FMT>
__global__ void __launch_bounds__(kPMThreads) OM_PM_EVAL_ATTR pm_eval_kernel(const PMK P)
{
    const int64_t n = *P.nitems;
    const int64_t used = cdiv(n, (int64_t)blockDim.x);           // blocks with items
    const int64_t per = cdiv(used, (int64_t)kXCD);                // of them per XCD
    const int xcd = blockIdx.x % kXCD, j = blockIdx.x / kXCD;
    if (j >= per)
        return;
    const int64_t i = ((int64_t)xcd * per + j) * blockDim.x + threadIdx.x;
    if (i < n)
        pm_eval_item<SM, FMT>(P, (P.sort_use && *P.sort_use) ? P.sorted : P.items, i);
}

template <int SM, int FMT>
__device__ __forceinline__ void pm_eval_item(const PMK &P, const int *items, int64_t i)
{
    const int it = items[i];
    const size_t p = (size_t)(it >> 4);
    const int slot = it & 15;
    const int y = (int)(p / P.W), x = (int)(p - (size_t)y * P.W);
    const size_t HW = (size_t)P.H * P.W;
    // the plane of the slot (the pixel's own, or neighbour slot's candidate), then one cost call
    double dp = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    bool ok;
    if (slot == 8) {
        const float d0 = P.depth[p];
        ok = isfinite(d0);
        dp = d0;
        b0 = P.normal[p];
        b1 = P.normal[HW + p];
        b2 = P.normal[2 * HW + p];
    } else {
        ok = pm_candidate(P, y, x, slot, dp, b0, b1, b2);
    }
    double cost = slot == 8 ? 1e9 : NAN, photo = -1.0;  // no plane: the reference's 1e9 / -1, or skip
    if (ok)
        cost = pm_cost<SM, FMT>(P, y, x, dp, b0, b1, b2, photo);
    P.cc[(p * 9 + slot) * 2] = cost;
    P.cc[(p * 9 + slot) * 2 + 1] = photo;
}

// (3) the reference's accept sequence over the memoised costs
__device__ __forceinline__ int pm_decide_px(const PMK &P, const PassK &Q, int64_t t)
{
    int y, x;
    if (!colour_pixel(P, Q, t, y, x))
        return 0;
    const size_t HW = (size_t)P.H * P.W, p = (size_t)y * P.W + x;
    if (!P.active[p])
        return 0;
    const double *c = P.cc + p * 18;
    double best = c[16], sc = c[17];
    int changed = 0, last = -1;
    for (int k = 0; k < 8; ++k) {
        const int yn = y + kOffY[k], xn = x + kOffX[k];
        if (yn < 0 || xn < 0 || yn >= P.H || xn >= P.W)  // never listed, so never written
            continue;
        const double cst = c[2 * k];
        if (!(cst < best))  // NaN: no candidate from this neighbour
            continue;
        best = cst;
        sc = c[2 * k + 1];
        last = k;
        changed += 1;
    }
    // Each accepted candidate overwrites the plane and depends only on its neighbour, which this
    // pass does not change, so only the last accepted one is formed and stored.
    if (last >= 0) {
        double dp, b0, b1, b2;
        pm_candidate(P, y, x, last, dp, b0, b1, b2);
        P.depth[p] = (float)dp;
        P.normal[p] = (float)b0;
        P.normal[HW + p] = (float)b1;
        P.normal[2 * HW + p] = (float)b2;
    }
    P.best[p] = best;
    P.sc[p] = sc;
    P.chg[p] = changed ? 1 : 0;
    return changed;
}

// the workgroup's sum of v, in thread 0
__device__ __forceinline__ int block_total(int v)
{
    __shared__ int part[kPassThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    int s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            s += part[w];
    return s;
}

__global__ void __launch_bounds__(kPassThreads) pm_decide_kernel(const PMK P, const PassK Q)
{
    const int n = pm_decide_px(P, Q, (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    count_changes(P, block_total(n));  // one counter update per workgroup
}

// ---------------------------------------------------------------- single-view kernels
template <int SM, int FMT>
__global__ void __launch_bounds__(kPMThreads) OM_PM_ATTR pm_propagate_kernel(const PMK P, const PassK Q)
{
    pm_propagate_px<SM, FMT>(P, Q, (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) pm_rank_rows_kernel(const PMK P, const PassK Q) { pm_rank_row(P, Q, blockIdx.x); }

__global__ void __launch_bounds__(256) pm_scan_rows_kernel(const PMK P, const PassK Q) { pm_scan_view(P, Q); }

template <int SM, int FMT>
__global__ void __launch_bounds__(kPMThreads) OM_PM_RANDOM_ATTR pm_random_kernel(const PMK P, const PassK Q)
{
    const int n = pm_random_px<SM, FMT>(P, Q, (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    count_changes(P, block_total(n));
}

// bounded random refinement (single view): try tr's list, then its evaluation
template <int SM>
__global__ void __launch_bounds__(kPassThreads) pm_try_list_kernel(const PMK P, const PassK Q, int tr)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    if (b == 0 && threadIdx.x == 0) {  // the lists of the other try: their evaluation has run
        P.nritems[1 - tr] = 0;
        if (tr == 1)
            P.nritems[2] = 0;
    }
    pm_try_list_px<SM>(P, Q, (int64_t)b * blockDim.x + threadIdx.x, tr);
}

// the pixels listed for both tries in sequence (costs above 10; in practice none)
template <int SM>
__global__ void __launch_bounds__(64) pm_try_fallback_kernel(const PMK P, const PassK Q)
{
    const int n = P.nritems[2];
    const int64_t ncol = (int64_t)P.H * ((P.W + 1) / 2);
    int c = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        c += pm_random_fallback<SM>(P, Q, (size_t)P.ritems[ncol - 1 - k]);
    count_changes(P, c);
}

// the listed tries spread evenly over the XCDs, each XCD a contiguous range of the list
template <int SM>
__global__ void __launch_bounds__(kPMThreads) OM_PM_TRY_ATTR pm_try_eval_kernel(const PMK P, const PassK Q, int tr)
{
    const int64_t n = P.nritems[tr];
    const int64_t used = cdiv(n, (int64_t)blockDim.x), per = cdiv(used, (int64_t)kXCD);
    const int xcd = blockIdx.x % kXCD, j = blockIdx.x / kXCD;
    if (j >= per)
        return;
    const int64_t i = ((int64_t)xcd * per + j) * blockDim.x + threadIdx.x;
    const int c = i < n ? pm_try_eval<SM>(P, Q, i, tr) : 0;
    count_changes(P, block_total(c));
}

__global__ void pm_bound_reset_kernel(const PMK P)
{
    if (threadIdx.x < 3)
        P.nritems[threadIdx.x] = 0;
    if (threadIdx.x < 2)
        P.absmax[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256) pm_absmax_kernel(const PMK P) { pm_absmax_view(P, blockIdx.x, gridDim.x); }

__global__ void pm_eps_kernel(const PMK P) { pm_eps_view(P); }

__global__ void pm_init_kernel(uint32_t *st, uint32_t seed, unsigned long long *changed, const PMCamArgs A,
                               PMCam *cam)
{
    st[0] = seed | 1u;
    *changed = 0ull;
    for (int s = 0; s < A.S; ++s)
        cam[s] = A.cam[s];
}

// ---------------------------------------------------------------- batched kernels (many views)
// blocks of all views in one grid: view v owns logical blocks [blk0[v], blk0[v+1])
__device__ __forceinline__ int view_of(const int *blk0, int nviews, int b)
{
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (as_const(blk0)[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int SM>
__global__ void __launch_bounds__(kPMThreads) OM_PM_ATTR pm_propagate_batch(const PMK *views, const int *blk0, int nviews,
                                                                 const PassK Q)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int v = view_of(blk0, nviews, b);
    pm_propagate_px<SM, 0>(uniform_ref(views + v), Q, (int64_t)(b - as_const(blk0)[v]) * blockDim.x + threadIdx.x);
}

template <int SM>
__global__ void __launch_bounds__(kPMThreads) OM_PM_RANDOM_ATTR pm_random_batch(const PMK *views, const int *blk0, int nviews,
                                                              const PassK Q)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int v = view_of(blk0, nviews, b);
    const PMK &P = uniform_ref(views + v);
    const int n = pm_random_px<SM, 0>(P, Q, (int64_t)(b - as_const(blk0)[v]) * blockDim.x + threadIdx.x);
    count_changes(P, block_total(n));
}

// bounded random refinement over the views of a source bucket: try tr's lists (blocks of
// kPassThreads threads, view v's from blk0[v]) ...
template <int SM>
__global__ void __launch_bounds__(kPassThreads) pm_try_list_batch(const PMK *views, const int *blk0, int nviews,
                                                                  const PassK Q, int tr)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int v = view_of(blk0, nviews, b);
    const PMK &P = uniform_ref(views + v);
    const int lb = b - as_const(blk0)[v];
    if (lb == 0 && threadIdx.x == 0) {  // the lists of the other try: their evaluation has run
        P.nritems[1 - tr] = 0;
        if (tr == 1)
            P.nritems[2] = 0;
    }
    pm_try_list_px<SM>(P, Q, (int64_t)lb * blockDim.x + threadIdx.x, tr);
}

// the pixels listed for both tries in sequence, view blockIdx.y (costs above 10; in practice none)
template <int SM>
__global__ void __launch_bounds__(64) pm_try_fallback_batch(const PMK *views, const PassK Q)
{
    const PMK &P = uniform_ref(views + blockIdx.y);
    const int n = P.nritems[2];
    const int64_t ncol = (int64_t)P.H * ((P.W + 1) / 2);
    int c = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        c += pm_random_fallback<SM>(P, Q, (size_t)P.ritems[ncol - 1 - k]);
    count_changes(P, c);
}

// ... and their evaluation: the views' listed tries in blocks of kPMThreads as one sequence (view
// order), each XCD a contiguous range of it; a block finds its view from the list counts
template <int SM>
__global__ void __launch_bounds__(kPMThreads) OM_PM_TRY_ATTR pm_try_eval_batch(const PMK *views, int nviews,
                                                                           const PassK Q, int tr)
{
    __shared__ int s_v, s_lb;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int total = 0;
        for (int v0 = 0; v0 < nviews; v0 += 64) {
            const int v = v0 + lane;
            int u = v < nviews ? (int)cdiv((int64_t)views[v].nritems[tr], (int64_t)kPMThreads) : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                u += __shfl_xor(u, o, 64);
            total += u;
        }
        const int per = (total + kXCD - 1) / kXCD;
        const int xcd = blockIdx.x % kXCD, j = blockIdx.x / kXCD;
        const int L = j < per ? xcd * per + j : total;  // this block's place in the sequence
        int fv = -1, fl = 0, acc = 0;
        for (int v0 = 0; v0 < nviews && fv < 0 && L < total; v0 += 64) {
            const int v = v0 + lane;
            const int u = v < nviews ? (int)cdiv((int64_t)views[v].nritems[tr], (int64_t)kPMThreads) : 0;
            int incl = u;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int w = __shfl_up(incl, o, 64);
                if (lane >= o)
                    incl += w;
            }
            const int start = acc + incl - u;
            const unsigned long long m = __ballot(u > 0 && L >= start && L < start + u);
            if (m) {
                const int l = __builtin_ctzll(m);
                fv = v0 + l;
                fl = L - __shfl(start, l, 64);
            }
            acc += __shfl(incl, 63, 64);
        }
        if (lane == 0) {
            s_v = fv;
            s_lb = fl;
        }
    }
    __syncthreads();
    const int v = __builtin_amdgcn_readfirstlane(s_v);
    if (v < 0)
        return;
    const PMK &P = uniform_ref(views + v);
    const int64_t i = (int64_t)s_lb * blockDim.x + threadIdx.x;
    const int c = i < (int64_t)P.nritems[tr] ? pm_try_eval<SM>(P, Q, i, tr) : 0;
    count_changes(P, block_total(c));
}

__global__ void __launch_bounds__(256) pm_absmax_batch(const PMK *views)
{
    pm_absmax_view(uniform_ref(views + blockIdx.y), blockIdx.x, gridDim.x);
}

__global__ void pm_eps_batch(const PMK *views, int nviews)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nviews)
        pm_eps_view(views[v]);
}

// item counters of every view; with clear_chg also the evaluation counts and the changed flags of
// every pixel (px0: first pixel of each view)
__global__ void pm_memo_reset_batch(const PMK *views, const int64_t *px0, int nviews, int clear_chg)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nviews) {
        *views[t].nitems = 0;
        if (clear_chg)
            *views[t].nevals = 0ull;
    }
    if (!clear_chg || t >= px0[nviews])
        return;
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (px0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    views[lo].chg[t - px0[lo]] = 0;
}

__global__ void __launch_bounds__(256) pm_rank_rows_batch(const PMK *views, const int *row0, int nviews,
                                                          const PassK Q)
{
    const int v = view_of(row0, nviews, blockIdx.x);
    pm_rank_row(uniform_ref(views + v), Q, blockIdx.x - as_const(row0)[v]);
}

__global__ void __launch_bounds__(256) pm_scan_rows_batch(const PMK *views, const PassK Q)
{
    pm_scan_view(uniform_ref(views + blockIdx.x), Q);
}

__global__ void pm_init_batch(const PMK *views, int nviews)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nviews) {
        views[v].st[0] = (uint32_t)views[v].seed | 1u;
        *views[v].changed = 0ull;
        views[v].nritems[0] = views[v].nritems[1] = views[v].nritems[2] = 0;
        views[v].absmax[0] = views[v].absmax[1] = 0u;
    }
}

// workspace layout
struct PMWs {
    double *best, *sc;
    int *rank, *row_cnt, *row_off;
    uint32_t *st;  // 2 states
    unsigned long long *changed;
    PMCam *cam;    // 16 source tables
    int *sorted, *bins;  // first-iteration item order
    double *cc;    // memo (single view): (H, W, 9, 2)
    uint8_t *chg;  // (H, W)
    int *items, *nitems;
    unsigned long long *nevals;
    double *eps;   // bounded random refinement: margin, image maxima, stream states, try list
    uint32_t *absmax, *pst;
    int *ritems, *nritems;
};

// The memo path lists items as pixel * 16 + slot in an int, so it serves images below 2^27 pixels;
// larger ones take the direct per-pixel kernels.  OM_PM_MEMO_MAX_PIXELS (environment, read at each
// call) lowers that limit: the tests use it to run the direct path on small images.
static bool pm_memo_ok(int H, int W)
{
    int64_t lim = 1LL << 27;
    if (const char *e = getenv("OM_PM_MEMO_MAX_PIXELS")) {
        const long long v = atoll(e);
        if (v >= 0 && v < lim)
            lim = v;
    }
    return (int64_t)H * W < lim;
}

// tiles of source 0 for the first-iteration item order: 16 pixels, larger for images beyond
// kSortBins tiles
static void pm_sort_tiles(PMK &P, int iw0, int ih0)
{
    int sh = 4;
    while (cdiv(iw0, 1 << sh) * cdiv(ih0, 1 << sh) > kSortBins)
        ++sh;
    P.sort_shift = sh;
    P.sort_tx = (int)cdiv(iw0, 1 << sh);
    P.sort_ty = (int)cdiv(ih0, 1 << sh);
    int rs = 4;
    while (cdiv(P.W, 1 << rs) * cdiv(P.H, 1 << rs) > kSortBins)
        ++rs;
    P.sort_rshift = rs;
    P.sort_rx = (int)cdiv(P.W, 1 << rs);
    P.sort_ry = (int)cdiv(P.H, 1 << rs);
}

// The random refinement rejects tries by the geometric bound unless OM_PM_RANDOM_BOUND=0
// (environment, read at each call; the tests use it to run the sequential kernel on the same inputs).
static bool pm_bound_on()
{
    const char *e = getenv("OM_PM_RANDOM_BOUND");
    return !(e && e[0] == '0');
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static size_t pm_ws_layout(int H, int W, char *base, PMWs *w, bool memo = false)
{
    const size_t HW = (size_t)H * W;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o += align256(bytes);
        return p;
    };
    char *b = take(HW * 8), *s = take(HW * 8), *rk = take(HW * 4), *rc = take((size_t)H * 4),
         *ro = take((size_t)H * 4), *st = take(16), *ch = take(8), *cm = take(sizeof(PMCam) * 16);
    char *ep = take(8), *am = take(8), *ps = take(HW * 4), *ri = take((size_t)H * ((W + 1) / 2) * 4),
         *nr = take(16);
    char *cc = nullptr, *cg = nullptr, *it = nullptr, *ni = nullptr, *ne = nullptr, *so = nullptr, *bn = nullptr;
    if (memo) {
        const size_t ncol = (size_t)H * ((W + 1) / 2);
        cc = take(HW * 18 * sizeof(double));
        cg = take(HW);
        it = take(ncol * 9 * sizeof(int));
        ni = take(sizeof(int));
        ne = take(sizeof(unsigned long long));
        so = take(ncol * 9 * sizeof(int));
        bn = take((size_t)(2 * kSortBins + 2) * sizeof(int));
    }
    if (w) {
        w->sorted = (int *)so;
        w->bins = (int *)bn;
        w->cc = (double *)cc;
        w->chg = (uint8_t *)cg;
        w->items = (int *)it;
        w->nitems = (int *)ni;
        w->nevals = (unsigned long long *)ne;
        w->best = (double *)b;
        w->sc = (double *)s;
        w->rank = (int *)rk;
        w->row_cnt = (int *)rc;
        w->row_off = (int *)ro;
        w->st = (uint32_t *)st;
        w->changed = (unsigned long long *)ch;
        w->cam = (PMCam *)cm;
        w->eps = (double *)ep;
        w->absmax = (uint32_t *)am;
        w->pst = (uint32_t *)ps;
        w->ritems = (int *)ri;
        w->nritems = (int *)nr;
    }
    return o;
}

template <int SM, int FMT>
static int pm_run(PMK &P, const PMCamArgs &A, int iters, hipStream_t stream, uint32_t seed, const PMWs &w)
{
    const int64_t ncol = (int64_t)P.H * ((P.W + 1) / 2);
    const unsigned gA = (unsigned)cdiv(ncol, kPMThreads), gN = (unsigned)cdiv(ncol, kPassThreads);
    const unsigned gT = (unsigned)cdiv(need_threads(P.H, P.W), kPassThreads);
    const unsigned gR = (unsigned)cdiv(need_threads(P.H, P.W), kPMThreads);
    P.cam = w.cam;
    P.st = w.st;
    const bool memo = w.cc != nullptr;
    P.cc = w.cc;
    P.chg = w.chg;
    P.items = w.items;
    P.nitems = w.nitems;
    P.nevals = w.nevals;
    P.sorted = w.sorted;
    P.bins = w.bins;
    P.eps = w.eps;
    P.absmax = w.absmax;
    P.pst = w.pst;
    P.ritems = w.ritems;
    P.nritems = w.nritems;
    pm_sort_tiles(P, A.cam[0].iw, A.cam[0].ih);
    const int64_t HW = (int64_t)P.H * P.W;
    // one block per 128 possible items, rounded up to whole rounds over the XCDs
    const int64_t eval_grid = cdiv(cdiv(9 * ncol, (int64_t)kPMThreads), (int64_t)kXCD) * kXCD;
    const bool bounded = FMT == 0 && P.geo_w > 0 && pm_bound_on();
    const unsigned gL = (unsigned)cdiv(need_threads(P.H, P.W), kPassThreads);
    const unsigned gE = (unsigned)(cdiv(cdiv(ncol, (int64_t)kPMThreads), (int64_t)kXCD) * kXCD);
    hipLaunchKernelGGL(pm_init_kernel, dim3(1), dim3(1), 0, stream, w.st, seed, w.changed, A, w.cam);
    if (memo)
        hipLaunchKernelGGL(pm_memo_reset_kernel, dim3((unsigned)cdiv(HW, 256)), dim3(256), 0, stream, P, HW, 1);
    if (bounded) {
        hipLaunchKernelGGL(pm_bound_reset_kernel, dim3(1), dim3(64), 0, stream, P);
        hipLaunchKernelGGL(pm_absmax_kernel, dim3(256), dim3(256), 0, stream, P);
        hipLaunchKernelGGL(pm_eps_kernel, dim3(1), dim3(1), 0, stream, P);
    }
    int pass = 0;
    for (int it = 0; it < iters; ++it) {
        for (int colour = 0; colour < 2; ++colour, ++pass) {
            PassK Q;
            Q.colour = colour;
            Q.parity = pass & 1;
            Q.dsc = 0.05 / (it + 1);
            Q.nsc = 0.3 / (it + 1);
            Q.first = it == 0;
            if (memo) {
                ProfScope ps("pm_propagate", stream);
                if (pass > 0)
                    hipLaunchKernelGGL(pm_memo_reset_kernel, dim3(1), dim3(64), 0, stream, P, HW, 0);
                hipLaunchKernelGGL(pm_need_kernel, dim3(gT), dim3(kPassThreads), 0, stream, P, Q);
                const PMK Pe = Q.first ? pm_sort_items(P, ncol, stream) : P;
                {
                    ProfScope pe("pm_eval", stream);
                    hipLaunchKernelGGL((pm_eval_kernel<SM, FMT>), dim3((unsigned)eval_grid), dim3(kPMThreads), 0,
                                       stream, Pe);
                }
                hipLaunchKernelGGL(pm_decide_kernel, dim3(gN), dim3(kPassThreads), 0, stream, P, Q);
            } else {
                ProfScope ps("pm_propagate", stream);
                hipLaunchKernelGGL((pm_propagate_kernel<SM, FMT>), dim3(gA), dim3(kPMThreads), 0, stream, P, Q);
            }
            {
                ProfScope ps("pm_rank_rows", stream);
                hipLaunchKernelGGL(pm_rank_rows_kernel, dim3((unsigned)P.H), dim3(256), 0, stream, P, Q);
            }
            {
                ProfScope ps("pm_scan_rows", stream);
                hipLaunchKernelGGL(pm_scan_rows_kernel, dim3(1), dim3(256), 0, stream, P, Q);
            }
            {
                ProfScope ps("pm_random", stream);
                if (bounded) {
                    for (int tr = 0; tr < 2; ++tr) {
                        hipLaunchKernelGGL(pm_try_list_kernel<SM>, dim3(gL), dim3(kPassThreads), 0, stream, P, Q, tr);
                        hipLaunchKernelGGL(pm_try_eval_kernel<SM>, dim3(gE), dim3(kPMThreads), 0, stream, P, Q, tr);
                        if (tr == 0)
                            hipLaunchKernelGGL(pm_try_fallback_kernel<SM>, dim3(4), dim3(64), 0, stream, P, Q);
                    }
                } else {
                    hipLaunchKernelGGL((pm_random_kernel<SM, FMT>), dim3(gR), dim3(kPMThreads), 0, stream, P, Q);
                }
            }
        }
    }
    return launch_check("pm kernels");
}

}  // namespace om

using namespace om;

extern "C" size_t om_pm_workspace_bytes(int32_t H, int32_t W)
{
    if (H <= 0 || W <= 0)
        return 256;
    return pm_ws_layout(H, W, nullptr, nullptr, pm_memo_ok(H, W));
}

static int pm_refine_impl(const float *ref, int32_t H, int32_t W, const float *ray, const double *Cr, int32_t S,
                          const void *const *imgs, const float *const *deps, const int32_t *ih, const int32_t *iw,
                          const double *par, float *depth, float *normal, float *score, const uint8_t *active,
                          int32_t r, int32_t step, int32_t top_k, double geo_w, double geo_tol, int32_t iters,
                          double dmin, double dmax, uint32_t seed, void *workspace, size_t workspace_bytes,
                          int64_t *changed, void *stream, int fmt)
{
    if (changed)
        *changed = 0;
    if (S <= 0 || S > 16)  // the reference returns 0 without work (_dense.pyx:587-588)
        return OM_OK;
    if (H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    if (step <= 0)
        return fail(OM_EINVAL, "pm_refine needs step >= 1");
    if (top_k < 1)
        return fail(OM_EINVAL, "pm_refine needs top_k >= 1");
    if (!imgs || !deps || !ih || !iw || !par || !Cr)
        return fail(OM_EINVAL, "null host array");
    if ((int64_t)H * W == 0 || iters <= 0)
        return OM_OK;
    if (!tap_size_ok(H, W))
        return fail(OM_ESHAPE, "pm_refine: reference images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported");
    const bool memo = pm_memo_ok(H, W);
    const size_t need = pm_ws_layout(H, W, nullptr, nullptr, memo);
    if (!workspace || workspace_bytes < need)
        return fail(OM_EINVAL, "workspace of %zu bytes needed (got %zu)", need, workspace_bytes);
    PMWs w;
    pm_ws_layout(H, W, (char *)workspace, &w, memo);
    PMK P;
    memset(&P, 0, sizeof P);
    P.ref = ref;
    P.ray = ray;
    P.active = active;
    P.depth = depth;
    P.normal = normal;
    P.score = score;
    P.H = H;
    P.W = W;
    P.S = S;
    P.r = r;
    P.step = step;
    P.kk = top_k < S ? top_k : S;
    P.cr0 = Cr[0];
    P.cr1 = Cr[1];
    P.cr2 = Cr[2];
    P.geo_w = geo_w;
    P.geo_tol = geo_tol;
    P.dmin = dmin;
    P.dmax = dmax;
    P.best = w.best;
    P.sc = w.sc;
    P.rank = w.rank;
    P.row_cnt = w.row_cnt;
    P.row_off = w.row_off;
    P.changed = w.changed;
    P.seed = (int)seed;
    P.pinhole = 1;
    PMCamArgs A;
    memset(&A, 0, sizeof A);
    A.S = S;
    for (int s = 0; s < S; ++s) {
        if (ih[s] < 1 || iw[s] < 1)
            return fail(OM_ESHAPE, "source %d has an empty image", s);
        if (!tap_size_ok(ih[s], iw[s]))
            return fail(OM_ESHAPE, "source %d: images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported", s);
        PMCam &c = A.cam[s];
        c.img = (const float *)imgs[s];
        c.dep = deps[s];
        c.ih = ih[s];
        c.iw = iw[s];
        c.wl = (double)iw[s] - 1.001;
        c.hl = (double)ih[s] - 1.001;
        for (int t = 0; t < 18; ++t)
            c.par[t] = par[18 * s + t];
        if (par[18 * s + 13] != 0.0 || par[18 * s + 14] != 0.0 || par[18 * s + 15] != 0.0)
            P.pinhole = 0;
        for (int t = 0; t < 9; ++t)
            c.rf[t] = (float)c.par[t];
        for (int t = 0; t < 3; ++t)
            c.bf[t] = (float)(Cr[t] - c.par[9 + t]);
        c.ff = (float)c.par[12];
        c.cxf = (float)c.par[16];
        c.cyf = (float)c.par[17];
        c.wlf = (float)c.wl;
        c.hlf = (float)c.hl;
        c.k1f = (float)c.par[13];
        c.k2f = (float)c.par[14];
        c.k3f = (float)c.par[15];
    }
    hipStream_t st = as_stream(stream);
    int rc;
    using RunFn = int (*)(PMK &, const PMCamArgs &, int, hipStream_t, uint32_t, const PMWs &);
    static const RunFn runs[4][4] = {{pm_run<4, 0>, pm_run<8, 0>, pm_run<12, 0>, pm_run<16, 0>},
                                     {pm_run<4, OM_FMT_F16>, pm_run<8, OM_FMT_F16>, pm_run<12, OM_FMT_F16>,
                                      pm_run<16, OM_FMT_F16>},
                                     {pm_run<4, OM_FMT_BF16>, pm_run<8, OM_FMT_BF16>, pm_run<12, OM_FMT_BF16>,
                                      pm_run<16, OM_FMT_BF16>},
                                     {pm_run<4, OM_FMT_F32>, pm_run<8, OM_FMT_F32>, pm_run<12, OM_FMT_F32>,
                                      pm_run<16, OM_FMT_F32>}};
    const int bucket = S <= 4 ? 0 : (S <= 8 ? 1 : (S <= 12 ? 2 : 3));
    rc = runs[fmt][bucket](P, A, iters, st, seed, w);
    if (rc)
        return rc;
    if (changed) {
        unsigned long long h = 0;
        if ((rc = hip_check(hipMemcpyAsync(&h, w.changed, sizeof h, hipMemcpyDeviceToHost, st), "copy changed")))
            return rc;
        if ((rc = hip_check(hipStreamSynchronize(st), "pm_refine")))
            return rc;
        *changed = (int64_t)h;
    }
    return OM_OK;
}

extern "C" int om_pm_workspace_counts(const void *workspace, int32_t H, int32_t W, int64_t *evaluated)
{
    if (!workspace || !evaluated)
        return fail(OM_EINVAL, "null argument");
    *evaluated = -1;
    if (H <= 0 || W <= 0 || !pm_memo_ok(H, W))
        return OM_OK;
    PMWs w;
    pm_ws_layout(H, W, (char *)workspace, &w, true);
    unsigned long long h = 0;
    int rc = hip_check(hipMemcpy(&h, w.nevals, sizeof h, hipMemcpyDeviceToHost), "copy counts");
    if (!rc)
        *evaluated = (int64_t)h;
    return rc;
}

extern "C" int om_pm_refine(const float *ref, int32_t H, int32_t W, const float *ray, const double *Cr, int32_t S,
                            const float *const *imgs, const float *const *deps, const int32_t *ih, const int32_t *iw,
                            const double *par, float *depth, float *normal, float *score, const uint8_t *active,
                            int32_t r, int32_t step, int32_t top_k, double geo_w, double geo_tol, int32_t iters,
                            double dmin, double dmax, uint32_t seed, void *workspace, size_t workspace_bytes,
                            int64_t *changed, void *stream)
{
    return pm_refine_impl(ref, H, W, ray, Cr, S, (const void *const *)imgs, deps, ih, iw, par, depth, normal, score,
                          active, r, step, top_k, geo_w, geo_tol, iters, dmin, dmax, seed, workspace, workspace_bytes,
                          changed, stream, 0);
}

extern "C" int om_pm_refine_lp(const float *ref, int32_t H, int32_t W, const float *ray, const double *Cr, int32_t S,
                               const void *const *imgs16, const float *const *deps, const int32_t *ih,
                               const int32_t *iw, const double *par, float *depth, float *normal, float *score,
                               const uint8_t *active, int32_t r, int32_t step, int32_t top_k, double geo_w,
                               double geo_tol, int32_t iters, double dmin, double dmax, uint32_t seed, void *workspace,
                               size_t workspace_bytes, int64_t *changed, int32_t fmt, void *stream)
{
    if (fmt != OM_FMT_F16 && fmt != OM_FMT_BF16 && fmt != OM_FMT_F32)
        return fail(OM_EINVAL, "unknown storage format %d", fmt);
    return pm_refine_impl(ref, H, W, ray, Cr, S, imgs16, deps, ih, iw, par, depth, normal, score, active, r, step,
                          top_k, geo_w, geo_tol, iters, dmin, dmax, seed, workspace, workspace_bytes, changed, stream,
                          fmt);
}

// ------------------------------------------------------------------------------------------
// normals from a depth map (densify._init_normals) and batched PatchMatch (om_batch)
// ------------------------------------------------------------------------------------------
#include "batch.h"

namespace om {

// densify.py:481-496 in NumPy's float32 operation order: points = ray * depth, 2-px central
// differences, np.cross (each product rounded, then subtracted), norm = sqrt((x0^2 + x1^2) + x2^2),
// (0,0,1) where undefined, flipped to face the camera.
__device__ __forceinline__ void normal_px(const float *depth, const float *ray, int H, int W, float *normal, int64_t p)
{
    const int64_t HW = (int64_t)H * W;
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    auto P = [&](int c, int yy, int xx) {
        const float d = depth[(int64_t)yy * W + xx];
        const float dn = isfinite(d) ? d : NAN;
        return ray[c * HW + (int64_t)yy * W + xx] * dn;
    };
    float tu[3], tv[3];
    const bool hx = x >= 2 && x < W - 2, hy = y >= 2 && y < H - 2;
    for (int c = 0; c < 3; ++c) {
        tu[c] = hx ? P(c, y, x + 2) - P(c, y, x - 2) : NAN;
        tv[c] = hy ? P(c, y + 2, x) - P(c, y - 2, x) : NAN;
    }
    float n[3];
    float t;
    n[0] = tu[1] * tv[2];
    t = tu[2] * tv[1];
    n[0] = n[0] - t;
    n[1] = tu[2] * tv[0];
    t = tu[0] * tv[2];
    n[1] = n[1] - t;
    n[2] = tu[0] * tv[1];
    t = tu[1] * tv[0];
    n[2] = n[2] - t;
    float ss = n[0] * n[0];
    ss = ss + n[1] * n[1];
    ss = ss + n[2] * n[2];
    const float nn = sqrtf(ss);
    const bool bad = !isfinite(nn) || nn < 1e-12f;
    if (bad) {
        n[0] = 0.f;
        n[1] = 0.f;
        n[2] = 1.f;
    } else {
        n[0] = n[0] / nn;
        n[1] = n[1] / nn;
        n[2] = n[2] / nn;
    }
    float dot = n[0] * ray[p];
    dot = dot + n[1] * ray[HW + p];
    dot = dot + n[2] * ray[2 * HW + p];
    const float sgn = dot > 0.f ? -1.f : 1.f;
    for (int c = 0; c < 3; ++c)
        normal[c * HW + p] = n[c] * sgn;
}

__global__ void init_normals_kernel(const float *depth, const float *ray, int H, int W, float *normal)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (int64_t)H * W)
        normal_px(depth, ray, H, W, normal, p);
}

// batched start of PatchMatch, per pixel of all views: copy the start depth / score maps in
// (the pass-1 maps stay untouched for the neighbours' geometric term, densify.py:620-634) and
// optionally set the normals from the start depth (densify._init_normals, densify.py:481-496)
__global__ void pm_prepare_batch(const PMK *views, const int64_t *px0, int nviews, int normals)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= px0[nviews])
        return;
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (px0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const PMK &P = views[lo];
    const int64_t p = t - px0[lo];
    if (P.depth_from)
        P.depth[p] = P.depth_from[p];
    if (P.score_from)
        P.score[p] = P.score_from[p];
    if (normals)
        normal_px(P.depth_from ? P.depth_from : P.depth, P.ray, P.H, P.W, P.normal, p);
}

static int sm_bucket(int S) { return S <= 4 ? 4 : (S <= 8 ? 8 : (S <= 12 ? 12 : 16)); }

int batch_pm_set(om_batch *b, const om_pm_view *views, int nviews, int r, int step, int top_k, double geo_w,
                 int iters, int init_normals)
{
    if (nviews < 1 || !views)
        return fail(OM_EINVAL, "batch needs at least one view");
    if (step <= 0 || top_k < 1 || iters < 0)
        return fail(OM_EINVAL, "pm_refine needs step >= 1, top_k >= 1, iters >= 0");
    std::vector<PMK> P(nviews);
    bool memo = true;   // the memo path lists a view's items as pixel * 16 + slot in an int
    for (int v = 0; v < nviews; ++v)
        memo = memo && pm_memo_ok(views[v].H, views[v].W);
    b->pm_changed = (unsigned long long *)b->dev_alloc(sizeof(unsigned long long) * nviews);
    b->pm_nevals.assign(memo ? nviews : 0, nullptr);
    if (!b->pm_changed)
        return fail(OM_ENOMEM, "batch: counters");
    for (int v = 0; v < nviews; ++v) {
        const om_pm_view &V = views[v];
        if (V.S < 1 || V.S > 16)
            return fail(OM_ENSRC, "view %d: PatchMatch needs 1..16 sources", v);
        if (V.H < 1 || V.W < 1)
            return fail(OM_ESHAPE, "view %d: empty image", v);
        if (!tap_size_ok(V.H, V.W))
            return fail(OM_ESHAPE, "view %d: reference images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported", v);
        char *ws = (char *)b->dev_alloc(pm_ws_layout(V.H, V.W, nullptr, nullptr, memo));
        if (!ws)
            return fail(OM_ENOMEM, "batch: PatchMatch workspace");
        PMWs w;
        pm_ws_layout(V.H, V.W, ws, &w, memo);
        PMCam cams[16];
        memset(cams, 0, sizeof cams);
        PMK &Q = P[v];
        memset(&Q, 0, sizeof Q);
        Q.pinhole = 1;
        for (int s = 0; s < V.S; ++s) {
            if (V.ih[s] < 1 || V.iw[s] < 1)
                return fail(OM_ESHAPE, "view %d source %d has an empty image", v, s);
            if (!tap_size_ok(V.ih[s], V.iw[s]))
                return fail(OM_ESHAPE, "view %d source %d: images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported", v, s);
            cams[s].img = V.imgs[s];
            cams[s].dep = V.deps[s];
            cams[s].ih = V.ih[s];
            cams[s].iw = V.iw[s];
            cams[s].wl = (double)V.iw[s] - 1.001;
            cams[s].hl = (double)V.ih[s] - 1.001;
            for (int t = 0; t < 18; ++t)
                cams[s].par[t] = V.par[s][t];
            if (V.par[s][13] != 0.0 || V.par[s][14] != 0.0 || V.par[s][15] != 0.0)
                Q.pinhole = 0;
        }
        int rc = hip_check(hipMemcpy(w.cam, cams, sizeof(PMCam) * 16, hipMemcpyHostToDevice), "upload");
        if (rc)
            return rc;
        Q.ref = V.ref;
        Q.ray = V.ray;
        Q.active = V.active;
        Q.depth = V.depth;
        Q.normal = V.normal;
        Q.score = V.score;
        Q.H = V.H;
        Q.W = V.W;
        Q.S = V.S;
        Q.r = r;
        Q.step = step;
        Q.kk = top_k < V.S ? top_k : V.S;
        Q.seed = (int)V.seed;
        Q.cr0 = V.Cr[0];
        Q.cr1 = V.Cr[1];
        Q.cr2 = V.Cr[2];
        Q.geo_w = geo_w;
        Q.geo_tol = V.geo_tol;
        Q.dmin = V.dmin;
        Q.dmax = V.dmax;
        Q.best = w.best;
        Q.sc = w.sc;
        Q.rank = w.rank;
        Q.row_cnt = w.row_cnt;
        Q.row_off = w.row_off;
        Q.st = w.st;
        Q.changed = b->pm_changed + v;
        Q.cam = w.cam;
        Q.depth_from = V.depth_from;
        Q.score_from = V.score_from;
        Q.cc = w.cc;
        Q.chg = w.chg;
        Q.items = w.items;
        Q.nitems = w.nitems;
        Q.nevals = w.nevals;
        Q.sorted = w.sorted;
        Q.bins = w.bins;
        Q.eps = w.eps;
        Q.absmax = w.absmax;
        Q.pst = w.pst;
        Q.ritems = w.ritems;
        Q.nritems = w.nritems;
        pm_sort_tiles(Q, V.iw[0], V.ih[0]);
        if (memo)
            b->pm_nevals[v] = w.nevals;
    }
    // views ordered by source bucket: the propagation / random kernels run once per bucket with
    // the register footprint of that bucket (4, 8, 12 or 16 sources)
    std::vector<int> order(nviews);
    for (int v = 0; v < nviews; ++v)
        order[v] = v;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return sm_bucket(P[a].S) < sm_bucket(P[c].S); });
    std::vector<PMK> Ps(nviews);
    std::vector<int> row0(nviews + 1, 0), gblk0, glist0;
    std::vector<int64_t> px0(nviews + 1, 0);
    b->pm_ngroups = 0;
    for (int i = 0; i < nviews; ++i) {
        const PMK &Q = P[order[i]];
        Ps[i] = Q;
        row0[i + 1] = row0[i] + Q.H;
        px0[i + 1] = px0[i] + (int64_t)Q.H * Q.W;
        const int sm = sm_bucket(Q.S);
        if (b->pm_ngroups == 0 || b->pm_groups[b->pm_ngroups - 1].sm != sm) {
            PMGroup &g = b->pm_groups[b->pm_ngroups++];
            g.sm = sm;
            g.first = i;
            g.nv = 0;
            g.blk = 0;
            g.blk0_off = (int)gblk0.size();
            gblk0.push_back(0);
            g.blk_list = 0;
            g.blk_eval = 0;
            g.list0_off = (int)glist0.size();
            glist0.push_back(0);
        }
        PMGroup &g = b->pm_groups[b->pm_ngroups - 1];
        g.blk += (int)cdiv(need_threads(Q.H, Q.W), kPMThreads);  // random refinement walks tiles
        g.nv += 1;
        gblk0.push_back(g.blk);
        g.blk_list += (int)cdiv(need_threads(Q.H, Q.W), kPassThreads);
        glist0.push_back(g.blk_list);
        g.blk_eval += (int)cdiv((int64_t)Q.H * ((Q.W + 1) / 2), (int64_t)kPMThreads);
    }
    for (int gi = 0; gi < b->pm_ngroups; ++gi)  // whole rounds over the XCDs
        b->pm_groups[gi].blk_eval = (int)(cdiv((int64_t)b->pm_groups[gi].blk_eval, (int64_t)kXCD) * kXCD);
    b->pm_views = b->dev_alloc(sizeof(PMK) * nviews);
    b->pm_blk0 = (int *)b->dev_alloc(sizeof(int) * gblk0.size());
    b->pm_list0 = (int *)b->dev_alloc(sizeof(int) * glist0.size());
    b->pm_row0 = (int *)b->dev_alloc(sizeof(int) * (nviews + 1));
    b->pm_px0 = (int64_t *)b->dev_alloc(sizeof(int64_t) * (nviews + 1));
    if (!b->pm_views || !b->pm_blk0 || !b->pm_list0 || !b->pm_row0 || !b->pm_px0)
        return fail(OM_ENOMEM, "batch: PatchMatch tables");
    int rc;
    if ((rc = hip_check(hipMemcpy(b->pm_views, Ps.data(), sizeof(PMK) * nviews, hipMemcpyHostToDevice), "upload")) ||
        (rc = hip_check(hipMemcpy(b->pm_blk0, gblk0.data(), sizeof(int) * gblk0.size(), hipMemcpyHostToDevice),
                        "upload")) ||
        (rc = hip_check(hipMemcpy(b->pm_list0, glist0.data(), sizeof(int) * glist0.size(), hipMemcpyHostToDevice),
                        "upload")) ||
        (rc = hip_check(hipMemcpy(b->pm_row0, row0.data(), sizeof(int) * (nviews + 1), hipMemcpyHostToDevice),
                        "upload")) ||
        (rc = hip_check(hipMemcpy(b->pm_px0, px0.data(), sizeof(int64_t) * (nviews + 1), hipMemcpyHostToDevice),
                        "upload")))
        return rc;
    b->pm_host.assign((const char *)Ps.data(), (const char *)Ps.data() + sizeof(PMK) * nviews);
    b->pm_n = nviews;
    b->pm_iters = iters;
    b->pm_rows = row0[nviews];
    b->pm_px = px0[nviews];
    b->pm_normals = init_normals != 0;
    b->pm_memo = memo;
    b->pm_bounded = geo_w > 0 && pm_bound_on();
    b->pm_prepare = b->pm_normals;
    for (int v = 0; v < nviews; ++v)
        b->pm_prepare = b->pm_prepare || views[v].depth_from || views[v].score_from;
    // side streams of the memoised propagation (default priority, non-blocking)
    const int nside = memo && nviews > 1 ? std::min(OM_PM_SIDE_STREAMS, 4) : 0;
    if (nside > 1) {
        if ((rc = hip_check(hipEventCreateWithFlags(&b->pm_fork, hipEventDisableTiming), "hipEventCreate")))
            return rc;
        for (int k = 0; k < nside; ++k) {
            if ((rc = hip_check(hipStreamCreateWithFlags(&b->pm_side[k], hipStreamNonBlocking), "hipStreamCreate")) ||
                (rc = hip_check(hipEventCreateWithFlags(&b->pm_join[k], hipEventDisableTiming), "hipEventCreate")))
                return rc;
        }
        b->pm_nside = nside;
    }
    return OM_OK;
}

// the memoised propagation of one source bucket, view by view: the views of a bucket together hold
// more source images and depth maps than the last-level cache, one view's fit; the launches are
// recorded in the batch's graph, so launching per view costs little
// the memoised propagation of each view of group g: view i's chain (need, first-iteration sort,
// eval, decide) on streams[(*next)++ % ns]; a view's chain touches only its own workspace, so
// chains of different views may run concurrently
static void pm_group_memo(const om_batch *b, const PMGroup &g, const PassK &Q, const hipStream_t *streams, int ns,
                          int *next)
{
    const PMK *host = (const PMK *)b->pm_host.data();
    for (int i = g.first; i < g.first + g.nv; ++i) {
        const hipStream_t st = streams[(*next)++ % ns];
        const PMK &P = host[i];
        const int64_t ncol = (int64_t)P.H * ((P.W + 1) / 2);
        const unsigned gN = (unsigned)cdiv(ncol, kPassThreads);
        const unsigned gT = (unsigned)cdiv(need_threads(P.H, P.W), kPassThreads);
        const unsigned ge = (unsigned)(cdiv(cdiv(9 * ncol, (int64_t)kPMThreads), (int64_t)kXCD) * kXCD);
        hipLaunchKernelGGL(pm_need_kernel, dim3(gT), dim3(kPassThreads), 0, st, P, Q);
        const PMK Pe = Q.first ? pm_sort_items(P, ncol, st) : P;
        switch (g.sm) {
        case 4: hipLaunchKernelGGL((pm_eval_kernel<4, 0>), dim3(ge), dim3(kPMThreads), 0, st, Pe); break;
        case 8: hipLaunchKernelGGL((pm_eval_kernel<8, 0>), dim3(ge), dim3(kPMThreads), 0, st, Pe); break;
        case 12: hipLaunchKernelGGL((pm_eval_kernel<12, 0>), dim3(ge), dim3(kPMThreads), 0, st, Pe); break;
        default: hipLaunchKernelGGL((pm_eval_kernel<16, 0>), dim3(ge), dim3(kPMThreads), 0, st, Pe); break;
        }
        hipLaunchKernelGGL(pm_decide_kernel, dim3(gN), dim3(kPassThreads), 0, st, P, Q);
    }
}

static void pm_group_launch(const om_batch *b, const PMGroup &g, bool propagate, const PassK &Q, hipStream_t st)
{
    const PMK *V = (const PMK *)b->pm_views + g.first;
    const int *blk0 = b->pm_blk0 + g.blk0_off;
#define OM_PM_GROUP_CASE(SM_)                                                                                          \
    case SM_:                                                                                                          \
        if (propagate)                                                                                                 \
            hipLaunchKernelGGL(pm_propagate_batch<SM_>, dim3((unsigned)g.blk), dim3(kPMThreads), 0, st, V, blk0, g.nv, \
                               Q);                                                                                     \
        else                                                                                                           \
            hipLaunchKernelGGL(pm_random_batch<SM_>, dim3((unsigned)g.blk), dim3(kPMThreads), 0, st, V, blk0, g.nv, Q); \
        break;
    switch (g.sm) {
        OM_PM_GROUP_CASE(4)
        OM_PM_GROUP_CASE(8)
        OM_PM_GROUP_CASE(12)
        OM_PM_GROUP_CASE(16)
    }
#undef OM_PM_GROUP_CASE
}

// the bounded random refinement of one source bucket: try tr's lists, then their evaluation
static void pm_group_try(const om_batch *b, const PMGroup &g, int tr, const PassK &Q, hipStream_t st)
{
    const PMK *V = (const PMK *)b->pm_views + g.first;
    const int *list0 = b->pm_list0 + g.list0_off;
#define OM_PM_TRY_CASE(SM_)                                                                                            \
    case SM_:                                                                                                          \
        hipLaunchKernelGGL(pm_try_list_batch<SM_>, dim3((unsigned)g.blk_list), dim3(kPassThreads), 0, st, V, list0,    \
                           g.nv, Q, tr);                                                                               \
        hipLaunchKernelGGL(pm_try_eval_batch<SM_>, dim3((unsigned)g.blk_eval), dim3(kPMThreads), 0, st, V, g.nv, Q,    \
                           tr);                                                                                        \
        if (tr == 0)                                                                                                   \
            hipLaunchKernelGGL(pm_try_fallback_batch<SM_>, dim3(4, (unsigned)g.nv), dim3(64), 0, st, V, Q);            \
        break;
    switch (g.sm) {
        OM_PM_TRY_CASE(4)
        OM_PM_TRY_CASE(8)
        OM_PM_TRY_CASE(12)
        OM_PM_TRY_CASE(16)
    }
#undef OM_PM_TRY_CASE
}

static void pm_batch_passes(om_batch *b, hipStream_t st)
{
    const PMK *V = (const PMK *)b->pm_views;
    int pass = 0;
    for (int it = 0; it < b->pm_iters; ++it) {
        for (int colour = 0; colour < 2; ++colour, ++pass) {
            PassK Q;
            Q.colour = colour;
            Q.parity = pass & 1;
            Q.dsc = 0.05 / (it + 1);
            Q.nsc = 0.3 / (it + 1);
            Q.first = it == 0;
            {
                ProfScope ps("pm_propagate_batch", st);
                if (b->pm_memo) {
                    if (pass > 0)
                        hipLaunchKernelGGL(pm_memo_reset_batch, dim3((unsigned)cdiv(b->pm_n, 64)), dim3(64), 0, st, V,
                                           b->pm_px0, b->pm_n, 0);
                    int next = 0;
                    if (b->pm_nside > 1) {
                        // fork: the side streams start after everything enqueued so far
                        // (enqueue errors surface at the launch check of the caller's next kernel)
                        (void)hipEventRecord(b->pm_fork, st);
                        for (int k = 0; k < b->pm_nside; ++k)
                            (void)hipStreamWaitEvent(b->pm_side[k], b->pm_fork, 0);
                        for (int gi = 0; gi < b->pm_ngroups; ++gi)
                            pm_group_memo(b, b->pm_groups[gi], Q, b->pm_side, b->pm_nside, &next);
                        // join: the launch stream continues once every chain has finished
                        for (int k = 0; k < b->pm_nside; ++k) {
                            (void)hipEventRecord(b->pm_join[k], b->pm_side[k]);
                            (void)hipStreamWaitEvent(st, b->pm_join[k], 0);
                        }
                    } else {
                        for (int gi = 0; gi < b->pm_ngroups; ++gi)
                            pm_group_memo(b, b->pm_groups[gi], Q, &st, 1, &next);
                    }
                } else {
                    for (int gi = 0; gi < b->pm_ngroups; ++gi)
                        pm_group_launch(b, b->pm_groups[gi], true, Q, st);
                }
            }
            hipLaunchKernelGGL(pm_rank_rows_batch, dim3((unsigned)b->pm_rows), dim3(256), 0, st, V, b->pm_row0,
                               b->pm_n, Q);
            hipLaunchKernelGGL(pm_scan_rows_batch, dim3((unsigned)b->pm_n), dim3(256), 0, st, V, Q);
            {
                ProfScope ps("pm_random_batch", st);
                for (int gi = 0; gi < b->pm_ngroups; ++gi) {
                    if (b->pm_bounded) {
                        pm_group_try(b, b->pm_groups[gi], 0, Q, st);
                        pm_group_try(b, b->pm_groups[gi], 1, Q, st);
                    } else {
                        pm_group_launch(b, b->pm_groups[gi], false, Q, st);
                    }
                }
            }
        }
    }
}

int batch_pm_enqueue(om_batch *b, hipStream_t st)
{
    if (b->pm_n == 0)
        return OM_OK;
    ProfScope ps("pm_batch", st);
    const PMK *V = (const PMK *)b->pm_views;
    hipLaunchKernelGGL(pm_init_batch, dim3((unsigned)cdiv(b->pm_n, 64)), dim3(64), 0, st, V, b->pm_n);
    if (b->pm_prepare)
        hipLaunchKernelGGL(pm_prepare_batch, dim3((unsigned)cdiv(b->pm_px, 256)), dim3(256), 0, st, V, b->pm_px0,
                           b->pm_n, (int)b->pm_normals);
    if (b->pm_memo)
        hipLaunchKernelGGL(pm_memo_reset_batch, dim3((unsigned)cdiv(b->pm_px > b->pm_n ? b->pm_px : b->pm_n, 256)),
                           dim3(256), 0, st, V, b->pm_px0, b->pm_n, 1);
    if (b->pm_bounded) {  // the margin of every view from its images' largest values
        hipLaunchKernelGGL(pm_absmax_batch, dim3(64, (unsigned)b->pm_n), dim3(256), 0, st, V);
        hipLaunchKernelGGL(pm_eps_batch, dim3((unsigned)cdiv(b->pm_n, 64)), dim3(64), 0, st, V, b->pm_n);
    }
    pm_batch_passes(b, st);
    return launch_check("pm batch kernels");
}

}  // namespace om

extern "C" int om_batch_pm_evaluated(om_batch *b, int64_t *evaluated)
{
    if (!b || !evaluated)
        return fail(OM_EINVAL, "null argument");
    if (b->pm_nevals.empty()) {
        for (int v = 0; v < b->pm_n; ++v)
            evaluated[v] = -1;
        return OM_OK;
    }
    for (int v = 0; v < b->pm_n; ++v) {
        unsigned long long h = 0;
        if (int rc = hip_check(hipMemcpy(&h, b->pm_nevals[v], sizeof h, hipMemcpyDeviceToHost), "copy evaluations"))
            return rc;
        evaluated[v] = (int64_t)h;
    }
    return OM_OK;
}

extern "C" int om_init_normals(const float *depth, const float *ray, int32_t H, int32_t W, float *normal, void *stream)
{
    if (H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    const int64_t n = (int64_t)H * W;
    if (n == 0)
        return OM_OK;
    hipLaunchKernelGGL(init_normals_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), depth, ray,
                       H, W, normal);
    return launch_check("init_normals_kernel");
}
