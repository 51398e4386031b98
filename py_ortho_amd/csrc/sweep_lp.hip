// Reduced-precision plane sweep for gfx950 (the precision variant of the depth-map path).
//
// Same algorithm and outputs as the FP64 sweep (sweep.hip; reference: _dense.sweep_hypothesis,
// orthomosaic/_dense.pyx:740-838, with the row kernels of orthomosaic/_sweep.h:33-214), with the
// arithmetic re-planned for speed instead of bit parity:
//   - reference and source images are stored as FP16 or BF16 (half the bytes of every tap);
//   - per plane the depth 1/inv, per pixel the warp b + d R ray, the projection and the bilinear
//     tap are FP32 (the plane and camera parameters are formed in FP64 and rounded once);
//   - the window sums of V, J, J^2, J*I and the vertical sums are FP32 running sums;
//   - the NCC is FP32 except the variance m2 - mj^2 and its guard (> 1e-6), which are FP64, so
//     the cancellation does not decide which pixels count as textured.
// The winner state (best, prev, s_prev_best, s_next_best, idx) has the FP64 path's layout and
// meaning, so om_sweep_finalize turns it into depth as before.  Errors against the FP64 path are
// measured by the tests and the benchmark (NCC within ~1e-3, identical winners wherever the two
// best scores differ by more than that).
//
// LDS rows are padded to an odd number of words so that the horizontal pass (lanes spread over
// rows) and the vertical pass (lanes spread over columns) read distinct banks.  Constant planes are
// warped two at a time, and the window loops of radii 2 and 5 are compiled with the radius known.
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "om_common.h"

namespace om {
namespace lp {

// 32 x 32 output tiles: at r = 5 the 42 x 42 halo is warped in 7 rounds of the workgroup's lanes
// and the horizontal sums cover 42 rows per 32 output rows, where 64 x 16 tiles warp 75 x 26 in 8
// rounds and sum 26 rows per 16 (a 4000x3000 view, 12 sources, 96 planes: 85.0 against 94.4 ms,
// same bits).  An odd halo stride (43) would add a warp round (89.2 ms); the rows are 32 + 2r words
// apart instead, with the few pad words of halo32_pad between them.
constexpr int kTW = 32;        // output columns per tile
constexpr int kThreads = 256;  // 4 wave64
constexpr int kVR = 4;         // output rows per thread in the vertical pass
constexpr int kTH = kThreads / kTW * kVR;  // tile height = 8 row segments x kVR
__host__ __device__ constexpr int lp_halo_stride(int r) { return kTW + 2 * r; }
// Pad words before halo row hy (rows hc = 32 + 2r words apart, hc even): chosen from hc mod 8 so that
// the 8 rows of every aligned group start in 8 distinct classes mod 8.  The horizontal pass gives a
// half-wave 8 such rows x 4 runs 8 words apart, so its 32 lanes read 32 distinct LDS banks (r = 5:
// 5 pad words, still 7 warp rounds; SQ_LDS_BANK_CONFLICT down by three quarters and the 4000x3000
// view's FP16 sweep 85.3 -> 82.9 ms, BF16 84.2 -> 83.2 ms, same bits).
__host__ __device__ constexpr int halo32_pad(int hy, int hc)
{
    return hc % 8 == 0 ? hy : (hc % 8 == 4 ? hy >> 1 : (hy + 4) >> 3);
}
__host__ __device__ constexpr int lp_row0(int hy, int r)
{
    return hy * lp_halo_stride(r) + halo32_pad(hy, lp_halo_stride(r));
}
__host__ __device__ constexpr int lp_halo_words(int hr, int r) { return lp_row0(hr - 1, r) + lp_halo_stride(r); }
constexpr int kRun = 8;        // columns per horizontal running sum
constexpr int kMaxR = 7;

struct SrcLP {
    const void *img;          // FP16 / BF16 patterns or FP32 values (FMT)
    int h, w;
    float k1, k2, k3;  // radial distortion (0 for a pinhole)
    float g[12];  // Rs (9, row-major) and b = Rs (C_ref - C_src) (3), rounded from FP64
    float f, cx, cy, wl, hl;  // focal, principal point, w - 1.001, h - 1.001
};

struct SweepLPK {
    const void *ref;
    const float *mu, *sd, *ray;
    int H, W, n, r, k, mode;
    const float *ha, *hb;
    int hoff, i0, nh, tiles_x, tiles_y;
    int pinhole;              // every source has k1 = k2 = k3 = 0
    float *best, *prev, *spb, *snb;
    int *idx;
    SrcLP src[16];
};

// element i of an image stored in format FMT, as float
template <int FMT>
__device__ __forceinline__ float ld_img(const void *base, size_t i)
{
    if (FMT == OM_FMT_F32)
        return ((const float *)base)[i];
    const uint16_t v = ((const uint16_t *)base)[i];
    if (FMT == OM_FMT_BF16)
        return __uint_as_float((uint32_t)v << 16);
    return (float)__builtin_bit_cast(_Float16, v);
}

__device__ __forceinline__ float hyp_lp(const SweepLPK &P, int i, int p, float a, float b)
{
    switch (P.mode) {
    case OM_HYP_LINEAR: {
        const float t = (float)i * b;
        return a + t;
    }
    case OM_HYP_REFINE: {
        float v = (float)(i - P.hoff) * b;
        v = a + v;
        return (v >= 1e-6f || v != v) ? v : 1e-6f;
    }
    case OM_HYP_CONST:
        return P.ha[i - P.i0];
    default:
        return P.ha[(int64_t)(i - P.i0) * ((int64_t)P.H * P.W) + p];
    }
}

template <int K>
__device__ __forceinline__ void topk_ins(float (&t)[K], float v)
{
#pragma unroll
    for (int a = K - 1; a >= 0; --a) {
        const bool lt = t[a] < v;
        const float up = (a > 0) ? t[a - 1] : v;
        const bool prev_lt = (a > 0) ? (t[a - 1] < v) : true;
        if (lt)
            t[a] = prev_lt ? up : v;
    }
}

// Window radii with kernels of their own (loops unrolled, a horizontal run's samples kept in
// registers): 2 and 5; other radii <= kMaxR run the generic kernel (R = -1, radius from P.r).
__host__ __device__ constexpr int lp_radius_class(int r) { return r == 2 ? 0 : (r == 5 ? 1 : 2); }

// Constant planes are warped two at a time: the plane-independent term fmaf(g0, rx, fmaf(g1, ry,
// g2 rz)) of each coordinate is formed once per halo pixel and source for both.
constexpr int kPlaneGroupLP = 2;

__host__ __device__ constexpr int lp_group(bool cd) { return cd ? kPlaneGroupLP : 1; }

// bytes of the valid-sample counts of the horizontal windows (one byte each, at most 2 r + 1)
__host__ __device__ constexpr size_t lp_count_bytes(int hr) { return ((size_t)hr * (kTW + 1) + 3) & ~(size_t)3; }

// LDS bytes: per halo pixel its index, the reference, one warped value per plane of the group and
// (per-pixel hypotheses) the plane's depth; three float planes of horizontal sums and their counts
// The radius-5 kernel keeps each lane's horizontal task's 8 + 2r reference values in registers
// instead of a reference plane in LDS (at most one task per lane: 4 runs x (32 + 2r) rows <= 256):
// 39 248 bytes of LDS, so four workgroups share a CU (122 VGPRs, no spills).  A 4000x3000 view, 12
// sources, 96 planes: FP16 82.6 -> 70.1 ms, BF16 83.0 -> 69.0 ms, same bits.  At r = 2 four fit
// already.
__host__ __device__ constexpr bool lp_iref(int r) { return r == 5; }

__host__ __device__ constexpr size_t lp_lds(int r, bool cd)
{
    return (size_t)lp_halo_words(kTH + 2 * r, r) * 4 * ((lp_iref(r) ? 1 : 2) + lp_group(cd) + (cd ? 0 : 1)) +
           (size_t)3 * (kTH + 2 * r) * (kTW + 1) * 4 + lp_count_bytes(kTH + 2 * r);
}

// FP32 horizontal running sums of one 8-column run (counts as bytes); with a compile-time radius
// each sample's value and its reference are read once and kept until it leaves the window
template <int R>
__device__ __forceinline__ void lp_hsum_run(const float *Jr, const float *Ir, int r, float *h1, float *h2, float *h3,
                                            uint8_t *hc)
{
    if constexpr (R >= 0) {
        constexpr int NE = kRun + 2 * R;
        float ej[NE], ei[NE];
        int eok[NE];
        auto elem = [&](int e) {
            const float J = Jr[e];
            const bool ok = J == J;
            ej[e] = ok ? J : 0.f;
            ei[e] = Ir[e];
            eok[e] = ok;
        };
        int sv = 0;
        float sj = 0.f, sjj = 0.f, sij = 0.f;
#pragma unroll
        for (int a = 0; a <= 2 * R; ++a) {
            elem(a);
            sv += eok[a];
            sj += ej[a];
            sjj = fmaf(ej[a], ej[a], sjj);
            sij = fmaf(ej[a], ei[a], sij);
        }
        hc[0] = (uint8_t)sv;
        h1[0] = sj;
        h2[0] = sjj;
        h3[0] = sij;
#pragma unroll
        for (int c = 1; c < kRun; ++c) {
            const int en = c + 2 * R, eo = c - 1;
            elem(en);
            const float jn = ej[en], jo = ej[eo];
            sv += eok[en] - eok[eo];
            sj += jn - jo;
            sjj = fmaf(jn, jn, fmaf(-jo, jo, sjj));
            sij = fmaf(jn, ei[en], fmaf(-jo, ei[eo], sij));
            hc[c] = (uint8_t)sv;
            h1[c] = sj;
            h2[c] = sjj;
            h3[c] = sij;
        }
    } else {
        int sv = 0;
        float sj = 0.f, sjj = 0.f, sij = 0.f;
        for (int a = 0; a <= 2 * r; ++a) {
            const float J = Jr[a];
            const bool ok = J == J;
            const float jv = ok ? J : 0.f;
            sv += ok;
            sj += jv;
            sjj = fmaf(jv, jv, sjj);
            sij = fmaf(jv, Ir[a], sij);
        }
        hc[0] = (uint8_t)sv;
        h1[0] = sj;
        h2[0] = sjj;
        h3[0] = sij;
#pragma unroll
        for (int c = 1; c < kRun; ++c) {
            const float Jn = Jr[c + 2 * r], Jo = Jr[c - 1];
            const bool okn = Jn == Jn, oko = Jo == Jo;
            const float jn = okn ? Jn : 0.f, jo = oko ? Jo : 0.f;
            sv += (int)okn - (int)oko;
            sj += jn - jo;
            sjj = fmaf(jn, jn, fmaf(-jo, jo, sjj));
            sij = fmaf(jn, Ir[c + 2 * r], fmaf(-jo, Ir[c - 1], sij));
            hc[c] = (uint8_t)sv;
            h1[c] = sj;
            h2[c] = sjj;
            h3[c] = sij;
        }
    }
}

// CD: constant planes (OM_HYP_CONST): one depth per plane, warped in groups of kPlaneGroupLP.
// R: the window radius when it has kernels of its own, else -1.
template <int K, int FMT, bool CD, int R>
__global__ void __launch_bounds__(kThreads) sweep_lp_kernel(const SweepLPK P)
{
    extern __shared__ float lds_f[];
    constexpr int G = lp_group(CD);
    const int r = R >= 0 ? R : P.r;
    const int HR = kTH + 2 * r, HC = kTW + 2 * r;
    const int HCP = lp_halo_stride(r);  // halo row stride
    constexpr int kTWP = kTW + 1;    // padded row stride of the horizontal sums
    const int NH = lp_halo_words(HR, r);  // halo words; pad words are outside the image
    float *sD = lds_f;                             // (NH) depth of the current plane (!CD)
    int *sP = (int *)(CD ? lds_f : sD + NH);      // (NH) pixel index, -1 outside
    constexpr bool IREG = R >= 0 && lp_iref(R);
    static_assert(!IREG || (kTH + 2 * R) * (kTW / kRun) <= kThreads, "one horizontal task per lane");
    float *sI = (float *)(sP + NH);  // (NH) reference (!IREG)
    float *sJ = IREG ? sI : sI + NH;  // (G, NH) warped source, NaN = invalid
    float *sH = sJ + G * NH;         // (3, HR, kTWP) horizontal sums of J, J^2, J I
    uint8_t *sC = (uint8_t *)(sH + 3 * HR * kTWP);  // (HR, kTWP) valid samples of each window

    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int tx0 = (tile % P.tiles_x) * kTW, ty0 = (tile / P.tiles_x) * kTH;
    const int H = P.H, W = P.W;
    const size_t HW = (size_t)H * W;
    const int tid = threadIdx.x;
    const bool gen_maps = P.mode == OM_HYP_LINEAR || P.mode == OM_HYP_REFINE;

    // the halo: pixel index (-1 outside the image and in pad words) and reference value
    for (int o = tid; o < NH; o += kThreads) {
        int hy = o / HCP;
        if (o < lp_row0(hy, r))
            --hy;  // the pads (fewer than HCP words) put o in row hy - 1 or just after it
        const int hx = o - lp_row0(hy, r);
        const int gy = ty0 - r + hy, gx = tx0 - r + hx;
        const bool in = hx < HC && gy >= 0 && gy < H && gx >= 0 && gx < W;
        const int p = in ? gy * W + gx : -1;
        sP[o] = p;
        if (!IREG)
            sI[o] = in ? ld_img<FMT>(P.ref, (size_t)p) : 0.f;
    }
    // IREG: the reference values of this lane's horizontal task (row tid / 4, run tid % 4)
    constexpr int NEI = IREG ? kRun + 2 * (R > 0 ? R : 0) : 1;
    float Ireg[NEI];
    if (IREG) {
        const int hr = tid / (kTW / kRun), c0 = (tid % (kTW / kRun)) * kRun;
        const int gy = ty0 - r + hr;
#pragma unroll
        for (int e = 0; e < NEI; ++e) {
            const int gx = tx0 - r + c0 + e;
            const bool in = hr < HR && gy >= 0 && gy < H && gx >= 0 && gx < W;
            Ireg[e] = in ? ld_img<FMT>(P.ref, (size_t)gy * W + gx) : 0.f;
        }
    }

    const int vc = tid % kTW, vy0 = (tid / kTW) * kVR;
    const int ox = tx0 + vc;
    float mu_[kVR], sd_[kVR], st_best[kVR], st_prev[kVR], st_spb[kVR], st_snb[kVR];
    int st_idx[kVR];
    float tk[G][kVR][K];
#pragma unroll
    for (int j = 0; j < kVR; ++j) {
        const int oy = ty0 + vy0 + j;
        const bool in = oy < H && ox < W;
        const size_t p = in ? (size_t)oy * W + ox : 0;
        mu_[j] = in ? P.mu[p] : 0.f;
        sd_[j] = in ? P.sd[p] : 1.f;
        st_best[j] = in ? P.best[p] : 0.f;
        st_prev[j] = in ? P.prev[p] : 0.f;
        st_spb[j] = in ? P.spb[p] : 0.f;
        st_snb[j] = in ? P.snb[p] : 0.f;
        st_idx[j] = in ? P.idx[p] : 0;
    }
    const float win = (float)((2 * r + 1) * (2 * r + 1));
    __syncthreads();

    for (int i = P.i0; i < P.i0 + P.nh; i += G) {
        const int ng = min(G, P.i0 + P.nh - i);   // planes of this group (wave-uniform)
        float dc[G];
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
            dc[gg] = 0.f;
        if (CD) {
#pragma unroll
            for (int gg = 0; gg < G; ++gg)
                if (gg < ng)
                    dc[gg] = (float)(1.0 / (double)P.ha[i + gg - P.i0]);  // plane parameter in FP64, rounded once
        } else {
            for (int o = tid; o < NH; o += kThreads) {
                const int p = sP[o];
                float d = NAN;
                if (p >= 0) {
                    const float a = gen_maps ? P.ha[p] : 0.f, b = gen_maps ? P.hb[p] : 0.f;
                    d = (float)(1.0 / (double)hyp_lp(P, i, p, a, b));
                }
                sD[o] = d;
            }
        }
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
#pragma unroll
            for (int j = 0; j < kVR; ++j)
#pragma unroll
                for (int a = 0; a < K; ++a)
                    tk[gg][j][a] = -INFINITY;
        if (!CD)
            __syncthreads();

        for (int s = 0; s < P.n; ++s) {
            const SrcLP &S = P.src[s];
            const float *g = S.g;
            // ---- W: warp the halo (FP32 projection, FP32 bilinear of 16-bit taps) for every plane
            // of the group; each step loads the ray of its next halo pixel (pixel 0's outside the
            // image) before it works on its own
            const char *rayx = (const char *)P.ray, *rayy = (const char *)(P.ray + HW),
                       *rayz = (const char *)(P.ray + 2 * HW);
            int pn = tid < NH ? sP[tid] : -1;
            uint32_t on = (uint32_t)max(pn, 0) << 2;
            float rxn = *(const float *)(rayx + on), ryn = *(const float *)(rayy + on),
                  rzn = *(const float *)(rayz + on);
            for (int o = tid; o < NH; o += kThreads) {
                const int p = pn;
                const float rx = rxn, ry = ryn, rz = rzn;
                pn = o + kThreads < NH ? sP[o + kThreads] : -1;
                on = (uint32_t)max(pn, 0) << 2;
                rxn = *(const float *)(rayx + on);
                ryn = *(const float *)(rayy + on);
                rzn = *(const float *)(rayz + on);
                float J[G];
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    J[gg] = NAN;
                if (p >= 0) {
                    const float ax = fmaf(g[0], rx, fmaf(g[1], ry, g[2] * rz));
                    const float ay = fmaf(g[3], rx, fmaf(g[4], ry, g[5] * rz));
                    const float az = fmaf(g[6], rx, fmaf(g[7], ry, g[8] * rz));
#pragma unroll
                    for (int gg = 0; gg < G; ++gg) {
                        if (gg >= ng)
                            break;
                        const float d = CD ? dc[gg] : sD[o];
                        const float X = fmaf(d, ax, g[9]);
                        const float Y = fmaf(d, ay, g[10]);
                        const float Z = fmaf(d, az, g[11]);
                        if (Z > 1e-6f) {
                            const float iz = __builtin_amdgcn_rcpf(Z);
                            const float nx = X * iz, ny = Y * iz;
                            float dd = S.f;   // radial model of _sweep.h:55-56 in FP32
                            if (!P.pinhole) {
                                const float r2 = fmaf(nx, nx, ny * ny);
                                dd = S.f * fmaf(r2, fmaf(r2, fmaf(r2, S.k3, S.k2), S.k1), 1.0f);
                            }
                            const float u = fmaf(dd, nx, S.cx), v = fmaf(dd, ny, S.cy);
                            if (u >= 0.f && v >= 0.f && u <= S.wl && v <= S.hl) {
                                const int x0 = (int)u, y0 = (int)v;
                                const float fx = u - (float)x0, fy = v - (float)y0;
                                const size_t t = (size_t)y0 * S.w + x0;
                                const float a00 = ld_img<FMT>(S.img, t), a01 = ld_img<FMT>(S.img, t + 1);
                                const float a10 = ld_img<FMT>(S.img, t + S.w), a11 = ld_img<FMT>(S.img, t + S.w + 1);
                                const float t0 = fmaf(fx, a01 - a00, a00), t1 = fmaf(fx, a11 - a10, a10);
                                const float val = fmaf(fy, t1 - t0, t0);
                                if (isfinite(val))
                                    J[gg] = val;
                            }
                        }
                    }
                }
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    sJ[gg * NH + o] = J[gg];
            }
            __syncthreads();
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
                if (gg >= ng)
                    break;
                if (gg > 0)
                    __syncthreads();  // the previous plane's V pass has read sH
                const float *sJg = sJ + gg * NH;
                // ---- H: FP32 running sums over 8-column runs; a wave's lanes take 16 rows x 4 runs.
                for (int t = tid; t < HR * (kTW / kRun); t += kThreads) {
                    const int hr = t / (kTW / kRun), c0 = (t % (kTW / kRun)) * kRun;
                    const int oh = hr * kTWP + c0;
                    const int h0 = lp_row0(hr, r) + c0;
                    lp_hsum_run<R>(sJg + h0, IREG ? Ireg : sI + h0, r, sH + oh, sH + HR * kTWP + oh,
                                   sH + 2 * HR * kTWP + oh, sC + oh);
                }
                __syncthreads();
                // ---- V: FP32 vertical sums, NCC (variance and its guard in FP64)
                {
                    const float *h1 = sH + vc, *h2 = h1 + HR * kTWP, *h3 = h2 + HR * kTWP;
                    const uint8_t *h0 = sC + vc;
                    int c0 = 0;  // valid samples: a small integer, exact either way
                    float c1 = 0.f, c2 = 0.f, c3 = 0.f;
#pragma unroll
                    for (int a = 0; a <= 2 * r; ++a) {
                        const int o = (vy0 + a) * kTWP;
                        c0 += h0[o];
                        c1 += h1[o];
                        c2 += h2[o];
                        c3 += h3[o];
                    }
#pragma unroll
                    for (int j = 0; j < kVR; ++j) {
                        if (j > 0) {
                            const int on = (vy0 + j + 2 * r) * kTWP, oo = (vy0 + j - 1) * kTWP;
                            c0 += (int)h0[on] - (int)h0[oo];
                            c1 += h1[on] - h1[oo];
                            c2 += h2[on] - h2[oo];
                            c3 += h3[on] - h3[oo];
                        }
                        float out = -1.0f;
                        if ((float)c0 > 0.9f * win) {  // >= 90 % of the window valid
                            const float ic = __builtin_amdgcn_rcpf((float)c0);
                            const float mj = c1 * ic, m2 = c2 * ic, mij = c3 * ic;
                            const double vj = (double)m2 - (double)mj * (double)mj;
                            const float sj = vj > 1e-6 ? (float)sqrt(vj) : 1e-3f;
                            const float q = (mij - mu_[j] * mj) * __builtin_amdgcn_rcpf(sd_[j] * sj);
                            out = q > 1.0f ? 1.0f : (q < -1.0f ? -1.0f : q);
                        }
                        topk_ins<K>(tk[gg][j], out);
                    }
                }
            }
        }
        // ---- combine + streaming winner of each plane of the group, in order (_sweep.h:181-214)
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            if (gg >= ng)
                break;
            const int ii = i + gg;
#pragma unroll
            for (int j = 0; j < kVR; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int a = 0; a < K; ++a)
                    s = s + tk[gg][j][a];
                s = s / (float)K;
                if (st_idx[j] == ii - 1)
                    st_snb[j] = s;
                if (s > st_best[j]) {
                    st_best[j] = s;
                    st_spb[j] = st_prev[j];
                    st_snb[j] = -2.0f;
                    st_idx[j] = ii;
                }
                st_prev[j] = s;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kVR; ++j) {
        const int oy = ty0 + vy0 + j;
        if (oy < H && ox < W) {
            const size_t p = (size_t)oy * W + ox;
            P.best[p] = st_best[j];
            P.prev[p] = st_prev[j];
            P.spb[p] = st_spb[j];
            P.snb[p] = st_snb[j];
            P.idx[p] = st_idx[j];
        }
    }
}

__global__ void to_lp_kernel(const float *src, int64_t n, int fmt, void *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float x = src[i];
    if (fmt == OM_FMT_F32) {
        ((float *)out)[i] = x;
        return;
    }
    uint16_t *dst = (uint16_t *)out;
    if (fmt == OM_FMT_BF16) {
        const uint32_t u = __float_as_uint(x);
        dst[i] = (x != x) ? (uint16_t)0x7FC0 : (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
    } else {
        dst[i] = __builtin_bit_cast(uint16_t, (_Float16)x);
    }
}

}  // namespace lp
}  // namespace om

using namespace om;
using namespace om::lp;

extern "C" int om_to_lp(const float *src, int64_t n, int32_t fmt, void *dst, void *stream)
{
    if (fmt != OM_FMT_F16 && fmt != OM_FMT_BF16 && fmt != OM_FMT_F32)
        return fail(OM_EINVAL, "unknown storage format %d", fmt);
    if (n < 0)
        return fail(OM_ESHAPE, "negative size");
    if (n == 0)
        return OM_OK;
    hipLaunchKernelGGL(to_lp_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), src, n, fmt,
                       dst);
    return launch_check("to_lp_kernel");
}

extern "C" int om_sweep_lp(const om_sweep_desc *d, int32_t fmt, const void *ref16, void *stream)
{
    if (!d || !ref16)
        return fail(OM_EINVAL, "om_sweep_lp: null argument");
    if (fmt != OM_FMT_F16 && fmt != OM_FMT_BF16 && fmt != OM_FMT_F32)
        return fail(OM_EINVAL, "unknown storage format %d", fmt);
    if (d->n_src < 1 || d->n_src > 16)
        return fail(OM_ENSRC, "sweep needs 1..16 sources (got %d)", d->n_src);
    if (d->r < 0 || d->r > kMaxR)
        return fail(OM_EUNSUP, "reduced-precision sweep supports window radius 0..%d (got %d)", kMaxR, d->r);
    if (d->H < 0 || d->W < 0 || d->nh < 0 || d->i0 < 0)
        return fail(OM_ESHAPE, "negative size");
    if (!tap_size_ok(d->H, d->W))
        return fail(OM_ESHAPE, "sweep: reference images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported");
    if (d->top_k < 1)
        return fail(OM_EINVAL, "top_k must be >= 1");
    const int k = d->top_k < d->n_src ? d->top_k : d->n_src;
    if (k > 4)
        return fail(OM_EUNSUP, "reduced-precision sweep supports min(top_k, n) <= 4 (got %d)", k);
    if (d->hyp_mode < OM_HYP_MAPS || d->hyp_mode > OM_HYP_CONST)
        return fail(OM_EINVAL, "unknown hypothesis generator %d", d->hyp_mode);
    SweepLPK P;
    memset(&P, 0, sizeof P);
    P.pinhole = 1;
    P.ref = ref16;
    P.mu = d->mu;
    P.sd = d->sd;
    P.ray = d->ray;
    P.H = d->H;
    P.W = d->W;
    P.n = d->n_src;
    P.r = d->r;
    P.k = k;
    P.mode = d->hyp_mode;
    P.ha = d->hyp_a;
    P.hb = d->hyp_b;
    P.hoff = d->hyp_offset;
    P.i0 = d->i0;
    P.nh = d->nh;
    P.best = d->best;
    P.prev = d->prev;
    P.spb = d->s_prev_best;
    P.snb = d->s_next_best;
    P.idx = d->idx;
    for (int s = 0; s < P.n; ++s) {
        const om_src &S = d->src[s];
        if (S.h < 1 || S.w < 1 || !S.img)
            return fail(OM_ESHAPE, "source %d is empty", s);
        if (!tap_size_ok(S.h, S.w))
            return fail(OM_ESHAPE, "source %d: images of 2^30 pixels or more, or with a side of 2^24 or more, are "
                                   "not supported", s);
        SrcLP &Q = P.src[s];
        Q.k1 = (float)S.geo[13];
        Q.k2 = (float)S.geo[14];
        Q.k3 = (float)S.geo[15];
        if (S.geo[13] != 0.0 || S.geo[14] != 0.0 || S.geo[15] != 0.0)
            P.pinhole = 0;
        Q.img = S.img;
        Q.h = S.h;
        Q.w = S.w;
        for (int t = 0; t < 12; ++t)
            Q.g[t] = (float)S.geo[t];
        Q.f = (float)S.geo[12];
        Q.cx = (float)S.geo[16];
        Q.cy = (float)S.geo[17];
        Q.wl = (float)((double)S.w - 1.001);
        Q.hl = (float)((double)S.h - 1.001);
    }
    P.tiles_x = (int)cdiv(P.W, kTW);
    P.tiles_y = (int)cdiv(P.H, kTH);
    const int64_t tiles = (int64_t)P.tiles_x * P.tiles_y;
    if (tiles == 0 || P.nh == 0)
        return OM_OK;
    const bool cd = P.mode == OM_HYP_CONST;
    const size_t lds = lp_lds(P.r, cd);
    using KFn = void (*)(const SweepLPK);
#define OM_LP_R(K_, FMT_, CD_)                                                                                    \
    {sweep_lp_kernel<K_, FMT_, CD_, 2>, sweep_lp_kernel<K_, FMT_, CD_, 5>, sweep_lp_kernel<K_, FMT_, CD_, -1>}
#define OM_LP_K(K_, FMT_) {OM_LP_R(K_, FMT_, false), OM_LP_R(K_, FMT_, true)}
#define OM_LP_ROW(FMT_) {OM_LP_K(1, FMT_), OM_LP_K(2, FMT_), OM_LP_K(3, FMT_), OM_LP_K(4, FMT_)}
    static const KFn table[3][4][2][3] = {OM_LP_ROW(OM_FMT_F16), OM_LP_ROW(OM_FMT_BF16), OM_LP_ROW(OM_FMT_F32)};
#undef OM_LP_ROW
#undef OM_LP_K
#undef OM_LP_R
    hipStream_t st = as_stream(stream);
    ProfScope ps("sweep_lp", st);
    hipLaunchKernelGGL(table[fmt - OM_FMT_F16][k - 1][cd][lp_radius_class(P.r)], dim3((unsigned)tiles),
                       dim3(kThreads), lds, st, P);
    return launch_check("sweep_lp_kernel");
}
