// Fused plane sweep for gfx950: every hypothesis of one reference view in one launch.
//
// Reference semantics: _dense.sweep_hypothesis (orthomosaic/_dense.pyx:740-838) with the row
// kernels of orthomosaic/_sweep.h:33-214, looped over hypotheses by densify._sweep
// (orthomosaic/densify.py:336-342).
//
// Work decomposition.  A workgroup (256 threads, 4 wavefronts) owns a 64-column x TH-row tile of
// output pixels and a contiguous chunk of hypotheses.  For each (hypothesis, source) it
//   W: warps the tile plus an r-pixel halo into LDS (FP64 projection + FP64 bilinear, the
//      reference's exact operation sequence; NaN marks an invalid sample),
//   H: forms the horizontal window sums of V, J, J^2, J*I for every halo row (FP64 running sums
//      over 32-column runs, stored as float32 exactly where the reference rounds them),
//   V: forms the vertical sums in FP64 and the NCC of each output pixel, and folds it into a
//      per-pixel top-k kept in registers.
// After the last source of a hypothesis the per-pixel winner state (best, prev, s_prev_best,
// s_next_best, idx) is updated in registers; it is read once and written once per launch.
//
// Exactness.  The library is compiled with -ffp-contract=off and uses IEEE division and square
// root, so every sample, window sum and score is computed with the reference's operations in the
// reference's order.  The reference forms window sums as running sums along the whole row
// (_sweep.h:96-113) and down the whole column (_sweep.h:139-146); the kernel forms them over a
// 32-column run and a TH/4-row run.  Both equal the exact sum whenever no partial sum rounds
// (float32 inputs, FP64 accumulators: any window of values whose magnitudes span less than
// 2^29), so the float32 outputs are bit-identical to the reference in practice and within an
// FP64 rounding of it otherwise.
//
// Hypothesis chunking.  Small images give too few tiles to fill 256 CUs, so the hypotheses are
// split into chunks processed by different workgroups.  Each chunk records its local winner and
// the scores at its ends; om_sweep_merge replays the streaming winner rule over the chunks in
// order, which reproduces the sequential result exactly.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "om_common.h"

namespace om {

constexpr int kThreads = 256;                     // threads per workgroup
constexpr int kMaxR = 7;                          // largest window radius of the fused kernel
// Tile shapes (output columns x rows), named by their height: 32 x 32 (the default), 64 x 16 and
// 64 x 8 (less LDS and fewer registers, more workgroups).  At r = 5 the square tile's 42 x 42 halo
// is warped in 7 rounds of the workgroup's lanes where 64 x 16's 75 x 26 takes 8, and its
// horizontal pass fits in three waves: the 48-view block's sweep 1324 against 1431 ms, a 640x480
// view 1.72 against 1.84 ms (both at three workgroups per CU); at r = 2 the two shapes tie
// (1068 ms; 1.43 and 1.44 ms).  The
// vertical pass gives each thread VR = 4 (or, at 8 rows, 2) output rows of one column: TW columns
// x 256 / TW row segments.
constexpr int kTHDefault = 32;
__host__ __device__ constexpr int tile_w(int th) { return th == 32 ? 32 : 64; }
__host__ __device__ constexpr int vseg_of(int tw) { return kThreads / tw; }
__host__ __device__ constexpr int twp_of(int tw) { return tw + 1; }  // row stride of the horizontal sums
constexpr int kPartPlanes = 6;                    // chunk partials: best, idx, prev-of-best, next-of-best, first, last

struct SrcK {
    const float *img;
    int h, w;
    double g[18];
};

struct SweepK {
    const float *ref, *mu, *sd, *ray;
    int H, W, n, r, k;        // k = min(top_k, n)
    int mode;                 // OM_HYP_*
    const float *ha, *hb;
    int hoff;
    int i0, nh, hc, chunks;   // hypotheses [i0, i0+nh), hc per chunk
    int TH, tiles_x, tiles_y;
    int pinhole;              // every source has k1 = k2 = k3 = 0
    float *best, *prev, *spb, *snb;
    int *idx;
    float *part;              // (chunks, 6, H, W) when chunks > 1
    float *depth_out;         // batched launches: parabola depth of the view, or null
    int fin_n;                // hypotheses of the whole sweep (the parabola's n): i0 + nh
    SrcK src[16];
};

template <class D>  // SweepK, or SweepK in the constant address space (batched launches)
__device__ __forceinline__ float hyp_value(const D &P, int i, int p, float a, float b)
{
    // a, b: the generator's per-pixel inputs at p (base/step or iu/ds)
    switch (P.mode) {
    case OM_HYP_LINEAR: {
        const float t = (float)i * b;  // _fast.pyx:638-640, no contraction
        return a + t;
    }
    case OM_HYP_REFINE: {
        float v = (float)(i - P.hoff) * b;  // _fast.pyx:654-658
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
__device__ __forceinline__ void topk_insert(float (&t)[K], float v)
{
    // keep t sorted descending; equal values keep their order (irrelevant for the sum)
#pragma unroll
    for (int a = K - 1; a >= 0; --a) {
        const bool lt = t[a] < v;
        const float up = (a > 0) ? t[a - 1] : v;
        const bool prev_lt = (a > 0) ? (t[a - 1] < v) : true;
        if (lt)
            t[a] = prev_lt ? up : v;
    }
}

// The horizontal sums' rows are padded to an odd number of words, and so are the 64-wide tiles'
// halo rows, so that neither the horizontal pass (each half-wave over 8 rows x 4 runs) nor the
// vertical pass (lanes spread over columns) has LDS bank conflicts.  The 32-wide tiles' halo rows
// stay unpadded: a pad column adds a warp round at r = 5 (42 x 43 > 7 x 256 pixels), which cost
// more than the horizontal pass's two-way conflicts save (48-view block: 1397 against 1324 ms).
// Pad words between rows that put each half-wave's rows in distinct banks without the extra round
// (the precision kernel's layout) cut SQ_LDS_BANK_CONFLICT by three quarters here but not the time
// (r = 5: 1324.7 against 1321.5 ms; r = 2: 1073 against 1067): this kernel waits on its VALU.

// The horizontal pass: one task per (halo row, run of kHRun columns), spread over all four waves
// (the passes are separated by barriers, so a pass on fewer waves lengthens every workgroup's
// step: 32-column runs on one wave issue a quarter fewer instructions yet made the 48-view block's
// sweep 19 % slower).
constexpr int kHRun = 8;

__host__ __device__ constexpr int halo_stride(int r, int tw) { return tw == 64 ? (tw + 2 * r) | 1 : tw + 2 * r; }

// Constant planes are warped in groups of kPlaneGroup per pass over the halo: the plane-independent
// part of the projection, Rs * ray (_sweep.h:51-53 forms it as its own parenthesised term), is
// computed once per halo pixel and source for the group.
#ifndef OM_SWEEP_GROUP
#define OM_SWEEP_GROUP 2
#endif
constexpr int kPlaneGroup = OM_SWEEP_GROUP;

__host__ __device__ constexpr int plane_group(bool constd) { return constd ? kPlaneGroup : 1; }

// bytes of the valid-sample counts of the horizontal windows (one byte each, at most 2 r + 1)
__host__ __device__ constexpr size_t count_bytes(int hr, int tw) { return ((size_t)hr * twp_of(tw) + 3) & ~(size_t)3; }

// LDS bytes of one workgroup.  Per halo pixel its image pixel index (except at r = 5), the
// reference and one warped source value per plane of the group, plus (hypotheses that vary per
// pixel) the depth of the current plane; three float planes of horizontal sums (J, J^2, J I) and
// their byte counts.  At r = 5 with constant planes that is 39 188 bytes for 32 x 32 tiles (four
// workgroups per CU) and 45 372 for 64 x 16 (three).
// Window radii with kernels of their own (the window loops unrolled, a horizontal run's samples
// kept in registers): 2 (5x5, densify's default) and 5 (11x11); any other radius <= kMaxR runs the
// generic kernels (R = -1, radius read from the descriptor).
__host__ __device__ constexpr int radius_class(int r) { return r == 2 ? 0 : (r == 5 ? 1 : 2); }

// Whether the kernels of a radius class keep each halo pixel's image index in LDS.  The radius-5
// kernels recompute it instead (for a tile whose halo lies inside the image, base + hp + hy (W - HC):
// 4 instructions), which brings a 32 x 32 tile's LDS from 46 244 to 39 188 bytes, so four
// workgroups share a CU: the 48-view block's sweep 1322 -> 1218 ms, same bits (with the index
// recomputed but three workgroups: 1393 ms).  Keeping instead each lane's horizontal task's
// reference values in registers (the other 7 KB) spilled hot values: 1430 ms.  Kernels that keep
// a top-3 or top-4 per pixel need more registers, and 64-wide tiles at r = 5 need 45 372 bytes
// even without the indices: both keep the index in LDS at three workgroups.
__host__ __device__ constexpr bool idx_lds(int k, int rclass, int tw) { return rclass != 1 || k > 2 || tw != 32; }

__host__ __device__ constexpr size_t sweep_lds(int th, int r, int k, bool constd)
{
    return (size_t)(th + 2 * r) * halo_stride(r, tile_w(th)) *
               ((idx_lds(k, radius_class(r), tile_w(th)) ? 4 : 0) + (constd ? 4 + 4 * plane_group(true) : 16)) +
           (size_t)3 * (th + 2 * r) * twp_of(tile_w(th)) * 4 + count_bytes(th + 2 * r, tile_w(th));
}

// Waves per SIMD the fused sweep kernels are compiled for, that is 256-thread workgroups per CU.
// The radius-2 kernels and the radius-5 kernels of 32 x 32 tiles, of top-1 and top-2: four (at
// most 128 VGPRs; their LDS fits four times into 160 KB).  They then spill 32-124 bytes a lane,
// mostly the per-pixel winner state that is touched once per plane group, and still gain: a
// 640x480 view at r = 2 1.43 -> 1.23 ms, the 48-view block at r = 2 1070 -> 924 ms.  The generic
// kernels (other radii, up to 7), the radius-5 kernels of 64-wide tiles (whose LDS fits three
// times) and those of top-3 and top-4: three (at most 168 VGPRs).
#ifndef OM_SWEEP_WAVES
#define OM_SWEEP_WAVES 3
#endif
#ifndef OM_SWEEP_WAVES_R2
#define OM_SWEEP_WAVES_R2 4
#endif
#ifndef OM_SWEEP_WAVES_R5
#define OM_SWEEP_WAVES_R5 4
#endif
__host__ __device__ constexpr int sweep_waves(int K, int TH, int R)
{
    return K > 2 ? OM_SWEEP_WAVES
                 : (R == 2 ? OM_SWEEP_WAVES_R2 : (R == 5 && TH == 32 ? OM_SWEEP_WAVES_R5 : OM_SWEEP_WAVES));
}
#define OM_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(sweep_waves(K, TH, R), sweep_waves(K, TH, R))))

// Horizontal window sums of one run (_sweep.h:85-114): columns c0 .. c0 + kHRun - 1 of halo row
// Jr / Ir, sums of V (count), J, J^2 and J I over 2 r + 1 columns, formed in FP64 and stored as
// float exactly where the reference rounds them.  Each column adds the entering sample, then
// subtracts the leaving one (the reference's order).  CACHE (a compile-time radius R >= 0): each
// sample's validity and products are formed once and kept in registers until it leaves the
// window; else the leaving sample is read again.
template <int R, bool CACHE>
__device__ __forceinline__ void hsum_run(const float *Jr, const float *Ir, int r, float *h1, float *h2, float *h3,
                                         uint8_t *hc)
{
    if constexpr (R >= 0 && CACHE) {
        constexpr int NE = kHRun + 2 * R;
        float ej[NE], ejj[NE], eji[NE];
        int eok[NE];
        auto elem = [&](int e) {
            const float J = Jr[e];
            const bool ok = J == J;
            const float jv = ok ? J : 0.f;
            ej[e] = jv;
            ejj[e] = jv * jv;
            eji[e] = jv * Ir[e];
            eok[e] = ok;
        };
        int sv = 0;
        double sj = 0.0, sjj = 0.0, sij = 0.0;
#pragma unroll
        for (int e = 0; e <= 2 * R; ++e) {
            elem(e);
            sv += eok[e];
            sj = sj + (double)ej[e];
            sjj = sjj + (double)ejj[e];
            sij = sij + (double)eji[e];
        }
        hc[0] = (uint8_t)sv;
        h1[0] = (float)sj;
        h2[0] = (float)sjj;
        h3[0] = (float)sij;
#pragma unroll
        for (int c = 1; c < kHRun; ++c) {
            const int en = c + 2 * R, eo = c - 1;
            elem(en);
            sv += eok[en] - eok[eo];
            sj = sj + (double)ej[en];
            sjj = sjj + (double)ejj[en];
            sij = sij + (double)eji[en];
            sj = sj - (double)ej[eo];
            sjj = sjj - (double)ejj[eo];
            sij = sij - (double)eji[eo];
            hc[c] = (uint8_t)sv;
            h1[c] = (float)sj;
            h2[c] = (float)sjj;
            h3[c] = (float)sij;
        }
    } else {
        int sv = 0;
        double sj = 0.0, sjj = 0.0, sij = 0.0;
        for (int t = 0; t <= 2 * r; ++t) {
            const float J = Jr[t];
            const bool ok = J == J;
            const float jv = ok ? J : 0.f;
            sv += ok;
            sj = sj + (double)jv;
            sjj = sjj + (double)(jv * jv);
            sij = sij + (double)(jv * Ir[t]);
        }
        hc[0] = (uint8_t)sv;
        h1[0] = (float)sj;
        h2[0] = (float)sjj;
        h3[0] = (float)sij;
        for (int c = 1; c < kHRun; ++c) {
            const float Jn = Jr[c + 2 * r], Jo = Jr[c - 1];
            const bool okn = Jn == Jn, oko = Jo == Jo;
            const float jn = okn ? Jn : 0.f, jo = oko ? Jo : 0.f;
            sv += (int)okn - (int)oko;
            sj = sj + (double)jn;
            sjj = sjj + (double)(jn * jn);
            sij = sij + (double)(jn * Ir[c + 2 * r]);
            sj = sj - (double)jo;
            sjj = sjj - (double)(jo * jo);
            sij = sij - (double)(jo * Ir[c - 1]);
            hc[c] = (uint8_t)sv;
            h1[c] = (float)sj;
            h2[c] = (float)sjj;
            h3[c] = (float)sij;
        }
    }
}

// One (tile, hypothesis chunk) work item of one view; P may live in the kernel arguments or in
// device memory (batched launches).  CD: the hypotheses are constant planes (OM_HYP_CONST), so the
// depth of a plane is one value for the whole tile and needs no LDS; they are warped in groups of
// G = kPlaneGroup.  TH: the tile's height (8, 16 or 32 rows), which sets its width.  R: the window
// radius when it has kernels of its own, else -1 (P.r).
template <int K, int TH, bool CD, int R, class D>
__device__ __forceinline__ void sweep_tile(const D &P, int tile, int chunk, double *lds_d)
{
    constexpr int G = plane_group(CD);
    constexpr int TW = tile_w(TH), VR = TH / vseg_of(TW), kTWP = twp_of(TW);
    const int r = R >= 0 ? R : P.r;
    const int HR = TH + 2 * r;          // halo rows
    const int HC = TW + 2 * r;          // halo columns
    const int HCP = halo_stride(r, TW); // row stride
    const int NHP = HR * HCP;           // padded halo; the pad column is outside the image
    double *sD = lds_d;                                          // (HR, HCP) depth of the current hypothesis (!CD)
    constexpr bool IDX = idx_lds(K, R >= 0 ? radius_class(R) : 2, TW);
    // The horizontal pass keeps its samples in registers except in the four-wave radius-5 kernels:
    // at 128 VGPRs the cache made them spill more than it saved (the 48-view block's sweep 1205 ->
    // 1182 ms without it, spills 312 -> 112 bytes; at r = 2 the cache still pays, 926 -> 899 ms).
    constexpr bool HCACHE = R >= 0 && (R == 2 || sweep_waves(K, TH, R) < 4);
    int *sP = CD ? (int *)lds_d : (int *)(sD + NHP);             // (HR, HCP) pixel index, -1 outside
    float *sI = IDX ? (float *)(sP + NHP) : (float *)sP;       // (HR, HCP) reference
    float *sJ = sI + NHP;                                        // (G, HR, HCP) warped source, NaN = invalid
    float *sH = sJ + G * NHP;                                    // (3, HR, kTWP) horizontal sums of J, J^2, J I
    uint8_t *sC = (uint8_t *)(sH + 3 * HR * kTWP);               // (HR, kTWP) valid samples of each window

    const int tid = threadIdx.x;
    const int tx0 = (tile % P.tiles_x) * TW;
    const int ty0 = (tile / P.tiles_x) * TH;
    const int ia = P.i0 + chunk * P.hc;
    const int ib = min(P.i0 + P.nh, ia + P.hc);
    const int H = P.H, W = P.W;
    const size_t HW = (size_t)H * W;
    const bool gen_maps = P.mode == OM_HYP_LINEAR || P.mode == OM_HYP_REFINE;
    const int gx0 = tx0 - r, gy0 = ty0 - r;  // image position of halo pixel (0, 0)
    const bool inner = HCP == HC && gx0 >= 0 && gy0 >= 0 && gx0 + HC <= W && gy0 + HR <= H;
    const int pin0 = inner ? gy0 * W + gx0 : 0;
    // image pixel of the halo pixel hp, -1 outside the image (or in the pad column)
    auto halo_px = [&](int hp) -> int {
        const int hy = hp / HCP, hx = hp - hy * HCP;
        const int gy = gy0 + hy, gx = gx0 + hx;
        return (hx < HC && gy >= 0 && gy < H && gx >= 0 && gx < W) ? gy * W + gx : -1;
    };

    // ---- the tile + halo: pixel index and reference value
    for (int hp = tid; hp < NHP; hp += kThreads) {
        const int p = halo_px(hp);
        if (IDX)
            sP[hp] = p;
        sI[hp] = p >= 0 ? P.ref[p] : 0.f;
    }

    // ---- output pixels of this thread in the vertical pass: column vc, rows vy0 .. vy0+VR-1
    const int vc = tid % TW, vseg = tid / TW, vy0 = vseg * VR;
    const int ox = tx0 + vc;
    float mu_[VR], sd_[VR];
    float st_best[VR], st_prev[VR], st_spb[VR], st_snb[VR], st_first[VR];
    int st_idx[VR];
    float tk[G][VR][K];
    const bool chunked = P.chunks > 1;
#pragma unroll
    for (int j = 0; j < VR; ++j) {
        const int oy = ty0 + vy0 + j;
        const bool in = oy < H && ox < W;
        const size_t p = in ? (size_t)oy * W + ox : 0;
        mu_[j] = in ? P.mu[p] : 0.f;
        sd_[j] = in ? P.sd[p] : 1.f;
        if (chunked) {
            st_best[j] = -INFINITY;
            st_prev[j] = NAN;
            st_spb[j] = NAN;
            st_snb[j] = NAN;
            st_idx[j] = INT_MIN;
        } else {
            st_best[j] = in ? P.best[p] : 0.f;
            st_prev[j] = in ? P.prev[p] : 0.f;
            st_spb[j] = in ? P.spb[p] : 0.f;
            st_snb[j] = in ? P.snb[p] : 0.f;
            st_idx[j] = in ? P.idx[p] : 0;
        }
        st_first[j] = NAN;
    }
    const double kk = 1.0 / (double)((2 * r + 1) * (2 * r + 1));
    __syncthreads();

    for (int i = ia; i < ib; i += G) {
        const int ng = min(G, ib - i);   // planes of this group (wave-uniform)
        // inverse depth -> depth, per halo pixel (shared by all sources) or once per constant
        // plane; the previous hypothesis's last warp pass finished reading sD before the barrier
        // after it
        double dc[G];
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
            dc[gg] = 0.0;
        if (CD) {
#pragma unroll
            for (int gg = 0; gg < G; ++gg)
                if (gg < ng)
                    dc[gg] = 1.0 / (double)P.ha[i + gg - P.i0];
        } else {
            for (int hp = tid; hp < NHP; hp += kThreads) {
                const int p = IDX ? sP[hp] : halo_px(hp);
                double d = NAN;
                if (p >= 0) {
                    const float a = gen_maps ? P.ha[p] : 0.f, b = gen_maps ? P.hb[p] : 0.f;
                    d = 1.0 / (double)hyp_value(P, i, p, a, b);
                }
                sD[hp] = d;
            }
        }
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
#pragma unroll
            for (int j = 0; j < VR; ++j)
#pragma unroll
                for (int a = 0; a < K; ++a)
                    tk[gg][j][a] = -INFINITY;
        if (!CD)
            __syncthreads();

        for (int s = 0; s < P.n; ++s) {
            const auto &S = P.src[s];
            const auto *g = S.g;
            const double wl = (double)S.w - 1.001, hl = (double)S.h - 1.001;
            const float *img = S.img;
            const int sw = S.w;
            // ---- W: warp the halo into LDS (_sweep.h:49-80) for every plane of the group.  Each
            // step loads the ray of its next halo pixel (pixel 0's outside the image) before it
            // works on its own.
            const char *rayx = (const char *)P.ray, *rayy = (const char *)(P.ray + HW),
                       *rayz = (const char *)(P.ray + 2 * HW);
            auto hpx = [&](int hp) -> int {
                if (IDX)
                    return sP[hp];
                if (inner) {  // the whole halo in the image, no pad column: no bounds to check
                    const int hy = hp / HCP;
                    return pin0 + hp + hy * (W - HCP);
                }
                return halo_px(hp);
            };
            int pn = tid < NHP ? hpx(tid) : -1;
            int pn2 = tid + kThreads < NHP ? hpx(tid + kThreads) : -1;
            uint32_t on = (uint32_t)max(pn, 0) << 2;
            float rxn = *(const float *)(rayx + on), ryn = *(const float *)(rayy + on),
                  rzn = *(const float *)(rayz + on);
            for (int hp = tid; hp < NHP; hp += kThreads) {
                float J[G];
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    J[gg] = NAN;
                const int p = pn;
                const double rx = rxn, ry = ryn, rz = rzn;
                pn = pn2;
                pn2 = hp + 2 * kThreads < NHP ? hpx(hp + 2 * kThreads) : -1;
                on = (uint32_t)max(pn, 0) << 2;
                rxn = *(const float *)(rayx + on);
                ryn = *(const float *)(rayy + on);
                rzn = *(const float *)(rayz + on);
                if (p >= 0) {
                    // Rs * ray, the plane-independent term of b + d * (Rs * ray)
                    const double ax = (g[0] * rx + g[1] * ry) + g[2] * rz;
                    const double ay = (g[3] * rx + g[4] * ry) + g[5] * rz;
                    const double az = (g[6] * rx + g[7] * ry) + g[8] * rz;
#pragma unroll
                    for (int gg = 0; gg < G; ++gg) {
                        if (gg < ng) {
                            const double d = CD ? dc[gg] : sD[hp];
                            const double X = g[9] + d * ax;
                            const double Y = g[10] + d * ay;
                            const double Z = g[11] + d * az;
                            if (!(Z <= 1e-6)) {
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
                                if (u >= 0.0 && v >= 0.0 && u <= wl && v <= hl) {
                                    const double val = bil_f64(img, sw, u, v);
                                    if (isfinite(val))
                                        J[gg] = (float)val;
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    sJ[gg * NHP + hp] = J[gg];
            }
            __syncthreads();
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
                if (gg >= ng)
                    break;
                if (gg > 0)
                    __syncthreads();  // the previous plane's V pass has read sH
                const float *sJg = sJ + gg * NHP;
                // ---- H: horizontal window sums of one 8-column run per task.  Each half-wave
                // takes 8 rows x 4 runs (64-wide tiles, lane l: row (l / 4) % 8, run l % 4 +
                // 4 (l / 32); 32-wide: row (l / 4) % 8 + 8 (l / 32), run l % 4), and with the odd
                // row strides every half-wave's addresses fall in distinct LDS banks
                // (SQ_LDS_BANK_CONFLICT of a 2000x1500, S = 8, 128-plane sweep: 0 against 1.1e9
                // cycles with 8 rows x 8 runs per wave; the single-view sweep 0.4-1.7 % faster, the
                // 48-view block at half scale 368.5 against 368.8 ms).
                constexpr int RPW = TW == 64 ? 8 : 16;  // halo rows a wave's 64 tasks cover
                const int ntask = (HR + RPW - 1) / RPW * 64;
                for (int t = tid; t < ntask; t += kThreads) {
                    const int l = t & 63;
                    const int hr = (t >> 6) * RPW + ((l >> 2) & 7) + (TW == 64 ? 0 : 8 * (l >> 5));
                    const int c0 = ((l & 3) + (TW == 64 ? 4 * (l >> 5) : 0)) * kHRun;
                    if (hr >= HR)
                        continue;
                    const int o = hr * kTWP + c0;
                    hsum_run<R, HCACHE>(sJg + hr * HCP + c0, sI + hr * HCP + c0, r, sH + o, sH + HR * kTWP + o,
                                sH + 2 * HR * kTWP + o, sC + o);
                }
                __syncthreads();
                // ---- V: vertical sums and NCC (_sweep.h:116-178)
                {
                    const float *h1 = sH + vc, *h2 = h1 + HR * kTWP, *h3 = h2 + HR * kTWP;
                    const uint8_t *h0 = sC + vc;
                    int c0 = 0;  // valid samples: a small integer, exact either way
                    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
#pragma unroll
                    for (int t = 0; t <= 2 * r; ++t) {
                        const int o = (vy0 + t) * kTWP;
                        c0 += h0[o];
                        c1 = c1 + (double)h1[o];
                        c2 = c2 + (double)h2[o];
                        c3 = c3 + (double)h3[o];
                    }
#pragma unroll
                    for (int j = 0; j < VR; ++j) {
                        if (j > 0) {
                            const int on = (vy0 + j + 2 * r) * kTWP, oo = (vy0 + j - 1) * kTWP;
                            double t;
                            c0 += (int)h0[on] - (int)h0[oo];
                            t = c1 + (double)h1[on];
                            c1 = t - (double)h1[oo];
                            t = c2 + (double)h2[on];
                            c2 = t - (double)h2[oo];
                            t = c3 + (double)h3[on];
                            c3 = t - (double)h3[oo];
                        }
                        float out;
                        const double cov = (double)c0 * kk;
                        if (cov <= 0.9) {
                            out = -1.0f;
                        } else {
                            const double num[3] = {c1 * kk, c2 * kk, c3 * kk};
                            double mq[3];
                            quotients_shared<3>(num, cov, mq);  // cov > 0.9: IEEE quotients
                            const double mj = mq[0], m2 = mq[1], mij = mq[2];
                            double sj = m2 - mj * mj;
                            sj = sj > 1e-6 ? (sj < INFINITY ? sqrt_normal(sj) : sqrt(sj)) : 1e-3;
                            const double q = (mij - (double)mu_[j] * mj) / ((double)sd_[j] * sj);
                            out = q > 1.0 ? 1.0f : (q < -1.0 ? -1.0f : (float)q);
                        }
                        topk_insert<K>(tk[gg][j], out);
                    }
                }
            }
            // the next source's W pass writes sJ only (last read before the barrier above) and
            // its H pass writes sH after the next barrier, so no barrier is needed here
        }
        // ---- combine + streaming winner of each plane of the group, in order (_sweep.h:181-214)
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            if (gg >= ng)
                break;
            const int ii = i + gg;
#pragma unroll
            for (int j = 0; j < VR; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int a = 0; a < K; ++a)
                    s = s + tk[gg][j][a];
                s = s / (float)K;
                if (ii == ia)
                    st_first[j] = s;
                if (st_idx[j] == ii - 1)
                    st_snb[j] = s;
                if (s > st_best[j]) {
                    st_best[j] = s;
                    st_spb[j] = st_prev[j];
                    st_snb[j] = chunked ? NAN : -2.0f;
                    st_idx[j] = ii;
                }
                st_prev[j] = s;
            }
        }
    }

#pragma unroll
    for (int j = 0; j < VR; ++j) {
        const int oy = ty0 + vy0 + j;
        if (oy < H && ox < W) {
            const size_t p = (size_t)oy * W + ox;
            if (chunked) {
                float *pc = P.part + (size_t)chunk * kPartPlanes * HW;
                pc[p] = st_best[j];
                pc[HW + p] = __int_as_float(st_idx[j]);
                pc[2 * HW + p] = st_spb[j];
                pc[3 * HW + p] = st_snb[j];
                pc[4 * HW + p] = st_first[j];
                pc[5 * HW + p] = st_prev[j];
            } else {
                P.best[p] = st_best[j];
                P.prev[p] = st_prev[j];
                P.spb[p] = st_spb[j];
                P.snb[p] = st_snb[j];
                P.idx[p] = st_idx[j];
            }
        }
    }
}

template <int K, int TH, bool CD, int R>
__global__ void __launch_bounds__(kThreads) OM_SWEEP_ATTR sweep_kernel(const SweepK P)
{
    extern __shared__ double lds_d[];
    const int lb = xcd_block(blockIdx.x, gridDim.x);  // tile-major: an XCD sweeps a band of tiles
    sweep_tile<K, TH, CD, R>(P, lb / P.chunks, lb % P.chunks, lds_d);
}

// Persistent batch: work items (view, tile, chunk) of the views vidx[0..nv) from one queue; item0[j]
// is the first item of view vidx[j].  Every wave of a workgroup reads the same item index from LDS,
// so all of them leave the loop together once the queue is empty.
template <int K, int TH, bool CD, int R>
__global__ void __launch_bounds__(kThreads) OM_SWEEP_ATTR sweep_batch_kernel(const SweepK *views, const int *vidx, const int *item0,
                                                               int nv, int total, int *counter)
{
    extern __shared__ double lds_d[];
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0)
            s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);  // the same for every lane
        if (it >= total)
            break;
        int lo = 0, hi = nv - 1;  // last view whose first item is <= it
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (as_const(item0)[mid] <= it)
                lo = mid;
            else
                hi = mid - 1;
        }
        const OM_CONST_AS SweepK &P = as_const(views)[as_const(vidx)[lo]];
        const int local = it - as_const(item0)[lo];
        sweep_tile<K, TH, CD, R>(P, local / P.chunks, local % P.chunks, lds_d);
    }
}

// Replays the streaming winner rule (_sweep.h:204-212) over the chunks in hypothesis order.
__device__ __forceinline__ void merge_px(const float *__restrict__ part, int chunks, int64_t HW, float *best,
                                         float *prev, float *spb, float *snb, int *idx, int i0, int hc, int64_t p)
{
    float gb = best[p], gp = prev[p], gs = spb[p], gn = snb[p];
    int gi = idx[p];
    for (int c = 0; c < chunks; ++c) {
        const float *pc = part + (size_t)c * kPartPlanes * HW;
        const float cb = pc[p], cp = pc[2 * HW + p], cn = pc[3 * HW + p], cf = pc[4 * HW + p], cl = pc[5 * HW + p];
        const int ci = __float_as_int(pc[HW + p]);
        const int a = i0 + c * hc;
        if (gi == a - 1)
            gn = cf;
        if (cb > gb) {
            gb = cb;
            gi = ci;
            gs = (cp != cp) ? gp : cp;
            if (cn != cn)
                gn = (c + 1 < chunks) ? part[(size_t)(c + 1) * kPartPlanes * HW + 4 * HW + p] : -2.0f;
            else
                gn = cn;
        }
        gp = cl;
    }
    best[p] = gb;
    prev[p] = gp;
    spb[p] = gs;
    snb[p] = gn;
    idx[p] = gi;
}

__global__ void sweep_merge_kernel(const float *__restrict__ part, int chunks, int64_t HW, float *best, float *prev,
                                   float *spb, float *snb, int *idx, int i0, int hc)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < HW)
        merge_px(part, chunks, HW, best, prev, spb, snb, idx, i0, hc, p);
}

// parabola of densify._sweep (densify.py:354-361), NumPy's float32 / float64 promotion kept
template <class D>
__device__ __forceinline__ void finalize_px(const D &P, int64_t p, float *depth)
{
    const float a = (P.mode == OM_HYP_LINEAR || P.mode == OM_HYP_REFINE) ? P.ha[p] : 0.f;
    const float b = (P.mode == OM_HYP_LINEAR || P.mode == OM_HYP_REFINE) ? P.hb[p] : 0.f;
    const float inv0 = hyp_value(P, 0, (int)p, a, b);
    const int n = P.fin_n;
    const float dinv = n > 1 ? hyp_value(P, 1, (int)p, a, b) - inv0 : 0.0f;
    const float best = P.best[p], spb = P.spb[p], snb = P.snb[p];
    const int idx = P.idx[p];
    const float den = (spb - 2.0f * best) + snb;
    const bool interior = idx > 0 && idx < n - 1 && den < -1e-6f;
    float off = interior ? (0.5f * (spb - snb)) / den : 0.0f;
    off = off < -0.5f ? -0.5f : (off > 0.5f ? 0.5f : off);
    const double inv = (double)inv0 + ((double)idx + (double)off) * (double)dinv;
    const double m = (inv > 1e-9 || inv != inv) ? inv : 1e-9;
    depth[p] = (float)(1.0 / m);
}

__global__ void sweep_finalize_kernel(SweepK P, float *depth)
{
    const int64_t HW = (int64_t)P.H * P.W;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < HW)
        finalize_px(P, p, depth);
}

// batched merge + parabola over the pixels of all views: view v owns [px0[v], px0[v+1])
__device__ __forceinline__ int view_of_px(const int64_t *px0, int nviews, int64_t t)
{
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (px0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void sweep_batch_post_kernel(const SweepK *views, const int64_t *px0, int nviews, int do_merge,
                                        int do_final)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= px0[nviews])
        return;
    const int v = view_of_px(px0, nviews, t);
    const SweepK &P = views[v];
    const int64_t p = t - px0[v];
    const int64_t HW = (int64_t)P.H * P.W;
    if (do_merge && P.chunks > 1)
        merge_px(P.part, P.chunks, HW, P.best, P.prev, P.spb, P.snb, P.idx, P.i0, P.hc, p);
    if (do_final && P.depth_out)
        finalize_px(P, p, P.depth_out);
}

// parabola of the views fidx[0..nf): view fidx[j] owns pixels [fpx0[j], fpx0[j+1])
__global__ void sweep_multi_finalize_kernel(const SweepK *views, const int *fidx, const int64_t *fpx0, int nf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fpx0[nf])
        return;
    const int j = view_of_px(fpx0, nf, t);
    const OM_CONST_AS SweepK &P = as_const(views)[as_const(fidx)[j]];
    finalize_px(P, t - fpx0[j], P.depth_out);
}

__global__ void reset_counters_kernel(int *c, int n)
{
    if ((int)threadIdx.x < n)
        c[threadIdx.x] = 0;
}

// a batch run starts every view from the initial winner state of densify._sweep
// (densify.py:331-335): best = prev = s_prev_best = s_next_best = -2, idx = -1
__global__ void sweep_batch_reset_kernel(const SweepK *views, const int64_t *px0, int nviews)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= px0[nviews])
        return;
    const int v = view_of_px(px0, nviews, t);
    const SweepK &P = views[v];
    const int64_t p = t - px0[v];
    P.best[p] = -2.0f;
    P.prev[p] = -2.0f;
    P.spb[p] = -2.0f;
    P.snb[p] = -2.0f;
    P.idx[p] = -1;
}

// fused: the limits of the fused kernels (1..16 sources, r <= kMaxR, min(top_k, n) <= 4); without
// them (the parabola, which reads only the winner state) any source count of om_sweep_rows is taken
// and the sources are not copied
static int build_params(const om_sweep_desc *d, SweepK &P, bool fused = true)
{
    if (!d)
        return fail(OM_EINVAL, "om_sweep: null descriptor");
    if (d->n_src < 1 || d->n_src > (fused ? 16 : 256))
        return fail(OM_ENSRC, "sweep needs 1..%d sources (got %d)", fused ? 16 : 256, d->n_src);
    if (d->r < 0)
        return fail(OM_EINVAL, "negative window radius");
    if (d->H < 0 || d->W < 0 || d->nh < 0)
        return fail(OM_ESHAPE, "negative size");
    if (!tap_size_ok(d->H, d->W))
        return fail(OM_ESHAPE, "sweep: reference images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported");
    if (fused && d->r > kMaxR)
        return fail(OM_EUNSUP, "fused sweep supports window radius <= %d (got %d; om_sweep_rows takes any)", kMaxR,
                    d->r);
    if (d->top_k < 1)
        return fail(OM_EINVAL, "top_k must be >= 1");
    const int k = d->top_k < d->n_src ? d->top_k : d->n_src;
    if (fused && k > 4)
        return fail(OM_EUNSUP, "fused sweep supports min(top_k, n) <= 4 (got %d; om_sweep_rows takes any)", k);
    if (d->hyp_mode < OM_HYP_MAPS || d->hyp_mode > OM_HYP_CONST)
        return fail(OM_EINVAL, "unknown hypothesis generator %d", d->hyp_mode);
    if (!d->src)
        return fail(OM_EINVAL, "null source list");
    memset(&P, 0, sizeof P);
    P.ref = d->ref;
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
    P.fin_n = d->i0 + d->nh;
    P.nh = d->nh;
    P.best = d->best;
    P.prev = d->prev;
    P.spb = d->s_prev_best;
    P.snb = d->s_next_best;
    P.idx = d->idx;
    P.pinhole = 1;
    for (int s = 0; s < (fused ? d->n_src : 0); ++s) {
        const om_src &o = d->src[s];
        if (!tap_size_ok(o.h, o.w))
            return fail(OM_ESHAPE, "source %d: images of 2^30 pixels or more, or with a side of 2^24 or more, are not supported", s);
        P.src[s].img = o.img;
        P.src[s].h = o.h;
        P.src[s].w = o.w;
        for (int t = 0; t < 18; ++t)
            P.src[s].g[t] = o.geo[t];
        if (o.geo[13] != 0.0 || o.geo[14] != 0.0 || o.geo[15] != 0.0)
            P.pinhole = 0;
    }
    if (d->tile_h != 0 && d->tile_h != 8 && d->tile_h != 16 && d->tile_h != 32)
        return fail(OM_EINVAL, "tile height must be 8, 16 or 32 rows (got %d)", d->tile_h);
    P.TH = d->tile_h ? d->tile_h : kTHDefault;
    P.tiles_x = (int)cdiv(d->W, tile_w(P.TH));
    P.tiles_y = (int)cdiv(d->H, P.TH);
    return OM_OK;
}

// kernel instances: top-k 1..4, tile height 8, 16 or 32, constant planes or not, radius class
using SingleFn = void (*)(const SweepK);
using BatchFn = void (*)(const SweepK *, const int *, const int *, int, int, int *);

template <int K, int TH, bool CD>
static SingleFn single_r(int rc)
{
    return rc == 0 ? sweep_kernel<K, TH, CD, 2> : (rc == 1 ? sweep_kernel<K, TH, CD, 5> : sweep_kernel<K, TH, CD, -1>);
}

template <int K, int TH, bool CD>
static BatchFn batch_r(int rc)
{
    return rc == 0 ? sweep_batch_kernel<K, TH, CD, 2>
                   : (rc == 1 ? sweep_batch_kernel<K, TH, CD, 5> : sweep_batch_kernel<K, TH, CD, -1>);
}

template <int K>
static SingleFn single_k(int th, bool cd, int rc)
{
    if (th == 32)
        return cd ? single_r<K, 32, true>(rc) : single_r<K, 32, false>(rc);
    if (th == 16)
        return cd ? single_r<K, 16, true>(rc) : single_r<K, 16, false>(rc);
    return cd ? single_r<K, 8, true>(rc) : single_r<K, 8, false>(rc);
}

template <int K>
static BatchFn batch_k(int th, bool cd, int rc)
{
    if (th == 32)
        return cd ? batch_r<K, 32, true>(rc) : batch_r<K, 32, false>(rc);
    if (th == 16)
        return cd ? batch_r<K, 16, true>(rc) : batch_r<K, 16, false>(rc);
    return cd ? batch_r<K, 8, true>(rc) : batch_r<K, 8, false>(rc);
}

static SingleFn single_kernel(int k, int th, bool cd, int r)
{
    const int rc = radius_class(r);
    switch (k) {
    case 1: return single_k<1>(th, cd, rc);
    case 2: return single_k<2>(th, cd, rc);
    case 3: return single_k<3>(th, cd, rc);
    default: return single_k<4>(th, cd, rc);
    }
}

static BatchFn batch_kernel(int k, int th, bool cd, int r)
{
    const int rc = radius_class(r);
    switch (k) {
    case 1: return batch_k<1>(th, cd, rc);
    case 2: return batch_k<2>(th, cd, rc);
    case 3: return batch_k<3>(th, cd, rc);
    default: return batch_k<4>(th, cd, rc);
    }
}

// Hypothesis chunks of a single-view launch (desc.chunks when set).  All workgroups of a launch
// do the same work, so tiles x c of them run in ceil(tiles c / slots) rounds, slots = resident
// workgroups of the chip, each round as long as one chunk of ceil(nh / c) planes plus the tile's
// set-up and its partials (about one plane): the count with the fewest round-planes, the fewest
// chunks among equals.  640x480 (300 tiles, 768 slots, 128 planes): 5 chunks, 2 full rounds of 26
// planes, where 4 chunks left the second of 2 rounds half empty.  Only counts whose partials fit
// in `limit` bytes are considered (one chunk needs none).
static size_t part_bytes(int chunks, const SweepK &P);

// Resident workgroups per CU of a kernel at an LDS size on the current device, remembered per
// (device, kernel, LDS) so that a sweep call does not query the runtime each time.  fallback when
// the query fails.
static int blocks_per_cu(int dev, const void *fn, size_t lds, int fallback)
{
    struct Entry {
        int dev;
        const void *fn;
        size_t lds;
        int per_cu;
    };
    static std::mutex mu;
    static std::vector<Entry> seen;
    {
        std::lock_guard<std::mutex> g(mu);
        for (const Entry &e : seen)
            if (e.dev == dev && e.fn == fn && e.lds == lds)
                return e.per_cu;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, lds) != hipSuccess || per_cu < 1)
        return fallback;  // not remembered: a later call asks again
    std::lock_guard<std::mutex> g(mu);
    seen.push_back({dev, fn, lds, per_cu});
    return per_cu;
}

static int choose_chunks(const om_sweep_desc *d, const SweepK &P, size_t limit = SIZE_MAX)
{
    const int64_t tiles = std::max<int64_t>((int64_t)P.tiles_x * P.tiles_y, 1);
    const int nh = std::max(d->nh, 1);
    int c = d->chunks;
    if (c <= 0) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        const bool cd = P.mode == OM_HYP_CONST;
        const int per_cu = blocks_per_cu(dev, (const void *)single_kernel(P.k, P.TH, cd, P.r),
                                         sweep_lds(P.TH, P.r, P.k, cd), 3);
        const int64_t slots = (int64_t)cus * per_cu;
        // Among the counts that give at least two full rounds of resident workgroups (when any
        // does), the fewest round-planes.  With fewer workgroups than that, CUs run below full
        // occupancy for much of the launch: a 600x800 view, S = 12, 192 planes, at four workgroups
        // per CU took 8.5 ms as 2 chunks (one round, the fewest round-planes) and 6.6 ms as 6.
        for (int pass = 0; pass < 2; ++pass) {
            int64_t best = INT64_MAX;
            int pick = 0;
            for (int t = 1; t <= std::min(64, nh); ++t) {
                if (t > 1 && part_bytes(t, P) > limit)
                    break;
                if (pass == 0 && tiles * t < 2 * slots)
                    continue;
                const int64_t cost = cdiv(tiles * t, slots) * (cdiv(nh, t) + 1);
                if (cost < best) {
                    best = cost;
                    pick = t;
                }
            }
            c = pick;
            if (pick > 0)
                break;
        }
        c = std::max(c, 1);
    }
    c = std::max(1, std::min(c, nh));
    return c;
}

// The sweep of one view over hypotheses [i0, i0 + nh) in `chunks` hypothesis chunks, whose
// partial winner states go to `part` (chunks * 6 * H * W floats) when chunks > 1.
static int sweep_launch(SweepK &Pl, int chunks, float *part, hipStream_t st, bool prof = true)
{
    int rc;
    Pl.hc = (int)cdiv(Pl.nh, chunks);
    Pl.chunks = (int)cdiv(Pl.nh, Pl.hc);
    Pl.part = part;
    const bool cd = Pl.mode == OM_HYP_CONST;
    const size_t lds = sweep_lds(Pl.TH, Pl.r, Pl.k, cd);
    const dim3 grid((unsigned)((int64_t)Pl.tiles_x * Pl.tiles_y * Pl.chunks));
    ProfScope ps(prof ? "sweep" : nullptr, st);
    hipLaunchKernelGGL(single_kernel(Pl.k, Pl.TH, cd, Pl.r), grid, dim3(kThreads), lds, st, Pl);
    if ((rc = launch_check("sweep_kernel")))
        return rc;
    if (Pl.chunks > 1) {
        const int64_t HW = (int64_t)Pl.H * Pl.W;
        ProfScope pm(prof ? "sweep_merge" : nullptr, st);
        hipLaunchKernelGGL(sweep_merge_kernel, dim3((unsigned)cdiv(HW, 256)), dim3(256), 0, st, Pl.part, Pl.chunks,
                           HW, Pl.best, Pl.prev, Pl.spb, Pl.snb, Pl.idx, Pl.i0, Pl.hc);
        if ((rc = launch_check("sweep_merge_kernel")))
            return rc;
    }
    return OM_OK;
}

static size_t part_bytes(int chunks, const SweepK &P)
{
    return chunks > 1 ? (size_t)chunks * kPartPlanes * (size_t)P.H * P.W * sizeof(float) : 0;
}

// Chunk partials of om_sweep_noalloc: one array per GPU (each device that loads the library holds
// its own copy until the process ends).  Host code must not pass the array's name to a kernel (it
// is a host-side stand-in); it uses the device address from hipGetSymbolAddress, looked up once per
// device on the first call there.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_ptr(void **out)
{
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc)
        return rc;
    if (dev < 0 || dev >= 64)
        return fail(OM_EUNSUP, "om_sweep_noalloc: device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if ((rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")))
            return rc;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OM_OK;
}

}  // namespace om

using namespace om;

extern "C" size_t om_sweep_workspace_bytes(const om_sweep_desc *d)
{
    SweepK P;
    if (build_params(d, P) != OM_OK)
        return 0;
    return part_bytes(choose_chunks(d, P), P);
}

extern "C" int om_sweep(const om_sweep_desc *d, void *stream)
{
    SweepK Pl;
    int rc = build_params(d, Pl);
    if (rc)
        return rc;
    if (Pl.H == 0 || Pl.W == 0 || Pl.nh == 0)
        return OM_OK;
    // automatic chunks: the best count whose partials fit the caller's workspace
    int chunks = choose_chunks(d, Pl, d->workspace ? d->workspace_bytes : 0);
    if (chunks > 1 && (!d->workspace || d->workspace_bytes < part_bytes(chunks, Pl)))
        chunks = 1;  // no room for the requested partials: one chunk
    return sweep_launch(Pl, chunks, (float *)d->workspace, as_stream(stream));
}

extern "C" size_t om_sweep_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int om_sweep_noalloc(const om_sweep_desc *d, void *stream)
{
    SweepK Pl;
    int rc = build_params(d, Pl);
    if (rc)
        return rc;
    const int chunks = choose_chunks(d, Pl, kNoAllocScratchBytes);  // automatic counts fit the scratch
    const size_t need = part_bytes(chunks, Pl);
    if (need > kNoAllocScratchBytes)
        return fail(OM_ESHAPE,
                    "om_sweep_noalloc: %d hypothesis chunks of a %dx%d view need %zu bytes of partials, more "
                    "than its %zu-byte scratch (pass fewer chunks, or om_sweep with a workspace)",
                    chunks, Pl.H, Pl.W, need, kNoAllocScratchBytes);
    if (Pl.H == 0 || Pl.W == 0 || Pl.nh == 0)
        return OM_OK;
    void *part = nullptr;
    if (chunks > 1 && (rc = scratch_ptr(&part)))
        return rc;
    // no profiling scopes here: they create HIP events
    return sweep_launch(Pl, chunks, (float *)part, as_stream(stream), false);
}

extern "C" int om_sweep_hypothesis(const float *ref, const float *mu, const float *sd, int32_t H, int32_t W,
                                   int32_t n, const float *const *imgs, const int32_t *img_hw, const float *ray,
                                   const double *geo, const float *inv, int32_t r, int32_t top_k, int32_t i,
                                   float *best, float *prev, float *s_prev_best, float *s_next_best, int32_t *idx,
                                   void *stream)
{
    if (n < 1 || n > 16)
        return fail(OM_ENSRC, "sweep_hypothesis needs 1..16 sources");
    if (!imgs || !img_hw || !geo)
        return fail(OM_EINVAL, "null source arrays");
    om_src src[16];
    for (int s = 0; s < n; ++s) {
        src[s].img = imgs[s];
        src[s].h = img_hw[2 * s];
        src[s].w = img_hw[2 * s + 1];
        memcpy(src[s].geo, geo + 18 * s, sizeof(double) * 18);
    }
    om_sweep_desc d;
    memset(&d, 0, sizeof d);
    d.ref = ref;
    d.mu = mu;
    d.sd = sd;
    d.ray = ray;
    d.H = H;
    d.W = W;
    d.n_src = n;
    d.src = src;
    d.r = r;
    d.top_k = top_k;
    d.hyp_mode = OM_HYP_MAPS;
    d.hyp_a = inv;
    d.i0 = i;
    d.nh = 1;
    d.best = best;
    d.prev = prev;
    d.s_prev_best = s_prev_best;
    d.s_next_best = s_next_best;
    d.idx = idx;
    d.chunks = 1;
    return om_sweep(&d, stream);
}

extern "C" int om_sweep_finalize(const om_sweep_desc *d, float *depth, void *stream)
{
    SweepK P;
    int rc = build_params(d, P, false);
    if (rc)
        return rc;
    if (d->i0 != 0)
        return fail(OM_EINVAL, "om_sweep_finalize expects the full hypothesis range (i0 = 0)");
    if (d->nh < 1)
        return fail(OM_EINVAL, "the parabola needs at least one hypothesis");
    if (d->hyp_mode == OM_HYP_MAPS && d->nh > 1 && !d->hyp_a)
        return fail(OM_EINVAL, "null hypothesis maps");
    const int64_t HW = (int64_t)P.H * P.W;
    if (HW == 0)
        return OM_OK;
    hipLaunchKernelGGL(sweep_finalize_kernel, dim3((unsigned)cdiv(HW, 256)), dim3(256), 0, as_stream(stream), P,
                       depth);
    return launch_check("sweep_finalize_kernel");
}

// ------------------------------------------------------------------------------------------
// batched sweep (om_batch): descriptors of all views in device memory, one persistent launch
// ------------------------------------------------------------------------------------------
#include "batch.h"

namespace om {

// persistent launches of a batch: one per (tile height, constant planes) pair
static int plan_groups(const SweepK *P, int nviews, int k, int r, int cus, int grid_blocks, std::vector<int> &vidx,
                       std::vector<int> &item0, SweepGroup *groups, int &ng, bool skip_empty)
{
    ng = 0;
    for (int th : {32, 16, 8}) {
        for (int cdi = 1; cdi >= 0; --cdi) {
            const bool cd = cdi != 0;
            SweepGroup &g = groups[ng];
            g.th = th;
            g.cd = cd;
            g.first = (int)vidx.size();
            int64_t total = 0;
            for (int v = 0; v < nviews; ++v) {
                if (P[v].TH != th || (P[v].mode == OM_HYP_CONST) != cd || (skip_empty && P[v].nh < 1))
                    continue;
                vidx.push_back(v);
                item0.push_back((int)total);
                total += (int64_t)P[v].tiles_x * P[v].tiles_y * P[v].chunks;
                if (total >= INT_MAX)
                    return fail(OM_EUNSUP, "batch has too many work items");
            }
            g.nv = (int)vidx.size() - g.first;
            if (g.nv == 0 || total == 0)
                continue;
            g.total = (int)total;
            g.lds = sweep_lds(th, r, k, cd);
            int per_cu = 1;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)batch_kernel(k, th, cd, r), kThreads,
                                                             g.lds) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            g.grid = grid_blocks > 0 ? grid_blocks : (int)std::min<int64_t>((int64_t)per_cu * cus, total);
            ++ng;
        }
    }
    return OM_OK;
}

int batch_sweep_set(om_batch *b, const om_sweep_desc *views, int nviews, float *const *depth_out, int chunks,
                    int tile_h, int grid_blocks)
{
    if (nviews < 1 || !views)
        return fail(OM_EINVAL, "batch needs at least one view");
    if (tile_h != 0 && tile_h != 8 && tile_h != 16 && tile_h != 32)
        return fail(OM_EINVAL, "tile height must be 8, 16 or 32 rows (got %d)", tile_h);
    std::vector<SweepK> P(nviews);
    int k = 0, r = -1;
    int64_t tiles = 0;
    for (int v = 0; v < nviews; ++v) {
        om_sweep_desc d = views[v];
        if (tile_h)
            d.tile_h = tile_h;
        int rc = build_params(&d, P[v]);
        if (rc)
            return fail(rc, "view %d: %s", v, om_last_error());
        if (v == 0) {
            k = P[v].k;
            r = P[v].r;
        } else if (P[v].k != k || P[v].r != r) {
            return fail(OM_EINVAL, "all views of a batch need the same window radius and min(top_k, n)");
        }
        if (d.i0 != 0)  // every run starts each view from the initial winner state
            return fail(OM_EINVAL, "view %d: a batch sweeps hypotheses from i0 = 0", v);
        if (depth_out && depth_out[v] && d.nh < 1)
            return fail(OM_EINVAL, "view %d: the parabola needs at least one hypothesis", v);
        tiles += (int64_t)P[v].tiles_x * P[v].tiles_y;
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    // chunks: the batch argument for all views, else a view's own desc.chunks, else enough chunks
    // for about 4 workgroups per CU over the whole batch
    const int c_auto = (int)std::min<int64_t>(std::max<int64_t>(cdiv(4LL * cus, std::max<int64_t>(tiles, 1)), 1), 64);
    std::vector<int64_t> px0(nviews + 1, 0);
    bool merge = false, final_ = false;
    for (int v = 0; v < nviews; ++v) {
        SweepK &Q = P[v];
        const int c = chunks > 0 ? chunks : (views[v].chunks > 0 ? views[v].chunks : c_auto);
        const int cv = std::max(1, std::min(c, std::max(Q.nh, 1)));
        Q.hc = (int)cdiv(std::max(Q.nh, 1), cv);
        Q.chunks = (int)cdiv(std::max(Q.nh, 1), Q.hc);
        const int64_t HW = (int64_t)Q.H * Q.W;
        if (Q.chunks > 1) {
            Q.part = (float *)b->dev_alloc((size_t)Q.chunks * kPartPlanes * HW * sizeof(float));
            if (!Q.part)
                return fail(OM_ENOMEM, "batch: chunk partials");
            merge = true;
        }
        Q.depth_out = depth_out ? depth_out[v] : nullptr;
        final_ = final_ || Q.depth_out;
        px0[v + 1] = px0[v] + HW;
    }
    // one persistent launch per (tile height, constant planes), each with its own queue
    std::vector<int> vidx, item0;
    int rc = plan_groups(P.data(), nviews, k, r, cus, grid_blocks, vidx, item0, b->sw_groups, b->sw_ngroups, false);
    if (rc)
        return rc;
    b->sw_views = b->dev_alloc(sizeof(SweepK) * nviews);
    b->sw_vidx = (int *)b->dev_alloc(sizeof(int) * nviews);
    b->sw_item0 = (int *)b->dev_alloc(sizeof(int) * nviews);
    b->sw_px0 = (int64_t *)b->dev_alloc(sizeof(int64_t) * (nviews + 1));
    b->sw_counter = (int *)b->dev_alloc(sizeof(int) * kSweepGroups);
    if (!b->sw_views || !b->sw_vidx || !b->sw_item0 || !b->sw_px0 || !b->sw_counter)
        return fail(OM_ENOMEM, "batch: descriptor tables");
    if ((rc = hip_check(hipMemcpy(b->sw_views, P.data(), sizeof(SweepK) * nviews, hipMemcpyHostToDevice), "upload")) ||
        (rc = hip_check(hipMemcpy(b->sw_vidx, vidx.data(), sizeof(int) * nviews, hipMemcpyHostToDevice), "upload")) ||
        (rc = hip_check(hipMemcpy(b->sw_item0, item0.data(), sizeof(int) * nviews, hipMemcpyHostToDevice), "upload")) ||
        (rc = hip_check(hipMemcpy(b->sw_px0, px0.data(), sizeof(int64_t) * (nviews + 1), hipMemcpyHostToDevice),
                        "upload")))
        return rc;
    b->sw_n = nviews;
    b->sw_k = k;
    b->sw_r = r;
    b->sw_px = px0[nviews];
    b->sw_merge = merge;
    b->sw_final = final_;
    return OM_OK;
}

int batch_sweep_enqueue(om_batch *b, hipStream_t st)
{
    if (b->sw_n == 0)
        return OM_OK;
    ProfScope ps("sweep_batch", st);
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, st, b->sw_counter, kSweepGroups);
    hipLaunchKernelGGL(sweep_batch_reset_kernel, dim3((unsigned)cdiv(b->sw_px, 256)), dim3(256), 0, st,
                       (const SweepK *)b->sw_views, b->sw_px0, b->sw_n);
    for (int gi = 0; gi < b->sw_ngroups; ++gi) {
        const SweepGroup &g = b->sw_groups[gi];
        ProfScope pk("sweep_batch_kernel", st);
        hipLaunchKernelGGL(batch_kernel(b->sw_k, g.th, g.cd, b->sw_r), dim3((unsigned)g.grid), dim3(kThreads), g.lds, st,
                           (const SweepK *)b->sw_views, b->sw_vidx + g.first, b->sw_item0 + g.first, g.nv, g.total,
                           b->sw_counter + gi);
    }
    int rc = launch_check("sweep_batch_kernel");
    if (rc)
        return rc;
    if (b->sw_merge || b->sw_final) {
        hipLaunchKernelGGL(sweep_batch_post_kernel, dim3((unsigned)cdiv(b->sw_px, 256)), dim3(256), 0, st,
                           (const SweepK *)b->sw_views, b->sw_px0, b->sw_n, (int)b->sw_merge, (int)b->sw_final);
        rc = launch_check("sweep_batch_post_kernel");
    }
    return rc;
}

}  // namespace om

// ------------------------------------------------------------------------------------------
// one-shot multi-view sweep step (om_sweep_multi): hypotheses [i0, i0 + nh) of every view,
// continuing each view's winner state -- the step of a continuous-batching server
// ------------------------------------------------------------------------------------------
namespace om {

struct MultiLayout {
    size_t views, vidx, item0, fidx, fpx0, counter, total;
};

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

static MultiLayout multi_layout(int n)
{
    MultiLayout L;
    size_t o = 0;
    L.views = o;
    o += al256(sizeof(SweepK) * n);
    L.vidx = o;
    o += al256(sizeof(int) * n);
    L.item0 = o;
    o += al256(sizeof(int) * n);
    L.fidx = o;
    o += al256(sizeof(int) * n);
    L.fpx0 = o;
    o += al256(sizeof(int64_t) * (n + 1));
    L.counter = o;
    o += 256;
    L.total = o;
    return L;
}

}  // namespace om

extern "C" size_t om_sweep_multi_bytes(int32_t nviews)
{
    return nviews < 1 ? 0 : multi_layout(nviews).total;
}

extern "C" int om_sweep_multi(const om_sweep_desc *views, int32_t nviews, float *const *depth_out, int32_t tile_h,
                              void *table, size_t table_bytes, void *staging, void *stream)
{
    if (nviews < 1 || !views)
        return fail(OM_EINVAL, "om_sweep_multi needs at least one view");
    if (tile_h != 0 && tile_h != 8 && tile_h != 16 && tile_h != 32)
        return fail(OM_EINVAL, "tile height must be 8, 16 or 32 rows (got %d)", tile_h);
    const MultiLayout L = multi_layout(nviews);
    if (!table || table_bytes < L.total || !staging)
        return fail(OM_EINVAL, "om_sweep_multi needs a device table and host staging of %zu bytes", L.total);
    char *host = (char *)staging;
    SweepK *P = (SweepK *)(host + L.views);
    int k = 0, r = -1, nf = 0;
    int *fidx = (int *)(host + L.fidx);
    int64_t *fpx0 = (int64_t *)(host + L.fpx0);
    fpx0[0] = 0;
    for (int v = 0; v < nviews; ++v) {
        om_sweep_desc d = views[v];
        if (tile_h)
            d.tile_h = tile_h;
        int rc = build_params(&d, P[v]);
        if (rc)
            return fail(rc, "view %d: %s", v, om_last_error());
        if (v == 0) {
            k = P[v].k;
            r = P[v].r;
        } else if (P[v].k != k || P[v].r != r) {
            return fail(OM_EINVAL, "all views of a step need the same window radius and min(top_k, n)");
        }
        SweepK &Q = P[v];
        Q.hc = std::max(Q.nh, 1);
        Q.chunks = 1;
        Q.part = nullptr;
        Q.depth_out = depth_out ? depth_out[v] : nullptr;
        if (Q.depth_out) {
            if (Q.fin_n < 1)
                return fail(OM_EINVAL, "view %d: the parabola needs at least one hypothesis", v);
            fidx[nf] = v;
            fpx0[nf + 1] = fpx0[nf] + (int64_t)Q.H * Q.W;
            ++nf;
        }
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    SweepGroup groups[kSweepGroups];
    int ng = 0;
    std::vector<int> gv, gi0;
    int rc0 = plan_groups(P, nviews, k, r, cus, 0, gv, gi0, groups, ng, true);
    if (rc0)
        return rc0;
    std::copy(gv.begin(), gv.end(), (int *)(host + L.vidx));
    std::copy(gi0.begin(), gi0.end(), (int *)(host + L.item0));
    hipStream_t st = as_stream(stream);
    char *dtab = (char *)table;
    int rc = hip_check(hipMemcpyAsync(dtab, host, L.counter, hipMemcpyHostToDevice, st), "table upload");
    if (rc)
        return rc;
    const SweepK *dviews = (const SweepK *)(dtab + L.views);
    int *counter = (int *)(dtab + L.counter);
    ProfScope ps("sweep_multi", st);
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, st, counter, kSweepGroups);
    for (int gi = 0; gi < ng; ++gi) {
        const SweepGroup &g = groups[gi];
        hipLaunchKernelGGL(batch_kernel(k, g.th, g.cd, r), dim3((unsigned)g.grid), dim3(kThreads), g.lds, st, dviews,
                           (const int *)(dtab + L.vidx) + g.first, (const int *)(dtab + L.item0) + g.first, g.nv,
                           g.total, counter + gi);
    }
    if ((rc = launch_check("sweep_batch_kernel")))
        return rc;
    if (nf) {
        hipLaunchKernelGGL(sweep_multi_finalize_kernel, dim3((unsigned)cdiv(fpx0[nf], 256)), dim3(256), 0, st, dviews,
                           (const int *)(dtab + L.fidx), (const int64_t *)(dtab + L.fpx0), nf);
        rc = launch_check("sweep_multi_finalize_kernel");
    }
    return rc;
}

