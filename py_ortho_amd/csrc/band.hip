// Per-pixel inverse-depth search band of a depth map's coarse sweep on gfx950: densify._depth_band
// (orthomosaic/densify.py:233-274) and the linear hypothesis step of densify.depth_map
// (densify.py:409-411).
//
// Input: the view's own tie points already projected into the coarse camera (u, v, z) in float64
// (z > 0).  Three kernels:
//   clear  - per block of `block` x `block` pixels: lowest depth, highest depth, point count;
//   bin    - one thread per point: the block of the point (NumPy's floor division of the clipped
//            coordinates), then 64-bit atomic min / max of z's bit pattern (z > 0, so the bit
//            patterns order like the values) and a count;
//   expand - one thread per pixel: the 3x3 block neighbourhood of its block (blocks with >= 5
//            points only), widened by margin + 5 % of the span and clipped to the view range, or
//            the view range itself; then inv_lo = 1 / hi, inv_hi = 1 / max(lo, 1e-6) rounded to
//            float32, and the step (inv_hi - inv_lo) / max(nd - 1, 1) in float32.
// The FP64 operations are the reference's, in its order, so the maps are bit-identical.
#include <cmath>

#include "om_common.h"

namespace om {
namespace band {

constexpr uint64_t kPosInfBits = 0x7FF0000000000000ull;

__global__ void clear_kernel(uint64_t *zlo, uint64_t *zhi, int *cnt, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) {
        zlo[b] = kPosInfBits;
        zhi[b] = 0;   // +0.0: below every z > 0
        cnt[b] = 0;
    }
}

// a // b for a >= 0, b > 0 as NumPy's float floor_divide computes it (npy_divmod)
__device__ __forceinline__ double np_floordiv(double a, double b)
{
    const double mod = fmod(a, b);
    const double div = (a - mod) / b;
    double fl = floor(div);
    if (div - fl > 0.5)
        fl += 1.0;
    return fl;
}

__global__ void bin_kernel(const double *__restrict__ uvz, int64_t n, int H, int W, int block, int gw,
                           uint64_t *zlo, uint64_t *zhi, int *cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double u = uvz[3 * i], v = uvz[3 * i + 1], z = uvz[3 * i + 2];
    if (!(u >= 0.0 && v >= 0.0 && u < (double)W && v < (double)H))
        return;
    const double vc = fmin(fmax(v, 0.0), (double)(H - 1)), uc = fmin(fmax(u, 0.0), (double)(W - 1));
    const int by = (int)np_floordiv(vc, (double)block), bx = (int)np_floordiv(uc, (double)block);
    const int b = by * gw + bx;
    const uint64_t bits = (uint64_t)__double_as_longlong(z);
    atomicMin((unsigned long long *)&zlo[b], (unsigned long long)bits);
    atomicMax((unsigned long long *)&zhi[b], (unsigned long long)bits);
    atomicAdd(&cnt[b], 1);
}

struct ExpandK {
    const uint64_t *zlo, *zhi;
    const int *cnt;
    int H, W, block, gh, gw;
    double lo_g, hi_g, margin;
    int nd;
    int uniform;              // fewer than 50 own observations: the view range everywhere
    float *inv_lo, *inv_hi, *step;
};

__global__ void expand_kernel(const ExpandK P)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)P.H * P.W)
        return;
    const int y = (int)(p / P.W), x = (int)(p - (int64_t)y * P.W);
    double lo_b = P.lo_g, hi_b = P.hi_g;
    float ilo, ihi;
    if (P.uniform) {
        ilo = (float)(1.0 / P.hi_g);
        ihi = (float)(1.0 / P.lo_g);
    } else {
        const int by = y / P.block, bx = x / P.block;
        const int b = by * P.gw + bx;
        if (P.cnt[b] >= 5) {
            double lo3 = INFINITY, hi3 = -INFINITY;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int yy = by + dy, xx = bx + dx;
                    if (yy < 0 || yy >= P.gh || xx < 0 || xx >= P.gw)
                        continue;
                    const int q = yy * P.gw + xx;
                    if (P.cnt[q] >= 5) {
                        lo3 = fmin(lo3, __longlong_as_double((long long)P.zlo[q]));
                        hi3 = fmax(hi3, __longlong_as_double((long long)P.zhi[q]));
                    }
                }
            // the block itself has points, so both are finite
            const double pad = P.margin + 0.05 * (hi3 - lo3);
            lo_b = fmin(fmax(lo3 - pad, P.lo_g), P.hi_g);
            hi_b = fmin(fmax(hi3 + pad, P.lo_g), P.hi_g);
        }
        ilo = (float)(1.0 / hi_b);
        ihi = (float)(1.0 / fmax(lo_b, 1e-6));
    }
    P.inv_lo[p] = ilo;
    if (P.inv_hi)
        P.inv_hi[p] = ihi;
    if (P.step) {
        const float d = ihi - ilo;
        P.step[p] = d / (float)(P.nd > 1 ? P.nd - 1 : 1);
    }
}

size_t ws_bytes(int H, int W, int block)
{
    const int64_t gh = cdiv(H, block), gw = cdiv(W, block);
    return (size_t)(gh * gw) * (8 + 8 + 4) + 256;
}

}  // namespace band
}  // namespace om

using namespace om;

extern "C" size_t om_depth_band_workspace_bytes(int32_t H, int32_t W, int32_t block)
{
    if (H < 1 || W < 1 || block < 1)
        return 0;
    return band::ws_bytes(H, W, block);
}

extern "C" int om_depth_band(const double *uvz, int64_t n, int32_t n_obs, int32_t H, int32_t W, int32_t block,
                             double lo_g, double hi_g, double margin, int32_t nd, float *inv_lo, float *inv_hi,
                             float *step, void *workspace, size_t workspace_bytes, void *stream)
{
    if (H < 0 || W < 0 || n < 0)
        return fail(OM_ESHAPE, "depth_band: negative size");
    if (block < 1)
        return fail(OM_EINVAL, "depth_band: block must be >= 1");
    if (nd < 1)
        return fail(OM_EINVAL, "depth_band: nd must be >= 1");
    if (!inv_lo)
        return fail(OM_EINVAL, "depth_band: null inv_lo");
    if ((int64_t)H * W == 0)
        return OM_OK;
    const bool uniform = n_obs < 50;
    const int gh = (int)cdiv(H, block), gw = (int)cdiv(W, block);
    const size_t need = band::ws_bytes(H, W, block);
    if (!uniform && (!workspace || workspace_bytes < need))
        return fail(OM_EINVAL, "depth_band: workspace of %zu bytes needed", need);
    if (!uniform && n > 0 && !uvz)
        return fail(OM_EINVAL, "depth_band: null points");
    hipStream_t st = as_stream(stream);
    band::ExpandK P{};
    const int nb = gh * gw;
    if (!uniform) {
        auto *zlo = (uint64_t *)workspace;
        auto *zhi = zlo + nb;
        auto *cnt = (int *)(zhi + nb);
        hipLaunchKernelGGL(band::clear_kernel, dim3((unsigned)cdiv(nb, 256)), dim3(256), 0, st, zlo, zhi, cnt, nb);
        int rc = launch_check("depth_band clear");
        if (rc)
            return rc;
        if (n > 0) {
            hipLaunchKernelGGL(band::bin_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, uvz, n, H, W, block,
                               gw, zlo, zhi, cnt);
            if ((rc = launch_check("depth_band bin")))
                return rc;
        }
        P.zlo = zlo;
        P.zhi = zhi;
        P.cnt = cnt;
    }
    P.H = H;
    P.W = W;
    P.block = block;
    P.gh = gh;
    P.gw = gw;
    P.lo_g = lo_g;
    P.hi_g = hi_g;
    P.margin = margin;
    P.nd = nd;
    P.uniform = uniform;
    P.inv_lo = inv_lo;
    P.inv_hi = inv_hi;
    P.step = step;
    hipLaunchKernelGGL(band::expand_kernel, dim3((unsigned)cdiv((int64_t)H * W, 256)), dim3(256), 0, st, P);
    return launch_check("depth_band expand");
}
