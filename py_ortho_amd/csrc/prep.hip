// Matching-image preparation on gfx950: _fast.depth_view (_fast.pyx:577-616), which densify's
// load_view (densify.py:528-535) runs on every decoded view: uint8 RGB x gain + bias in float32,
// grey = ((r + g + b) / 3 - 128) / 64 (NaN where the mask is False), colour = uint8(clip(v + 0.5,
// 0, 255)).  The reference's float32 operations in its order; the decoded image goes to the device
// as 3 bytes a pixel instead of a 4-byte grey.
//
// Four pixels a thread: 12 bytes of RGB and 4 of mask in, 16 bytes of grey and 12 of colour out,
// as dword loads and stores (HBM-bound: 15 bytes read or written a pixel).
#include "om_common.h"

namespace om {
namespace prep {

struct ViewK {
    const uint8_t *rgb;
    const uint8_t *mask;   // NULL: no mask
    float *grey;
    uint8_t *rgb8;
    int64_t n;             // pixels
    float g[3], b[3];
    int has_g, has_b;
};

__device__ __forceinline__ uint8_t clip_u8(float x)
{
    const float c = x < 0.0f ? 0.0f : (x > 255.0f ? 255.0f : x);   // _clip32 (_fast.pyx:357-362)
    return (uint8_t)c;
}

__device__ __forceinline__ void one_pixel(const ViewK &p, uint8_t r, uint8_t g, uint8_t b, bool masked, float &grey,
                                          uint8_t out[3])
{
    float v[3] = {(float)r, (float)g, (float)b};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (p.has_g)
            v[c] = v[c] * p.g[c];
        if (p.has_b)
            v[c] = v[c] + p.b[c];
    }
    float s = v[0] + v[1];
    s = s + v[2];
    s = s / 3.0f;
    s = s - 128.0f;
    s = s / 64.0f;
    grey = masked ? __builtin_nanf("") : s;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[c] = clip_u8(v[c] + 0.5f);
}

// VEC: four pixels with dword accesses (pointers 4-byte aligned, grey 16-byte aligned)
template <bool VEC>
__global__ __launch_bounds__(256) void depth_view_kernel(ViewK p)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t p0 = q * 4;
    if (p0 >= p.n)
        return;
    if (VEC && p0 + 4 <= p.n) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(p.rgb + 3 * p0);
        const uint32_t w0 = src[0], w1 = src[1], w2 = src[2];
        const uint8_t px[12] = {(uint8_t)w0, (uint8_t)(w0 >> 8), (uint8_t)(w0 >> 16), (uint8_t)(w0 >> 24),
                                (uint8_t)w1, (uint8_t)(w1 >> 8), (uint8_t)(w1 >> 16), (uint8_t)(w1 >> 24),
                                (uint8_t)w2, (uint8_t)(w2 >> 8), (uint8_t)(w2 >> 16), (uint8_t)(w2 >> 24)};
        const uint32_t mw = p.mask ? *reinterpret_cast<const uint32_t *>(p.mask + p0) : 0x01010101u;
        float4 gv;
        uint8_t o[12];
        float *gp = &gv.x;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            one_pixel(p, px[3 * i], px[3 * i + 1], px[3 * i + 2], ((mw >> (8 * i)) & 0xFF) == 0, gp[i], &o[3 * i]);
        *reinterpret_cast<float4 *>(p.grey + p0) = gv;
        uint32_t *dst = reinterpret_cast<uint32_t *>(p.rgb8 + 3 * p0);
        dst[0] = o[0] | (uint32_t)o[1] << 8 | (uint32_t)o[2] << 16 | (uint32_t)o[3] << 24;
        dst[1] = o[4] | (uint32_t)o[5] << 8 | (uint32_t)o[6] << 16 | (uint32_t)o[7] << 24;
        dst[2] = o[8] | (uint32_t)o[9] << 8 | (uint32_t)o[10] << 16 | (uint32_t)o[11] << 24;
        return;
    }
    for (int64_t i = p0; i < p0 + 4 && i < p.n; ++i) {
        float gr;
        uint8_t o[3];
        one_pixel(p, p.rgb[3 * i], p.rgb[3 * i + 1], p.rgb[3 * i + 2], p.mask && p.mask[i] == 0, gr, o);
        p.grey[i] = gr;
        p.rgb8[3 * i] = o[0];
        p.rgb8[3 * i + 1] = o[1];
        p.rgb8[3 * i + 2] = o[2];
    }
}

}  // namespace prep
}  // namespace om

using namespace om;
using namespace om::prep;

extern "C" int om_depth_view(const uint8_t *rgb, int32_t h, int32_t w, const float *gain, const float *bias,
                             const uint8_t *mask, float *grey, uint8_t *rgb8, void *stream)
{
    if (h < 0 || w < 0)
        return fail(OM_ESHAPE, "om_depth_view: negative size");
    const int64_t n = (int64_t)h * w;
    if (n == 0)
        return OM_OK;
    if (!rgb || !grey || !rgb8)
        return fail(OM_EINVAL, "om_depth_view: null pointer");
    ViewK p{};
    p.rgb = rgb;
    p.mask = mask;
    p.grey = grey;
    p.rgb8 = rgb8;
    p.n = n;
    p.has_g = gain != nullptr;
    p.has_b = bias != nullptr;
    for (int c = 0; c < 3; ++c) {
        p.g[c] = gain ? gain[c] : 1.0f;
        p.b[c] = bias ? bias[c] : 0.0f;
    }
    const auto al = [](const void *q, uintptr_t a) { return (reinterpret_cast<uintptr_t>(q) & (a - 1)) == 0; };
    const bool vec = al(rgb, 4) && al(rgb8, 4) && al(grey, 16) && (!mask || al(mask, 4));
    const dim3 grid((unsigned)cdiv(cdiv(n, 4), 256));
    if (vec)
        hipLaunchKernelGGL(depth_view_kernel<true>, grid, dim3(256), 0, as_stream(stream), p);
    else
        hipLaunchKernelGGL(depth_view_kernel<false>, grid, dim3(256), 0, as_stream(stream), p);
    return launch_check("om_depth_view");
}
