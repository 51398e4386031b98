// Depth-map storage formats on gfx950: the per-view file set of densify's pass 1 and 2
// (densify.py:590-592, :613-614, :639-641).
//
// - float32 -> float16 as NumPy's astype(np.float16) (densify.py:591): round to nearest even,
//   overflow to infinity, subnormal halves; a NaN keeps its sign and the top ten payload bits, and
//   stays a NaN when those are zero (NumPy's rule), so the stored bytes equal the reference's.
// - float16 -> float32 (densify.py:613): exact; NaN payloads shifted into place as NumPy does.
// - the final filter D[(S < min_ncc) | low] = NaN (densify.py:641) with S the re-widened float16
//   score and the threshold compared in float32, as NumPy compares a float32 array with a
//   Python float.
#include "om_common.h"

namespace om {
namespace fmt {

__device__ __forceinline__ uint16_t f32_to_f16_np(float x)
{
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {   // NaN
        uint16_t h = (uint16_t)(0x7C00u | ((u & 0x007FFFFFu) >> 13));
        if (h == 0x7C00u)
            h = 0x7C01u;
        return (uint16_t)(((u >> 16) & 0x8000u) | h);
    }
    return __builtin_bit_cast(uint16_t, (_Float16)x);                // IEEE round to nearest even
}

__device__ __forceinline__ float f16_to_f32_np(uint16_t h)
{
    if ((h & 0x7C00u) == 0x7C00u && (h & 0x03FFu))                   // NaN: sign, payload << 13
        return __uint_as_float(((uint32_t)(h & 0x8000u) << 16) | 0x7F800000u | ((uint32_t)(h & 0x03FFu) << 13));
    return (float)__builtin_bit_cast(_Float16, h);                   // exact
}

__global__ void f32_to_f16_kernel(const float *src, int64_t n, uint16_t *dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = f32_to_f16_np(src[i]);
}

__global__ void f16_to_f32_kernel(const uint16_t *src, int64_t n, float *dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = f16_to_f32_np(src[i]);
}

__global__ void filter_depth_kernel(float *depth, const uint16_t *score16, const uint8_t *low, int64_t n, float thr)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float s = f16_to_f32_np(score16[i]);
    if (s < thr || (low && low[i]))
        depth[i] = __builtin_nanf("");
}

}  // namespace fmt
}  // namespace om

using namespace om;
using namespace om::fmt;

extern "C" int om_f32_to_f16(const float *src, int64_t n, uint16_t *dst, void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "om_f32_to_f16: negative size");
    if (n == 0)
        return OM_OK;
    if (!src || !dst)
        return fail(OM_EINVAL, "om_f32_to_f16: null pointer");
    hipLaunchKernelGGL(f32_to_f16_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), src, n, dst);
    return launch_check("om_f32_to_f16");
}

extern "C" int om_f16_to_f32(const uint16_t *src, int64_t n, float *dst, void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "om_f16_to_f32: negative size");
    if (n == 0)
        return OM_OK;
    if (!src || !dst)
        return fail(OM_EINVAL, "om_f16_to_f32: null pointer");
    hipLaunchKernelGGL(f16_to_f32_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), src, n, dst);
    return launch_check("om_f16_to_f32");
}

extern "C" int om_filter_depth(float *depth, const uint16_t *score16, const uint8_t *low, int64_t n, double min_ncc,
                               void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "om_filter_depth: negative size");
    if (n == 0)
        return OM_OK;
    if (!depth || !score16)
        return fail(OM_EINVAL, "om_filter_depth: null pointer");
    hipLaunchKernelGGL(filter_depth_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), depth,
                       score16, low, n, (float)min_ncc);
    return launch_check("om_filter_depth");
}
