// Shared host/device helpers of libom355 (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/om355.h"

namespace om {

// thread-local message behind om_last_error()
void set_error(const char *fmt, ...);

inline int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_error("%s", buf);
    return code;
}

inline int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return OM_OK;
    return fail(e == hipErrorOutOfMemory ? OM_ENOMEM : OM_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// launch check: configuration errors surface at launch time
inline int launch_check(const char *what) { return hip_check(hipGetLastError(), what); }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with its own L2.
// xcd_block maps the hardware block index to a logical one so that every XCD walks a contiguous
// range of logical blocks (a band of the image), which keeps its working set in its own L2.
constexpr int kXCD = 8;
__device__ __forceinline__ int xcd_block(int b, int nblocks)
{
#ifdef OM_NO_XCD_REMAP
    return b;
#endif
    const int xcd = b % kXCD, k = b / kXCD;
    const int per = nblocks / kXCD, extra = nblocks % kXCD;
    return xcd < extra ? xcd * (per + 1) + k : extra * (per + 1) + (xcd - extra) * per + k;
}

// Optional per-launch timing (om_profile_enable): HIP events recorded on the launch stream
// around each instrumented kernel, read back by om_profile_read.  Off by default; when off a
// scope costs one relaxed atomic load.
bool prof_on();
void prof_push(const char *name, hipEvent_t a, hipEvent_t b);

// Device tables that no kernel writes (view descriptors, camera tables) are read through the
// constant address space, so that loads at a wave-uniform address become scalar loads instead of
// 64 lanes fetching the same words.
#define OM_CONST_AS __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ const OM_CONST_AS T *as_const(const T *p)
{
    return (const OM_CONST_AS T *)p;
}

// a reference to *p whose loads the compiler emits as constant-address-space (scalar) loads
template <class T>
__device__ __forceinline__ const T &uniform_ref(const T *p)
{
    return *(const T *)as_const(p);
}

// x[i] / z for z > 1e-6 (or NaN), each bit-identical to IEEE division.  When every numerator lies
// in [2^-900, 2^100] in magnitude and z <= 2^100, the quotients share one reciprocal of z and run
// the steps of the compiler's own f64 division (v_rcp_f64, two Newton steps, quotient, FMA
// residual, FMA correction): in that range its operand scaling (v_div_scale_f64) and special-case
// fix-up (v_div_fixup_f64) return their inputs unchanged, so the bits agree.  Zeros and tiny or
// huge operands take plain division; a NaN operand gives a NaN quotient on either path.
template <int N>
__device__ __forceinline__ void quotients_shared(const double (&x)[N], double z, double (&q)[N])
{
    double lo = fabs(x[0]), hi = fabs(x[0]);
#pragma unroll
    for (int i = 1; i < N; ++i) {
        lo = fmin(lo, fabs(x[i]));
        hi = fmax(hi, fabs(x[i]));
    }
    if (lo >= 0x1p-900 && hi <= 0x1p100 && z <= 0x1p100) {
        double r = __builtin_amdgcn_rcp(z);
        r = fma(r, fma(-z, r, 1.0), r);
        r = fma(r, fma(-z, r, 1.0), r);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double q0 = x[i] * r;
            q[i] = fma(fma(-z, q0, x[i]), r, q0);
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
            q[i] = x[i] / z;
    }
}

// the projection's pair: xc / zc and yc / zc
__device__ __forceinline__ void div2_shared(double xc, double yc, double zc, double &nx, double &ny)
{
    const double x[2] = {xc, yc};
    double q[2];
    quotients_shared<2>(x, zc, q);
    nx = q[0];
    ny = q[1];
}

// IEEE square root of x for 2^-767 <= x < inf: the compiler's own f64 sqrt sequence (v_rsq_f64,
// then Newton steps on the root g and the half reciprocal root h) without its operand scaling
// (x < 2^-767) and its zero / infinity fix-up, which leave every x of that range unchanged.
__device__ __forceinline__ double sqrt_normal(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// Bilinear tap of a float32 image at 0 <= u <= w - 1.001, 0 <= v <= h - 1.001, in FP64 in the
// order (a00 (1 - fx) + a01 fx) (1 - fy) + (a10 (1 - fx) + a11 fx) fy with x0 = (int)u and
// fx = u - x0.  That difference is exact and equals u - floor(u), which v_fract_f64 returns; the
// two rows are addressed by 32-bit byte offsets from the image base (callers check that images
// hold fewer than 2^30 pixels).
__device__ __forceinline__ void bil_load(const float *im, int w, double u, double v, float &a00, float &a01, float &a10,
                                         float &a11, double &fx, double &fy)
{
    const uint32_t x0 = (uint32_t)(int)u, y0 = (uint32_t)(int)v;
    fx = __builtin_amdgcn_fract(u);
    fy = __builtin_amdgcn_fract(v);
    const uint32_t o0 = (__umul24(y0, (uint32_t)w) + x0) << 2, o1 = o0 + ((uint32_t)w << 2);
    const float *r0 = (const float *)((const char *)im + o0), *r1 = (const float *)((const char *)im + o1);
    a00 = r0[0];
    a01 = r0[1];
    a10 = r1[0];
    a11 = r1[1];
}

// the bilinear's arithmetic on taps loaded by bil_load
__device__ __forceinline__ double bil_mix(double a00, double a01, double a10, double a11, double fx, double fy)
{
    return (a00 * (1.0 - fx) + a01 * fx) * (1.0 - fy) + (a10 * (1.0 - fx) + a11 * fx) * fy;
}

__device__ __forceinline__ double bil_f64(const float *im, int w, double u, double v)
{
    float a00, a01, a10, a11;
    double fx, fy;
    bil_load(im, w, u, v, a00, a01, a10, a11, fx, fy);
    return bil_mix(a00, a01, a10, a11, fx, fy);
}

// most pixels an image read by bil_f64 may hold (32-bit byte offsets)
constexpr int64_t kMaxTapPixels = int64_t(1) << 30;

// Sizes an image tapped through 32-bit offsets may have: fewer than 2^30 pixels, and each side
// below 2^24 (row offsets are formed with 24-bit multiplies, __umul24).
inline bool tap_size_ok(int64_t h, int64_t w)
{
    return h < (int64_t(1) << 24) && w < (int64_t(1) << 24) && h * w < kMaxTapPixels;
}

struct ProfScope {
    const char *name = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(const char *n, hipStream_t s)
    {
        if (!n || !prof_on())
            return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            return;
        name = n;
        st = s;
        if (hipEventRecord(a, st) != hipSuccess)
            name = nullptr;
    }
    ~ProfScope()
    {
        if (name && hipEventRecord(b, st) == hipSuccess)
            prof_push(name, a, b);
        else {
            if (a)
                (void)hipEventDestroy(a);
            if (b)
                (void)hipEventDestroy(b);
        }
    }
};

}  // namespace om
