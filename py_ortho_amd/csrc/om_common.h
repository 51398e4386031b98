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

struct ProfScope {
    const char *name = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(const char *n, hipStream_t s)
    {
        if (!prof_on())
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
