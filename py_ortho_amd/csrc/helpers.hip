// Library entry points and the depth-map helper kernels (hypothesis maps, window means, pyramid).
//
// Reference semantics: orthomosaic/_fast.pyx:628-713 (hyp_linear, hyp_refine, box_zero_mean) and
// orthomosaic/densify.py:127-130 (_half), :327-328 (mu / sd of _sweep).  These helpers reproduce
// the reference's float32 rounding sequence exactly (the library is built with -ffp-contract=off).
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "om_common.h"

namespace om {

static thread_local char g_err[512] = "";

static std::atomic<int> g_prof{0};
static std::mutex g_prof_mu;
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};
static std::vector<ProfRec> g_prof_recs;

bool prof_on() { return g_prof.load(std::memory_order_relaxed) != 0; }

void prof_push(const char *name, hipEvent_t a, hipEvent_t b)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_recs.push_back({name, a, b});
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

__global__ void hyp_linear_kernel(const float *base, const float *step, int64_t n, float fi, float *out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        const float t = fi * step[p];
        out[p] = base[p] + t;
    }
}

__global__ void hyp_refine_kernel(const float *iu, const float *ds, int64_t n, float ft, float *out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        float v = ft * ds[p];
        v = iu[p] + v;
        out[p] = (v >= 1e-6f || v != v) ? v : 1e-6f;
    }
}

// 2x2 box downsample: ((a + b) + c) + d then * 0.25 in float32 (NumPy's order, densify.py:130)
__global__ void half_kernel(const float *g, int H, int W, float *out)
{
    const int w = W / 2, h = H / 2;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)h * w)
        return;
    const int y = (int)(p / w), x = (int)(p - (int64_t)y * w);
    const float a = g[(int64_t)(2 * y) * W + 2 * x], b = g[(int64_t)(2 * y + 1) * W + 2 * x];
    const float c = g[(int64_t)(2 * y) * W + 2 * x + 1], d = g[(int64_t)(2 * y + 1) * W + 2 * x + 1];
    float s = a + b;
    s = s + c;
    s = s + d;
    out[p] = s * 0.25f;
}

// box_zero_mean pass 1: per column, float32 cumulative sum down the padded column
// (r+1 zero rows above, r below), written as the full (H+k, W) table.  One thread per column
// keeps NumPy's sequential summation order.
__global__ void box_colsum_kernel(const float *a, int H, int W, int r, float *tab, int square)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W)
        return;
    const int k = 2 * r + 1, PH = H + k;
    float col = 0.f;
    for (int t = 0; t < PH; ++t) {
        const int ay = t - (r + 1);
        float v = 0.f;
        if (ay >= 0 && ay < H) {
            v = a[(int64_t)ay * W + x];
            if (square)
                v = v * v;
        }
        col = (t == 0) ? v : col + v;
        tab[(int64_t)t * W + x] = col;
    }
}

// pass 2: per output row, the column-window differences c[y+k] - c[y] padded (r+1 left, r right)
// and summed cumulatively along the row in float32; written as the (H, W+k) row table.
__global__ void box_rowsum_kernel(const float *tab, int H, int W, int r, float *rows)
{
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H)
        return;
    const int k = 2 * r + 1, PW = W + k;
    const float *cur = tab + (int64_t)(y + k) * W, *old = tab + (int64_t)y * W;
    float *row = rows + (int64_t)y * PW;
    float acc = 0.f;
    for (int x = 0; x < PW; ++x) {
        const float v = (x < r + 1 || x >= r + 1 + W) ? 0.f : cur[x - r - 1] - old[x - r - 1];
        acc = (x == 0) ? v : acc + v;
        row[x] = acc;
    }
}

// pass 3: window difference along the row times float32(1/k^2); stats = 1 also forms
// sd = sqrt(max(box(a^2) - mu^2, 1e-6)) from two row tables (densify.py:327-328)
__global__ void box_out_kernel(const float *rows, const float *rows2, int H, int W, int r, float scale, float *out,
                               float *sd)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)H * W)
        return;
    const int k = 2 * r + 1, PW = W + k;
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    const float *row = rows + (int64_t)y * PW;
    const float m = (row[x + k] - row[x]) * scale;
    out[p] = m;
    if (rows2) {
        const float *row2 = rows2 + (int64_t)y * PW;
        const float m2 = (row2[x + k] - row2[x]) * scale;
        const float mm = m * m;
        float v = m2 - mm;
        v = (v >= 1e-6f || v != v) ? v : 1e-6f;  // np.maximum(., 1e-6) in float32
        sd[p] = sqrtf(v);
    }
}

// World ray of every pixel with camera-axis depth 1 (densify._rays_world, densify.py:288-297):
// 8 fixed-point undistortion steps in FP64 (_fast.pyx:145-164, one rounding per operation), then
// R^T (x, y, 1) summed over j = 0, 1, 2 in order, rounded to float32.
struct RaysK {
    double R[9];
    double cx, cy, f, k1, k2, k3;
    int H, W;
};

__global__ void rays_world_kernel(const RaysK P, float *out)
{
    const int64_t HW = (int64_t)P.H * P.W;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= HW)
        return;
    const int v = (int)(p / P.W), u = (int)(p - (int64_t)v * P.W);
    const double nx = ((double)u - P.cx) / P.f, ny = ((double)v - P.cy) / P.f;
    double x = nx, y = ny;
    for (int it = 0; it < 8; ++it) {
        double q = x * x;
        const double t = y * y;
        q = q + t;
        double d = q * P.k3;
        d = P.k2 + d;
        d = q * d;
        d = P.k1 + d;
        d = q * d;
        d = 1.0 + d;
        x = nx / d;
        y = ny / d;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double a = P.R[i] * x, b = P.R[3 + i] * y, c = P.R[6 + i] * 1.0;
        out[i * HW + p] = (float)((a + b) + c);
    }
}

static size_t box_tables_bytes(int H, int W, int r)
{
    const int64_t k = 2 * (int64_t)r + 1;
    return sizeof(float) * (size_t)((H + k) * W + (int64_t)H * (W + k));
}

}  // namespace om

using namespace om;

extern "C" int om_version(void) { return 10000; }

extern "C" int om_profile_enable(int32_t on)
{
    g_prof.store(on ? 1 : 0);
    return OM_OK;
}

extern "C" int om_profile_read(const char *kernel, double *total_ms, int64_t *launches)
{
    double tot = 0.0;
    int64_t n = 0;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    std::vector<ProfRec> keep;
    for (const ProfRec &r : g_prof_recs) {
        if (kernel && strcmp(kernel, r.name) != 0) {
            keep.push_back(r);
            continue;
        }
        float ms = 0.f;
        int rc = hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
        if (!rc)
            rc = hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        if (rc)
            return rc;
        tot += ms;
        n += 1;
    }
    g_prof_recs.swap(keep);
    if (total_ms)
        *total_ms = tot;
    if (launches)
        *launches = n;
    return OM_OK;
}

extern "C" const char *om_last_error(void) { return g_err; }

extern "C" int om_device_info(char *name, int32_t name_len, int32_t *cus, int32_t *clock_khz, int64_t *mem_bytes)
{
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc)
        return rc;
    hipDeviceProp_t prop;
    if ((rc = hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties")))
        return rc;
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cus)
        *cus = prop.multiProcessorCount;
    if (clock_khz)
        *clock_khz = prop.clockRate;
    if (mem_bytes)
        *mem_bytes = (int64_t)prop.totalGlobalMem;
    return OM_OK;
}

extern "C" int om_hyp_linear(const float *base, const float *step, int64_t n, int64_t i, float *out, void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "negative size");
    if (n == 0)
        return OM_OK;
    hipLaunchKernelGGL(hyp_linear_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), base, step, n,
                       (float)i, out);
    return launch_check("hyp_linear_kernel");
}

extern "C" int om_hyp_refine(const float *iu, const float *ds, int64_t n, int64_t t, float *out, void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "negative size");
    if (n == 0)
        return OM_OK;
    hipLaunchKernelGGL(hyp_refine_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), iu, ds, n,
                       (float)t, out);
    return launch_check("hyp_refine_kernel");
}

extern "C" int om_half(const float *g, int32_t H, int32_t W, float *out, void *stream)
{
    if (H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    const int64_t n = (int64_t)(H / 2) * (W / 2);
    if (n == 0)
        return OM_OK;
    hipLaunchKernelGGL(half_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), g, H, W, out);
    return launch_check("half_kernel");
}

extern "C" int om_rays_world(int32_t H, int32_t W, const double *R, double cx, double cy, double f, double k1,
                             double k2, double k3, float *out, void *stream)
{
    if (H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    if (!R)
        return fail(OM_EINVAL, "null rotation");
    const int64_t n = (int64_t)H * W;
    if (n == 0)
        return OM_OK;
    RaysK P;
    for (int i = 0; i < 9; ++i)
        P.R[i] = R[i];
    P.cx = cx;
    P.cy = cy;
    P.f = f;
    P.k1 = k1;
    P.k2 = k2;
    P.k3 = k3;
    P.H = H;
    P.W = W;
    hipLaunchKernelGGL(rays_world_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), P, out);
    return launch_check("rays_world_kernel");
}

extern "C" size_t om_box_workspace_bytes(int32_t H, int32_t W, int32_t r)
{
    if (H <= 0 || W <= 0 || r < 0)
        return 0;
    return box_tables_bytes(H, W, r);
}

static int box_impl(const float *a, int H, int W, int r, float *out, float *sd, void *ws, size_t ws_bytes,
                    hipStream_t st)
{
    if (r < 0)
        return fail(OM_EINVAL, "box_zero_mean needs r >= 0");
    if (H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    if (H == 0 || W == 0)
        return OM_OK;
    const size_t one = box_tables_bytes(H, W, r);
    const size_t need = sd ? 2 * one : one;
    if (!ws || ws_bytes < need)
        return fail(OM_EINVAL, "workspace of %zu bytes needed (got %zu)", need, ws_bytes);
    const int64_t k = 2 * (int64_t)r + 1;
    float *tab = (float *)ws;
    float *rows = tab + (H + k) * W;
    float *rows2 = nullptr;
    const float scale = (float)(1.0 / (double)(k * k));
    int rc;
    hipLaunchKernelGGL(box_colsum_kernel, dim3((unsigned)cdiv(W, 64)), dim3(64), 0, st, a, H, W, r, tab, 0);
    hipLaunchKernelGGL(box_rowsum_kernel, dim3((unsigned)cdiv(H, 64)), dim3(64), 0, st, tab, H, W, r, rows);
    if (sd) {
        float *tab2 = (float *)((char *)ws + one);
        rows2 = tab2 + (H + k) * W;
        hipLaunchKernelGGL(box_colsum_kernel, dim3((unsigned)cdiv(W, 64)), dim3(64), 0, st, a, H, W, r, tab2, 1);
        hipLaunchKernelGGL(box_rowsum_kernel, dim3((unsigned)cdiv(H, 64)), dim3(64), 0, st, tab2, H, W, r, rows2);
    }
    const int64_t n = (int64_t)H * W;
    hipLaunchKernelGGL(box_out_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, rows, rows2, H, W, r, scale,
                       out, sd);
    if ((rc = launch_check("box kernels")))
        return rc;
    return OM_OK;
}

extern "C" int om_box_zero_mean(const float *a, int32_t H, int32_t W, int32_t r, float *out, void *workspace,
                                size_t workspace_bytes, void *stream)
{
    return box_impl(a, H, W, r, out, nullptr, workspace, workspace_bytes, as_stream(stream));
}

extern "C" int om_ref_stats(const float *ref, int32_t H, int32_t W, int32_t r, float *mu, float *sd, void *workspace,
                            size_t workspace_bytes, void *stream)
{
    if (!sd)
        return fail(OM_EINVAL, "null sd");
    return box_impl(ref, H, W, r, mu, sd, workspace, workspace_bytes, as_stream(stream));
}
