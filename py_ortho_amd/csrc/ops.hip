// Height-sweep array operations of the backend interface (orthomosaic/backend.py:67-99), for the
// ground-grid sweep that the reference routes through its compute backends (mvs.py:254-336).
//
//   om_sample_view   _mvs.sample_view (orthomosaic/_mvs.pyx:13-54): ground cell -> camera (FP64,
//                    radial k1..k3) -> bilinear with edge clamp; one launch for a (D, H, W) stack.
//   om_box_sat_mean  _fast.box_sat_mean (orthomosaic/_fast.pyx:897-943): float64 summed-area
//                    mean with the reference's two cumulative sums and four-corner order.
//   om_topk_mean     CPUBackend.topk_mean (backend.py:96-99): mean of the k largest along axis 0.
//   om_argminmax0    CPUBackend.argmin0 / argmax0 (backend.py:88-94): first index, NumPy NaN rule.
#include <cmath>

#include "om_common.h"

namespace om {

struct SampleK {
    double R[9], C[3];
    double f, k1, k2, k3, cx, cy, X0, Y0, gsd;
    int h, w, nch, H, W;
    int64_t planes;
};

__global__ void sample_view_kernel(const SampleK P, const float *img, const float *Z, float *out, uint8_t *valid)
{
    const int64_t n = P.planes * P.H * P.W;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int64_t plane = (int64_t)P.H * P.W;
    const int64_t rc = t % plane;
    const int r = (int)(rc / P.W), c = (int)(rc - (int64_t)r * P.W);
    const double Y = P.Y0 - (r + 0.5) * P.gsd;
    const double X = P.X0 + (c + 0.5) * P.gsd;
    const double dx = X - P.C[0], dy = Y - P.C[1], dz = (double)Z[t] - P.C[2];
    const double zc = P.R[6] * dx + P.R[7] * dy + P.R[8] * dz;
    valid[t] = 0;
    if (zc <= 1e-6)
        return;
    const double xc = P.R[0] * dx + P.R[1] * dy + P.R[2] * dz;
    const double yc = P.R[3] * dx + P.R[4] * dy + P.R[5] * dz;
    const double nx = xc / zc, ny = yc / zc;
    const double q = nx * nx + ny * ny;
    const double d = 1.0 + q * (P.k1 + q * (P.k2 + q * P.k3));
    const double u = P.f * d * nx + P.cx, v = P.f * d * ny + P.cy;
    if (!(u >= 0 && v >= 0 && u <= P.w - 1 && v <= P.h - 1))
        return;
    const int x0 = (int)u, y0 = (int)v;
    const int x1 = x0 + 1 < P.w ? x0 + 1 : x0, y1 = y0 + 1 < P.h ? y0 + 1 : y0;
    const double fx = u - x0, fy = v - y0;
    const int nch = P.nch;
    const float *a = img + ((size_t)y0 * P.w + x0) * nch, *b = img + ((size_t)y0 * P.w + x1) * nch;
    const float *e = img + ((size_t)y1 * P.w + x0) * nch, *g = img + ((size_t)y1 * P.w + x1) * nch;
    float *o = out + (size_t)t * nch;
    for (int ch = 0; ch < nch; ++ch)
        o[ch] = (float)((1 - fy) * ((1 - fx) * (double)a[ch] + fx * (double)b[ch]) +
                        fy * ((1 - fx) * (double)e[ch] + fx * (double)g[ch]));
    valid[t] = 1;
}

// SAT pass 1: column cumulative sums (FP64) down the padded rows, one thread per (plane, column)
__global__ void sat_col_kernel(const float *a, int64_t B, int H, int W, int r, double *col)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * W)
        return;
    const int64_t b = t / W;
    const int x = (int)(t - b * W);
    const int PH = H + 2 * r + 1;
    double s = 0.0;
    for (int i = 0; i < PH; ++i) {
        const int ay = i - r - 1;
        const double v = (ay >= 0 && ay < H) ? (double)a[(b * H + ay) * W + x] : 0.0;
        s = s + v;
        col[(b * PH + i) * W + x] = s;
    }
}

// SAT pass 2: running sum of the column sums along each padded row, one thread per (plane, row)
__global__ void sat_row_kernel(const double *col, int64_t B, int H, int W, int r, double *sat)
{
    const int PH = H + 2 * r + 1, PW = W + 2 * r + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * PH)
        return;
    const double *c = col + t * W;
    double *o = sat + t * PW;
    for (int x = 0; x <= r; ++x)
        o[x] = 0.0;
    double tot = 0.0;
    for (int x = 0; x < W; ++x) {
        tot = tot + c[x];
        o[x + r + 1] = tot;
    }
    for (int x = r + 1 + W; x < PW; ++x)
        o[x] = tot;
}

// SAT pass 3: ((S[y+k][x+k] - S[y][x+k]) - S[y+k][x]) + S[y][x], / k^2, to float32
__global__ void sat_out_kernel(const double *sat, int64_t B, int H, int W, int r, float *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * H * W)
        return;
    const int k = 2 * r + 1, PH = H + k, PW = W + k;
    const int64_t b = t / ((int64_t)H * W);
    const int64_t rem = t - b * H * W;
    const int y = (int)(rem / W), x = (int)(rem - (int64_t)y * W);
    const double *cur = sat + (b * PH + y + k) * PW, *old = sat + (b * PH + y) * PW;
    double v = cur[x + k] - old[x + k];
    v = v - cur[x];
    v = v + old[x];
    out[t] = (float)(v / ((double)k * (double)k));
}

__global__ void topk_mean_kernel(const float *st, int n, int64_t M, int k, float *out)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M)
        return;
    float t[16];
    for (int a = 0; a < 16; ++a)
        t[a] = -INFINITY;
    bool nan = false;
    for (int i = 0; i < n; ++i) {
        const float v = st[(int64_t)i * M + m];
        if (v != v) {
            nan = true;
            continue;
        }
        for (int a = k - 1; a >= 0; --a) {  // sorted insertion into the k best
            if (t[a] < v) {
                t[a] = (a > 0 && t[a - 1] < v) ? t[a - 1] : v;
            }
        }
    }
    float s = 0.0f;
    for (int a = 0; a < k; ++a)
        s += t[a];
    out[m] = nan ? NAN : s / (float)k;
}

__global__ void argminmax0_kernel(const float *a, int n, int64_t M, int want_max, int64_t *idx, float *val)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M)
        return;
    int bi = 0;
    float bv = a[m];
    if (bv == bv) {
        for (int i = 1; i < n; ++i) {
            const float v = a[(int64_t)i * M + m];
            if (v != v) {  // NumPy: the first NaN wins
                bi = i;
                bv = v;
                break;
            }
            if (want_max ? (v > bv) : (v < bv)) {
                bi = i;
                bv = v;
            }
        }
    }
    idx[m] = bi;
    val[m] = bv;
}

}  // namespace om

using namespace om;

extern "C" int om_sample_view(const float *img, int32_t h, int32_t w, int32_t nch, const double *R, const double *C,
                              double f, double k1, double k2, double k3, double cx, double cy, double X0, double Y0,
                              double gsd, const float *Z, int64_t planes, int32_t H, int32_t W, float *out,
                              uint8_t *valid, void *stream)
{
    if (h < 1 || w < 1 || nch < 1 || H < 0 || W < 0 || planes < 0)
        return fail(OM_ESHAPE, "sample_view: bad sizes");
    const int64_t n = planes * H * W;
    if (n == 0)
        return OM_OK;
    SampleK P;
    for (int i = 0; i < 9; ++i)
        P.R[i] = R[i];
    for (int i = 0; i < 3; ++i)
        P.C[i] = C[i];
    P.f = f;
    P.k1 = k1;
    P.k2 = k2;
    P.k3 = k3;
    P.cx = cx;
    P.cy = cy;
    P.X0 = X0;
    P.Y0 = Y0;
    P.gsd = gsd;
    P.h = h;
    P.w = w;
    P.nch = nch;
    P.H = H;
    P.W = W;
    P.planes = planes;
    hipLaunchKernelGGL(sample_view_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), P, img, Z,
                       out, valid);
    return launch_check("sample_view_kernel");
}

extern "C" size_t om_box_sat_workspace_bytes(int64_t B, int32_t H, int32_t W, int32_t r)
{
    if (B <= 0 || H <= 0 || W <= 0 || r < 0)
        return 0;
    const int64_t PH = H + 2 * (int64_t)r + 1, PW = W + 2 * (int64_t)r + 1;
    return (size_t)(B * PH * ((int64_t)W + PW)) * sizeof(double);
}

extern "C" int om_box_sat_mean(const float *a, int64_t B, int32_t H, int32_t W, int32_t r, float *out,
                               void *workspace, size_t workspace_bytes, void *stream)
{
    if (r < 0)
        return fail(OM_EINVAL, "radius must be nonnegative");
    if (B < 0 || H < 0 || W < 0)
        return fail(OM_ESHAPE, "negative size");
    if (B == 0 || H == 0 || W == 0)
        return OM_OK;
    const size_t need = om_box_sat_workspace_bytes(B, H, W, r);
    if (!workspace || workspace_bytes < need)
        return fail(OM_EINVAL, "workspace of %zu bytes needed", need);
    const int64_t PH = H + 2 * (int64_t)r + 1;
    double *col = (double *)workspace;
    double *sat = col + B * PH * W;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(sat_col_kernel, dim3((unsigned)cdiv(B * W, 64)), dim3(64), 0, st, a, B, H, W, r, col);
    hipLaunchKernelGGL(sat_row_kernel, dim3((unsigned)cdiv(B * PH, 64)), dim3(64), 0, st, col, B, H, W, r, sat);
    hipLaunchKernelGGL(sat_out_kernel, dim3((unsigned)cdiv(B * H * W, 256)), dim3(256), 0, st, sat, B, H, W, r, out);
    return launch_check("box_sat kernels");
}

extern "C" int om_topk_mean(const float *st, int32_t n, int64_t M, int32_t k, float *out, void *stream)
{
    if (n < 1 || k < 1 || k > n)
        return fail(OM_EINVAL, "topk_mean needs 1 <= k <= n");
    if (k > 16)
        return fail(OM_EUNSUP, "topk_mean supports k <= 16");
    if (M <= 0)
        return OM_OK;
    hipLaunchKernelGGL(topk_mean_kernel, dim3((unsigned)cdiv(M, 256)), dim3(256), 0, as_stream(stream), st, n, M, k,
                       out);
    return launch_check("topk_mean_kernel");
}

extern "C" int om_argminmax0(const float *a, int32_t n, int64_t M, int32_t want_max, int64_t *idx, float *val,
                             void *stream)
{
    if (n < 1)
        return fail(OM_EINVAL, "attempt to get argmin/argmax of an empty sequence");
    if (M <= 0)
        return OM_OK;
    hipLaunchKernelGGL(argminmax0_kernel, dim3((unsigned)cdiv(M, 256)), dim3(256), 0, as_stream(stream), a, n, M,
                       want_max, idx, val);
    return launch_check("argminmax0_kernel");
}

namespace om {
__global__ void fp64_steps_kernel(const double *a, const double *b, const double *c, int64_t n, double *qa, double *qb,
                                  double *qs, double *fr, double *sq)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double x, y;
    div2_shared(a[i], b[i], c[i], x, y);
    qa[i] = x;
    qb[i] = y;
    const double num[3] = {b[i], a[i] + b[i], a[i]};
    double q[3];
    quotients_shared<3>(num, c[i], q);
    qs[i] = q[1];
    fr[i] = __builtin_amdgcn_fract(a[i]);
    // the square root the sweep's NCC takes for 1e-6 < c < inf
    const double s = c[i];
    sq[i] = (s > 1e-6 && s < INFINITY) ? sqrt_normal(s) : sqrt(s);
}
}  // namespace om

extern "C" int om_fp64_steps(const double *a, const double *b, const double *c, int64_t n, double *qa, double *qb,
                             double *qs, double *fr, double *sq, void *stream)
{
    if (n < 0)
        return fail(OM_ESHAPE, "om_fp64_steps: negative size");
    if (n == 0)
        return OM_OK;
    if (!a || !b || !c || !qa || !qb || !qs || !fr || !sq)
        return fail(OM_EINVAL, "om_fp64_steps: null pointer");
    hipLaunchKernelGGL(fp64_steps_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, as_stream(stream), a, b, c, n,
                       qa, qb, qs, fr, sq);
    return launch_check("fp64_steps_kernel");
}
