// Geometric consistency and sequential depth-map fusion on gfx950.
//
// densify._consistency (densify.py:441-478) counts, per pixel of a depth map, the neighbour depth
// maps that agree with it: the pixel's 3D point is projected into neighbour j, j's depth at the
// nearest pixel is back-projected, and the result must land within px_tol pixels of the source
// pixel with a relative depth difference within rel_tol[j] (_fast.pyx:201-274).
//
// densify._fuse_loop (densify.py:688-762) walks the views in order; for view k every finite, not
// yet used pixel of the stride grid averages its point with the agreeing neighbours' points, is
// kept when at least min_views images agree, and marks the grid pixel of each agreeing neighbour
// as used, so that later views do not emit the same surface sample again.  Views therefore run
// one after another (one launch pair per view); inside a view every grid pixel is independent.
//
// Arithmetic: FP64 in the reference's order (undistort, project, hypot, compare), 3x3 products as
// NumPy's OpenBLAS dgemm forms them (s = a0*b0; s = fma(a1, b1, s); s = fma(a2, b2, s)).  Points are
// emitted in the reference's order (views in the given order, grid pixels row-major) through a
// per-view scan of block counts: eval counts the kept points of each block; emit adds up the counts
// of the blocks before its own (or, for views of more than kInlinePrefix blocks, a one-workgroup
// scan kernel does) and writes each kept point at its offset.
//
// Layout: depth (h, w) float32, rgb (h, w, 3) uint8, used flags on the stride grid
// (ceil(h / stride), ceil(w / stride)) uint8 -- the reference's used map is only ever set and read
// on grid pixels ((vi // st) * st, (ui // st) * st; densify.py:749-750).
#include "om_common.h"

#include <cmath>
#include <vector>

namespace om {
namespace fu {

constexpr int kMaxNb = 16;
constexpr int kBlock = 256;

// _fast._undistort (_fast.pyx:145-164).  With k1 = k2 = k3 = 0 and a finite radius every iteration
// computes d = 1 exactly and x = nx / 1 = nx, so the loop is skipped without changing a bit.
__device__ __forceinline__ void undistort(double u, double v, const om_cam &c, double &xo, double &yo)
{
    const double nx = (u - c.cx) / c.f, ny = (v - c.cy) / c.f;
    double x = nx, y = ny;
    if (!(c.k1 == 0.0 && c.k2 == 0.0 && c.k3 == 0.0 && isfinite(nx * nx + ny * ny))) {
        for (int it = 0; it < 8; ++it) {
            double r2 = x * x;
            const double t = y * y;
            r2 = r2 + t;
            double d = r2 * c.k3;
            d = c.k2 + d;
            d = r2 * d;
            d = c.k1 + d;
            d = r2 * d;
            d = 1.0 + d;
            x = nx / d;
            y = ny / d;
        }
    }
    xo = x;
    yo = y;
}

// _fast._project (_fast.pyx:167-185): z = max(xc2, 1e-6), NaN propagating
__device__ __forceinline__ void project(double x0, double x1, double x2, const om_cam &c, double &uo, double &vo)
{
    const double z = (x2 >= 1e-6 || isnan(x2)) ? x2 : 1e-6;
    const double x = x0 / z, y = x1 / z;
    double r2 = x * x, t = y * y;
    r2 = r2 + t;
    double d = r2 * c.k3;
    d = c.k2 + d;
    d = r2 * d;
    d = c.k1 + d;
    d = r2 * d;
    d = 1.0 + d;
    d = c.f * d;
    t = d * x;
    uo = t + c.cx;
    t = d * y;
    vo = t + c.cy;
}

// one output of a NumPy (n, 3) @ (3, 3) float64 product as OpenBLAS dgemm forms it
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2)
{
    double s = a0 * b0;
    s = fma(a1, b1, s);
    return fma(a2, b2, s);
}

// X = C + [x d, y d, d] @ R (ref_points + the product of densify.py:452-453 / 709-710)
__device__ __forceinline__ void ref_point(double d, double u, double v, const om_cam &c, double X[3])
{
    double x, y;
    undistort(u, v, c, x, y);
    const double m0 = x * d, m1 = y * d, m2 = d;
    for (int i = 0; i < 3; ++i)
        X[i] = c.C[i] + dot3(m0, m1, m2, c.R[i], c.R[3 + i], c.R[6 + i]);
}

// densify._project_neighbour (densify.py:171-179) for one point: neighbour_sample, Xj = Cj + M @ Rj,
// xk = (Xj - Cr) @ Rr^T and agreement.  Returns ok; (a, b) is the neighbour pixel and Xj its
// back-projected point when ok.  ok needs a finite neighbour depth, which also excludes every
// pixel outside the neighbour image (the reference reads NaN there), so those return early.
__device__ __forceinline__ bool neighbour(const double X[3], const om_cam &cj, const float *Dj, const om_cam &cr,
                                          double uu, double vv, double px_tol, double rel_tol, int &a, int &b,
                                          double Xj[3])
{
    const double d0 = X[0] - cj.C[0], d1 = X[1] - cj.C[1], d2 = X[2] - cj.C[2];
    const double xc0 = dot3(d0, d1, d2, cj.R[0], cj.R[1], cj.R[2]);
    const double xc1 = dot3(d0, d1, d2, cj.R[3], cj.R[4], cj.R[5]);
    const double xc2 = dot3(d0, d1, d2, cj.R[6], cj.R[7], cj.R[8]);
    double u, v;
    project(xc0, xc1, xc2, cj, u, v);
    // the reference converts rint(u) to int64 first: out-of-range and NaN values fail these tests
    // either way
    const double ru = rint(u), rv = rint(v);
    if (!(xc2 > 0 && ru >= 0 && rv >= 0 && ru < (double)cj.w && rv < (double)cj.h))
        return false;
    a = (int)ru;
    b = (int)rv;
    const double dd = (double)Dj[(int64_t)b * cj.w + a];
    if (!isfinite(dd))
        return false;
    double x, y;
    undistort(ru, rv, cj, x, y);
    const double m0 = x * dd, m1 = y * dd, m2 = dd;
    for (int i = 0; i < 3; ++i)
        Xj[i] = cj.C[i] + dot3(m0, m1, m2, cj.R[i], cj.R[3 + i], cj.R[6 + i]);
    const double e0 = Xj[0] - cr.C[0], e1 = Xj[1] - cr.C[1], e2 = Xj[2] - cr.C[2];
    const double xk0 = dot3(e0, e1, e2, cr.R[0], cr.R[1], cr.R[2]);
    const double xk1 = dot3(e0, e1, e2, cr.R[3], cr.R[4], cr.R[5]);
    const double xk2 = dot3(e0, e1, e2, cr.R[6], cr.R[7], cr.R[8]);
    double ub, vb;
    project(xk0, xk1, xk2, cr, ub, vb);
    const double t = rel_tol * dd;
    return hypot(ub - uu, vb - vv) <= px_tol && fabs(dd - xc2) <= t;
}

// ------------------------------------------------------------------------------------------
// consistency
// ------------------------------------------------------------------------------------------
struct ConsNb {
    const float *depth;
    double tol;
    om_cam cam;
};

struct ConsK {
    const float *depth;
    int16_t *cnt;
    double px_tol;
    int32_t n_nb;
    om_cam cam;
    ConsNb nb[kMaxNb];
};

__global__ __launch_bounds__(kBlock) void consistency_kernel(ConsK p)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t W = p.cam.w, n = (int64_t)p.cam.h * W;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int v = (int)(i / W), u = (int)(i - (int64_t)v * W);
    const float d = p.depth[i];
    int c = 0;
    if (isfinite(d)) {
        double X[3];
        ref_point((double)d, (double)u, (double)v, p.cam, X);
        for (int s = 0; s < p.n_nb; ++s) {
            double Xj[3];
            int a, b;
            c += neighbour(X, p.nb[s].cam, p.nb[s].depth, p.cam, (double)u, (double)v, p.px_tol, p.nb[s].tol, a, b,
                           Xj);
        }
    }
    p.cnt[i] = (int16_t)c;
}

// ------------------------------------------------------------------------------------------
// fusion: eval (points, marks, per-block counts) + emit (offsets, ordered compaction)
// ------------------------------------------------------------------------------------------
struct FuseNb {
    const float *depth;
    uint8_t *used;
    double tol;
    int32_t gw;
    om_cam cam;
};

struct FuseWs {
    uint8_t *keep;          // per grid pixel: min(count, 255) when kept, else 0
    double *px, *py, *pz;   // per grid pixel: the averaged point when kept
    int32_t *blkcnt;        // kept points per block
    int64_t *blkoff;        // output offset of each block
    int64_t *bases;         // output offset of the current view's first point (2, alternating)
    int64_t *total;         // total[0]: points so far; total[1]: candidates evaluated (finite,
                            // unused grid pixels)
};

struct FuseK {
    const float *depth;
    const uint8_t *used;
    om_cam cam;
    int32_t gh, gw, st, n_nb, min_views, nblk;
    double px_tol;
    FuseWs ws;
    FuseNb nb[kMaxNb];
};

__global__ void fuse_init_kernel(int64_t *bases, int64_t *total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        bases[0] = 0;
        total[0] = 0;
        total[1] = 0;
    }
}

// One thread per grid pixel.  (A view of 1600 x 1200 at stride 2 is 1875 workgroups, about seven
// waves per SIMD; splitting a pixel's neighbours over the waves of a workgroup, with the ordered
// sum through LDS, measured slower: 49 against 33 microseconds per view.)
__global__ __launch_bounds__(kBlock) void fuse_eval_kernel(FuseK p)
{
    __shared__ int32_t smark[kMaxNb][kBlock];
    const int tid = threadIdx.x;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t G = (int64_t)p.gh * p.gw;
    const int64_t g = (int64_t)lb * kBlock + tid;
    int kept = 0, cand = 0;
    if (g < G) {
        const int gv = (int)(g / p.gw), gu = (int)(g - (int64_t)gv * p.gw);
        const int v = gv * p.st, u = gu * p.st;
        const float d = p.depth[(int64_t)v * p.cam.w + u];
        uint8_t kc = 0;
        if (isfinite(d) && !p.used[g]) {
            cand = 1;
            double X[3];
            ref_point((double)d, (double)u, (double)v, p.cam, X);
            double a0 = X[0], a1 = X[1], a2 = X[2];
            int cnt = 1;
            uint32_t okm = 0;
            // neighbours in list order: the sums are order dependent (densify.py:718-724)
            for (int s = 0; s < p.n_nb; ++s) {
                double Xj[3];
                int a, b;
                if (neighbour(X, p.nb[s].cam, p.nb[s].depth, p.cam, (double)u, (double)v, p.px_tol, p.nb[s].tol, a, b,
                              Xj)) {
                    a0 = a0 + Xj[0];
                    a1 = a1 + Xj[1];
                    a2 = a2 + Xj[2];
                    ++cnt;
                    okm |= 1u << s;
                    smark[s][tid] = (b / p.st) * p.nb[s].gw + a / p.st;
                }
            }
            if (cnt >= p.min_views) {
                kept = 1;
                for (int s = 0; s < p.n_nb; ++s)
                    if (okm >> s & 1u)
                        p.nb[s].used[smark[s][tid]] = 1;
                const double c = (double)cnt;
                p.ws.px[g] = a0 / c;
                p.ws.py[g] = a1 / c;
                p.ws.pz[g] = a2 / c;
                kc = (uint8_t)(cnt < 255 ? cnt : 255);
            }
        }
        p.ws.keep[g] = kc;
    }
    const int ncand = __syncthreads_count(cand);
    const int count = __syncthreads_count(kept);
    if (tid == 0) {
        p.ws.blkcnt[lb] = count;
        if (ncand)
            atomicAdd((unsigned long long *)&p.ws.total[1], (unsigned long long)ncand);
    }
}

// exclusive scan of one view's block counts into output offsets after the points of earlier views
// (bases[par]); writes the next view's base.  One workgroup, for views with many blocks; smaller
// views take their block offsets inside the emit kernel.
constexpr int kScan = 1024;
constexpr int kInlinePrefix = 2048;   // up to this many blocks, emit sums the counts before it

__global__ __launch_bounds__(kScan) void fuse_scan_kernel(FuseWs ws, int32_t nblk, int32_t par)
{
    __shared__ int64_t wtot[kScan / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t carry = ws.bases[par];
    for (int base = 0; base < nblk; base += kScan) {
        const int i = base + tid;
        const int64_t x = i < nblk ? ws.blkcnt[i] : 0;
        int64_t y = x;                               // inclusive scan inside the wave
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t z = __shfl_up(y, o, 64);
            if (lane >= o)
                y += z;
        }
        if (lane == 63)
            wtot[wave] = y;
        __syncthreads();
        int64_t before = 0, all = 0;
        for (int w = 0; w < kScan / 64; ++w) {
            before += w < wave ? wtot[w] : 0;
            all += wtot[w];
        }
        if (i < nblk)
            ws.blkoff[i] = carry + before + y - x;
        carry += all;
        __syncthreads();
    }
    if (tid == 0) {
        ws.bases[par ^ 1] = carry;
        ws.total[0] = carry;
    }
}

// sum over the kBlock threads of a workgroup (every thread gets it)
__device__ __forceinline__ int64_t block_sum(int64_t x, int64_t *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        x += __shfl_down(x, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = x;
    __syncthreads();
    int64_t s = 0;
    for (int w = 0; w < kBlock / 64; ++w)
        s += sh[w];
    return s;
}

struct EmitK {
    FuseWs ws;
    const uint8_t *rgb;
    int32_t W, gh, gw, st, owner, nblk, par, inline_prefix;
    int64_t cap;
    double *pts;
    uint8_t *cols, *nv;
    int32_t *own;
};

__global__ __launch_bounds__(kBlock) void fuse_emit_kernel(EmitK e)
{
    __shared__ int wsum[kBlock / 64];
    __shared__ int64_t red[kBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t G = (int64_t)e.gh * e.gw;
    const int64_t g = (int64_t)lb * kBlock + tid;
    const int kc = g < G ? e.ws.keep[g] : 0;
    const uint64_t bal = __ballot(kc != 0);
    if (lane == 0)
        wsum[wave] = __popcll(bal);
    int64_t boff;
    if (e.inline_prefix) {
        // this block's offset: the view's base plus the kept points of the blocks before it
        int64_t x = 0;
        for (int i = tid; i < lb; i += kBlock)
            x += e.ws.blkcnt[i];
        boff = e.ws.bases[e.par] + block_sum(x, red);
        if (lb == 0) {                               // the next view's base
            int64_t y = 0;
            for (int i = tid; i < e.nblk; i += kBlock)
                y += e.ws.blkcnt[i];
            const int64_t next = e.ws.bases[e.par] + block_sum(y, red);
            if (tid == 0) {
                e.ws.bases[e.par ^ 1] = next;
                e.ws.total[0] = next;
            }
        }
    } else {
        boff = e.ws.blkoff[lb];
    }
    __syncthreads();
    if (kc == 0)
        return;
    int64_t off = boff + __popcll(bal & ((1ull << lane) - 1));
    for (int w = 0; w < wave; ++w)
        off += wsum[w];
    if (off >= e.cap)
        return;
    e.pts[3 * off] = e.ws.px[g];
    e.pts[3 * off + 1] = e.ws.py[g];
    e.pts[3 * off + 2] = e.ws.pz[g];
    if (e.cols) {
        const int gv = (int)(g / e.gw), gu = (int)(g - (int64_t)gv * e.gw);
        const uint8_t *c = e.rgb + 3 * ((int64_t)gv * e.st * e.W + (int64_t)gu * e.st);
        e.cols[3 * off] = c[0];
        e.cols[3 * off + 1] = c[1];
        e.cols[3 * off + 2] = c[2];
    }
    e.nv[off] = (uint8_t)kc;
    e.own[off] = e.owner;
}

inline int64_t grid_n(int32_t n, int32_t st) { return cdiv(n, st); }

struct WsLayout {
    size_t keep, px, py, pz, blkcnt, blkoff, bases, bytes;
};

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

inline WsLayout ws_layout(int64_t maxG)
{
    const int64_t nblk = cdiv(maxG, kBlock);
    WsLayout L;
    size_t o = 0;
    L.keep = o;
    o = up256(o + (size_t)maxG);
    L.px = o;
    o = up256(o + 8 * (size_t)maxG);
    L.py = o;
    o = up256(o + 8 * (size_t)maxG);
    L.pz = o;
    o = up256(o + 8 * (size_t)maxG);
    L.blkcnt = o;
    o = up256(o + 4 * (size_t)nblk);
    L.blkoff = o;
    o = up256(o + 8 * (size_t)nblk);
    L.bases = o;
    o = up256(o + 16);
    L.bytes = o;
    return L;
}

inline bool cam_ok(const om_cam &c)
{
    return c.w > 0 && c.h > 0 && std::isfinite(c.f) && c.f != 0.0;
}

}  // namespace fu
}  // namespace om

using namespace om;
using namespace om::fu;

extern "C" int om_consistency(const float *depth, const om_cam *cam, int32_t n_nb, const float *const *nb_depth,
                              const om_cam *nb_cam, const double *rel_tol, double px_tol, int16_t *cnt, void *stream)
{
    if (!depth || !cam || !cnt || (n_nb > 0 && (!nb_depth || !nb_cam || !rel_tol)))
        return fail(OM_EINVAL, "om_consistency: null pointer");
    if (n_nb < 0 || n_nb > kMaxNb)
        return fail(OM_ENSRC, "om_consistency: %d neighbours (0..%d supported)", n_nb, kMaxNb);
    if (!cam_ok(*cam))
        return fail(OM_EINVAL, "om_consistency: bad reference camera (w=%d h=%d f=%g)", cam->w, cam->h, cam->f);
    ConsK p{};
    p.depth = depth;
    p.cnt = cnt;
    p.px_tol = px_tol;
    p.n_nb = n_nb;
    p.cam = *cam;
    for (int s = 0; s < n_nb; ++s) {
        if (!nb_depth[s] || !cam_ok(nb_cam[s]))
            return fail(OM_EINVAL, "om_consistency: neighbour %d has no depth map or a bad camera", s);
        p.nb[s].depth = nb_depth[s];
        p.nb[s].tol = rel_tol[s];
        p.nb[s].cam = nb_cam[s];
    }
    const int64_t n = (int64_t)cam->h * cam->w;
    if (n > ((int64_t)1 << 31) - kBlock)
        return fail(OM_EUNSUP, "om_consistency: %lld pixels", (long long)n);
    ProfScope ps("consistency", as_stream(stream));
    hipLaunchKernelGGL(consistency_kernel, dim3((unsigned)cdiv(n, kBlock)), dim3(kBlock), 0, as_stream(stream), p);
    return launch_check("om_consistency");
}

extern "C" size_t om_fuse_workspace_bytes(const om_fuse_view *views, int32_t n_views, int32_t stride)
{
    if (!views || n_views <= 0)
        return 0;
    const int32_t st = stride < 1 ? 1 : stride;
    int64_t maxG = 1;
    for (int k = 0; k < n_views; ++k)
        if (views[k].depth && views[k].cam.w > 0 && views[k].cam.h > 0)
            maxG = std::max(maxG, grid_n(views[k].cam.h, st) * grid_n(views[k].cam.w, st));
    return ws_layout(maxG).bytes;
}

extern "C" int om_fuse(int32_t n_order, const int32_t *order, int32_t n_views, const om_fuse_view *views,
                       const int32_t *nbrs, int32_t max_nb, const double *rel_tol, double px_tol, int32_t min_views,
                       int32_t stride, double *pts, uint8_t *cols, uint8_t *nv, int32_t *owner, int64_t cap,
                       int64_t *total, void *workspace, size_t workspace_bytes, void *stream)
{
    if (n_order < 0 || n_views < 0 || max_nb < 0 || cap < 0)
        return fail(OM_EINVAL, "om_fuse: negative count");
    if ((n_order > 0 && (!order || !views)) || (max_nb > 0 && (!nbrs || !rel_tol)) || !total || !workspace ||
        (cap > 0 && (!pts || !nv || !owner)))
        return fail(OM_EINVAL, "om_fuse: null pointer");
    const int32_t st = stride < 1 ? 1 : stride;
    // validate everything before the first launch
    int64_t maxG = 1;
    for (int32_t o = 0; o < n_order; ++o) {
        const int32_t k = order[o];
        if (k < 0 || k >= n_views)
            return fail(OM_EINVAL, "om_fuse: order[%d] = %d outside 0..%d", o, k, n_views - 1);
        const om_fuse_view &V = views[k];
        if (!V.depth || !V.used || !cam_ok(V.cam))
            return fail(OM_EINVAL, "om_fuse: view %d has no depth map, no used map or a bad camera", k);
        if (cols && !V.rgb)
            return fail(OM_EINVAL, "om_fuse: view %d has no colours", k);
        const int64_t G = grid_n(V.cam.h, st) * grid_n(V.cam.w, st);
        if (G > ((int64_t)1 << 31) - kBlock)
            return fail(OM_EUNSUP, "om_fuse: view %d has %lld grid pixels", k, (long long)G);
        maxG = std::max(maxG, G);
        int nb = 0;
        for (int s = 0; s < max_nb; ++s) {
            const int32_t j = nbrs[(int64_t)k * max_nb + s];
            if (j < 0)
                continue;
            if (j >= n_views || j == k)
                return fail(OM_EINVAL, "om_fuse: view %d lists neighbour %d", k, j);
            if (!views[j].depth || !views[j].used || !cam_ok(views[j].cam))
                return fail(OM_EINVAL, "om_fuse: neighbour %d of view %d has no depth map, no used map or a bad camera",
                            j, k);
            ++nb;
        }
        if (nb > kMaxNb)
            return fail(OM_ENSRC, "om_fuse: view %d has %d neighbours (at most %d)", k, nb, kMaxNb);
    }
    const WsLayout L = ws_layout(maxG);
    if (workspace_bytes < L.bytes)
        return fail(OM_EINVAL, "om_fuse: workspace %zu bytes, %zu needed", workspace_bytes, L.bytes);
    char *w = static_cast<char *>(workspace);
    FuseWs ws;
    ws.keep = reinterpret_cast<uint8_t *>(w + L.keep);
    ws.px = reinterpret_cast<double *>(w + L.px);
    ws.py = reinterpret_cast<double *>(w + L.py);
    ws.pz = reinterpret_cast<double *>(w + L.pz);
    ws.blkcnt = reinterpret_cast<int32_t *>(w + L.blkcnt);
    ws.blkoff = reinterpret_cast<int64_t *>(w + L.blkoff);
    ws.bases = reinterpret_cast<int64_t *>(w + L.bases);
    ws.total = total;
    hipStream_t s = as_stream(stream);
    ProfScope ps("fuse", s);
    hipLaunchKernelGGL(fuse_init_kernel, dim3(1), dim3(64), 0, s, ws.bases, ws.total);
    if (int rc = launch_check("om_fuse init"))
        return rc;
    for (int32_t o = 0; o < n_order; ++o) {
        const int32_t k = order[o];
        const om_fuse_view &V = views[k];
        FuseK p{};
        p.depth = V.depth;
        p.used = V.used;
        p.cam = V.cam;
        p.gh = (int32_t)grid_n(V.cam.h, st);
        p.gw = (int32_t)grid_n(V.cam.w, st);
        p.st = st;
        p.min_views = min_views;
        p.px_tol = px_tol;
        p.ws = ws;
        for (int sl = 0; sl < max_nb; ++sl) {
            const int32_t j = nbrs[(int64_t)k * max_nb + sl];
            if (j < 0)
                continue;
            FuseNb &q = p.nb[p.n_nb++];
            q.depth = views[j].depth;
            q.used = views[j].used;
            q.tol = rel_tol[(int64_t)k * max_nb + sl];
            q.gw = (int32_t)grid_n(views[j].cam.w, st);
            q.cam = views[j].cam;
        }
        const int64_t G = (int64_t)p.gh * p.gw;
        p.nblk = (int32_t)cdiv(G, kBlock);
        {
            ProfScope pe("fuse_eval", s);
            hipLaunchKernelGGL(fuse_eval_kernel, dim3((unsigned)p.nblk), dim3(kBlock), 0, s, p);
        }
        const int par = o & 1, inl = p.nblk <= kInlinePrefix;
        if (!inl)
            hipLaunchKernelGGL(fuse_scan_kernel, dim3(1), dim3(kScan), 0, s, ws, p.nblk, par);
        EmitK e{};
        e.ws = ws;
        e.rgb = V.rgb;
        e.W = V.cam.w;
        e.gh = p.gh;
        e.gw = p.gw;
        e.st = st;
        e.owner = k;
        e.cap = cap;
        e.pts = pts;
        e.cols = cols;
        e.nv = nv;
        e.own = owner;
        e.nblk = p.nblk;
        e.par = par;
        e.inline_prefix = inl;
        hipLaunchKernelGGL(fuse_emit_kernel, dim3((unsigned)p.nblk), dim3(kBlock), 0, s, e);
        if (int rc = launch_check("om_fuse"))
            return rc;
    }
    return OM_OK;
}
