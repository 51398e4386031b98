/*
 * om_oracle.c -- CPU restatement of pyOrthomosaic's per-view depth-map kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels of py_ortho_amd and
 * the CPU baseline that bench.py times beside them.  The product library (libom355.so) never
 * links, loads or calls it.
 *
 * Every function restates one reference kernel (file:line given at each function) with the same
 * IEEE operations in the same order and the same float32 / float64 conversion points.  The
 * library is built with -O3 -fno-math-errno -ffp-contract=off and no -march flag, like an x86-64
 * build of the reference (setup.py:10-17): no fused multiply-add is formed anywhere, so results
 * are bit-identical to the reference built on that platform.
 *
 * Parity pins (what checks each function, beyond retyping it):
 *   orc_sweep_hypothesis  bit-exact vs oracle/_ref (the reference's own _sweep.h row kernels,
 *                         compiled from /root/reference) and vs orc_ncc_warp + orc_combine (the
 *                         reference's unfused spec, the pairing of test_native_parity.py:97-112);
 *                         FP64 brute-force windows + NCC invariants on tiny inputs.
 *   orc_ncc_warp/combine  brute-force window NCC (numpy, FP64) and top-k/argmax brute force.
 *   orc_pm_refine         golden vectors written by the compiled reference (tests/golden), plus
 *                         the planar-scene fixed point and the xorshift32 known sequence.
 *   orc_box_zero_mean,
 *   orc_hyp_linear/refine the NumPy closed forms of test_native_parity.py:401-422.
 *   orc_rays_grid         the pinhole closed form and the project/undistort round trip.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

enum { ORC_OK = 0, ORC_EINVAL = -1, ORC_ENOMEM = -3 };

/* ------------------------------------------------------------------------------------------ */
/* plane sweep: one source view as seen by the sweep                                           */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const float *img;
    int64_t h, w;
    const double *g; /* 18 values: Rs row-major (9), b = Rs (C_ref - C_src) (3), f k1 k2 k3 cx cy */
    float *J, *V;    /* (W,) warped row and its validity */
    float *ring;     /* (2r+2) x 4 x W horizontal sums */
    double *cs;      /* 4 x W column sums */
    float *ncc;      /* (W,) NCC of the output row */
} orc_src;

/* Warp of one reference row into one source: _sweep.h:33-82 (semantics _dense.pyx:34-66).
 * A non-finite projected coordinate (possible only for inv == 0) is treated as outside the image;
 * the reference reads out of bounds there. */
static void sw_warp_row(const float *inv, const float *r0, const float *r1, const float *r2, int64_t W,
                        orc_src *s)
{
    const double *g = s->g;
    const double lim_u = (double)s->w - 1.001, lim_v = (double)s->h - 1.001;
    for (int64_t x = 0; x < W; x++) {
        const double d = 1.0 / (double)inv[x];
        const double rx = r0[x], ry = r1[x], rz = r2[x];
        const double X = g[9] + d * ((g[0] * rx + g[1] * ry) + g[2] * rz);
        const double Y = g[10] + d * ((g[3] * rx + g[4] * ry) + g[5] * rz);
        const double Z = g[11] + d * ((g[6] * rx + g[7] * ry) + g[8] * rz);
        s->J[x] = 0.0f;
        s->V[x] = 0.0f;
        if (Z <= 1e-6)
            continue;
        const double nx = X / Z, ny = Y / Z;
        const double rr = nx * nx + ny * ny;
        const double dd = g[12] * (1.0 + rr * (g[13] + rr * (g[14] + rr * g[15])));
        const double u = dd * nx + g[16], v = dd * ny + g[17];
        if (!(u >= 0.0 && v >= 0.0 && u <= lim_u && v <= lim_v))
            continue;
        const int64_t x0 = (int64_t)u, y0 = (int64_t)v;
        const double fx = u - (double)x0, fy = v - (double)y0;
        const float *p = s->img + y0 * s->w + x0;
        const double a00 = p[0], a01 = p[1], a10 = p[s->w], a11 = p[s->w + 1];
        const double val = (a00 * (1.0 - fx) + a01 * fx) * (1.0 - fy) + (a10 * (1.0 - fx) + a11 * fx) * fy;
        if (isfinite(val)) {
            s->J[x] = (float)val;
            s->V[x] = 1.0f;
        }
    }
}

/* Horizontal running sums of V, J, J^2, J*I over a (2r+1) window: _sweep.h:85-114. */
static void sw_hsum_row(const float *I, const orc_src *s, int64_t W, int r, float *hs)
{
    const float *J = s->J, *V = s->V;
    double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
    const int64_t pre = r < W ? r : W;
    for (int64_t x = 0; x < pre; x++) {
        a = a + V[x];
        b = b + J[x];
        c = c + (double)(J[x] * J[x]);
        e = e + (double)(J[x] * I[x]);
    }
    for (int64_t x = 0; x < W; x++) {
        const int64_t xi = x + r, xo = x - r - 1;
        if (xi < W) {
            a = a + V[xi];
            b = b + J[xi];
            c = c + (double)(J[xi] * J[xi]);
            e = e + (double)(J[xi] * I[xi]);
        }
        if (xo >= 0) {
            a = a - V[xo];
            b = b - J[xo];
            c = c - (double)(J[xo] * J[xo]);
            e = e - (double)(J[xo] * I[xo]);
        }
        hs[x] = (float)a;
        hs[W + x] = (float)b;
        hs[2 * W + x] = (float)c;
        hs[3 * W + x] = (float)e;
    }
}

/* NCC of one output row from the column sums: _sweep.h:149-178. */
static void sw_ncc_row(const double *cs, const float *mu, const float *sd, double kk, int64_t W, float *out)
{
    for (int64_t x = 0; x < W; x++) {
        const double cov = cs[x] * kk;
        if (cov <= 0.9) {
            out[x] = -1.0f;
            continue;
        }
        const double mj = cs[W + x] * kk / cov;
        const double m2 = cs[2 * W + x] * kk / cov;
        const double mij = cs[3 * W + x] * kk / cov;
        double sj = m2 - mj * mj;
        sj = sj > 1e-6 ? sqrt(sj) : 1e-3;
        const double q = (mij - (double)mu[x] * mj) / ((double)sd[x] * sj);
        out[x] = q > 1.0 ? 1.0f : (q < -1.0 ? -1.0f : (float)q);
    }
}

/* Descending insertion sort of n <= 16 values, mean of the first k, streaming winner update of
 * one pixel: _sweep.h:181-214 (= _dense.pyx:113-144). */
static inline void sw_combine_px(float *buf, int n, int k, int i, float *best, float *prev, float *spb,
                                 float *snb, int32_t *idx)
{
    for (int a = 1; a < n; a++) {
        const float t = buf[a];
        int c = a - 1;
        while (c >= 0 && buf[c] < t) {
            buf[c + 1] = buf[c];
            c--;
        }
        buf[c + 1] = t;
    }
    float s = 0.0f;
    for (int a = 0; a < k; a++)
        s = s + buf[a];
    s = s / (float)k;
    if (*idx == i - 1)
        *snb = s;
    if (s > *best) {
        *best = s;
        *spb = *prev;
        *snb = -2.0f;
        *idx = i;
    }
    *prev = s;
}

/* One hypothesis of the plane sweep for all sources: _dense.pyx:740-838 (validation :759-773,
 * row order :804-835).  Return 0, ORC_EINVAL (same conditions as the reference's ValueError) or
 * ORC_ENOMEM. */
ORC_EXPORT int orc_sweep_hypothesis(const float *ref, const float *mu, const float *sd, int64_t H, int64_t W,
                                    int n, const float *const *imgs, const int64_t *img_hw, const float *ray,
                                    const double *geo, const float *inv, int r, int top_k, int i, float *best,
                                    float *prev, float *spb, float *snb, int32_t *idx)
{
    if (n < 1 || n > 16 || r < 0)
        return ORC_EINVAL;
    if (H == 0 || W == 0)
        return ORC_OK;
    const int64_t HW = H * W, R2 = 2 * (int64_t)r + 2, win = 2 * (int64_t)r + 1;
    const double kk = 1.0 / (double)(win * win);
    const int k = top_k < n ? top_k : n;
    orc_src src[16];
    const size_t fl = (size_t)(3 * W + R2 * 4 * W);
    float *fb = malloc(sizeof(float) * fl * (size_t)n);
    double *db = calloc((size_t)(4 * W) * (size_t)n, sizeof(double));
    if (!fb || !db) {
        free(fb);
        free(db);
        return ORC_ENOMEM;
    }
    for (int s = 0; s < n; s++) {
        src[s].img = imgs[s];
        src[s].h = img_hw[2 * s];
        src[s].w = img_hw[2 * s + 1];
        src[s].g = geo + 18 * s;
        src[s].J = fb + fl * (size_t)s;
        src[s].V = src[s].J + W;
        src[s].ncc = src[s].V + W;
        src[s].ring = src[s].ncc + W;
        src[s].cs = db + (size_t)(4 * W) * (size_t)s;
    }
    /* the first r rows enter the column sums before any output row */
    for (int64_t y = 0; y < r && y < H; y++) {
        for (int s = 0; s < n; s++) {
            float *hs = src[s].ring + (y % R2) * 4 * W;
            sw_warp_row(inv + y * W, ray + y * W, ray + HW + y * W, ray + 2 * HW + y * W, W, &src[s]);
            sw_hsum_row(ref + y * W, &src[s], W, r, hs);
            for (int64_t x = 0; x < 4 * W; x++)
                src[s].cs[x] = src[s].cs[x] + hs[x];
        }
    }
    float buf[16];
    for (int64_t y = 0; y < H; y++) {
        const int64_t yi = y + r, yo = y - r - 1;
        for (int s = 0; s < n; s++) {
            double *cs = src[s].cs;
            if (yi < H) {
                float *hi = src[s].ring + (yi % R2) * 4 * W;
                sw_warp_row(inv + yi * W, ray + yi * W, ray + HW + yi * W, ray + 2 * HW + yi * W, W, &src[s]);
                sw_hsum_row(ref + yi * W, &src[s], W, r, hi);
                if (yo >= 0) {
                    const float *ho = src[s].ring + (yo % R2) * 4 * W;
                    for (int64_t x = 0; x < 4 * W; x++) {
                        const double t = cs[x] + hi[x]; /* add before subtract, per element */
                        cs[x] = t - ho[x];
                    }
                } else {
                    for (int64_t x = 0; x < 4 * W; x++)
                        cs[x] = cs[x] + hi[x];
                }
            } else if (yo >= 0) {
                const float *ho = src[s].ring + (yo % R2) * 4 * W;
                for (int64_t x = 0; x < 4 * W; x++)
                    cs[x] = cs[x] - ho[x];
            }
            sw_ncc_row(cs, mu + y * W, sd + y * W, kk, W, src[s].ncc);
        }
        const int64_t o = y * W;
        for (int64_t x = 0; x < W; x++) {
            for (int s = 0; s < n; s++)
                buf[s] = src[s].ncc[x];
            sw_combine_px(buf, n, k, i, best + o + x, prev + o + x, spb + o + x, snb + o + x, idx + o + x);
        }
    }
    free(fb);
    free(db);
    return ORC_OK;
}

/* Unfused semantic spec, one source: _dense.pyx:13-110.  J, V: (H,W) scratch, hs: (4,H,W). */
ORC_EXPORT int orc_ncc_warp(const float *ref, const float *mu, const float *sd, int64_t H, int64_t W,
                            const float *img, int64_t h, int64_t w, const float *ray, const double *g18,
                            const float *inv, int r, float *out, float *Jm, float *Vm, float *hs)
{
    if (r < 0)
        return ORC_EINVAL;
    const int64_t HW = H * W, win = 2 * (int64_t)r + 1;
    const double kk = 1.0 / (double)(win * win);
    orc_src s;
    s.img = img;
    s.h = h;
    s.w = w;
    s.g = g18;
    for (int64_t y = 0; y < H; y++) {
        s.J = Jm + y * W;
        s.V = Vm + y * W;
        sw_warp_row(inv + y * W, ray + y * W, ray + HW + y * W, ray + 2 * HW + y * W, W, &s);
    }
    float *row = malloc(sizeof(float) * 4 * (size_t)(W ? W : 1));
    double *cs = calloc((size_t)(4 * (W ? W : 1)), sizeof(double));
    if (!row || !cs) {
        free(row);
        free(cs);
        return ORC_ENOMEM;
    }
    for (int64_t y = 0; y < H; y++) {
        s.J = Jm + y * W;
        s.V = Vm + y * W;
        sw_hsum_row(ref + y * W, &s, W, r, row);
        for (int c = 0; c < 4; c++)
            memcpy(hs + c * HW + y * W, row + c * W, sizeof(float) * (size_t)W);
    }
    /* column accumulators: _dense.pyx:84-110 */
    for (int64_t y = 0; y < r && y < H; y++)
        for (int c = 0; c < 4; c++)
            for (int64_t x = 0; x < W; x++)
                cs[c * W + x] = cs[c * W + x] + hs[c * HW + y * W + x];
    for (int64_t y = 0; y < H; y++) {
        for (int64_t x = 0; x < W; x++) {
            for (int c = 0; c < 4; c++) {
                if (y + r < H)
                    cs[c * W + x] = cs[c * W + x] + hs[c * HW + (y + r) * W + x];
                if (y - r - 1 >= 0)
                    cs[c * W + x] = cs[c * W + x] - hs[c * HW + (y - r - 1) * W + x];
            }
        }
        sw_ncc_row(cs, mu + y * W, sd + y * W, kk, W, out + y * W);
    }
    free(row);
    free(cs);
    return ORC_OK;
}

/* Unfused combine: _dense.pyx:113-144.  S: (n,H,W).  The reference sorts into `float buf[16]`,
 * which its n > 16 caller (densify.py:343-353) overruns; this restatement sorts all n sources (up
 * to 256), as the loop is written. */
ORC_EXPORT int orc_combine(const float *S, int n, int64_t H, int64_t W, int top_k, int i, float *best, float *prev,
                           float *spb, float *snb, int32_t *idx)
{
    if (n < 1 || n > 256)
        return ORC_EINVAL;
    const int64_t HW = H * W;
    const int k = top_k < n ? top_k : n;
    float buf[256];
    for (int64_t p = 0; p < HW; p++) {
        for (int a = 0; a < n; a++)
            buf[a] = S[a * HW + p];
        sw_combine_px(buf, n, k, i, best + p, prev + p, spb + p, snb + p, idx + p);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* depth helpers (_fast.pyx, compiled there with -ffp-contract=off)                            */
/* ------------------------------------------------------------------------------------------ */

/* (base + i*step).astype(float32): _fast.pyx:628-641 */
ORC_EXPORT void orc_hyp_linear(const float *base, const float *step, int64_t n, int64_t i, float *out)
{
    const float fi = (float)i;
    for (int64_t p = 0; p < n; p++) {
        const float t = fi * step[p];
        out[p] = base[p] + t;
    }
}

/* np.maximum(iu + t*ds, 1e-6) in float32, NaN propagating: _fast.pyx:644-658 */
ORC_EXPORT void orc_hyp_refine(const float *iu, const float *ds, int64_t n, int64_t t, float *out)
{
    const float ft = (float)t, lo = (float)1e-6;
    for (int64_t p = 0; p < n; p++) {
        float v = ft * ds[p];
        v = iu[p] + v;
        out[p] = (v >= lo || isnan(v)) ? v : lo;
    }
}

/* Zero-padded (2r+1)^2 window mean from float32 cumulative sums: _fast.pyx:661-713 (the
 * densify._box NumPy expression densify.py:99-103). */
ORC_EXPORT int orc_box_zero_mean(const float *a, int64_t H, int64_t W, int r, float *out)
{
    if (r < 0)
        return ORC_EINVAL;
    if (H == 0 || W == 0)
        return ORC_OK;
    const int64_t k = 2 * (int64_t)r + 1, PW = W + k, PH = H + k, R = k + 1;
    float *col = malloc(sizeof(float) * (size_t)W);
    float *tab = malloc(sizeof(float) * (size_t)(R * W));
    float *row = malloc(sizeof(float) * (size_t)PW);
    if (!col || !tab || !row) {
        free(col);
        free(tab);
        free(row);
        return ORC_ENOMEM;
    }
    const float scale = (float)(1.0 / (double)(k * k));
    for (int64_t t = 0; t < PH; t++) {
        /* table row t = cumulative sum (down the columns) of padded rows 0..t (r+1 zero rows on top) */
        const int64_t ay = t - (r + 1);
        const int inside = ay >= 0 && ay < H;
        float *cur = tab + (t % R) * W;
        for (int64_t x = 0; x < W; x++) {
            const float v = inside ? a[ay * W + x] : 0.0f;
            col[x] = (t == 0) ? v : col[x] + v;
            cur[x] = col[x];
        }
        const int64_t y = t - k;
        if (y < 0)
            continue;
        const float *old = tab + (y % R) * W;
        /* along the row: r+1 zeros left, r right, cumulative sum, window difference */
        for (int64_t x = 0; x < PW; x++) {
            const float v = (x < r + 1 || x >= r + 1 + W) ? 0.0f : cur[x - r - 1] - old[x - r - 1];
            row[x] = (x == 0) ? v : row[x - 1] + v;
        }
        for (int64_t x = 0; x < W; x++) {
            const float d = row[x + k] - row[x];
            out[y * W + x] = d * scale;
        }
    }
    free(col);
    free(tab);
    free(row);
    return ORC_OK;
}

/* _Cam.rays over the pixel grid (8 fixed-point undistortion steps): _fast.pyx:145-198. */
ORC_EXPORT void orc_rays_grid(int64_t H, int64_t W, double cx, double cy, double f, double k1, double k2, double k3,
                              double *xo, double *yo)
{
    for (int64_t r = 0; r < H; r++) {
        for (int64_t c = 0; c < W; c++) {
            const double nx = ((double)c - cx) / f, ny = ((double)r - cy) / f;
            double x = nx, y = ny;
            for (int it = 0; it < 8; it++) {
                double q = x * x;
                const double t = y * y;
                q = q + t;
                double d = q * k3;
                d = k2 + d;
                d = q * d;
                d = k1 + d;
                d = q * d;
                d = 1.0 + d;
                x = nx / d;
                y = ny / d;
            }
            xo[r * W + c] = x;
            yo[r * W + c] = y;
        }
    }
}

/* 2x2 box downsample in float32 (NumPy order ((a + b) + c) + d, then * 0.25): densify.py:127-130.
 * out: (H/2, W/2). */
ORC_EXPORT void orc_half(const float *g, int64_t H, int64_t W, float *out)
{
    const int64_t h = H / 2, w = W / 2;
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) {
            const float a = g[(2 * y) * W + 2 * x], b = g[(2 * y + 1) * W + 2 * x];
            const float c = g[(2 * y) * W + 2 * x + 1], d = g[(2 * y + 1) * W + 2 * x + 1];
            float s = a + b;
            s = s + c;
            s = s + d;
            out[y * w + x] = s * 0.25f;
        }
}

/* Ground cell (X0 + (c+.5) gsd, Y0 - (r+.5) gsd, Z[r,c]) -> camera -> bilinear sample of an
 * (h, w, nch) image with edge clamp: _mvs.sample_view, _mvs.pyx:13-54.  out must be zeroed by the
 * caller (the reference's CPU backend allocates it with np.zeros, backend.py:71). */
ORC_EXPORT void orc_sample_view(const float *img, int64_t h, int64_t w, int64_t nch, const double *R, const double *C,
                                double f, double k1, double k2, double k3, double cx, double cy, double X0, double Y0,
                                double gsd, const float *Z, int64_t H, int64_t W, float *out, uint8_t *valid)
{
    for (int64_t r = 0; r < H; r++) {
        const double Y = Y0 - ((double)r + 0.5) * gsd;
        for (int64_t c = 0; c < W; c++) {
            const double X = X0 + ((double)c + 0.5) * gsd;
            const double dx = X - C[0], dy = Y - C[1], dz = (double)Z[r * W + c] - C[2];
            const double zc = R[6] * dx + R[7] * dy + R[8] * dz;
            valid[r * W + c] = 0;
            if (zc <= 1e-6)
                continue;
            const double xc = R[0] * dx + R[1] * dy + R[2] * dz;
            const double yc = R[3] * dx + R[4] * dy + R[5] * dz;
            const double nx = xc / zc, ny = yc / zc;
            const double q = nx * nx + ny * ny;
            const double d = 1.0 + q * (k1 + q * (k2 + q * k3));
            const double u = f * d * nx + cx, v = f * d * ny + cy;
            if (!(u >= 0 && v >= 0 && u <= (double)(w - 1) && v <= (double)(h - 1)))
                continue;
            const int64_t x0 = (int64_t)u, y0 = (int64_t)v;
            const int64_t x1 = x0 + 1 < w ? x0 + 1 : x0, y1 = y0 + 1 < h ? y0 + 1 : y0;
            const double fx = u - (double)x0, fy = v - (double)y0;
            for (int64_t ch = 0; ch < nch; ch++)
                out[(r * W + c) * nch + ch] =
                    (float)((1 - fy) * ((1 - fx) * (double)img[(y0 * w + x0) * nch + ch] +
                                        fx * (double)img[(y0 * w + x1) * nch + ch]) +
                            fy * ((1 - fx) * (double)img[(y1 * w + x0) * nch + ch] +
                                  fx * (double)img[(y1 * w + x1) * nch + ch]));
            valid[r * W + c] = 1;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* PatchMatch repair: _dense.pyx:405-666                                                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const float *ref;
    int H, W;
    const float *ray;
    double cr0, cr1, cr2;
    int S;
    const float *const *img;
    const float *const *dep;
    const int *iw, *ih;
    const double *par;
    int r, step, top_k;
    double geo_w, geo_tol;
} pm_ctx;

/* world point -> source pixel, par = R(9) C(3) f k1 k2 k3 cx cy: _dense.pyx:429-445 */
static inline int pm_proj(const double *p, double X, double Y, double Z, double *u, double *v, double *z)
{
    const double dx = X - p[9], dy = Y - p[10], dz = Z - p[11];
    const double xc = p[0] * dx + p[1] * dy + p[2] * dz;
    const double yc = p[3] * dx + p[4] * dy + p[5] * dz;
    const double zc = p[6] * dx + p[7] * dy + p[8] * dz;
    if (zc <= 1e-6)
        return 0;
    const double nx = xc / zc, ny = yc / zc;
    const double q = nx * nx + ny * ny;
    const double dd = p[12] * (1.0 + q * (p[13] + q * (p[14] + q * p[15])));
    *u = dd * nx + p[16];
    *v = dd * ny + p[17];
    *z = zc;
    return 1;
}

/* bilinear sample, NaN outside [0, w-1.001] x [0, h-1.001]: _dense.pyx:448-458 */
static inline double pm_bil(const float *im, int w, int h, double u, double v)
{
    if (!(u >= 0.0 && v >= 0.0 && u <= (double)w - 1.001 && v <= (double)h - 1.001))
        return NAN;
    const int x0 = (int)u, y0 = (int)v;
    const double fx = u - (double)x0, fy = v - (double)y0;
    const float *p = im + (int64_t)y0 * w + x0;
    return ((double)p[0] * (1.0 - fx) + (double)p[1] * fx) * (1.0 - fy)
           + ((double)p[w] * (1.0 - fx) + (double)p[w + 1] * fx) * fy;
}

/* slanted-window cost: _dense.pyx:461-556 */
static double pm_cost(const pm_ctx *c, int y, int x, double d, const double *n, double *photo_out)
{
    const int H = c->H, W = c->W;
    const int64_t HW = (int64_t)H * W, p = (int64_t)y * W + x;
    const double rpx = c->ray[p], rpy = c->ray[HW + p], rpz = c->ray[2 * HW + p];
    const double ndr = n[0] * rpx + n[1] * rpy + n[2] * rpz;
    double sI[16], sJ[16], sII[16], sJJ[16], sIJ[16], ncc[16];
    int cnt[16];
    int nsamp = 0;
    if (ndr >= -1e-6 || d <= 0)
        return 10.0;
    const double num = d * ndr;
    for (int s = 0; s < c->S; s++) {
        sI[s] = sJ[s] = sII[s] = sJJ[s] = sIJ[s] = 0.0;
        cnt[s] = 0;
    }
    for (int a = -c->r; a <= c->r; a += c->step) {
        const int yy = y + a;
        if (yy < 0 || yy >= H)
            continue;
        for (int b = -c->r; b <= c->r; b += c->step) {
            const int xx = x + b;
            if (xx < 0 || xx >= W)
                continue;
            const int64_t q = (int64_t)yy * W + xx;
            const double I = c->ref[q];
            if (!isfinite(I))
                continue;
            const double qx = c->ray[q], qy = c->ray[HW + q], qz = c->ray[2 * HW + q];
            const double den = n[0] * qx + n[1] * qy + n[2] * qz;
            if (den >= -1e-9)
                continue;
            const double t = num / den;
            const double X = c->cr0 + t * qx, Y = c->cr1 + t * qy, Z = c->cr2 + t * qz;
            nsamp += 1;
            for (int s = 0; s < c->S; s++) {
                double u, v, z;
                if (!pm_proj(c->par + 18 * s, X, Y, Z, &u, &v, &z))
                    continue;
                const double J = pm_bil(c->img[s], c->iw[s], c->ih[s], u, v);
                if (isfinite(J)) {
                    sI[s] += I;
                    sJ[s] += J;
                    sII[s] += I * I;
                    sJJ[s] += J * J;
                    sIJ[s] += I * J;
                    cnt[s] += 1;
                }
            }
        }
    }
    for (int s = 0; s < c->S; s++) {
        if (nsamp < 4 || cnt[s] < 0.7 * nsamp) {
            ncc[s] = -1.0;
            continue;
        }
        const double m = (double)cnt[s];
        const double mi = sI[s] / m, mj = sJ[s] / m;
        const double vi = sII[s] / m - mi * mi;
        const double vj = sJJ[s] / m - mj * mj;
        const double cv = sIJ[s] / m - mi * mj;
        ncc[s] = (vi < 1e-8 || vj < 1e-8) ? 0.0 : cv / sqrt(vi * vj);
    }
    for (int a = 1; a < c->S; a++) {
        const double t = ncc[a];
        int b = a - 1;
        while (b >= 0 && ncc[b] < t) {
            ncc[b + 1] = ncc[b];
            b--;
        }
        ncc[b + 1] = t;
    }
    const int kk = c->top_k < c->S ? c->top_k : c->S;
    double photo = 0.0;
    for (int a = 0; a < kk; a++)
        photo += ncc[a];
    photo /= kk;
    *photo_out = photo;
    double geo = 0.0;
    if (c->geo_w > 0) {
        const double X = c->cr0 + d * rpx, Y = c->cr1 + d * rpy, Z = c->cr2 + d * rpz;
        int ng = 0;
        for (int s = 0; s < c->S; s++) {
            if (!c->dep[s])
                continue;
            ng += 1;
            double e = 1.0, u, v, z;
            if (pm_proj(c->par + 18 * s, X, Y, Z, &u, &v, &z)) {
                const int ui = (int)(u + 0.5), vi = (int)(v + 0.5);
                if (ui >= 0 && vi >= 0 && ui < c->iw[s] && vi < c->ih[s]) {
                    const double dj = c->dep[s][(int64_t)vi * c->iw[s] + ui];
                    if (isfinite(dj) && dj > 0) {
                        e = fabs(z - dj) / dj / c->geo_tol;
                        if (e > 1)
                            e = 1;
                    }
                }
            }
            geo += e;
        }
        if (ng)
            geo /= ng;
    }
    return (1.0 - photo) + c->geo_w * geo;
}

/* xorshift32 -> [0,1) from the low 24 bits: _dense.pyx:559-563 */
static inline double pm_rnd(uint32_t *st)
{
    *st ^= *st << 13;
    *st ^= *st >> 17;
    *st ^= *st << 5;
    return (double)(*st & 0xFFFFFFu) / 16777216.0;
}

/* Red-black PatchMatch over the active pixels: _dense.pyx:566-666.  Sources: imgs[s] (ih[s],iw[s]),
 * deps[s] may be NULL.  Returns the number of accepted hypothesis changes (0 when S is not in
 * 1..16, like the reference). */
ORC_EXPORT int64_t orc_pm_refine(const float *ref, int H, int W, const float *ray, const double *Cr, int S,
                                 const float *const *imgs, const float *const *deps, const int *ih, const int *iw,
                                 const double *par, float *depth, float *normal, float *score,
                                 const uint8_t *active, int r, int step, int top_k, double geo_w, double geo_tol,
                                 int iters, double dmin, double dmax, uint32_t seed)
{
    static const int oy[8] = {-1, 1, 0, 0, -5, 5, 0, 0};
    static const int ox[8] = {0, 0, -1, 1, 0, 0, -5, 5};
    if (S <= 0 || S > 16)
        return 0;
    pm_ctx c;
    c.ref = ref;
    c.H = H;
    c.W = W;
    c.ray = ray;
    c.cr0 = Cr[0];
    c.cr1 = Cr[1];
    c.cr2 = Cr[2];
    c.S = S;
    c.img = imgs;
    c.dep = deps;
    c.iw = iw;
    c.ih = ih;
    c.par = par;
    c.r = r;
    c.step = step;
    c.top_k = top_k;
    c.geo_w = geo_w;
    c.geo_tol = geo_tol;
    const int64_t HW = (int64_t)H * W;
    uint32_t st = seed | 1u;
    int64_t changed = 0;
    /* pm_cost leaves the photometric score untouched when it returns early (degenerate plane), so
     * the caller then reads the value of the previous evaluation, as the reference does with its
     * function-scope `ncc` (_dense.pyx:579, :619-620); the reference starts it uninitialised. */
    double photo = 0.0;
    for (int it = 0; it < iters; it++) {
        const double dsc = 0.05 / (it + 1), nsc = 0.3 / (it + 1);
        for (int colour = 0; colour < 2; colour++) {
            for (int y = 0; y < H; y++) {
                for (int x = 0; x < W; x++) {
                    const int64_t p = (int64_t)y * W + x;
                    if (!active[p] || (x + y) % 2 != colour)
                        continue;
                    double n[3] = {normal[p], normal[HW + p], normal[2 * HW + p]};
                    double best, sc;
                    if (isfinite(depth[p])) {
                        best = pm_cost(&c, y, x, depth[p], n, &photo);
                        sc = photo;
                    } else {
                        best = 1e9;
                        sc = -1;
                    }
                    for (int k = 0; k < 8; k++) {
                        const int yn = y + oy[k], xn = x + ox[k];
                        if (yn < 0 || xn < 0 || yn >= H || xn >= W)
                            continue;
                        const int64_t q = (int64_t)yn * W + xn;
                        if (!isfinite(depth[q]))
                            continue;
                        const double nb[3] = {normal[q], normal[HW + q], normal[2 * HW + q]};
                        const double dn = nb[0] * ray[q] + nb[1] * ray[HW + q] + nb[2] * ray[2 * HW + q];
                        const double ndr = nb[0] * ray[p] + nb[1] * ray[HW + p] + nb[2] * ray[2 * HW + p];
                        if (ndr >= -1e-6)
                            continue;
                        const double dp = (double)depth[q] * dn / ndr;
                        if (dp < dmin || dp > dmax)
                            continue;
                        const double cst = pm_cost(&c, y, x, dp, nb, &photo);
                        if (cst < best) {
                            best = cst;
                            sc = photo;
                            depth[p] = (float)dp;
                            normal[p] = (float)nb[0];
                            normal[HW + p] = (float)nb[1];
                            normal[2 * HW + p] = (float)nb[2];
                            changed += 1;
                        }
                    }
                    if (isfinite(depth[p])) {
                        for (int tr = 0; tr < 2; tr++) {
                            const double dp = (double)depth[p] * (1 + dsc * (2 * pm_rnd(&st) - 1));
                            double nb[3];
                            nb[0] = (double)normal[p] + nsc * (2 * pm_rnd(&st) - 1);
                            nb[1] = (double)normal[HW + p] + nsc * (2 * pm_rnd(&st) - 1);
                            nb[2] = (double)normal[2 * HW + p] + nsc * (2 * pm_rnd(&st) - 1);
                            const double nn = sqrt(nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2]);
                            if (nn < 1e-9 || dp < dmin || dp > dmax)
                                continue;
                            nb[0] /= nn;
                            nb[1] /= nn;
                            nb[2] /= nn;
                            const double cst = pm_cost(&c, y, x, dp, nb, &photo);
                            if (cst < best) {
                                best = cst;
                                sc = photo;
                                depth[p] = (float)dp;
                                normal[p] = (float)nb[0];
                                normal[HW + p] = (float)nb[1];
                                normal[2 * HW + p] = (float)nb[2];
                                changed += 1;
                            }
                        }
                    }
                    score[p] = (float)sc;
                }
            }
        }
    }
    return changed;
}

/* one raw step of the stream behind pm_rnd (tests pin it against Marsaglia's published value) */
ORC_EXPORT uint32_t orc_xorshift32_step(uint32_t y)
{
    (void)pm_rnd(&y);
    return y;
}

/* xorshift32 stream helper for tests: n successive draws of pm_rnd from `seed | 1`. */
ORC_EXPORT void orc_xorshift_draws(uint32_t seed, int64_t n, double *out)
{
    uint32_t st = seed | 1u;
    for (int64_t i = 0; i < n; i++)
        out[i] = pm_rnd(&st);
}

/* ------------------------------------------------------------------------------------------ */
/* CPU baseline runners: independent copies of one job on `nthreads` threads, the way densify    */
/* runs one view per worker thread (densify.py:596).                                            */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    /* sweep over linear hypotheses inv_i = base + f32(i) * step (densify.py:411-414) */
    const float *ref, *mu, *sd, *ray, *base, *step;
    int64_t H, W;
    int n;
    const float *const *imgs;
    const int64_t *img_hw;
    const double *geo;
    int r, top_k, nd;
    int rc;
} sweep_job;

static void *sweep_job_run(void *arg)
{
    sweep_job *j = arg;
    const int64_t HW = j->H * j->W;
    float *st = malloc(sizeof(float) * (size_t)(5 * HW));
    int32_t *idx = malloc(sizeof(int32_t) * (size_t)HW);
    if (!st || !idx) {
        j->rc = ORC_ENOMEM;
        free(st);
        free(idx);
        return NULL;
    }
    for (int64_t p = 0; p < 4 * HW; p++)
        st[p] = -2.0f;
    for (int64_t p = 0; p < HW; p++)
        idx[p] = -1;
    float *inv = st + 4 * HW;
    j->rc = ORC_OK;
    for (int i = 0; i < j->nd && j->rc == ORC_OK; i++) {
        orc_hyp_linear(j->base, j->step, HW, i, inv);
        j->rc = orc_sweep_hypothesis(j->ref, j->mu, j->sd, j->H, j->W, j->n, j->imgs, j->img_hw, j->ray, j->geo,
                                     inv, j->r, j->top_k, i, st, st + HW, st + 2 * HW, st + 3 * HW, idx);
    }
    free(st);
    free(idx);
    return NULL;
}

/* Runs `nthreads` concurrent copies of a linear-hypothesis sweep; returns 0 or the first error. */
ORC_EXPORT int orc_bench_sweep(int nthreads, const float *ref, const float *mu, const float *sd, int64_t H, int64_t W,
                               int n, const float *const *imgs, const int64_t *img_hw, const float *ray,
                               const double *geo, const float *base, const float *step, int nd, int r, int top_k)
{
    if (nthreads < 1 || nthreads > 512)
        return ORC_EINVAL;
    pthread_t th[512];
    sweep_job jobs[512];
    for (int t = 0; t < nthreads; t++) {
        sweep_job j = {ref, mu, sd, ray, base, step, H, W, n, imgs, img_hw, geo, r, top_k, nd, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, sweep_job_run, &jobs[t]);
    }
    int rc = ORC_OK;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc && !rc)
            rc = jobs[t].rc;
    }
    return rc;
}

typedef struct {
    const float *ref, *ray, *depth0, *normal0;
    const double *Cr, *par;
    int H, W, S;
    const float *const *imgs;
    const float *const *deps;
    const int *ih, *iw;
    const uint8_t *active;
    int r, step, top_k, iters;
    double geo_w, geo_tol, dmin, dmax;
    uint32_t seed;
    int64_t changed;
    int rc;
} pm_job;

static void *pm_job_run(void *arg)
{
    pm_job *j = arg;
    const int64_t HW = (int64_t)j->H * j->W;
    float *d = malloc(sizeof(float) * (size_t)(5 * HW));
    if (!d) {
        j->rc = ORC_ENOMEM;
        return NULL;
    }
    memcpy(d, j->depth0, sizeof(float) * (size_t)HW);
    memcpy(d + HW, j->normal0, sizeof(float) * (size_t)(3 * HW));
    j->changed = orc_pm_refine(j->ref, j->H, j->W, j->ray, j->Cr, j->S, j->imgs, j->deps, j->ih, j->iw, j->par, d,
                               d + HW, d + 4 * HW, j->active, j->r, j->step, j->top_k, j->geo_w, j->geo_tol,
                               j->iters, j->dmin, j->dmax, j->seed);
    j->rc = ORC_OK;
    free(d);
    return NULL;
}

/* Runs `nthreads` concurrent copies of one PatchMatch refinement (inputs are copied per thread). */
ORC_EXPORT int orc_bench_pm(int nthreads, const float *ref, int H, int W, const float *ray, const double *Cr, int S,
                            const float *const *imgs, const float *const *deps, const int *ih, const int *iw,
                            const double *par, const float *depth0, const float *normal0, const uint8_t *active,
                            int r, int step, int top_k, double geo_w, double geo_tol, int iters, double dmin,
                            double dmax, uint32_t seed)
{
    if (nthreads < 1 || nthreads > 512)
        return ORC_EINVAL;
    pthread_t th[512];
    pm_job jobs[512];
    for (int t = 0; t < nthreads; t++) {
        pm_job j = {ref, ray, depth0, normal0, Cr, par, H, W, S, imgs, deps, ih, iw, active, r, step, top_k, iters,
                    geo_w, geo_tol, dmin, dmax, seed, 0, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, pm_job_run, &jobs[t]);
    }
    int rc = ORC_OK;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc && !rc)
            rc = jobs[t].rc;
    }
    return rc;
}
