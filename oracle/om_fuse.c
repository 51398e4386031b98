/*
 * om_fuse.c -- CPU restatement of pyOrthomosaic's geometric consistency check and sequential
 * depth-map fusion (linked into liborc.so beside om_oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP fusion kernels and the CPU baseline that
 * bench.py times beside them.  The product library never links, loads or calls it.
 *
 * The reference runs these steps as Cython loops (_fast.pyx:201-283, built with -ffp-contract=off)
 * around NumPy 3x3 matrix products.  NumPy hands (n, 3) @ (3, 3) float64 products to OpenBLAS
 * dgemm, whose kernels form each output as  s = a0*b0;  s = fma(a1, b1, s);  s = fma(a2, b2, s)
 * (measured bit for bit against exact rational arithmetic on OpenBLAS 0.3.29, for n >= 2 and for
 * both R and R.T operands); mat3 below restates that.  Everything else is plain IEEE double in the
 * reference's order.
 *
 * Parity pins: tests/golden/fusion_grid3x3.npz (written by the compiled reference): counts of
 * densify._consistency for every view and the points, colours, view counts, owners and used maps
 * of densify._fuse_loop for two settings, compared bit for bit; the undistort/project pair against
 * the pinhole closed form (zero distortion) and its round trip.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

enum { ORC_OK = 0, ORC_EINVAL = -1, ORC_ENOMEM = -3 };

/* camera of densify._Cam (densify.py:133-160) with float64 parameters */
typedef struct {
    double R[9], C[3], f, k1, k2, k3, cx, cy;
    int32_t w, h;
} orc_cam;

/* _fast._undistort (_fast.pyx:145-164) */
static void fu_undistort(double u, double v, const orc_cam *c, double *xo, double *yo)
{
    double nx = (u - c->cx) / c->f, ny = (v - c->cy) / c->f, x = nx, y = ny, r2, d, t;
    for (int it = 0; it < 8; ++it) {
        r2 = x * x;
        t = y * y;
        r2 = r2 + t;
        d = r2 * c->k3;
        d = c->k2 + d;
        d = r2 * d;
        d = c->k1 + d;
        d = r2 * d;
        d = 1.0 + d;
        x = nx / d;
        y = ny / d;
    }
    *xo = x;
    *yo = y;
}

/* _fast._project (_fast.pyx:167-185): z = max(xc2, 1e-6), NaN propagating */
static void fu_project(double x0, double x1, double x2, const orc_cam *c, double *uo, double *vo)
{
    double z = (x2 >= 1e-6 || isnan(x2)) ? x2 : 1e-6;
    double x = x0 / z, y = x1 / z, r2, t, d;
    r2 = x * x;
    t = y * y;
    r2 = r2 + t;
    d = r2 * c->k3;
    d = c->k2 + d;
    d = r2 * d;
    d = c->k1 + d;
    d = r2 * d;
    d = 1.0 + d;
    d = c->f * d;
    t = d * x;
    *uo = t + c->cx;
    t = d * y;
    *vo = t + c->cy;
}

/* Planted mistake for the tests (never set outside them): form the products as a plain
 * (a0*b0 + a1*b1) + a2*b2, to show that the golden vectors pin the product order. */
static int fu_plain_products = 0;

ORC_EXPORT void orc_fuse_plain_products(int on) { fu_plain_products = on; }

/* one row of an (n, 3) @ (3, 3) NumPy product: out_i = sum_k a_k B[k][i] as OpenBLAS forms it.
 * trans = 1 multiplies by B^T (out_i = sum_k a_k B[i][k]). */
static void mat3(const double a[3], const double *B, int trans, double out[3])
{
    for (int i = 0; i < 3; ++i) {
        double b0 = trans ? B[3 * i] : B[i], b1 = trans ? B[3 * i + 1] : B[3 + i], b2 = trans ? B[3 * i + 2] : B[6 + i];
        if (fu_plain_products) {
            double p0 = a[0] * b0, p1 = a[1] * b1, p2 = a[2] * b2;
            out[i] = (p0 + p1) + p2;
            continue;
        }
        double s = a[0] * b0;
        s = fma(a[1], b1, s);
        out[i] = fma(a[2], b2, s);
    }
}

/* _fast.ref_points (_fast.pyx:201-218) + X = C + (. @ R) (densify.py:452-453, 709-710) */
static void fu_ref_point(const float *D, int64_t W, int64_t v, int64_t u, const orc_cam *cr, double X[3])
{
    double d = D[v * W + u], x, y, M[3], P[3];
    fu_undistort((double)u, (double)v, cr, &x, &y);
    M[0] = x * d;
    M[1] = y * d;
    M[2] = d;
    mat3(M, cr->R, 0, P);
    for (int i = 0; i < 3; ++i)
        X[i] = cr->C[i] + P[i];
}

/* densify._project_neighbour (densify.py:171-179) for one point: neighbour_sample
 * (_fast.pyx:221-254), Xj = Cj + M @ Rj, xk = (Xj - Cr) @ Rr^T, agreement (_fast.pyx:257-274).
 * Returns ok; ui, vi and Xj as the reference computes them. */
static int fu_neighbour(const double X[3], const orc_cam *cj, const float *Dj, const orc_cam *cr, int64_t uu,
                        int64_t vv, double px_tol, double rel_tol, int64_t *ui, int64_t *vi, double Xj[3])
{
    double d[3], xc[3], u, v, x, y, dd, M[3], P[3], xk[3], ub, vb, t;
    for (int i = 0; i < 3; ++i)
        d[i] = X[i] - cj->C[i];
    mat3(d, cj->R, 1, xc);
    fu_project(xc[0], xc[1], xc[2], cj, &u, &v);
    /* (int64_t)rint: an out-of-range or NaN value lands outside the image either way */
    double ru = rint(u), rv = rint(v);
    int inside = xc[2] > 0 && ru >= 0 && rv >= 0 && ru < cj->w && rv < cj->h;
    int64_t a = inside ? (int64_t)ru : -1, b = inside ? (int64_t)rv : -1;
    *ui = a;
    *vi = b;
    dd = inside ? (double)Dj[b * (int64_t)cj->w + a] : NAN;
    fu_undistort((double)a, (double)b, cj, &x, &y);
    M[0] = x * dd;
    M[1] = y * dd;
    M[2] = dd;
    mat3(M, cj->R, 0, P);
    for (int i = 0; i < 3; ++i)
        Xj[i] = cj->C[i] + P[i];
    for (int i = 0; i < 3; ++i)
        d[i] = Xj[i] - cr->C[i];
    mat3(d, cr->R, 1, xk);
    fu_project(xk[0], xk[1], xk[2], cr, &ub, &vb);
    t = rel_tol * dd;
    return isfinite(dd) && hypot(ub - (double)uu, vb - (double)vv) <= px_tol && fabs(dd - xc[2]) <= t;
}

/* densify._consistency (densify.py:441-478): per pixel of D (H = cr->h, W = cr->w) the number of
 * neighbours that agree; 0 where D is not finite.  nb_depth[j] is neighbour j's depth map. */
ORC_EXPORT int orc_consistency(const float *D, const orc_cam *cr, int32_t n_nb, const float *const *nb_depth,
                               const orc_cam *nb_cam, const double *rel_tol, double px_tol, int16_t *cnt)
{
    int64_t H = cr->h, W = cr->w;
    memset(cnt, 0, (size_t)(H * W) * sizeof(int16_t));
    for (int64_t v = 0; v < H; ++v)
        for (int64_t u = 0; u < W; ++u) {
            if (!isfinite(D[v * W + u]))
                continue;
            double X[3], Xj[3];
            int64_t ui, vi;
            fu_ref_point(D, W, v, u, cr, X);
            int16_t c = 0;
            for (int j = 0; j < n_nb; ++j)
                c += (int16_t)fu_neighbour(X, &nb_cam[j], nb_depth[j], cr, u, v, px_tol, rel_tol[j], &ui, &vi, Xj);
            cnt[v * W + u] = c;
        }
    return ORC_OK;
}

/* densify._fuse_loop (densify.py:688-762) over views order[0..n_order): depth[k] (h, w) float32,
 * rgb[k] (h, w, 3) uint8, used[k] (h, w) uint8 (read and updated as the reference's used maps),
 * nbrs (n_views, max_nb) neighbour indices (-1 = none; views without a depth map must be given as
 * -1), rel_tol (n_views, max_nb).  Appends up to cap points to pts (n, 3) / cols (n, 3) / nv /
 * owner in the reference's order and returns the number of points (which may exceed cap), or a
 * negative error. */
ORC_EXPORT int64_t orc_fuse(int32_t n_order, const int32_t *order, int32_t n_views, const float *const *depth,
                            const orc_cam *cams, const uint8_t *const *rgb, uint8_t *const *used, const int32_t *nbrs,
                            int32_t max_nb, const double *rel_tol, double px_tol, int32_t min_views, int32_t stride,
                            double *pts, uint8_t *cols, uint8_t *nv, int32_t *owner, int64_t cap)
{
    if (stride < 1)
        stride = 1;
    int64_t total = 0;
    for (int32_t o = 0; o < n_order; ++o) {
        int32_t k = order[o];
        if (k < 0 || k >= n_views)
            return ORC_EINVAL;
        const orc_cam *cr = &cams[k];
        int64_t H = cr->h, W = cr->w;
        /* candidates: finite, not used, on the stride grid, row-major */
        int64_t n = 0;
        for (int64_t v = 0; v < H; v += stride)
            for (int64_t u = 0; u < W; u += stride)
                n += isfinite(depth[k][v * W + u]) && !used[k][v * W + u];
        if (n == 0)
            continue;
        int64_t *vv = malloc((size_t)n * sizeof(int64_t)), *uu = malloc((size_t)n * sizeof(int64_t));
        double *acc = malloc((size_t)n * 3 * sizeof(double));
        int32_t *cnt = malloc((size_t)n * sizeof(int32_t));
        int nb = 0;
        int32_t nbk[64];
        for (int32_t s = 0; s < max_nb && nb < 64; ++s)
            if (nbrs[k * max_nb + s] >= 0)
                nbk[nb++] = s;
        uint8_t *ok = malloc((size_t)n * (nb ? nb : 1));
        int64_t *mu = malloc((size_t)n * (nb ? nb : 1) * sizeof(int64_t) * 2);
        if (!vv || !uu || !acc || !cnt || !ok || !mu) {
            free(vv), free(uu), free(acc), free(cnt), free(ok), free(mu);
            return ORC_ENOMEM;
        }
        n = 0;
        for (int64_t v = 0; v < H; v += stride)
            for (int64_t u = 0; u < W; u += stride)
                if (isfinite(depth[k][v * W + u]) && !used[k][v * W + u]) {
                    vv[n] = v;
                    uu[n] = u;
                    ++n;
                }
        for (int64_t i = 0; i < n; ++i) {
            double X[3];
            fu_ref_point(depth[k], W, vv[i], uu[i], cr, X);
            memcpy(&acc[3 * i], X, sizeof X);
            cnt[i] = 1;
            /* neighbours in list order: the sums are order dependent (densify.py:718-724) */
            for (int s = 0; s < nb; ++s) {
                int32_t j = nbrs[k * max_nb + nbk[s]];
                double Xj[3];
                int64_t ui, vi;
                int r = fu_neighbour(X, &cams[j], depth[j], cr, uu[i], vv[i], px_tol, rel_tol[k * max_nb + nbk[s]],
                                     &ui, &vi, Xj);
                ok[i * nb + s] = (uint8_t)r;
                mu[2 * (i * nb + s)] = vi;
                mu[2 * (i * nb + s) + 1] = ui;
                if (r) {
                    acc[3 * i] = acc[3 * i] + Xj[0];
                    acc[3 * i + 1] = acc[3 * i + 1] + Xj[1];
                    acc[3 * i + 2] = acc[3 * i + 2] + Xj[2];
                    cnt[i] = cnt[i] + 1;
                }
            }
        }
        for (int64_t i = 0; i < n; ++i) {
            if (cnt[i] < min_views)
                continue;
            for (int s = 0; s < nb; ++s)
                if (ok[i * nb + s]) {
                    int32_t j = nbrs[k * max_nb + nbk[s]];
                    int64_t vi = mu[2 * (i * nb + s)], ui = mu[2 * (i * nb + s) + 1];
                    used[j][(vi / stride) * stride * (int64_t)cams[j].w + (ui / stride) * stride] = 1;
                }
            if (total < cap) {
                for (int c = 0; c < 3; ++c) {
                    pts[3 * total + c] = acc[3 * i + c] / (double)cnt[i];
                    cols[3 * total + c] = rgb[k][3 * (vv[i] * W + uu[i]) + c];
                }
                nv[total] = (uint8_t)(cnt[i] < 255 ? cnt[i] : 255);
                owner[total] = k;
            }
            ++total;
        }
        free(vv), free(uu), free(acc), free(cnt), free(ok), free(mu);
    }
    return total;
}
