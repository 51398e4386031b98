/*
 * ref_sweep_driver.c -- drives the reference's own fused sweep row kernels (_sweep.h, included
 * from where it lies under the reference tree at build time; see oracle/Makefile) so that the
 * restatement in om_oracle.c can be checked against the reference's compiled arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (built into oracle/_ref/, never shipped or linked by the product).
 * The row order below is the one of _dense.sweep_hypothesis, _dense.pyx:804-835; the per-row
 * arithmetic is entirely the reference's (_sweep.h:33-214).
 */
#include <stdint.h>
#include <stdlib.h>

#include "_sweep.h"

__attribute__((visibility("default"))) int ref_sweep_hypothesis(
    const float *ref, const float *mu, const float *sd, int64_t H, int64_t W, int n, const float *const *imgs,
    const int64_t *img_hw, const float *ray, const double *geo, const float *inv, int r, int top_k, int i,
    float *best, float *prev, float *spb, float *snb, int32_t *idx)
{
    if (n < 1 || n > 16 || r < 0)
        return -1;
    if (H == 0 || W == 0)
        return 0;
    const ptrdiff_t HW = H * W, R2 = 2 * r + 2, k = 2 * r + 1;
    const double kkd = 1.0 / (double)(k * k);
    const ptrdiff_t per_f = 3 * W + R2 * 4 * W;
    om_sw_src src[16];
    float *fblock = malloc(sizeof(float) * (size_t)(n * per_f));
    double *dblock = malloc(sizeof(double) * (size_t)(n * 4 * W + 3 * W));
    if (!fblock || !dblock) {
        free(fblock);
        free(dblock);
        return -3;
    }
    for (int s = 0; s < n; s++) {
        om_sw_src *p = &src[s];
        const double *g = geo + 18 * s;
        p->img = imgs[s];
        p->h = img_hw[2 * s];
        p->w = img_hw[2 * s + 1];
        p->r00 = g[0]; p->r01 = g[1]; p->r02 = g[2];
        p->r10 = g[3]; p->r11 = g[4]; p->r12 = g[5];
        p->r20 = g[6]; p->r21 = g[7]; p->r22 = g[8];
        p->b0 = g[9]; p->b1 = g[10]; p->b2 = g[11];
        p->f = g[12]; p->k1 = g[13]; p->k2 = g[14]; p->k3 = g[15];
        p->cx = g[16]; p->cy = g[17];
        p->Jrow = fblock + s * per_f;
        p->Vrow = p->Jrow + W;
        p->nrow = p->Vrow + W;
        p->ring = p->nrow + W;
        p->cs = dblock + s * 4 * W;
        for (ptrdiff_t x = 0; x < 4 * W; x++)
            p->cs[x] = 0;
    }
    const int kk = top_k < n ? top_k : n;
    double *wU = dblock + n * 4 * W, *wV = wU + W, *wD = wV + W;
    for (ptrdiff_t yy = 0; yy < r && yy < H; yy++) {
        om_sw_warp_row(inv + yy * W, ray + yy * W, ray + HW + yy * W, ray + 2 * HW + yy * W, src, n, W, wU, wV, wD);
        for (int s = 0; s < n; s++) {
            float *hs = src[s].ring + (yy % R2) * 4 * W;
            om_sw_hsum_row(ref + yy * W, &src[s], W, r, hs);
            om_sw_cs_add(src[s].cs, hs, W);
        }
    }
    for (ptrdiff_t y = 0; y < H; y++) {
        if (y + r < H) {
            const ptrdiff_t yy = y + r;
            om_sw_warp_row(inv + yy * W, ray + yy * W, ray + HW + yy * W, ray + 2 * HW + yy * W, src, n, W, wU, wV,
                           wD);
        }
        for (int s = 0; s < n; s++) {
            om_sw_src *p = &src[s];
            if (y + r < H) {
                float *hs = p->ring + ((y + r) % R2) * 4 * W;
                om_sw_hsum_row(ref + (y + r) * W, p, W, r, hs);
                if (y - r - 1 >= 0)
                    om_sw_cs_addsub(p->cs, hs, p->ring + ((y - r - 1) % R2) * 4 * W, W);
                else
                    om_sw_cs_add(p->cs, hs, W);
            } else if (y - r - 1 >= 0) {
                om_sw_cs_sub(p->cs, p->ring + ((y - r - 1) % R2) * 4 * W, W);
            }
            om_sw_ncc_row(p->cs, mu + y * W, sd + y * W, kkd, W, p->nrow);
        }
        om_sw_combine_row(src, n, kk, i, W, best + y * W, prev + y * W, spb + y * W, snb + y * W, idx + y * W);
    }
    free(fblock);
    free(dblock);
    return 0;
}
