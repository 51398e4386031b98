/*
 * om_sgm.c -- CPU restatement of pyOrthomosaic's semi-global matching (_mvs.sgm, _mvs.pyx:98-250;
 * linked into liborc.so beside om_oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP SGM kernels and the CPU baseline that bench.py
 * times beside them.  The product library never links, loads or calls it.
 *
 * Eight path directions in the reference's order ((0,1), (0,-1), (1,0), (-1,0), (1,1), (1,-1),
 * (-1,1), (-1,-1)); along each,
 *   L(p, d) = (C(p, d) + min(L(q, d), L(q, d-1) + P1, L(q, d+1) + P1, best + P2)) - best,
 *   best = min_k L(q, k),  q = p - r,
 * with the reference's float32 operations and comparison order (_mvs.pyx:55-95), and the sum
 * A = 0 + L_0 + ... + L_7 in direction order.
 *
 * Parity pins: tests/golden/sgm_volumes.npz (written by the compiled reference: the shapes of
 * test_native_parity.py:438-447 with NaN and zero costs, and larger volumes), bit for bit; a
 * brute-force Python restatement of the recurrence on small volumes.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#define ORC_EXPORT __attribute__((visibility("default")))

static const int sgm_dirs[8][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};

/* one path direction on (H, W, D) buffers (_mvs.pyx:55-95) */
static void sgm_dir(const float *C, float *L, int64_t D, int64_t H, int64_t W, int dy, int dx, float P1, float P2)
{
    for (int64_t yy = 0; yy < H; ++yy) {
        int64_t y = dy >= 0 ? yy : H - 1 - yy;
        for (int64_t xx = 0; xx < W; ++xx) {
            int64_t x = dx >= 0 ? xx : W - 1 - xx;
            int64_t py = y - dy, px = x - dx;
            const float *c = C + (y * W + x) * D;
            float *cur = L + (y * W + x) * D;
            if (py < 0 || py >= H || px < 0 || px >= W) {
                for (int64_t d = 0; d < D; ++d)
                    cur[d] = c[d];
                continue;
            }
            const float *prv = L + (py * W + px) * D;
            float best = prv[0];
            for (int64_t d = 1; d < D; ++d)
                if (prv[d] < best)
                    best = prv[d];
            for (int64_t d = 0; d < D; ++d) {
                float v = prv[d];
                if (d > 0) {
                    float a = prv[d - 1] + P1;
                    if (a < v)
                        v = a;
                }
                if (d < D - 1) {
                    float b = prv[d + 1] + P1;
                    if (b < v)
                        v = b;
                }
                if (best + P2 < v)
                    v = best + P2;
                float t = c[d] + v;
                cur[d] = t - best;
            }
        }
    }
}

/* _mvs.sgm: cost (D, H, W) float32 -> agg (D, H, W) float32.  Returns 0, or -3 without memory. */
ORC_EXPORT int orc_sgm(const float *cost, int64_t D, int64_t H, int64_t W, float P1, float P2, float *agg)
{
    int64_t n = D * H * W;
    if (n == 0)
        return 0;
    float *C = malloc((size_t)n * sizeof(float)), *L = malloc((size_t)n * sizeof(float)),
          *A = malloc((size_t)n * sizeof(float));
    if (!C || !L || !A) {
        free(C), free(L), free(A);
        return -3;
    }
    for (int64_t d = 0; d < D; ++d)
        for (int64_t i = 0; i < H * W; ++i)
            C[i * D + d] = cost[d * H * W + i];
    for (int64_t i = 0; i < n; ++i)
        A[i] = 0;
    for (int k = 0; k < 8; ++k) {
        sgm_dir(C, L, D, H, W, sgm_dirs[k][0], sgm_dirs[k][1], P1, P2);
        for (int64_t i = 0; i < n; ++i)
            A[i] += L[i];
    }
    for (int64_t d = 0; d < D; ++d)
        for (int64_t i = 0; i < H * W; ++i)
            agg[d * H * W + i] = A[i * D + d];
    free(C), free(L), free(A);
    return 0;
}
