// Semi-global matching on gfx950: _mvs.sgm (_mvs.pyx:98-250), the cost aggregation of the height
// sweep (mvs.py:325), over B cost volumes (D, H, W) at once.
//
// Along each of the eight path directions r (the reference's order (0,1), (0,-1), (1,0), (-1,0),
// (1,1), (1,-1), (-1,1), (-1,-1)):
//   L(p, d) = (C(p, d) + min(L(q, d), L(q, d-1) + P1, L(q, d+1) + P1, best + P2)) - best,
//   best = min_k L(q, k),  q = p - r  (L = C where q lies outside),
// and the result is 0 + L_0 + ... + L_7 in that order, all in float32 with the reference's
// comparison order (_mvs.pyx:55-95), so results are bit-identical.
//
// Layout: the cost is transposed once to (B, H, W, D) so that a path step reads one cell's D
// values contiguously; every path of every direction and volume is one wave (lanes hold
// contiguous runs of D), all in one launch; each direction's L goes to its own (B, H, W, D)
// buffer; a last kernel adds the eight in order and transposes back to (B, D, H, W).
//
// best = the reference's sequential minimum: prv[0] if it is NaN, otherwise the first minimal
// non-NaN value (NaN at d > 0 never replaces it).  The wave reduction maps NaN at d > 0 to +inf,
// takes the minimum, and when the minimum is zero takes the sign of the first zero in d order.
#include "om_common.h"

#include <cmath>

namespace om {
namespace sg {

constexpr int kPathWaves = 4;   // paths (waves) per workgroup
constexpr int kTile = 32;

__constant__ int kDirs[8][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};

__host__ __device__ inline int64_t paths_per_volume(int H, int W) { return 2LL * H + 2LL * W + 4LL * (H + W - 1); }

// (B, D, HW) -> (B, HW, D) through a 32 x 32 LDS tile
__global__ __launch_bounds__(256) void sgm_to_hwd_kernel(const float *__restrict__ cost, float *__restrict__ C,
                                                          int D, int64_t HW)
{
    __shared__ float t[kTile][kTile + 1];
    const int64_t b = blockIdx.z;
    const int64_t c0 = (int64_t)blockIdx.x * kTile;
    const int d0 = blockIdx.y * kTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const float *src = cost + b * D * HW;
    for (int j = ty; j < kTile; j += 8) {
        const int d = d0 + j;
        const int64_t c = c0 + tx;
        if (d < D && c < HW)
            t[j][tx] = src[(int64_t)d * HW + c];
    }
    __syncthreads();
    float *dst = C + b * D * HW;
    for (int j = ty; j < kTile; j += 8) {
        const int64_t c = c0 + j;
        const int d = d0 + tx;
        if (d < D && c < HW)
            dst[c * D + d] = t[tx][j];
    }
}

// agg (B, D, HW) = 0 + L_0 + ... + L_7, the L_k in (B, HW, D)
__global__ __launch_bounds__(256) void sgm_sum_kernel(const float *__restrict__ L, float *__restrict__ agg, int D,
                                                       int64_t HW, int64_t vol, int64_t B)
{
    __shared__ float t[kTile][kTile + 1];
    const int64_t b = blockIdx.z;
    const int64_t c0 = (int64_t)blockIdx.x * kTile;
    const int d0 = blockIdx.y * kTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t dir_stride = B * vol;
    for (int j = ty; j < kTile; j += 8) {
        const int64_t c = c0 + j;
        const int d = d0 + tx;
        if (d < D && c < HW) {
            const float *p = L + b * vol + c * D + d;
            float a = 0.0f;
            for (int k = 0; k < 8; ++k)
                a += p[k * dir_stride];
            t[tx][j] = a;
        }
    }
    __syncthreads();
    for (int j = ty; j < kTile; j += 8) {
        const int d = d0 + j;
        const int64_t c = c0 + tx;
        if (d < D && c < HW)
            agg[b * vol + (int64_t)d * HW + c] = t[j][tx];
    }
}

// DPP moves (VALU, no LDS round trip): quad permutes and row rotations inside 16-lane rows,
// whole-wave shifts by one lane (GFX9-family DPP)
template <int CTRL>
__device__ __forceinline__ float dpp(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, 0xF, 0xF, false));
}
constexpr int kQuadXor1 = 0xB1;    // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;    // quad_perm [2, 3, 0, 1]
constexpr int kRowRor4 = 0x124;
constexpr int kRowRor8 = 0x128;
constexpr int kWaveShr1 = 0x138;   // lane i <- lane i - 1
constexpr int kWaveShl1 = 0x130;   // lane i <- lane i + 1

__device__ __forceinline__ float readlane(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }

// the smaller by comparison and selection (no canonicalising min instruction touches the value)
__device__ __forceinline__ float lesser(float a, float b) { return b < a ? b : a; }

// minimum over the wave of NaN-free values (the sign of a zero minimum is fixed by the caller)
__device__ __forceinline__ float wave_min(float m)
{
    m = lesser(m, dpp<kQuadXor1>(m));
    m = lesser(m, dpp<kQuadXor2>(m));
    m = lesser(m, dpp<kRowRor4>(m));
    m = lesser(m, dpp<kRowRor8>(m));
    return lesser(lesser(readlane(m, 0), readlane(m, 16)), lesser(readlane(m, 32), readlane(m, 48)));
}

// one path per wave; lane l holds d in [l * DPL, l * DPL + DPL)
template <int DPL>
__global__ __launch_bounds__(64 * kPathWaves) void sgm_path_kernel(const float *__restrict__ C, float *__restrict__ L,
                                                                    int64_t B, int D, int H, int W, float P1, float P2)
{
    const int lane = threadIdx.x & 63;
    const int64_t P = paths_per_volume(H, W);
    // XCD-aware: each XCD walks a contiguous range of paths (whole volumes), so the eight passes
    // over a volume's costs hit that XCD's L2
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t id = blk * kPathWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (id >= B * P)
        return;
    const int64_t b = id / P;
    int64_t r = id - b * P;
    // direction k and its start cell
    int k, y, x;
    if (r < 2LL * H) {                  // (0, +-1): one path per row
        k = (int)(r / H);
        y = (int)(r - (int64_t)k * H);
        x = k == 0 ? 0 : W - 1;
    } else if ((r -= 2LL * H) < 2LL * W) {   // (+-1, 0): one path per column
        k = 2 + (int)(r / W);
        x = (int)(r - (int64_t)(k - 2) * W);
        y = k == 2 ? 0 : H - 1;
    } else {                            // diagonals: start on the entry row, then the entry column
        r -= 2LL * W;
        const int64_t n = H + W - 1;
        k = 4 + (int)(r / n);
        const int q = (int)(r - (int64_t)(k - 4) * n);
        const int dy = kDirs[k][0], dx = kDirs[k][1];
        if (q < W) {
            y = dy > 0 ? 0 : H - 1;
            x = q;
        } else {
            y = dy > 0 ? q - W + 1 : H - 1 - (q - W + 1);
            x = dx > 0 ? 0 : W - 1;
        }
    }
    const int dy = kDirs[k][0], dx = kDirs[k][1];
    const int64_t vol = (int64_t)H * W * D;
    const float *Cb = C + b * vol;
    float *Lb = L + ((int64_t)k * B + b) * vol;
    const int dbase = lane * DPL;
    float prv[DPL], c[DPL];
    const float inf = __builtin_inff();
    {
        const int64_t cell = (int64_t)y * W + x;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const int d = dbase + i;
            prv[i] = d < D ? Cb[cell * D + d] : inf;
            if (d < D)
                Lb[cell * D + d] = prv[i];
        }
    }
    y += dy;
    x += dx;
    if (y >= 0 && y < H && x >= 0 && x < W) {
        const int64_t cell = (int64_t)y * W + x;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const int d = dbase + i;
            c[i] = d < D ? Cb[cell * D + d] : 0.0f;
        }
    }
    while (y >= 0 && y < H && x >= 0 && x < W) {
        const int64_t cell = (int64_t)y * W + x;
        // prefetch the next cell's costs
        const int ny = y + dy, nx = x + dx;
        const bool more = ny >= 0 && ny < H && nx >= 0 && nx < W;
        float cn[DPL];
        if (more) {
            const int64_t ncell = (int64_t)ny * W + nx;
#pragma unroll
            for (int i = 0; i < DPL; ++i) {
                const int d = dbase + i;
                cn[i] = d < D ? Cb[ncell * D + d] : 0.0f;
            }
        }
        // best = min_k prv[k] with the reference's sequential semantics
        float m = inf;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const float v = (isnan(prv[i]) && dbase + i > 0) ? inf : prv[i];
            m = v < m ? v : m;
        }
        m = wave_min(m);
        if (m == 0.0f) {                  // the sign of the first zero in d order
            float z = 0.0f;
            int has = 0;
#pragma unroll
            for (int i = DPL - 1; i >= 0; --i)
                if (dbase + i < D && prv[i] == 0.0f) {
                    z = prv[i];
                    has = 1;
                }
            const uint64_t bal = __ballot(has);
            m = readlane(z, __ffsll((long long)bal) - 1);
        }
        const float p0 = readlane(prv[0], 0);
        const float best = isnan(p0) ? p0 : m;
        const float left = dpp<kWaveShr1>(prv[DPL - 1]);    // prv[dbase - 1] (lane 0: unused)
        const float right = dpp<kWaveShl1>(prv[0]);         // prv[dbase + DPL] (lane 63: unused)
        float cur[DPL];
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const int d = dbase + i;
            float v = prv[i];
            if (d > 0) {
                const float a = (i > 0 ? prv[i - 1] : left) + P1;
                if (a < v)
                    v = a;
            }
            if (d < D - 1) {
                const float bb = (i < DPL - 1 ? prv[i + 1] : right) + P1;
                if (bb < v)
                    v = bb;
            }
            if (best + P2 < v)
                v = best + P2;
            const float t = c[i] + v;
            cur[i] = d < D ? t - best : inf;
            if (d < D)
                Lb[cell * D + d] = cur[i];
        }
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            prv[i] = cur[i];
            c[i] = cn[i];
        }
        y = ny;
        x = nx;
    }
}

inline int dpl_of(int D) { return D <= 64 ? 1 : D <= 128 ? 2 : D <= 256 ? 4 : D <= 512 ? 8 : 0; }

}  // namespace sg
}  // namespace om

using namespace om;
using namespace om::sg;

extern "C" size_t om_sgm_workspace_bytes(int64_t B, int32_t D, int32_t H, int32_t W)
{
    if (B <= 0 || D <= 0 || H <= 0 || W <= 0)
        return 0;
    return (size_t)9 * (size_t)B * (size_t)D * (size_t)H * (size_t)W * sizeof(float);
}

extern "C" int om_sgm(const float *cost, int64_t B, int32_t D, int32_t H, int32_t W, float P1, float P2, float *agg,
                      void *workspace, size_t workspace_bytes, void *stream)
{
    if (B < 0 || D < 0 || H < 0 || W < 0)
        return fail(OM_EINVAL, "om_sgm: negative size");
    if (B == 0 || D == 0 || H == 0 || W == 0)
        return OM_OK;
    if (!cost || !agg || !workspace)
        return fail(OM_EINVAL, "om_sgm: null pointer");
    const int dpl = dpl_of(D);
    if (!dpl)
        return fail(OM_EUNSUP, "om_sgm: %d cost planes (at most 512)", D);
    if (workspace_bytes < om_sgm_workspace_bytes(B, D, H, W))
        return fail(OM_EINVAL, "om_sgm: workspace %zu bytes, %zu needed", workspace_bytes,
                    om_sgm_workspace_bytes(B, D, H, W));
    const int64_t HW = (int64_t)H * W, vol = HW * D;
    if (B > 65535 || cdiv(HW, kTile) > 0x7fffffff)
        return fail(OM_EUNSUP, "om_sgm: %lld volumes of %lld cells", (long long)B, (long long)HW);
    float *C = static_cast<float *>(workspace), *L = C + B * vol;
    hipStream_t s = as_stream(stream);
    ProfScope ps("sgm", s);
    const dim3 tgrid((unsigned)cdiv(HW, kTile), (unsigned)cdiv(D, kTile), (unsigned)B);
    hipLaunchKernelGGL(sgm_to_hwd_kernel, tgrid, dim3(256), 0, s, cost, C, D, HW);
    const int64_t paths = B * paths_per_volume(H, W);
    const dim3 pgrid((unsigned)cdiv(paths, kPathWaves)), pblk(64 * kPathWaves);
    {
        ProfScope pp("sgm_path", s);
        switch (dpl) {
        case 1: hipLaunchKernelGGL(sgm_path_kernel<1>, pgrid, pblk, 0, s, C, L, B, D, H, W, P1, P2); break;
        case 2: hipLaunchKernelGGL(sgm_path_kernel<2>, pgrid, pblk, 0, s, C, L, B, D, H, W, P1, P2); break;
        case 4: hipLaunchKernelGGL(sgm_path_kernel<4>, pgrid, pblk, 0, s, C, L, B, D, H, W, P1, P2); break;
        default: hipLaunchKernelGGL(sgm_path_kernel<8>, pgrid, pblk, 0, s, C, L, B, D, H, W, P1, P2); break;
        }
    }
    hipLaunchKernelGGL(sgm_sum_kernel, tgrid, dim3(256), 0, s, L, agg, D, HW, vol, B);
    return launch_check("om_sgm");
}
