/*
 * om355.h -- C ABI of libom355.so, the MI355X (gfx950) depth-map kernels of py_ortho_amd.
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned device memory (a torch
 * tensor's data_ptr() from Python); the library never frees caller memory.  `stream` is a
 * hipStream_t (NULL = the default stream).  Every entry point returns 0 on success or a negative
 * OM_E* code; om_last_error() then describes the failure (thread-local).  Calls only enqueue work
 * on `stream` unless stated otherwise; they are reentrant across threads and streams.
 *
 * Reference interfaces replaced (pyOrthomosaic 0.9.3):
 *   om_sweep_hypothesis   _dense.sweep_hypothesis          orthomosaic/_dense.pyx:740-838
 *   om_sweep              the hypothesis loop of _sweep      orthomosaic/densify.py:336-342
 *   om_sweep_rows         ncc_warp + combine (n > 16 path)   orthomosaic/_dense.pyx:13-144, densify.py:343-353
 *   om_sweep_finalize     the parabola of _sweep             orthomosaic/densify.py:354-361
 *   om_pm_refine          _dense.pm_refine                  orthomosaic/_dense.pyx:566-666
 *   om_box_zero_mean      _fast.box_zero_mean (densify._box) orthomosaic/_fast.pyx:661-713
 *   om_hyp_linear         _fast.hyp_linear                  orthomosaic/_fast.pyx:628-641
 *   om_hyp_refine         _fast.hyp_refine                  orthomosaic/_fast.pyx:644-658
 *   om_half               densify._half                     orthomosaic/densify.py:127-130
 *   om_ref_stats          mu / sd of _sweep                 orthomosaic/densify.py:327-328
 *   om_consistency        densify._consistency              orthomosaic/densify.py:441-478
 *   om_fuse               densify._fuse / _fuse_loop        orthomosaic/densify.py:676-762
 *   om_sgm                _mvs.sgm                          orthomosaic/_mvs.pyx:98-250
 *   om_depth_view         _fast.depth_view                  orthomosaic/_fast.pyx:577-616
 *   om_depth_band         densify._depth_band + linear step orthomosaic/densify.py:233-274, :409-411
 *   om_f32_to_f16 / om_f16_to_f32 / om_filter_depth  the depth-map file set  densify.py:590-592, :613-614, :641
 */
#ifndef OM355_H
#define OM355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OM_OK = 0,
    OM_EINVAL = -1,   /* bad argument (reference: ValueError) */
    OM_ESHAPE = -2,   /* shapes disagree (reference: ValueError) */
    OM_ENOMEM = -3,   /* allocation failed (reference: MemoryError) */
    OM_EHIP = -4,     /* HIP runtime error */
    OM_ENSRC = -5,    /* number of sources outside 1..16 (reference: ValueError) */
    OM_EUNSUP = -6    /* valid input the library does not handle */
};

/* hypothesis generators of om_sweep (inverse depth of hypothesis i at pixel p) */
enum {
    OM_HYP_MAPS = 0,   /* hyp_a: (nh, H, W) float32 maps, map j is hypothesis i0 + j */
    OM_HYP_LINEAR = 1, /* hyp_a = base, hyp_b = step: base + f32(i) * step (densify.py:411-413) */
    OM_HYP_REFINE = 2, /* hyp_a = iu, hyp_b = ds: max(iu + f32(i - hyp_offset) * ds, 1e-6) (densify.py:423-426) */
    OM_HYP_CONST = 3   /* hyp_a: (nh,) float32 table, constant over the image (fronto-parallel planes) */
};

/* One source view.  img: (h, w) float32 device image, NaN = masked.  geo: the reference's 18
 * values per source (_dense.pyx:748): Rs row-major (9), b = Rs (C_ref - C_src) (3), f, k1, k2, k3,
 * cx, cy. */
typedef struct om_src {
    const float *img;
    int32_t h, w;
    double geo[18];
} om_src;

typedef struct om_sweep_desc {
    const float *ref, *mu, *sd; /* (H, W) float32: reference image (masked = 0) and its window mean/std */
    const float *ray;           /* (3, H, W) float32 world rays, camera-axis depth 1 */
    int32_t H, W;
    int32_t n_src;              /* 1..16 */
    const om_src *src;          /* host array of n_src descriptors */
    int32_t r, top_k;           /* window radius (window 2r+1), score = mean of the best top_k sources */
    int32_t hyp_mode;           /* OM_HYP_* */
    const float *hyp_a, *hyp_b; /* generator inputs (device) */
    int32_t hyp_offset;         /* OM_HYP_REFINE: t = i - hyp_offset */
    int32_t i0, nh;             /* hypotheses i0 .. i0 + nh - 1 (absolute indices for the winner) */
    /* streaming winner state (H, W), updated in place: _sweep initialises best/prev/spb/snb to -2
     * and idx to -1 (densify.py:331-335) */
    float *best, *prev, *s_prev_best, *s_next_best;
    int32_t *idx;
    /* device workspace for hypothesis chunking; may be NULL (then one chunk) */
    void *workspace;
    size_t workspace_bytes;
    int32_t chunks;             /* 0 = choose from the image size; 1 = no chunking */
    int32_t tile_h;             /* tile rows: 32 (32 columns), 16 or 8 (64 columns); 0 = 32 */
} om_sweep_desc;

/* library */
int om_version(void);
/* measurement: when on, every kernel launch is bracketed by HIP events on its stream;
 * om_profile_read sums (and forgets) the recorded launches of `kernel` (NULL = all kernels). */
int om_profile_enable(int32_t on);
int om_profile_read(const char *kernel, double *total_ms, int64_t *launches);
const char *om_last_error(void);
/* name (>= 64 bytes), compute units, clock kHz, global memory bytes of the current device */
int om_device_info(char *name, int32_t name_len, int32_t *cus, int32_t *clock_khz, int64_t *mem_bytes);

/* fused plane sweep: hypotheses [i0, i0+nh) of one view in one launch (+ a merge when chunked).
 * d->chunks = 0 chooses the hypothesis chunks for the chip (the count whose rounds of workgroups
 * sweep the fewest planes) among those whose partials fit d->workspace; om_sweep_workspace_bytes
 * gives the bytes of that choice unconstrained.  Any chunk count gives the same results. */
size_t om_sweep_workspace_bytes(const om_sweep_desc *d);
int om_sweep(const om_sweep_desc *d, void *stream);
/* The reference's unfused sweep (_dense.ncc_warp per source + _dense.combine, _dense.pyx:13-144,
 * the path densify._sweep takes above 16 sources, densify.py:343-353): horizontal running sums along
 * whole rows and vertical ones down whole columns in the reference's order, so every result equals
 * the reference bit for bit for any input.  Any source count (1..256), window radius r >= 0 and
 * top_k >= 1; same descriptor as om_sweep, with d->chunks = hypotheses per pass (0 = choose) and
 * a required workspace of om_sweep_rows_workspace_bytes(d) bytes (the pass's window sums and
 * scores, 20 bytes per pixel, source and hypothesis). */
size_t om_sweep_rows_workspace_bytes(const om_sweep_desc *d);
int om_sweep_rows(const om_sweep_desc *d, void *stream);
/* om_sweep that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same descriptor and results as om_sweep; d->workspace is ignored and the hypothesis
 * chunks' partial winner states go to a library-owned static device array of
 * om_sweep_noalloc_max_bytes() bytes (64 MiB) per GPU; with d->chunks = 0 the count is chosen among
 * those whose partials fit there.  Returns OM_ESHAPE when an explicit d->chunks needs more (pass
 * chunks = 0 or 1; one chunk needs none).
 * Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
size_t om_sweep_noalloc_max_bytes(void);
int om_sweep_noalloc(const om_sweep_desc *d, void *stream);

/* _dense.sweep_hypothesis, one hypothesis i with an explicit (H, W) inverse-depth map `inv`:
 * imgs/img_hw: n device image pointers and their (h, w); geo: host (n, 18) float64. */
int om_sweep_hypothesis(const float *ref, const float *mu, const float *sd, int32_t H, int32_t W, int32_t n,
                        const float *const *imgs, const int32_t *img_hw, const float *ray, const double *geo,
                        const float *inv, int32_t r, int32_t top_k, int32_t i, float *best, float *prev,
                        float *s_prev_best, float *s_next_best, int32_t *idx, void *stream);

/* parabola of _sweep (densify.py:354-361): depth = 1 / max(inv0 + (idx + off) * dinv, 1e-9) as
 * float32, with inv0/dinv the first hypothesis map and the spacing of d's generator (n = d->nh
 * hypotheses starting at d->i0 = 0). */
int om_sweep_finalize(const om_sweep_desc *d, float *depth, void *stream);

/* PatchMatch repair of one view: _dense.pm_refine.  Device arrays: ref (H,W) NaN = masked,
 * ray (3,H,W), depth (H,W), normal (3,H,W), score (H,W) in/out, active (H,W) uint8; imgs/deps:
 * S device pointers (deps[s] may be NULL), ih/iw/par (S,18)/Cr(3): host.  *changed receives the
 * number of accepted hypothesis changes (the reference's return value); it synchronises `stream`
 * when changed != NULL.  workspace: om_pm_workspace_bytes(H, W) bytes of device memory.
 * With geo_w > 0 a random try whose geometric cost alone (less a bound on the photometric score's
 * rounding above 1) is not below the pixel's cost is rejected without its photometric part: the
 * reference's comparison rejects it whatever that part is, so the results are the same. */
size_t om_pm_workspace_bytes(int32_t H, int32_t W);
int om_pm_refine(const float *ref, int32_t H, int32_t W, const float *ray, const double *Cr, int32_t S,
                 const float *const *imgs, const float *const *deps, const int32_t *ih, const int32_t *iw,
                 const double *par, float *depth, float *normal, float *score, const uint8_t *active, int32_t r,
                 int32_t step, int32_t top_k, double geo_w, double geo_tol, int32_t iters, double dmin, double dmax,
                 uint32_t seed, void *workspace, size_t workspace_bytes, int64_t *changed, void *stream);

/* depth helpers */
int om_hyp_linear(const float *base, const float *step, int64_t n, int64_t i, float *out, void *stream);
int om_hyp_refine(const float *iu, const float *ds, int64_t n, int64_t t, float *out, void *stream);
int om_box_zero_mean(const float *a, int32_t H, int32_t W, int32_t r, float *out, void *workspace,
                     size_t workspace_bytes, void *stream);
size_t om_box_workspace_bytes(int32_t H, int32_t W, int32_t r);
/* mu = box(ref), sd = sqrt(max(box(ref*ref) - mu*mu, 1e-6)) (densify.py:327-328) */
int om_ref_stats(const float *ref, int32_t H, int32_t W, int32_t r, float *mu, float *sd, void *workspace,
                 size_t workspace_bytes, void *stream);
int om_half(const float *g, int32_t H, int32_t W, float *out, void *stream);
/* densify._depth_band (densify.py:233-274) on the device: uvz (n, 3) float64 = the view's own tie
 * points in front of the coarse camera, projected (u, v) with their camera-axis depth z > 0;
 * n_obs = the view's observation count (below 50: the range [lo_g, hi_g] everywhere).  Per block of
 * block x block pixels with >= 5 points: the depth span over the 3x3 block neighbourhood widened by
 * margin + 5 % of it, clipped to [lo_g, hi_g]; other blocks [lo_g, hi_g].  Writes inv_lo = f32(1 / hi)
 * and, when not NULL, inv_hi = f32(1 / max(lo, 1e-6)) and step = (inv_hi - inv_lo) / max(nd - 1, 1)
 * in float32 (the coarse linear generator, densify.py:411), all (H, W).  Workspace:
 * om_depth_band_workspace_bytes(H, W, block) bytes of device memory (unused when n_obs < 50). */
size_t om_depth_band_workspace_bytes(int32_t H, int32_t W, int32_t block);
int om_depth_band(const double *uvz, int64_t n, int32_t n_obs, int32_t H, int32_t W, int32_t block, double lo_g,
                  double hi_g, double margin, int32_t nd, float *inv_lo, float *inv_hi, float *step, void *workspace,
                  size_t workspace_bytes, void *stream);
/* densify._rays_world (densify.py:288-297): (3, H, W) float32 world rays, camera-axis depth 1;
 * R: host row-major world->camera rotation. */
int om_rays_world(int32_t H, int32_t W, const double *R, double cx, double cy, double f, double k1, double k2,
                  double k3, float *out, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Batches: many views' depth maps planned once and launched together (no counterpart call in
 * the reference, which runs one view per host thread, densify.py:596).  A batch owns device
 * descriptor tables and workspaces, allocated by the om_batch_set_* calls; om_batch_run only
 * enqueues kernels, so it can be captured in a HIP graph (om_batch_capture) and replayed.
 * ------------------------------------------------------------------------------------------- */
typedef struct om_batch om_batch;

/* one view of a batched PatchMatch: the arguments of om_pm_refine for that view */
typedef struct om_pm_view {
    const float *ref;              /* (H, W), NaN = masked */
    int32_t H, W;
    const float *ray;              /* (3, H, W) */
    double Cr[3];
    int32_t S;                     /* 1..16 */
    const float *imgs[16];         /* device images */
    const float *deps[16];         /* device depth maps or NULL */
    int32_t ih[16], iw[16];
    double par[16][18];            /* R (9), C (3), f, k1, k2, k3, cx, cy per source */
    float *depth, *normal, *score; /* in/out */
    const uint8_t *active;         /* (H, W) */
    double geo_tol, dmin, dmax;
    uint32_t seed;
    /* optional start maps (device, NULL = none): every run first copies depth_from (H, W) into
     * depth and score_from (H, W) into score; with init_normals the normals come from the start depth */
    const float *depth_from, *score_from;
} om_pm_view;

om_batch *om_batch_create(void);
int om_batch_destroy(om_batch *b);
/* sweep of nviews views sharing r and min(top_k, n_src) (each desc's workspace is ignored);
 * every run resets each view's winner state and sweeps its hypotheses [0, nh) (i0 must be 0);
 * depth_out[v] (array may be NULL, entries may be NULL) receives view v's parabola depth;
 * chunks: hypothesis chunks for every view (0 = each desc's chunks, 0 there = auto);
 * tile_h: 8, 16 or 32 for every view (0 = each desc's tile_h, 0 there = 32); views of one tile height
 * share one persistent launch; grid_blocks: workgroups per launch (0 = all resident). */
int om_batch_set_sweep(om_batch *b, const om_sweep_desc *views, int32_t nviews, float *const *depth_out,
                       int32_t chunks, int32_t tile_h, int32_t grid_blocks);
/* PatchMatch over nviews views after the sweep (same r, step, top_k, geo_w, iters for all);
 * init_normals: first set each view's normals from its depth (densify._init_normals, densify.py:481-496). */
int om_batch_set_pm(om_batch *b, const om_pm_view *views, int32_t nviews, int32_t r, int32_t step, int32_t top_k,
                    double geo_w, int32_t iters, int32_t init_normals);
/* enqueue the batch's kernels on stream: the sweep (+merge, parabola), then the PatchMatch passes */
int om_batch_run(om_batch *b, void *stream);
/* record om_batch_run into a HIP graph (stream capture on `stream`) and instantiate it */
int om_batch_capture(om_batch *b, void *stream);
int om_batch_replay(om_batch *b, void *stream);
/* accepted PatchMatch changes per view (synchronises the device) */
int om_batch_pm_changed(om_batch *b, int64_t *changed);
/* After a run: the propagation cost evaluations each view performed (caller's view order), as
 * om_pm_workspace_counts reports them for one view; -1 where the batch runs without the memo. */
int om_batch_pm_evaluated(om_batch *b, int64_t *evaluated);

/* One step of a continuous-batching server: for every view, hypotheses [i0, i0 + nh) continuing its
 * winner state (views of any size, source count and hypothesis generator; same r and min(top_k, n)),
 * in one persistent launch per tile height; views with depth_out[v] != NULL finish (their i0 + nh is
 * the whole sweep) and get the parabola depth.  The parabola reads hypotheses 0 and 1, so for a
 * finishing view with i0 > 0 and OM_HYP_CONST or OM_HYP_MAPS, hyp_a must point i0 entries (maps)
 * into a table that starts at hypothesis 0: hyp_a[-i0 ...] is read.  Equal, view by view, to om_sweep
 * over the same range followed by om_sweep_finalize.  The call allocates nothing: table is device memory and staging
 * pinned host memory (hipHostMalloc), both of om_sweep_multi_bytes(nviews); the table is uploaded from
 * staging on `stream`, so staging may be reused once the stream has passed this call.  Replaces the
 * per-view thread pool of densify.py:596 for a stream of requests. */
size_t om_sweep_multi_bytes(int32_t nviews);
int om_sweep_multi(const om_sweep_desc *views, int32_t nviews, float *const *depth_out, int32_t tile_h, void *table,
                   size_t table_bytes, void *staging, void *stream);

/* Storage formats of the precision variants (om_sweep_lp, om_pm_refine_lp) and om_to_lp: FP16,
 * BF16, or FP32 (the images as they are: the variants' FP32 arithmetic without the 16-bit rounding,
 * the baseline their storage error is stated against) */
#define OM_FMT_F16 1
#define OM_FMT_BF16 2
#define OM_FMT_F32 3

/* float32 -> FP16 or BF16 (round to nearest even, NaN kept), or a copy for FP32, n elements */
int om_to_lp(const float *src, int64_t n, int32_t fmt, void *dst, void *stream);

/* The precision variant of om_sweep (no reference counterpart: the reference computes in FP64 on
 * the CPU, _dense.pyx:740-838): reference (ref16) and source images (src[s].img) stored in FP16,
 * BF16 or FP32 (fmt), warp and bilinear in FP32, window sums and NCC in FP32, plane depth and the variance guard in
 * FP64.  Hypotheses [i0, i0 + nh) continue the desc's winner state (one chunk); the radial model
 * (k1..k3) is evaluated in FP32; the state is finished by om_sweep_finalize.  Errors against the FP64 path are stated in
 * the tests and benchmark. */
int om_sweep_lp(const om_sweep_desc *d, int32_t fmt, const void *ref16, void *stream);

/* After om_pm_refine / om_pm_refine_lp on `workspace`: the propagation cost evaluations actually
 * computed (the memoised path reuses a candidate's cost while its plane is unchanged; the reference
 * evaluates H*W*9 per iteration, _dense.pyx:614-645); -1 when the call had no memo (very large views). */
int om_pm_workspace_counts(const void *workspace, int32_t H, int32_t W, int64_t *evaluated);

/* The precision variant of om_pm_refine (no reference counterpart: the reference's cost is FP64,
 * _dense.pyx:461-556): source images imgs16[s] stored as FP16, BF16 or FP32 (fmt), the plane (num = d n.ray)
 * and the geometric term in FP64, per sample the ray-plane intersection, projection and bilinear tap
 * in FP32, window sums in FP32, the variances and their guard in FP64.  Same red-black order and random
 * stream as om_pm_refine; the radial model (k1..k3) in FP32; same workspace. */
int om_pm_refine_lp(const float *ref, int32_t H, int32_t W, const float *ray, const double *Cr, int32_t S,
                    const void *const *imgs16, const float *const *deps, const int32_t *ih, const int32_t *iw,
                    const double *par, float *depth, float *normal, float *score, const uint8_t *active, int32_t r,
                    int32_t step, int32_t top_k, double geo_w, double geo_tol, int32_t iters, double dmin,
                    double dmax, uint32_t seed, void *workspace, size_t workspace_bytes, int64_t *changed,
                    int32_t fmt, void *stream);

/* Camera of densify._Cam (densify.py:133-160) with float64 parameters: R row-major world->camera,
 * C the centre, pinhole f / (cx, cy) with radial k1, k2, k3, image w x h. */
typedef struct om_cam {
    double R[9];
    double C[3];
    double f, k1, k2, k3, cx, cy;
    int32_t w, h;
} om_cam;

/* densify._consistency (densify.py:441-478): cnt (cam.h, cam.w) int16 = for each finite pixel of
 * depth, the number of the n_nb (<= 16) neighbours that agree with it (forward-backward reprojection
 * within px_tol pixels and relative depth within rel_tol[s]); 0 where depth is not finite.
 * nb_depth[s]: device (nb_cam[s].h, nb_cam[s].w) float32 depth maps; nb_cam, rel_tol: host arrays.
 * FP64 in the reference's order (_fast.pyx:201-274), 3x3 products as NumPy/OpenBLAS form them. */
int om_consistency(const float *depth, const om_cam *cam, int32_t n_nb, const float *const *nb_depth,
                   const om_cam *nb_cam, const double *rel_tol, double px_tol, int16_t *cnt, void *stream);

/* One view of om_fuse: device depth (h, w) float32 (NaN = none), colours (h, w, 3) uint8 (may be
 * NULL when cols is NULL), and its used flags on the stride grid (ceil(h / stride), ceil(w / stride))
 * uint8, zero before the first call (the reference's used map, densify.py:657, is only set and read
 * on grid pixels). */
typedef struct om_fuse_view {
    const float *depth;
    const uint8_t *rgb;
    uint8_t *used;
    om_cam cam;
} om_fuse_view;

/* workspace of om_fuse: per-grid-pixel staging for the largest view plus block counts */
size_t om_fuse_workspace_bytes(const om_fuse_view *views, int32_t n_views, int32_t stride);

/* densify._fuse_loop (densify.py:688-762): sequential fusion of the views order[0..n_order) (host
 * array).  nbrs (n_views, max_nb) host int32 neighbour lists, -1 for none (list a view only if it has
 * a depth map, as `j in have` does); rel_tol (n_views, max_nb) host float64.  Appends the fused
 * points in the reference's order: pts (cap, 3) float64, cols (cap, 3) uint8 (NULL to skip),
 * nv (cap,) uint8 = min(views, 255), owner (cap,) int32 = the reference view.  total (device,
 * 2 int64): total[0] = the number of points when the stream reaches the end of the call (only the
 * first cap are written, so total[0] > cap means the output was too small), total[1] = the grid
 * pixels evaluated (finite and not used).  Updates the views' used maps.
 * Two launches per view on `stream` (three above 2048 x 256 grid pixels); the host arrays are read
 * before the call returns. */
int om_fuse(int32_t n_order, const int32_t *order, int32_t n_views, const om_fuse_view *views, const int32_t *nbrs,
            int32_t max_nb, const double *rel_tol, double px_tol, int32_t min_views, int32_t stride, double *pts,
            uint8_t *cols, uint8_t *nv, int32_t *owner, int64_t cap, int64_t *total, void *workspace,
            size_t workspace_bytes, void *stream);

/* _mvs.sgm (_mvs.pyx:98-250): semi-global matching of B cost volumes cost (B, D, H, W) float32 ->
 * agg (B, D, H, W) float32, the sum over the 8 path directions in the reference's order with its
 * float32 operations (bit-identical).  D <= 512 (OM_EUNSUP above).  workspace: om_sgm_workspace_bytes
 * (9 volumes: the transposed cost and one path volume per direction). */
size_t om_sgm_workspace_bytes(int64_t B, int32_t D, int32_t H, int32_t W);
int om_sgm(const float *cost, int64_t B, int32_t D, int32_t H, int32_t W, float P1, float P2, float *agg,
           void *workspace, size_t workspace_bytes, void *stream);

/* _fast.depth_view (_fast.pyx:577-616), densify's matching image (densify.py:528-535): rgb (h, w, 3)
 * uint8 device -> grey (h, w) float32 = ((r + g + b) / 3 - 128) / 64 of rgb x gain + bias (NaN where
 * mask (h, w) uint8 is 0) and rgb8 (h, w, 3) uint8 = clip(v + 0.5, 0, 255).  gain, bias: host arrays
 * of 3 floats or NULL; mask: device or NULL.  Bit-identical. */
int om_depth_view(const uint8_t *rgb, int32_t h, int32_t w, const float *gain, const float *bias, const uint8_t *mask,
                  float *grey, uint8_t *rgb8, void *stream);

/* The depth-map file set of densify's passes (densify.py:590-592, :613-614, :639-641).
 * om_f32_to_f16: score.astype(np.float16) bit for bit (round to nearest even, overflow to inf,
 * subnormals, NaN keeps its sign and top payload bits); om_f16_to_f32: .astype(np.float32), exact;
 * om_filter_depth: depth[(f32(score16) < f32(min_ncc)) | low] = NaN in place (low may be NULL).
 * All arrays device, n elements. */
int om_f32_to_f16(const float *src, int64_t n, uint16_t *dst, void *stream);
int om_f16_to_f32(const uint16_t *src, int64_t n, float *dst, void *stream);
int om_filter_depth(float *depth, const uint16_t *score16, const uint8_t *low, int64_t n, double min_ncc,
                    void *stream);

/* densify._init_normals (densify.py:481-496): normals (3, H, W) of a depth map, float32 */
int om_init_normals(const float *depth, const float *ray, int32_t H, int32_t W, float *normal, void *stream);

/* height-sweep backend operations (orthomosaic/backend.py:67-99) */
/* _mvs.sample_view (_mvs.pyx:13-54) over `planes` height maps Z (planes, H, W): img (h, w, nch)
 * float32; out (planes, H, W, nch) must be zeroed by the caller (untouched where invalid). */
int om_sample_view(const float *img, int32_t h, int32_t w, int32_t nch, const double *R, const double *C, double f,
                   double k1, double k2, double k3, double cx, double cy, double X0, double Y0, double gsd,
                   const float *Z, int64_t planes, int32_t H, int32_t W, float *out, uint8_t *valid, void *stream);
/* _fast.box_sat_mean (_fast.pyx:897-943) of B float32 planes (B, H, W) */
size_t om_box_sat_workspace_bytes(int64_t B, int32_t H, int32_t W, int32_t r);
int om_box_sat_mean(const float *a, int64_t B, int32_t H, int32_t W, int32_t r, float *out, void *workspace,
                    size_t workspace_bytes, void *stream);
/* CPUBackend.topk_mean: mean of the k largest of st (n, M) along axis 0 (k <= 16) */
int om_topk_mean(const float *st, int32_t n, int64_t M, int32_t k, float *out, void *stream);
/* CPUBackend.argmin0 / argmax0: first index (int64) and value along axis 0 of a (n, M) */
int om_argminmax0(const float *a, int32_t n, int64_t M, int32_t want_max, int64_t *idx, float *val, void *stream);

/* Diagnostic, no counterpart in the reference: the two arithmetic shortcuts of the FP64 kernels,
 * exposed so that tests can pin them against IEEE arithmetic.  Quotients sharing one reciprocal
 * of the divisor when the operands allow, plain division otherwise (meant for c > 1e-6): qa = a / c
 * and qb = b / c as the projection computes its pair, qs = (a + b) / c as the middle one of the
 * sweep NCC's three (b, a + b, a over c); fr = a - floor(a) as the bilinear tap's v_fract_f64;
 * sq = sqrt(c) as the sweep NCC takes it (its Newton sequence without operand scaling for
 * 1e-6 < c < inf, sqrt otherwise).  Device arrays of n doubles. */
int om_fp64_steps(const double *a, const double *b, const double *c, int64_t n, double *qa, double *qb, double *qs,
                  double *fr, double *sq, void *stream);

#ifdef __cplusplus
}
#endif
#endif
