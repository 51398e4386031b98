// C ABI of om_batch: plan many views once, launch them together, capture the launch in a HIP graph.
#include <new>

#include "batch.h"

namespace om {
int batch_pm_set(om_batch *b, const om_pm_view *views, int nviews, int r, int step, int top_k, double geo_w,
                 int iters, int init_normals);
}

using namespace om;

extern "C" om_batch *om_batch_create(void) { return new (std::nothrow) om_batch(); }

extern "C" int om_batch_destroy(om_batch *b)
{
    if (!b)
        return OM_OK;
    b->release();
    delete b;
    return OM_OK;
}

static void drop_graph(om_batch *b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

extern "C" int om_batch_set_sweep(om_batch *b, const om_sweep_desc *views, int32_t nviews, float *const *depth_out,
                                  int32_t chunks, int32_t tile_h, int32_t grid_blocks)
{
    if (!b)
        return fail(OM_EINVAL, "null batch");
    if (b->sw_n)
        return fail(OM_EINVAL, "the batch's sweep is already set");
    drop_graph(b);
    return batch_sweep_set(b, views, nviews, depth_out, chunks, tile_h, grid_blocks);
}

extern "C" int om_batch_set_pm(om_batch *b, const om_pm_view *views, int32_t nviews, int32_t r, int32_t step,
                               int32_t top_k, double geo_w, int32_t iters, int32_t init_normals)
{
    if (!b)
        return fail(OM_EINVAL, "null batch");
    if (b->pm_n)
        return fail(OM_EINVAL, "the batch's PatchMatch is already set");
    drop_graph(b);
    return batch_pm_set(b, views, nviews, r, step, top_k, geo_w, iters, init_normals);
}

extern "C" int om_batch_run(om_batch *b, void *stream)
{
    if (!b)
        return fail(OM_EINVAL, "null batch");
    hipStream_t st = as_stream(stream);
    int rc = batch_sweep_enqueue(b, st);
    if (rc)
        return rc;
    return batch_pm_enqueue(b, st);
}

extern "C" int om_batch_capture(om_batch *b, void *stream)
{
    if (!b)
        return fail(OM_EINVAL, "null batch");
    if (!stream)
        return fail(OM_EINVAL, "capture needs a non-default stream");
    drop_graph(b);
    hipStream_t st = as_stream(stream);
    int rc = hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc)
        return rc;
    const int run_rc = om_batch_run(b, stream);
    hipGraph_t g = nullptr;
    rc = hip_check(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
    if (run_rc) {
        if (g)
            (void)hipGraphDestroy(g);
        return run_rc;
    }
    if (rc)
        return rc;
    b->graph = g;
    return hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

extern "C" int om_batch_replay(om_batch *b, void *stream)
{
    if (!b || !b->exec)
        return fail(OM_EINVAL, "batch has no captured graph");
    return hip_check(hipGraphLaunch(b->exec, as_stream(stream)), "hipGraphLaunch");
}

extern "C" int om_batch_pm_changed(om_batch *b, int64_t *changed)
{
    if (!b || !changed)
        return fail(OM_EINVAL, "null argument");
    if (b->pm_n == 0)
        return OM_OK;
    std::vector<unsigned long long> h(b->pm_n);
    int rc = hip_check(hipMemcpy(h.data(), b->pm_changed, sizeof(unsigned long long) * b->pm_n, hipMemcpyDeviceToHost),
                       "copy changed");
    if (rc)
        return rc;
    for (int v = 0; v < b->pm_n; ++v)
        changed[v] = (int64_t)h[v];
    return OM_OK;
}
