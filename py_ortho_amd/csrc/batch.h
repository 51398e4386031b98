// Internal state of an om_batch: many views' depth maps planned once and launched together.
#pragma once
#include <vector>

#include "om_common.h"

struct om_batch {
    std::vector<void *> allocs;  // device buffers owned by the batch

    // plane sweep over all views: one persistent launch + one post-processing launch
    int sw_n = 0, sw_k = 0, sw_th = 16, sw_grid = 0, sw_total = 0;
    size_t sw_lds = 0;
    int64_t sw_px = 0;
    bool sw_merge = false, sw_final = false;
    void *sw_views = nullptr;     // SweepK[sw_n]
    int *sw_item0 = nullptr;      // first work item of each view
    int64_t *sw_px0 = nullptr;    // first pixel of each view (sw_n + 1)
    int *sw_counter = nullptr;    // work queue head

    // PatchMatch over all views: 4 launches per colour pass
    int pm_n = 0, pm_sm = 0, pm_iters = 0, pm_blk = 0, pm_rows = 0;
    bool pm_normals = false, pm_prepare = false;
    int64_t pm_px = 0;
    void *pm_views = nullptr;     // PMK[pm_n]
    int *pm_blk0 = nullptr, *pm_row0 = nullptr;
    int64_t *pm_px0 = nullptr;
    unsigned long long *pm_changed = nullptr;  // (pm_n,)

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void *dev_alloc(size_t bytes)
    {
        void *p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
            return nullptr;
        allocs.push_back(p);
        return p;
    }
    void release()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (void *p : allocs)
            (void)hipFree(p);
        allocs.clear();
    }
};

namespace om {
int batch_sweep_set(om_batch *b, const om_sweep_desc *views, int nviews, float *const *depth_out, int chunks,
                    int tile_h, int grid_blocks);
int batch_sweep_enqueue(om_batch *b, hipStream_t st);
int batch_pm_enqueue(om_batch *b, hipStream_t st);
}  // namespace om
