// Internal state of an om_batch: many views' depth maps planned once and launched together.
#pragma once
#include <vector>

#include "om_common.h"

// one persistent sweep launch: the views of one tile height and plane kind (constant or per pixel)
constexpr int kSweepGroups = 6;  // 3 tile heights x 2 plane kinds
struct SweepGroup {
    int th = 16, first = 0, nv = 0, total = 0, grid = 0;
    bool cd = false;
    size_t lds = 0;
};

// the PatchMatch views of one source bucket: table entries [first, first + nv)
struct PMGroup {
    int sm = 16, first = 0, nv = 0, blk = 0, blk0_off = 0;
    int blk_list = 0, list0_off = 0, blk_eval = 0;  // bounded random refinement: list and evaluation grids
};

struct om_batch {
    std::vector<void *> allocs;  // device buffers owned by the batch

    // plane sweep over all views: one persistent launch per tile height + one post-processing launch
    int sw_n = 0, sw_k = 0, sw_r = 0, sw_ngroups = 0;
    SweepGroup sw_groups[kSweepGroups];
    int64_t sw_px = 0;
    bool sw_merge = false, sw_final = false;
    void *sw_views = nullptr;     // SweepK[sw_n]
    int *sw_vidx = nullptr;       // views ordered by group
    int *sw_item0 = nullptr;      // first work item of each view within its group (order of sw_vidx)
    int64_t *sw_px0 = nullptr;    // first pixel of each view (sw_n + 1)
    int *sw_counter = nullptr;    // work queue heads, one per group (kSweepGroups)

    // PatchMatch over all views: per colour pass, propagation and random refinement once per
    // source bucket, row ranking and stream scan once for all views
    int pm_n = 0, pm_iters = 0, pm_rows = 0, pm_ngroups = 0;
    PMGroup pm_groups[4];
    bool pm_normals = false, pm_prepare = false;
    bool pm_memo = false;         // memoised propagation (every view small enough for its item list)
    bool pm_bounded = false;      // random refinement with the geometric bound
    int64_t pm_px = 0;
    void *pm_views = nullptr;     // PMK[pm_n], ordered by source bucket
    int *pm_blk0 = nullptr;       // per group: first block of each of its views (nv + 1 each)
    int *pm_list0 = nullptr;      // the same for the bounded refinement's list kernel
    int *pm_row0 = nullptr;       // first row of each view (pm_n + 1)
    int64_t *pm_px0 = nullptr;    // first pixel of each view (pm_n + 1)
    unsigned long long *pm_changed = nullptr;  // (pm_n,) in the caller's view order
    std::vector<unsigned long long *> pm_nevals;  // memo path: each view's evaluation count (device)
    std::vector<char> pm_host;    // host copy of the PMK table (memo path: launched view by view)
    // memo path: the views' propagation chains alternate between side streams, forked from and
    // joined back to the launch stream around every colour pass
    int pm_nside = 0;
    hipStream_t pm_side[4] = {};
    hipEvent_t pm_fork = nullptr, pm_join[4] = {};

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
        for (int k = 0; k < 4; ++k) {
            if (pm_side[k])
                (void)hipStreamDestroy(pm_side[k]);
            if (pm_join[k])
                (void)hipEventDestroy(pm_join[k]);
            pm_side[k] = nullptr;
            pm_join[k] = nullptr;
        }
        if (pm_fork)
            (void)hipEventDestroy(pm_fork);
        pm_fork = nullptr;
        pm_nside = 0;
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
