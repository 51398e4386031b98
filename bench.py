#!/usr/bin/env python
"""Throughput of the per-view depth-map path on one MI355X.

Workload (default, the headline configuration): the batched block -- 48 reference views of the
seeded synthetic terrain from an 8 x 6 nadir grid at 90 m (70 % forward / 60 % side overlap),
sizes drawn from 800x600, 1200x900, 1600x1200 and 2400x1800 (f = 0.9375 W), 4-12 nearest source
views each; per view 192 fronto-parallel planes (11x11 NCC window, top-2) + the parabola, then 6
PatchMatch iterations (9x9 window every 2 px) on every pixel with the neighbours' pass-1 depth maps
as the geometric term.  One step = one replay of the HIP graph holding the whole block (one
persistent sweep launch, the parabola, the PatchMatch passes of all views), inputs resident in HBM.

Metric: Mpixel-hypotheses/s = sum over views of H * W * (planes + 11 * iterations) / step time (one
sweep pixel-hypothesis = one plane at one pixel over all sources; 11 PatchMatch cost evaluations
per pixel and iteration: current + 8 propagated + 2 random).  Also printed: p50/p99 step latency,
the dominant kernel's roofline (HIP events around its launches), the plain single-view GPU forms
in the same run (fused per-view calls and the per-plane launch loop), and the oracle (C
restatement of the reference) timed on this host's cores on centre crops of the same workload.

Other workloads: c1 (640x480 sweep), c2 (2000x1500 PatchMatch), c4 (4000x3000 FP16/BF16), c5
(served requests), fuse, sgm, prep.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload c3|c1|c2|c4|c5|fuse|sgm|prep] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import queue
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpixel-hypotheses/s on 1 GPU; p50/p99 depth-map latency at fixed accuracy"
FP64_PEAK = 78.6e12       # MI355X published FP64 (vector and matrix) TFLOP/s
HBM_PEAK = 8.0e12
# bytes between L2 and memory of one full-block sweep_batch_kernel launch of the default workload
# (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, calibrated against a 256 MiB torch multiply;
# profiles/block_c3.md): taps fetched again through the last-level cache and registers spilled at
# four workgroups per CU
C3_SWEEP_TRAFFIC = 5.069e11
# share of the SIMD cycles in which that launch issues a VALU instruction (SQ_INSTS_VALU x 4 cycles over
# 1024 SIMDs x 2.4 GHz x the launch; profiles/block_c3.md): the kernel is issue-bound, and its FLOP
# fraction is low because a division or square root counts as one FLOP but takes ~11-18 instructions
C3_SWEEP_VALU_ISSUE = 0.85


def progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def pm_flops_per_eval(S: int, nsamp: int) -> int:
    """Algorithmic FLOPs of one PatchMatch cost evaluation (FMA = 2): per sample per source
    homography apply 15 + bilinear 11 + 3 accumulators 5, a shared ref term 3 per sample, and per
    source plane->homography 60 + NCC 15 + geometric term 19."""
    return nsamp * (31 * S + 3) + 94 * S + 5


def sweep_flops_per_hyp(S: int) -> int:
    """Per pixel-hypothesis: 55 per source (homography 15, bilinear 11, J^2 / J*I 2, separable box
    sums 16, NCC 11) + 6 (hypothesis 2, top-k mean 3, winner 1)."""
    return 55 * S + 6


def _cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if v else None


def profiled_traffic(name: str):
    """HBM bytes per launch of the dominant kernel from the committed counter summary (two
    rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, calibrated against a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    import re
    m = re.search(r"traffic_bytes_per_launch = (\d+)", open(path).read())
    return int(m.group(1)) if m else None


def prof_read(lib, name):
    tot, n = ctypes.c_double(0), ctypes.c_int64(0)
    rc = lib.om_profile_read(name.encode(), ctypes.byref(tot), ctypes.byref(n))
    if rc:
        raise RuntimeError(f"om_profile_read rc={rc}")
    return tot.value, n.value


# ------------------------------------------------------------------------------- C2 PatchMatch
def _scene_c2(torch, cache):
    """The C2 scene, rendered on the device or loaded from `cache` (written on first use), so that
    profiled runs launch only the depth-map kernels."""
    from py_ortho_amd import scene
    if cache and os.path.exists(cache):
        z = np.load(cache, allow_pickle=False)
        sc = scene.config_c2_cameras()
        views = [sc.ref] + sc.srcs
        for k, v in enumerate(views):
            v.img = torch.from_numpy(z[f"img{k}"]).cuda()
            v.depth = torch.from_numpy(z[f"dep{k}"]).cuda()
        return sc
    sc = scene.config_c2(device="cuda")
    if cache:
        views = [sc.ref] + sc.srcs
        np.savez(cache, **{f"img{k}": v.img.cpu().numpy() for k, v in enumerate(views)},
                 **{f"dep{k}": v.depth.cpu().numpy() for k, v in enumerate(views)})
    return sc


def setup_c2(torch, cache=None):
    from py_ortho_amd import scene
    sc = _scene_c2(torch, cache)
    H, W = sc.ref.img.shape
    ray = scene.rays_world(sc.ref.cam, "cuda")
    depth0, normal0 = scene.pm_init(3, H, W, ray, 80.0, 105.0)
    return dict(scene=sc, H=H, W=W, ref=sc.ref.img.contiguous(), ray=ray,
                srcs=[v.img.contiguous() for v in sc.srcs], deps=[v.depth.contiguous() for v in sc.srcs],
                par=sc.par(), Cr=sc.ref.cam.C.copy(), depth0=depth0.cuda(), normal0=normal0.cuda(),
                active=torch.ones((H, W), dtype=torch.uint8, device="cuda"),
                r=5, step=2, top_k=2, geo_w=0.5, geo_tol=0.01, iters=8, dmin=80.0, dmax=105.0,
                seed=2654435761)


def run_c2(args, torch):
    from py_ortho_amd import _lib, dense
    lib = _lib.lib()
    c = setup_c2(torch, args.scene_cache)
    H, W, S = c["H"], c["W"], len(c["srcs"])
    depth = torch.empty_like(c["depth0"])
    normal = torch.empty_like(c["normal0"])
    score = torch.empty((H, W), dtype=torch.float32, device="cuda")

    stats = {}

    def step():
        depth.copy_(c["depth0"])
        normal.copy_(c["normal0"])
        score.fill_(-1.0)
        return dense.pm_refine(c["ref"], c["ray"], c["Cr"], c["srcs"], c["deps"], c["par"], depth, normal, score,
                               c["active"], c["r"], c["step"], c["top_k"], c["geo_w"], c["geo_tol"], c["iters"],
                               c["dmin"], c["dmax"], c["seed"], stats=stats)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    for name in ("pm_propagate", "pm_eval", "pm_random"):
        prof_read(lib, name)  # drop warm-up records
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        changed = step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    prop_ms, prop_n = prof_read(lib, "pm_propagate")
    eval_ms, eval_n = prof_read(lib, "pm_eval")
    rand_ms, rand_n = prof_read(lib, "pm_random")
    lib.om_profile_read(None, None, None)

    evals = H * W * 11 * c["iters"]
    ms = wall / args.steps * 1e3
    nsamp = ((2 * c["r"]) // c["step"] + 1) ** 2
    fpe = pm_flops_per_eval(S, nsamp)
    # dominant kernel: propagation, 9 nominal evaluations per pixel of one colour
    # The propagation pass is memoised: pm_need lists the candidate costs whose plane changed since
    # the pixel's last visit, pm_eval_kernel computes them (the dominant kernel), pm_decide replays
    # the accept sequence.  Algorithmic FLOPs per pm_eval_kernel launch = the evaluations it performed
    # (counted on the device) x FLOPs per evaluation; the reference's nominal work per pass is
    # (H W / 2) x 9 evaluations.
    passes = 2 * c["iters"]
    evaluated = stats.get("evaluated", -1)
    nominal_evals = (H * W / 2) * 9 * passes
    eval_avg = eval_ms / max(eval_n, 1) / 1e3
    prop_avg = prop_ms / max(prop_n, 1) / 1e3
    prop_flops = (evaluated / passes if evaluated >= 0 else (H * W / 2) * 9) * fpe
    achieved = prop_flops / eval_avg if eval_avg > 0 else 0.0
    # accuracy of the refined map against the rendered truth
    truth = c["scene"].ref.depth
    rel = ((depth - truth).abs() / truth).float()
    fin = torch.isfinite(rel)
    acc = {"frac_within_1pct": float((rel[fin] < 0.01).float().mean()),
           "median_rel_err": float(rel[fin].median()), "changed_last_step": int(changed)}
    res = {
        "metric": METRIC, "value": evals / (ms / 1e3) / 1e6, "unit": "Mpixel-hyp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C2 single-view PatchMatch, 2000x1500 ref, 8 sources (2x4 grid, 20 m, 100 m alt, "
                               "f=1800), 11x11 window step 2 (36 samples), 8 red-black iterations, geo term",
                   "H": H, "W": W, "sources": S, "iterations": c["iters"], "pixel_hypotheses_per_step": evals},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "roofline": {"bound": "valu", "engine": "VALU FP64 (no MFMA: gather-and-reduce)", "achieved": achieved / 1e12,
                     "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK,
                     "traffic": profiled_traffic("pm_eval_c2.md"), "kernel": "pm_eval_kernel",
                     "kernel_avg_ms": eval_avg * 1e3, "launches": eval_n,
                     "flops_per_launch": prop_flops, "flops_per_eval": fpe,
                     "evaluated_per_step": evaluated, "nominal_evaluations_per_step": nominal_evals,
                     "propagation_pass_avg_ms": prop_avg * 1e3,
                     "nominal_tflops": (H * W / 2) * 9 * fpe / prop_avg / 1e12 if prop_avg else 0.0,
                     "pm_random_avg_ms": rand_ms / max(rand_n, 1)},
        "accuracy": acc,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_pm(c, torch, args)
    return res


def cpu_baseline_pm(c, torch, args):
    """The oracle's pm_refine (C restatement of _dense.pm_refine, one view per thread like
    densify.py:596) on a centre crop of the reference with the full sources, `cores` concurrent
    copies; Mpixel-hyp/s = copies * crop pixels * 11 * iterations / wall."""
    import oracle
    H, W = c["H"], c["W"]
    ref = c["ref"].cpu().numpy()
    ray = c["ray"].cpu().numpy()
    d0 = c["depth0"].cpu().numpy()
    n0 = c["normal0"].cpu().numpy()
    srcs = [s.cpu().numpy() for s in c["srcs"]]
    deps = [d.cpu().numpy() for d in c["deps"]]
    cores = _cores() if args.cpu_threads <= 0 else args.cpu_threads

    def crop(side, iters, threads):
        y0, x0 = (H - side) // 2, (W - side) // 2
        sl = np.s_[y0:y0 + side, x0:x0 + side]
        a = np.ones((side, side), np.uint8)
        t = time.perf_counter()
        oracle.bench_pm(threads, ref[sl], np.ascontiguousarray(ray[:, y0:y0 + side, x0:x0 + side]), c["Cr"], srcs,
                        deps, c["par"], np.ascontiguousarray(d0[sl]), np.ascontiguousarray(n0[:, y0:y0 + side,
                                                                                              x0:x0 + side]),
                        a, c["r"], c["step"], c["top_k"], c["geo_w"], c["geo_tol"], iters, c["dmin"], c["dmax"],
                        c["seed"])
        return time.perf_counter() - t

    # marginal seconds per pixel-iteration on one core (the difference of two probes cancels the
    # call's fixed cost, which made a single small probe pick too small a crop)
    t1, t2 = crop(24, 1, 1), crop(48, 1, 1)
    per_px_it = (t2 - t1) / (48 * 48 - 24 * 24) if t2 > t1 else t2 / (48 * 48)
    target = 20.0                                  # seconds of work per core (bounded sample)
    side = int(max(128, min(256, math.sqrt(target / (per_px_it * c["iters"])))))
    wall = crop(side, c["iters"], cores)
    evals = cores * side * side * 11 * c["iters"]
    return {"value": evals / wall / 1e6, "unit": "Mpixel-hyp/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent copies of the oracle pm_refine on the {side}x{side} centre crop, "
                      f"S={len(srcs)}, {c['iters']} iterations ({wall:.1f} s wall)",
            "single_core_latency_s_per_mpix": wall / (side * side) * 1e6}


# ------------------------------------------------------------------------------- C1 plane sweep
def run_c1(args, torch):
    from py_ortho_amd import _lib, dense, scene
    lib = _lib.lib()
    sc = scene.config_c1(device="cuda")
    H, W = sc.ref.img.shape
    r, k, D = 2, 2, 128
    ref = sc.ref.img.contiguous()
    ray = scene.rays_world(sc.ref.cam, "cuda")
    mu, sd = dense.ref_stats(ref, r)
    planes = torch.linspace(60.0, 85.0, D, dtype=torch.float64)
    inv = (1.0 / planes).to(torch.float32).cuda()
    srcs = [v.img.contiguous() for v in sc.srcs]
    geo = sc.geo()
    hyp = {"mode": "const", "inv": inv}

    def step():
        st = dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, r, k)
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    prof_read(lib, "sweep")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st = step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    sw_ms, sw_n = prof_read(lib, "sweep")
    lib.om_profile_read(None, None, None)
    ms = wall / args.steps * 1e3
    S = len(srcs)
    avg = sw_ms / max(sw_n, 1) / 1e3
    flops = H * W * D * sweep_flops_per_hyp(S)
    truth = sc.ref.depth
    dw = planes.cuda()[st.idx.long()].float()
    rel = (dw - truth).abs() / truth
    valid = (st.best >= 0.3) & (255.0 * sd >= 0.5)
    res = {
        "metric": METRIC, "value": H * W * D / (ms / 1e3) / 1e6, "unit": "Mpixel-hyp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C1 plane sweep 640x480, 4 sources, 128 fronto-parallel planes 60-85 m, 5x5 NCC, "
                               "top-2, WTA", "H": H, "W": W, "sources": S, "planes": D},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "roofline": {"bound": "valu", "engine": "VALU FP64", "achieved": flops / avg / 1e12 if avg else 0,
                     "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": (flops / avg / FP64_PEAK) if avg else 0,
                     "traffic": profiled_traffic("sweep_c1.md"), "kernel": "sweep_kernel", "kernel_avg_ms": avg * 1e3},
        "accuracy": {"valid_frac": float(valid.float().mean()),
                     "valid_within_1pct": float((rel[valid] < 0.01).float().mean())},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_sweep_rate(ref.cpu().numpy(), mu.cpu().numpy(), sd.cpu().numpy(),
                                             ray.cpu().numpy(), [x.cpu().numpy() for x in srcs], geo,
                                             inv.cpu().numpy(), r, k, args, f"the {W}x{H} view, S={S}")
    return res


# ------------------------------------------------------------------------------- C3 batched block
def _survey_c3(torch, cache, scale):
    """The C3 block, rendered on the device or loaded from `cache` (written on first use)."""
    from py_ortho_amd import scene
    sizes = [(int(round(w * scale)), int(round(h * scale))) for w, h in scene.C3_SIZES]
    cams, counts, ext = scene.survey_cameras(4, 8, 6, 90.0, sizes, 0.9375, 0.7, 0.6, (4, 12))
    if cache and os.path.exists(cache):
        z = np.load(cache, allow_pickle=False)
        views = [scene.View(c, torch.from_numpy(z[f"img{k}"]).cuda(), torch.from_numpy(z[f"dep{k}"]).cuda())
                 for k, c in enumerate(cams)]
        C = np.stack([c.C for c in cams])
        return scene.Survey(scene.make_terrain(4, ext, n_bumps=40), views,
                            scene.neighbours_by_distance(C, 90.0, counts), 90.0 - 17.0, 92.0)
    sv = scene.config_c3(device="cuda", scale=scale)
    if cache:
        np.savez(cache, **{f"img{k}": v.img.cpu().numpy() for k, v in enumerate(sv.views)},
                 **{f"dep{k}": v.depth.cpu().numpy() for k, v in enumerate(sv.views)})
    return sv


def _depth_accuracy(maps, truths):
    """Over all pixels of all views: fraction within 1 % of the rendered depth, median relative error
    of the finite ones, fraction finite."""
    import torch
    within, errs = 0.0, []
    for d, t in zip(maps, truths):
        rel = ((d - t).abs() / t).float()
        rel = rel[torch.isfinite(rel)]
        within += float((rel < 0.01).float().sum())
        errs.append(rel)
    allr = torch.cat(errs)
    n = sum(int(t.numel()) for t in truths)
    return {"frac_within_1pct": within / n, "median_rel_err": float(allr.median()),
            "frac_finite": float(allr.numel()) / n}


def run_c3(args, torch):
    from py_ortho_amd import _lib, autotune, block
    lib = _lib.lib()
    t_scene = time.perf_counter()
    progress("c3: rendering the 48-view block")
    sv = _survey_c3(torch, args.scene_cache, args.scale)
    torch.cuda.synchronize()
    t_scene = time.perf_counter() - t_scene
    n = len(sv.views)
    blk = block.BlockDepth([v.img.contiguous() for v in sv.views], [v.cam for v in sv.views], sv.nbrs,
                           [(sv.dmin, sv.dmax)] * n)
    o = blk.opt
    sw_ph, pm_ph = blk.pixel_hypotheses()
    t_tune = time.perf_counter()
    progress("c3: autotuning tile shapes per size class")
    tune, table = autotune.tune_block(blk)
    t_tune = time.perf_counter() - t_tune
    blk.plan(tune)
    blk.run()  # warm-up, eager
    torch.cuda.synchronize()
    # one eager run with HIP events around the dominant kernels
    lib.om_profile_enable(1)
    prof_read(lib, "sweep_batch_kernel")
    blk.run()
    torch.cuda.synchronize()
    lib.om_profile_enable(0)
    sw_ms, sw_n = prof_read(lib, "sweep_batch_kernel")
    pr_ms, pr_n = prof_read(lib, "pm_propagate_batch")
    rn_ms, rn_n = prof_read(lib, "pm_random_batch")
    lib.om_profile_read(None, None, None)
    # the timed steps: graph replays of the whole block (sweep, parabola, PatchMatch of all views)
    blk.capture()
    progress(f"c3: {args.warmup} warm-up + {args.steps} timed graph replays")
    for _ in range(args.warmup):
        blk.replay()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        blk.replay()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    ms = wall / args.steps * 1e3
    raw_b = [d.clone() for d in blk.raw]
    dep_b = [d.clone() for d in blk.depth]
    truths = [v.depth for v in sv.views]
    acc = {"sweep": _depth_accuracy(raw_b, truths), "patchmatch": _depth_accuracy(dep_b, truths)}
    # the plain per-view form: fused single-view sweep + parabola + single-view PatchMatch per view
    progress("c3: plain per-view calls")
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    ts = time.perf_counter()
    blk.run_single()
    torch.cuda.synchronize()
    single_ms = (time.perf_counter() - ts) * 1e3
    lib.om_profile_enable(0)
    single_parts = {name: prof_read(lib, name)[0] for name in ("sweep", "pm_propagate", "pm_random")}
    lib.om_profile_read(None, None, None)
    same = all(bool(torch.equal(a.view(torch.int32), b.view(torch.int32))) for a, b in zip(raw_b, blk.raw)) and \
        all(bool(torch.equal(a.view(torch.int32), b.view(torch.int32))) for a, b in zip(dep_b, blk.depth))
    plain = {"fused_per_view_ms": single_ms, "speedup_vs_fused_per_view": single_ms / ms,
             "fused_per_view_kernel_ms": single_parts}
    if not args.no_plain_per_hyp:
        progress("c3: per-plane launch loop (plain single-view kernel shape)")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        blk.run_single(per_hypothesis=True)
        torch.cuda.synchronize()
        ph_ms = (time.perf_counter() - ts) * 1e3
        plain.update({"per_hypothesis_ms": ph_ms, "speedup_vs_per_hypothesis": ph_ms / ms})
    flops = sum(px * o.num_depths * sweep_flops_per_hyp(len(nb)) for px, nb in zip(blk.pixels(), blk.nbrs))
    sw_avg = sw_ms / max(sw_n, 1) / 1e3
    sw_tot = sw_ms / 1e3
    achieved = flops / sw_tot if sw_tot > 0 else 0.0
    S_list = [len(nb) for nb in blk.nbrs]
    res = {
        "metric": METRIC, "value": (sw_ph + pm_ph) / (ms / 1e3) / 1e6, "unit": "Mpixel-hyp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"C3 batched block: {n} views (8x6 grid, 90 m, 70/60 pct overlap) of sizes "
                               f"800x600..2400x1800 x {args.scale:g}, S=U{{4..12}}, {o.num_depths} planes 11x11 NCC "
                               f"top-2 + parabola, then {o.pm_iters} PatchMatch iterations (9x9 window step 2) on "
                               "every pixel, one HIP graph",
                   "views": n, "megapixels": sum(blk.pixels()) / 1e6, "sources_mean": float(np.mean(S_list)),
                   "sweep_pixel_hypotheses": sw_ph, "pm_pixel_hypotheses": pm_ph,
                   "sweep_mpix_hyp_s": sw_ph / sw_tot / 1e6 if sw_tot else None,
                   "pm_mpix_hyp_s": pm_ph / ((pr_ms + rn_ms) / 1e3) / 1e6 if pr_ms + rn_ms > 0 else None,
                   "autotune": {f"{h}x{w}": list(v) for (h, w), v in tune.items()},
                   "autotune_ns_per_pixel_hyp": {f"{h}x{w}": {f"th{a}_c{b}": round(t, 4) for (a, b), t in d.items()}
                                                 for (h, w), d in table.items()},
                   "autotune_s": t_tune, "scene_s": t_scene},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "roofline": {"bound": "valu", "engine": "VALU FP64 (no MFMA: gather-and-reduce)",
                     "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": C3_SWEEP_TRAFFIC if args.scale == 1.0 else None,
                     "valu_issue_frac": C3_SWEEP_VALU_ISSUE if args.scale == 1.0 else None,
                     "kernel": "sweep_batch_kernel",
                     "kernel_avg_ms": sw_avg * 1e3, "launches": sw_n, "flops_per_step": flops,
                     "pm_propagate_ms": pr_ms, "pm_random_ms": rn_ms},
        "plain_single_view": plain, "batch_equals_single_view": same,
        "accuracy": acc,
    }
    if not args.no_cpu:
        progress("c3: CPU baseline (oracle on centre crops)")
        res["cpu_baseline"] = cpu_baseline_c3(blk, sv, torch, args)
    return res


def cpu_baseline_c3(blk, sv, torch, args):
    """The oracle's sweep (linear planes, same count) and pm_refine on a centre crop of the view with
    the median source count, `cores` concurrent copies each; Mpixel-hyp/s = cores * (D + 11 iters) /
    (sweep + PatchMatch wall time per crop pixel)."""
    import oracle
    o = blk.opt
    cores = _cores() if args.cpu_threads <= 0 else args.cpu_threads
    S_list = [len(nb) for nb in blk.nbrs]
    k = int(np.argsort(S_list)[len(S_list) // 2])
    H, W = blk.size_class(k)
    ref = blk.ref[k].cpu().numpy()
    mu, sd = (t.cpu().numpy() for t in blk.stats[k])
    ray = blk.rays[k].cpu().numpy()
    srcs = [blk.imgs[j].cpu().numpy() for j in blk.nbrs[k]]
    deps = [blk.raw[j].cpu().numpy() for j in blk.nbrs[k]]
    geo = np.stack([block_geo(blk, j, k) for j in blk.nbrs[k]])
    par = blk.pm_view(k).par
    planes = blk.planes_host[k]
    raw = blk.raw[k].cpu().numpy()
    nrm = blk.normal[k].cpu().numpy()

    def crop_sl(side):
        y0, x0 = (H - side) // 2, (W - side) // 2
        return y0, x0, np.s_[y0:y0 + side, x0:x0 + side]

    def sweep(side, nd, threads):
        y0, x0, sl = crop_sl(side)
        base = np.full((side, side), planes[0], np.float32)
        step = np.full((side, side), planes[1] - planes[0], np.float32)
        t = time.perf_counter()
        oracle.bench_sweep(threads, ref[sl], mu[sl], sd[sl], srcs, np.ascontiguousarray(ray[:, y0:y0 + side,
                                                                                         x0:x0 + side]),
                           geo, base, step, nd, o.window, o.top_k)
        return time.perf_counter() - t

    def pm(side, iters, threads):
        y0, x0, sl = crop_sl(side)
        a = np.ones((side, side), np.uint8)
        t = time.perf_counter()
        oracle.bench_pm(threads, ref[sl], np.ascontiguousarray(ray[:, y0:y0 + side, x0:x0 + side]),
                        np.asarray(blk.cams[k].C, np.float64), srcs, deps, par, np.ascontiguousarray(raw[sl]),
                        np.ascontiguousarray(nrm[:, y0:y0 + side, x0:x0 + side]), a, o.pm_window, o.pm_step, o.top_k,
                        o.pm_geo_weight, o.geo_tol, iters, sv.dmin, sv.dmax, 1)
        return time.perf_counter() - t

    # crops: 512x512 for the sweep and 256x256 for PatchMatch, each shrunk (not below 128) only
    # when the one-core estimate exceeds `budget` seconds
    budget = 15.0
    sw_px = sweep(16, 8, 1) / (256 * 8)              # seconds per pixel-plane, one core
    side_s = int(max(128, min(512, min(H, W), math.sqrt(budget / (sw_px * o.num_depths)))))
    pm_px = pm(16, 1, 1) / 256                       # seconds per pixel-iteration, one core
    side_p = int(max(128, min(256, min(H, W), math.sqrt(budget / (pm_px * o.pm_iters)))))
    ws = sweep(side_s, o.num_depths, cores)
    wp = pm(side_p, o.pm_iters, cores)
    per_px = ws / (side_s * side_s) + wp / (side_p * side_p)   # wall seconds per pixel for `cores` copies
    return {"value": cores * (o.num_depths + 11 * o.pm_iters) / per_px / 1e6, "unit": "Mpixel-hyp/s",
            "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent copies of the oracle sweep ({o.num_depths} planes, {side_s}x{side_s} "
                      f"centre crop) and pm_refine ({o.pm_iters} iterations, {side_p}x{side_p} crop) of view {k} "
                      f"({W}x{H}, S={S_list[k]}); {ws:.1f} s + {wp:.1f} s wall",
            "calibration": "the oracle's sweep restatement runs at 0.85x the reference's own compiled _sweep.h "
                           "(bit-identical, 640x480 view, 128 planes, 4 sources, one core of this build's "
                           "container)"}


def block_geo(blk, j, k):
    from py_ortho_amd.block import _geo18
    return _geo18(blk.cams[j], blk.cams[k])


# ------------------------------------------------------------------------------- C5 served depth maps
def _pools_c5(torch, cache, scale):
    """The C5 image pools, rendered on the device or loaded from `cache` (written on first use)."""
    from py_ortho_amd import scene
    if cache and os.path.exists(cache):
        z = np.load(cache, allow_pickle=False)
        pools = []
        for c, (w, h, _) in enumerate(scene.C5_CLASSES):
            size = (int(round(w * scale)), int(round(h * scale)))
            cams, counts, ext = scene.survey_cameras(6 + c, 4, 4, 90.0, [size], 0.9375, 0.7, 0.6, (12, 12))
            views = [scene.View(cm, torch.from_numpy(z[f"img{c}_{k}"]).cuda(), torch.from_numpy(z[f"dep{c}_{k}"]).cuda())
                     for k, cm in enumerate(cams)]
            C = np.stack([cm.C for cm in cams])
            pools.append(scene.Survey(None, views, scene.neighbours_by_distance(C, 90.0, counts), 73.0, 92.0))
        return pools
    pools = scene.config_c5_pools(device="cuda", scale=scale)
    if cache:
        np.savez(cache, **{f"img{c}_{k}": v.img.cpu().numpy() for c, sv in enumerate(pools) for k, v in enumerate(sv.views)},
                 **{f"dep{c}_{k}": v.depth.cpu().numpy() for c, sv in enumerate(pools) for k, v in enumerate(sv.views)})
    return pools


def c5_requests(pools, rate, duration, seed=6):
    """Poisson arrivals at `rate`/s for `duration` s: size class 60/30/10 %, S = U{4..12} nearest
    views, planes U{128..384} -> [(arrival offset s, DepthRequest)]."""
    from py_ortho_amd import scene
    from py_ortho_amd.serve import DepthRequest
    rng = np.random.default_rng(seed)
    probs = np.array([p for _, _, p in scene.C5_CLASSES])
    out, t, i = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        c = int(rng.choice(len(pools), p=probs / probs.sum()))
        sv = pools[c]
        k = int(rng.integers(len(sv.views)))
        S = int(rng.integers(4, 13))
        out.append((t, DepthRequest((c, k), [(c, j) for j in sv.nbrs[k][:S]], sv.dmin, sv.dmax,
                                    int(rng.integers(128, 385)), tag=i)))
        i += 1
    return out


def run_c5(args, torch):
    from py_ortho_amd import block, dense, serve
    t_scene = time.perf_counter()
    pools = _pools_c5(torch, args.scene_cache, args.scale)
    torch.cuda.synchronize()
    t_scene = time.perf_counter() - t_scene
    images = {(c, k): (v.img.cpu().numpy(), v.cam) for c, sv in enumerate(pools) for k, v in enumerate(sv.views)}
    truth = {(c, k): v.depth for c, sv in enumerate(pools) for k, v in enumerate(sv.views)}
    window, top_k = 5, 2
    ex = serve.GpuExecutor(window=window, top_k=top_k)
    store = serve.ImageStore(lambda key: images[key], args.budget_gb * 1e9, upload=ex.upload)
    srv = serve.DepthServer(store, ex, budget_bytes=args.budget_gb * 1e9, step_planes=args.step_planes,
                            max_active_pixels=args.active_mpix * 1e6).start()
    # warm-up: two small requests (code objects, allocator)
    warm = c5_requests(pools, 50.0, 0.05, seed=99)[:2]
    for _, r in warm:
        r.planes = 16
    for f in [srv.submit(r) for _, r in warm]:
        f.result()
    arrivals = c5_requests(pools, args.rate, args.duration)
    N = len(arrivals)
    n_plain = min(N, args.plain_requests)
    # results are summarised as they complete (a 120 s stream is thousands of depth maps): the latency
    # of every request, the accuracy of a sample of them, the full maps of the first n_plain only
    lat = [0.0] * N
    t_done = [0.0] * N
    kept = {}
    acc_every = max(1, N // 200)
    acc_q: "queue.Queue" = queue.Queue()
    rel_parts = []
    done = threading.Event()
    lock = threading.Lock()
    n_done = [0]

    def on_done(j, fut):
        r = fut.result()
        lat[j] = r.latency * 1e3
        t_done[j] = r.t_done
        if j < n_plain:
            kept[j] = r
        if j % acc_every == 0:
            acc_q.put((j, r.depth))
        with lock:
            n_done[0] += 1
            if n_done[0] == N:
                done.set()

    def accuracy_worker():
        while True:
            item = acc_q.get()
            if item is None:
                return
            j, dep = item
            t = truth[arrivals[j][1].ref].cpu().numpy()
            e = np.abs(dep - t) / t
            rel_parts.append(e[np.isfinite(e)])

    t0 = time.perf_counter()

    def feed():
        for j, (ta, req) in enumerate(arrivals):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            f = srv.submit(req, t_arrival=t0 + ta)
            f.add_done_callback(lambda fut, j=j: on_done(j, fut))

    from py_ortho_amd import _lib
    lib = _lib.lib()
    lib.om_profile_enable(1)
    prof_read(lib, "sweep_multi")
    aw = threading.Thread(target=accuracy_worker, daemon=True)
    aw.start()
    progress(f"c5: {N} requests arriving over {args.duration:g} s")
    th = threading.Thread(target=feed)
    th.start()
    while not done.wait(30.0):
        progress(f"c5: {n_done[0]}/{N} requests finished after {time.perf_counter() - t0:.0f} s")
    th.join()
    srv.stop()
    acc_q.put(None)
    aw.join()
    lib.om_profile_enable(0)
    step_ms, step_n = prof_read(lib, "sweep_multi")
    lib.om_profile_read(None, None, None)
    t_end = max(t_done)
    px_h = [images[q.ref][0].size * q.planes for _, q in arrivals]
    flops = sum(images[q.ref][0].size * q.planes * sweep_flops_per_hyp(len(q.srcs)) for _, q in arrivals)
    wall = t_end - t0
    served = sum(px_h) / wall / 1e6
    progress("c5: plain per-request comparison")
    # plain per-request runs (one request at a time: fused single-view sweep + parabola), same inputs
    torch.cuda.synchronize()
    tp = time.perf_counter()
    same = True
    for j in range(n_plain):
        q = arrivals[j][1]
        img, cam = images[q.ref]
        ref = torch.from_numpy(img).cuda()
        ref = torch.where(torch.isfinite(ref), ref, torch.zeros_like(ref)).contiguous()
        H, W = ref.shape
        mu, sd = dense.ref_stats(ref, window)
        ray = dense.rays_world(H, W, cam.R, cam.cx, cam.cy, cam.f, cam.k1, cam.k2, cam.k3)
        srcs = [store.pin(k) for k in q.srcs]
        geo = np.stack([block._geo18(images[k][1], cam) for k in q.srcs])
        hyp = {"mode": "const", "inv": torch.from_numpy(block.inverse_depth_planes(q.dmin, q.dmax, q.planes)).cuda()}
        st = dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, window, top_k)
        d = dense.sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp, window, top_k, st).cpu().numpy()
        for k in q.srcs:
            store.unpin(k)
        same = same and np.array_equal(d.view(np.uint32), kept[j].depth.view(np.uint32))
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - tp
    plain = sum(px_h[:n_plain]) / plain_s / 1e6
    rel = np.concatenate(rel_parts)
    classes = {}
    for j, (ta, q) in enumerate(arrivals):
        H, W = images[q.ref][0].shape
        classes.setdefault(f"{W}x{H}", []).append(lat[j])
    res_line = {
        "metric": METRIC, "value": served, "unit": "Mpixel-hyp/s", "n_gpus": 1, "steps": len(arrivals),
        "warmup": len(warm), "ms_per_step": wall * 1e3 / max(len(arrivals), 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"C5 served depth maps: Poisson arrivals {args.rate:g}/s for {args.duration:g} s, sizes "
                               "1200x900 / 2400x1800 / 4000x3000 (60/30/10 pct)"
                               f" x {args.scale:g}, S=U{{4..12}}, planes U{{128..384}}, 11x11 NCC top-2 + parabola, "
                               f"{args.budget_gb:g} GB admission budget, continuous batching ({args.step_planes} "
                               f"planes per step, <= {args.active_mpix:g} MP in flight)",
                   "requests": len(arrivals), "pixel_hypotheses": int(sum(px_h)), "wall_s": wall,
                   "offered_mpix_hyp_s": sum(px_h) / args.duration / 1e6, "steps": srv.steps,
                   "max_active": srv.stats["max_active"], "image_uploads": store.uploads, "scene_s": t_scene},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99),
                       "by_class_p50": {k: _percentile(v, 50) for k, v in classes.items()},
                       "by_class_p99": {k: _percentile(v, 99) for k, v in classes.items()}},
        "plain_per_request": {"requests": n_plain, "mpix_hyp_s": plain, "speedup": served / plain,
                              "served_equals_plain": bool(same)},
        "accuracy": {"frac_within_1pct": float((rel < 0.01).mean()), "median_rel_err": float(np.median(rel)),
                     "requests_sampled": len(rel_parts)},
        "roofline": {"bound": "valu", "engine": "VALU FP64 (no MFMA: gather-and-reduce)",
                     "achieved": flops / (step_ms / 1e3) / 1e12 if step_ms else 0.0, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": flops / (step_ms / 1e3) / FP64_PEAK if step_ms else 0.0,
                     "traffic": None, "kernel": "sweep_batch_kernel (om_sweep_multi steps)",
                     "kernel_avg_ms": step_ms / max(step_n, 1), "launches": step_n,
                     "device_busy_frac": step_ms / 1e3 / wall},
    }
    if not args.no_cpu:
        res_line["cpu_baseline"] = cpu_baseline_sweep(images, arrivals, window, top_k, args)
    return res_line


def cpu_baseline_sweep(images, arrivals, window, top_k, args):
    """The oracle's sweep on a centre crop of the first request with 8 sources (its planes)."""
    import oracle
    from py_ortho_amd import block
    q = next((q for _, q in arrivals if len(q.srcs) == 8), arrivals[0][1])
    img, cam = images[q.ref]
    H, W = img.shape
    ref = np.nan_to_num(img).astype(np.float32)
    mu, sd = oracle.ref_stats(ref, window)
    ray = oracle.rays_world(H, W, cam.R, cam.cx, cam.cy, cam.f)
    srcs = [images[k][0] for k in q.srcs]
    geo = np.stack([block._geo18(images[k][1], cam) for k in q.srcs])
    planes = block.inverse_depth_planes(q.dmin, q.dmax, q.planes)
    return cpu_sweep_rate(ref, mu, sd, ray, srcs, geo, planes, window, top_k, args,
                          f"a {W}x{H} request, S={len(q.srcs)}")


def cpu_sweep_rate(ref, mu, sd, ray, srcs, geo, planes, window, top_k, args, what):
    """The oracle's sweep (FP64, the reference's arithmetic) over all planes on a centre crop of
    256x256 to 512x512 (~12 s per core), `cores` concurrent copies (one view per thread, densify.py:596):
    Mpixel-hyp/s = cores * crop pixels * planes / wall."""
    import oracle
    cores = _cores() if args.cpu_threads <= 0 else args.cpu_threads
    H, W = ref.shape
    nplanes = len(planes)

    def sweep(side, nd, threads):
        y0, x0 = (H - side) // 2, (W - side) // 2
        sl = np.s_[y0:y0 + side, x0:x0 + side]
        base = np.full((side, side), planes[0], np.float32)
        step = np.full((side, side), planes[1] - planes[0], np.float32)
        t = time.perf_counter()
        oracle.bench_sweep(threads, ref[sl], mu[sl], sd[sl], srcs, np.ascontiguousarray(ray[:, y0:y0 + side,
                                                                                         x0:x0 + side]),
                           geo, base, step, nd, window, top_k)
        return time.perf_counter() - t

    per = sweep(16, 8, 1) / (256 * 8)
    side = int(max(min(256, H, W), min(min(H, W), 512, math.sqrt(12.0 / (per * nplanes)))))
    wall = sweep(side, nplanes, cores)
    return {"value": cores * side * side * nplanes / wall / 1e6, "unit": "Mpixel-hyp/s", "cores": cores,
            "kind": "port", "sample": f"{cores} concurrent copies of the oracle sweep ({nplanes} planes, "
                                      f"{2 * window + 1}x{2 * window + 1} window) on the {side}x{side} centre crop of "
                                      f"{what} ({wall:.1f} s wall)"}


# ------------------------------------------------------------------------------- C4 precision variant
def run_c4(args, torch):
    """4000x3000 reference, S = 12, 384 planes + parabola + 8 PatchMatch iterations, with FP16 or BF16
    image storage (--fmt); the same work in FP64, and with FP32 storage (the variant's FP32 arithmetic
    without the 16-bit rounding), is run once beside it to state the error against both."""
    from py_ortho_amd import _lib, batch, block, dense, scene
    lib = _lib.lib()
    t_scene = time.perf_counter()
    sc = scene.config_c4(device="cuda", scale=args.scale)
    torch.cuda.synchronize()
    t_scene = time.perf_counter() - t_scene
    ref = sc.ref.img.contiguous()
    H, W = ref.shape
    S = len(sc.srcs)
    r, k, D, iters, pr, pstep = 5, 2, 384, 8, 5, 2
    dmin, dmax = 73.0, 92.0
    ray = scene.rays_world(sc.ref.cam, "cuda")
    mu, sd = dense.ref_stats(ref, r)
    geo = sc.geo()
    hyp = {"mode": "const", "inv": torch.from_numpy(block.inverse_depth_planes(dmin, dmax, D)).cuda()}
    srcs = [v.img.contiguous() for v in sc.srcs]
    deps = [v.depth.contiguous() for v in sc.srcs]
    active = torch.ones((H, W), dtype=torch.uint8, device="cuda")
    fmt = args.fmt
    stored = {}  # the images in each storage format, converted once (inputs resident in HBM)

    def step(f):
        """the whole view with image storage f ("f16", "bf16", "f32") or in FP64 (None)"""
        st = dense.SweepState(H, W)
        if f:
            if f not in stored:
                stored[f] = (dense.to_lp(ref, f), [dense.to_lp(x, f) for x in srcs])
            ref_f, srcs_f = stored[f]
            dense.sweep_lp(ref_f, mu, sd, srcs_f, ray, geo, hyp, r, k, f, state=st)
        else:
            srcs_f = srcs
            dense.sweep(ref, mu, sd, srcs, ray, geo, hyp, r, k, state=st)
        d = dense.sweep_finalize(ref, mu, sd, srcs, ray, geo, hyp, r, k, st)
        d_pm = d.clone()
        nrm = batch.init_normals(d, ray)
        score = st.best.clone()
        dense.pm_refine(ref, ray, sc.ref.cam.C, srcs_f, deps, sc.par(), d_pm, nrm, score, active,
                        pr, pstep, k, 0.5, 0.01, iters, dmin, dmax, 2654435761, fmt=f)
        return st, d, d_pm

    for _ in range(args.warmup):
        step(fmt)
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    prof_read(lib, "sweep_lp")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st_lp, d_lp, pm_lp = step(fmt)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    sw_ms, sw_n = prof_read(lib, "sweep_lp")
    pr_ms, pr_n = prof_read(lib, "pm_propagate")
    lib.om_profile_read(None, None, None)
    ms = wall / args.steps * 1e3
    torch.cuda.synchronize()
    t64 = time.perf_counter()
    st64, d64, pm64 = step(None)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t64) * 1e3
    valid = (st64.best > 0.3) & (64.0 * sd > 0.5)
    both = torch.isfinite(pm_lp) & torch.isfinite(pm64)

    def errors(a, b):
        """per-pixel cost and depth error of run a against run b (valid pixels of the FP64 sweep)"""
        (st_a, d_a, pm_a), (st_b, d_b, pm_b) = a, b
        db = (st_a.best - st_b.best).abs()[valid]
        rel_sw = ((d_a - d_b).abs() / d_b)[valid]
        fin = torch.isfinite(pm_a) & torch.isfinite(pm_b)
        rel_pm = ((pm_a - pm_b).abs() / pm_b)[fin]
        sub = lambda t: t.float()[:: max(1, t.numel() // 1000000)]  # noqa: E731
        return {"sweep_same_winner": float((st_a.idx == st_b.idx)[valid].float().mean()),
                "sweep_abs_best_median": float(db.median()), "sweep_abs_best_p99": float(torch.quantile(sub(db), 0.99)),
                "sweep_depth_rel_p99": float(torch.quantile(sub(rel_sw), 0.99)),
                "pm_depth_within_0.2pct": float((rel_pm < 2e-3).float().mean())}

    run_lp, run64 = (st_lp, d_lp, pm_lp), (st64, d64, pm64)
    err64 = errors(run_lp, run64)
    if fmt != "f32":
        run32 = step("f32")
        err32 = errors(run_lp, run32)
        err32_64 = errors(run32, run64)
    else:
        err32, err32_64 = None, err64
    truth = sc.ref.depth
    px_h = H * W * (D + 11 * iters)
    flops = H * W * D * sweep_flops_per_hyp(S)
    sw_avg = sw_ms / max(sw_n, 1) / 1e3
    res = {
        "metric": METRIC, "value": px_h / (ms / 1e3) / 1e6, "unit": "Mpixel-hyp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"{fmt} storage, f32 arithmetic (f64 plane and variance guard)",
        "data": "synthetic",
        "config": {"workload": f"C4 precision variant: {W}x{H} reference, S={S}, {D} planes 11x11 NCC top-2 + "
                               f"parabola, then {iters} PatchMatch iterations (11x11 window step 2), {fmt} images",
                   "H": H, "W": W, "sources": S, "planes": D, "pm_iterations": iters, "pixel_hypotheses": px_h,
                   "scene_s": t_scene},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "fp64_same_work_ms": t64, "speedup_vs_fp64": t64 / ms,
        "error_vs_fp64": err64, "error_vs_f32_storage": err32, "f32_storage_error_vs_fp64": err32_64,
        "accuracy_vs_truth": {"pm_median_rel_err": float((((pm_lp - truth).abs() / truth)[both]).median()),
                              "pm_fp64_median_rel_err": float((((pm64 - truth).abs() / truth)[both]).median())},
        "roofline": {"bound": "valu", "engine": "VALU FP32 (no MFMA: gather-and-reduce)",
                     "achieved": flops / sw_avg / 1e12 if sw_avg else 0.0, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": flops / sw_avg / 157.3e12 if sw_avg else 0.0, "traffic": None, "kernel": "sweep_lp_kernel",
                     "kernel_avg_ms": sw_avg * 1e3, "pm_propagate_ms_per_step": pr_ms / max(args.steps, 1)},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_view(
            ref.cpu().numpy(), mu.cpu().numpy(), sd.cpu().numpy(), ray.cpu().numpy(), [x.cpu().numpy() for x in srcs],
            geo, sc.par(), np.asarray(sc.ref.cam.C, np.float64), [x.cpu().numpy() for x in deps], d64.cpu().numpy(),
            batch.init_normals(d64, ray).cpu().numpy(), hyp["inv"].cpu().numpy(), r, k, pr, pstep, iters, dmin, dmax,
            args)
    return res


def cpu_baseline_view(ref, mu, sd, ray, srcs, geo, par, Cr, deps, depth0, normal0, planes, r, k, pr, pstep, iters,
                      dmin, dmax, args):
    """The oracle (FP64, the reference's arithmetic) on centre crops of one view: the sweep over all
    planes and pm_refine for all iterations, `cores` concurrent copies each; Mpixel-hyp/s = cores *
    (planes + 11 iterations) / (sweep + PatchMatch wall time per crop pixel)."""
    import oracle
    cores = _cores() if args.cpu_threads <= 0 else args.cpu_threads
    H, W = ref.shape
    D = len(planes)

    def crop(side):
        y0, x0 = (H - side) // 2, (W - side) // 2
        return y0, x0, np.s_[y0:y0 + side, x0:x0 + side]

    def sweep(side, nd, threads):
        y0, x0, sl = crop(side)
        base = np.full((side, side), planes[0], np.float32)
        step = np.full((side, side), planes[1] - planes[0], np.float32)
        t = time.perf_counter()
        oracle.bench_sweep(threads, ref[sl], mu[sl], sd[sl], srcs, np.ascontiguousarray(ray[:, y0:y0 + side,
                                                                                         x0:x0 + side]),
                           geo, base, step, nd, r, k)
        return time.perf_counter() - t

    def pm(side, it, threads):
        y0, x0, sl = crop(side)
        t = time.perf_counter()
        oracle.bench_pm(threads, ref[sl], np.ascontiguousarray(ray[:, y0:y0 + side, x0:x0 + side]), Cr, srcs, deps,
                        par, np.ascontiguousarray(depth0[sl]), np.ascontiguousarray(normal0[:, y0:y0 + side,
                                                                                          x0:x0 + side]),
                        np.ones((side, side), np.uint8), pr, pstep, k, 0.5, 0.01, it, dmin, dmax, 1)
        return time.perf_counter() - t

    # 256x256 for the sweep and 128x128 for PatchMatch at least, larger up to ~12 s a part per core
    target = 12.0
    side_s = int(max(256, min(512, math.sqrt(target / (sweep(16, 8, 1) / (256 * 8) * D)))))
    side_p = int(max(128, min(256, math.sqrt(target / (pm(16, 1, 1) / 256 * iters)))))
    ws = sweep(side_s, D, cores)
    wp = pm(side_p, iters, cores)
    per_px = ws / (side_s * side_s) + wp / (side_p * side_p)
    return {"value": cores * (D + 11 * iters) / per_px / 1e6, "unit": "Mpixel-hyp/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent copies of the oracle sweep ({D} planes, {side_s}x{side_s} centre crop) and "
                      f"pm_refine ({iters} iterations, {side_p}x{side_p} crop), S={len(srcs)}; "
                      f"{ws:.1f} s + {wp:.1f} s wall"}


# ------------------------------------------------------------------------------- fusion
def fuse_flops_per_candidate(S: int) -> int:
    """Algorithmic FLOPs of one fusion candidate (each IEEE op 1, FMA 2, divide 1): reference point
    (undistort 7, scale 2, rotate + centre 18) + final average 3, and per neighbour 109: rotate into
    j 18, project 16, round 2, undistort 7, back-project 20, rotate into the reference 18, project
    16, agreement 9, accumulate 3."""
    return 30 + 109 * S


def _survey_fuse(torch, cache, scale):
    """The fusion block and its depth maps (rendered on the device or loaded from `cache`)."""
    from py_ortho_amd import scene
    if cache and os.path.exists(cache):
        z = np.load(cache)
        cams, counts, ext = scene.survey_cameras(11, 8, 6, 90.0, [(int(round(1600 * scale)), int(round(1200 * scale)))],
                                                 0.9375, 0.7, 0.6, (4, 4))
        nbrs = scene.neighbours_by_distance(np.stack([c.C for c in cams]), 90.0, counts)
        return cams, nbrs, [z[f"d{k}"] for k in range(len(cams))], [z[f"m{k}"] for k in range(len(cams))]
    sv = scene.config_fuse(device="cuda", scale=scale)
    cams = [v.cam for v in sv.views]
    depths = scene.noisy_depths(sv, 12)
    med = [float(np.median(v.depth.cpu().numpy())) for v in sv.views]
    if cache:
        np.savez(cache, **{f"d{k}": d for k, d in enumerate(depths)}, **{f"m{k}": np.float64(m) for k, m in enumerate(med)})
    return cams, sv.nbrs, depths, med


def run_fuse(args, torch):
    from py_ortho_amd import _lib, fusion
    lib = _lib.lib()
    cams, nbrs, depths, med = _survey_fuse(torch, args.scene_cache, args.scale)
    n = len(cams)
    px_tol, min_views, stride = 1.5, 3, 2
    tol = {k: {j: float(np.clip(px_tol * float(med[k]) / max(cams[k].f * float(np.linalg.norm(cams[k].C - cams[j].C)),
                                                             1e-9), 0.01, 0.05)) for j in nbrs[k]} for k in range(n)}
    rng = np.random.default_rng(13)
    rgb_h = [rng.integers(0, 256, (c.h, c.w, 3), dtype=np.uint8) for c in cams]
    dm = {k: torch.from_numpy(depths[k]).cuda() for k in range(n)}
    rgb = {k: torch.from_numpy(rgb_h[k]).cuda() for k in range(n)}
    order = list(range(n))
    fz = fusion.Fuser(order, dm, cams, nbrs, set(order), lambda k: tol[k], px_tol, min_views, stride, rgb)

    def step():
        fz.run(reset_used=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    prof_read(lib, "fuse_eval")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    ev_ms, ev_n = prof_read(lib, "fuse_eval")
    lib.om_profile_read(None, None, None)
    P, C, N, O = fz.result()
    npts, ncand = len(P), fz.evaluated()
    ms = wall / args.steps * 1e3
    S = int(np.median([len(x) for x in nbrs]))
    ev_s = ev_ms / 1e3 / max(args.steps, 1)          # eval time of one step (all views)
    flops = ncand * fuse_flops_per_candidate(S)
    # consistency of every view (the check of the repair pass, densify.py:617)
    cons = []
    torch.cuda.synchronize()
    tc = time.perf_counter()
    for k in order:
        cons.append(fusion.consistency(k, dm[k], nbrs[k], dm, cams, px_tol, tol[k]))
    torch.cuda.synchronize()
    cons_s = time.perf_counter() - tc
    pix = fz.pixels
    res = {
        "metric": "depth-map Mpixel/s fused on 1 GPU (sequential geometric-consistency fusion)",
        "value": pix / (ms / 1e3) / 1e6, "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fusion of a {n}-view block, {cams[0].w}x{cams[0].h} depth maps (rendered depth "
                               "x (1 + N(0, 2e-3)), 3% + patch holes), 4 neighbours, px_tol 1.5, min_views 3, "
                               "point_stride 2", "views": n, "W": cams[0].w, "H": cams[0].h, "neighbours": S},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "points": npts, "candidates": ncand, "points_per_s": npts / (ms / 1e3),
        "consistency_ms_all_views": cons_s * 1e3,
        "roofline": {"bound": "valu", "engine": "VALU FP64", "achieved": flops / ev_s / 1e12 if ev_s else 0,
                     "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": (flops / ev_s / FP64_PEAK) if ev_s else 0,
                     "traffic": profiled_traffic("fuse_eval_block.md"), "kernel": "fuse_eval_kernel",
                     "kernel_avg_ms": ev_ms / max(ev_n, 1), "launches_per_step": ev_n / max(args.steps, 1)},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_fuse(order, depths, cams, rgb_h, nbrs, tol, px_tol, min_views, stride,
                                                (P, C, N, O), args)
    return res


def cpu_baseline_fuse(order, depths, cams, rgb, nbrs, tol, px_tol, min_views, stride, gpu_out, args):
    """The oracle (C restatement of densify._fuse_loop, one core) over the first views of the
    order, about 10 s of work; its points must equal the GPU's first points bit for bit."""
    import oracle
    n = len(cams)
    nb = np.full((n, max(len(x) for x in nbrs)), -1, np.int32)
    tl = np.zeros(nb.shape)
    for k in range(n):
        for s_, j in enumerate(nbrs[k]):
            nb[k, s_], tl[k, s_] = j, tol[k][j]
    t0 = time.perf_counter()                 # one view, to size a prefix of about 10 s
    oracle.fuse(order[:1], depths, cams, rgb, nb, tl, px_tol, min_views, stride)
    m = max(1, min(n, int(10.0 / max(time.perf_counter() - t0, 1e-3))))
    t0 = time.perf_counter()
    pts, cols, nv, owner, _ = oracle.fuse(order[:m], depths, cams, rgb, nb, tl, px_tol, min_views, stride)
    dt = time.perf_counter() - t0
    P, C, N, O = (x.cpu().numpy() for x in gpu_out)
    same = (len(pts) <= len(P) and np.array_equal(pts.view(np.uint64), P[:len(pts)].view(np.uint64))
            and np.array_equal(nv, N[:len(pts)]) and np.array_equal(owner, O[:len(pts)]))
    pix = sum(int(cams[k].h) * int(cams[k].w) for k in order[:m])
    return {"value": pix / dt / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": "port",
            "sample": f"the first {m} of {n} views in fusion order ({dt:.1f} s)",
            "gpu_points_identical": bool(same)}


# ------------------------------------------------------------------------------- SGM
def run_sgm(args, torch):
    """Semi-global matching of the height sweep's coarse level (mvs.py:305-325): B tiles of
    (D, H, W) = (96, 48, 48) -- a 160-cell tile with its 16-cell pad at level 2 (DenseOptions
    tile 160, window 3, levels 3) and 96 height planes -- aggregated in one om_sgm call."""
    from py_ortho_amd import _lib
    lib = _lib.lib()
    B, D, H, W = args.sgm_tiles, 96, 48, 48
    rng = np.random.default_rng(17)
    # smooth surfaces with cliffs under noise, as 1 - NCC of the sweep gives them
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    cost = np.empty((B, D, H, W), np.float32)
    for b in range(B):
        z = 30 + 10 * np.sin(xx / (6 + b % 5)) + 8 * np.cos(yy / (7 + b % 3))
        z[:, W // 2 + (b % 7):] += 15
        d = np.arange(D, dtype=np.float32)[:, None, None]
        cost[b] = np.minimum(np.abs(d - z[None]) * 0.1, 1.0) + rng.uniform(0, 0.4, (D, H, W))
    c = torch.from_numpy(cost).cuda()
    agg = torch.empty_like(c)
    nb = lib.om_sgm_workspace_bytes(B, D, H, W)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    P1, P2 = 0.05, 0.6

    def step():
        _lib.check(lib.om_sgm(c.data_ptr(), B, D, H, W, P1, P2, agg.data_ptr(), ws.data_ptr(), nb, st), "om_sgm")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    prof_read(lib, "sgm_path")
    prof_read(lib, "sgm")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    path_ms, path_n = prof_read(lib, "sgm_path")
    all_ms, all_n = prof_read(lib, "sgm")
    lib.om_profile_read(None, None, None)
    ms = wall / args.steps * 1e3
    n = B * D * H * W
    # one tile per call, as the reference calls sgm once per tile
    c1 = c[:1].contiguous()
    a1 = torch.empty_like(c1)
    nb1 = lib.om_sgm_workspace_bytes(1, D, H, W)
    ws1 = torch.empty(nb1, dtype=torch.uint8, device="cuda")
    for _ in range(3):
        _lib.check(lib.om_sgm(c1.data_ptr(), 1, D, H, W, P1, P2, a1.data_ptr(), ws1.data_ptr(), nb1, st), "om_sgm")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 50
    for _ in range(reps):
        _lib.check(lib.om_sgm(c1.data_ptr(), 1, D, H, W, P1, P2, a1.data_ptr(), ws1.data_ptr(), nb1, st), "om_sgm")
    torch.cuda.synchronize()
    single_ms = (time.perf_counter() - t1) / reps * 1e3
    path_s = path_ms / 1e3 / max(path_n, 1)
    res = {
        "metric": "aggregated cost entries/s on 1 GPU (semi-global matching, 8 paths)",
        "value": n / (ms / 1e3) / 1e6, "unit": "Mentries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"SGM of {B} coarse-level height-sweep tiles, (D, H, W) = ({D}, {H}, {W}), "
                               "P1 0.05, P2 0.6, one call", "tiles": B, "D": D, "H": H, "W": W},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "one_tile_per_call_ms": single_ms,
        "roofline": {"bound": "hbm", "achieved": 8.0 * n / (all_ms / 1e3 / max(all_n, 1)) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": 8.0 * n / (all_ms / 1e3 / max(all_n, 1)) / HBM_PEAK,
                     "traffic": profiled_traffic("sgm_tiles.md"), "kernel": "om_sgm (transpose, paths, ordered sum)",
                     "kernel_avg_ms": all_ms / max(all_n, 1), "path_kernel_avg_ms": path_s * 1e3},
    }
    if not args.no_cpu:
        import oracle
        k = 0
        t0 = time.perf_counter()
        while True:
            ref = oracle.sgm(cost[k % B], P1, P2)
            k += 1
            if time.perf_counter() - t0 > 10.0 or k >= B:
                break
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(ref.view(np.uint32), agg[(k - 1) % B].cpu().numpy().view(np.uint32)))
        res["cpu_baseline"] = {"value": k * D * H * W / dt / 1e6, "unit": "Mentries/s", "cores": 1, "kind": "port",
                               "sample": f"{k} tiles one after another ({dt:.1f} s)", "gpu_identical": same}
    return res


# ------------------------------------------------------------------------------- view preparation
def run_prep(args, torch):
    """densify's per-view preparation (load_view, densify.py:528-535, and the pyramid of
    depth_map, :384-386) for 48 decoded 1600x1200 views: depth_view (uint8 RGB x gain + bias ->
    grey with mask NaN + uint8 colour) and two halvings, inputs resident in HBM; also the rate
    with each view's upload from pinned host memory."""
    from py_ortho_amd import _lib, prep
    lib = _lib.lib()
    n_views, h, w = 48, 1200, 1600
    rng = np.random.default_rng(19)
    rgb_h = [torch.from_numpy(rng.integers(0, 256, (h, w, 3), dtype=np.uint8)).pin_memory() for _ in range(n_views)]
    mask_h = [torch.from_numpy((rng.random((h, w)) < 0.98).view(np.uint8)).pin_memory() for _ in range(n_views)]
    rgb_d = [x.cuda() for x in rgb_h]
    mask_d = [x.cuda() for x in mask_h]
    gains = [(1.0 + 0.1 * rng.standard_normal(3)).astype(np.float32) for _ in range(n_views)]

    def step(upload: bool):
        out = []
        for k in range(n_views):
            r, m = ((rgb_h[k].cuda(non_blocking=True), mask_h[k].cuda(non_blocking=True)) if upload
                    else (rgb_d[k], mask_d[k]))
            out.append(prep.prepare_view(r, gains[k], None, m, levels=3))
        return out

    for _ in range(args.warmup):
        step(False)
        step(True)
    torch.cuda.synchronize()
    lib.om_profile_enable(1)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step(False)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    lib.om_profile_enable(0)
    lib.om_profile_read(None, None, None)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall_up = time.perf_counter() - t1
    ms = wall / args.steps * 1e3
    pix = n_views * h * w
    # HBM bytes a pixel: depth_view reads 3 + 1 and writes 4 + 3; the halvings read 4 + 1 and
    # write 1 + 0.25
    bpp = 3 + 1 + 4 + 3 + 4 + 1 + 1 + 0.25
    res = {
        "metric": "views prepared per second on 1 GPU (matching image + colour + pyramid)",
        "value": pix / (ms / 1e3) / 1e6, "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n_views} decoded {w}x{h} views: depth_view (gain, mask) + 2 halvings",
                   "views": n_views, "W": w, "H": h},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
        "with_upload_mpix_s": pix / (wall_up / args.steps) / 1e6,
        "roofline": {"bound": "hbm", "achieved": bpp * pix / (ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": bpp * pix / (ms / 1e3) / HBM_PEAK, "traffic": None,
                     "kernel": "depth_view_kernel + half_kernel (step time, launches included)"},
    }
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 5.0 and k < n_views:
            g0, _ = oracle.depth_view(rgb_h[k].numpy(), gains[k], None, mask_h[k].numpy().view(bool))
            oracle.half(oracle.half(g0))
            k += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": k * h * w / dt / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": "port",
                               "sample": f"{k} views ({dt:.1f} s): NumPy depth_view + C halvings"}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c3", "c1", "c2", "c4", "c5", "fuse", "sgm", "prep"], default="c3")
    ap.add_argument("--sgm-tiles", type=int, default=256, help="sgm: tiles aggregated in one call")
    ap.add_argument("--fmt", choices=["f16", "bf16", "f32"], default="f16", help="C4: image storage format")
    ap.add_argument("--rate", type=float, default=40.0, help="C5: Poisson arrival rate (requests/s)")
    ap.add_argument("--duration", type=float, default=3.0, help="C5: seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="C5: admission budget (GB of HBM)")
    ap.add_argument("--step-planes", type=int, default=32, help="C5: planes per request per step")
    ap.add_argument("--plain-requests", type=int, default=24, help="C5: requests timed one at a time")
    ap.add_argument("--active-mpix", type=float, default=48.0,
                    help="C5: megapixels of admitted requests in flight (latency under overload)")
    ap.add_argument("--scale", type=float, default=1.0, help="C3, fuse: image size scale (1 = the full block)")
    ap.add_argument("--no-plain-per-hyp", action="store_true",
                    help="C3: skip timing the per-view form with one sweep launch per plane")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0 = all cores)")
    ap.add_argument("--scene-cache", default=None, help="npz file holding the rendered scene (created if absent)")
    ap.add_argument("--calib-mb", type=int, default=0,
                    help="first run one plain elementwise kernel over this many MiB (read N, write N) as the "
                         "reference point for hardware byte counters")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (views shard across "
              "GPUs without communication; see the sharding notes)", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    if args.calib_mb > 0:
        n = args.calib_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        del a, b
    res = {"c2": run_c2, "c1": run_c1, "c3": run_c3, "c4": run_c4, "c5": run_c5,
           "fuse": run_fuse, "sgm": run_sgm, "prep": run_prep}[args.workload](args, torch)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
