"""The densify drop-in (py_ortho_amd.densify) against the reference's own densify.densify run on a
seeded 3 x 3 survey (tests/golden/densify_grid3x3.npz, make_densify_golden.py): the dense cloud,
the pass-1 maps and the final filtered maps, bit for bit, with PatchMatch repair on and off."""
import os
from types import SimpleNamespace

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "densify_grid3x3.npz")
SETTINGS = {
    "pm": dict(num_depths=32, window=3, neighbors=4, patchmatch=True, pm_iters=2, pm_window=3, point_stride=2),
    "plain": dict(num_depths=24, window=2, neighbors=3, patchmatch=False, point_stride=2, min_views=2),
}


def _inputs(z):
    f, cx, cy, W, H = z["intr"]
    W, H = int(W), int(H)
    N = len(z["R"])
    intr = [SimpleNamespace(f=float(f), cx=float(cx), cy=float(cy), k1=0.0, k2=0.0, k3=0.0, width=W, height=H)]
    rec = SimpleNamespace(used=list(range(N)), R=z["R"], C=z["C"], X=z["X"], obs_cam=z["obs_cam"],
                          obs_pt=z["obs_pt"], intr=intr, cam_group=np.zeros(N, np.int32), rms_px=float(z["rms_px"]))
    frames = [SimpleNamespace(width=W, height=H, path=f"view{i}.jpg", index=i) for i in range(N)]
    ar = SimpleNamespace(frames=frames, workers=2)
    gains = {1: z["gain_1"], 4: z["gain_4"]}
    biases = {4: z["bias_4"], 6: z["bias_6"]}
    rgb = z["rgb"]
    masks = {"view2.jpg": z["mask2"]}
    return ar, rec, gains, biases, (lambda fr, s: rgb[fr.index]), (lambda path, w, h: masks.get(path))


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", list(SETTINGS))
def test_densify_matches_reference(name, tmp_path):
    """The default (fused sweep) densify against the reference under the parity bar: every depth-map
    file equal to the reference's except at a few pixels (an FP64 window sum rounding differently
    moves a score by ~1e-8, which can move a parabola step and what PatchMatch does from there), depth
    within 0.2 % on >= 99 % of the pixels finite in both, and >= 98 % of the reference's points
    present bit for bit in the cloud."""
    from py_ortho_amd import densify, mvs
    z = np.load(GOLDEN)
    ar, rec, gains, biases, load_rgb, load_mask = _inputs(z)
    opt = mvs.DepthOptions(keep_depthmaps=True, workers=2, **SETTINGS[name])
    cloud = densify.densify(ar, rec, gains, biases, str(tmp_path), opt, load_rgb=load_rgb, load_mask=load_mask)
    torch.cuda.synchronize()
    wd = tmp_path / "depthmaps"
    stems = [k.split("_", 1)[1] for k in z.files if k.startswith(name + "_")]
    stems = [s for s in stems if s.isdigit() or (s[0] in "pst" and s[1:].isdigit())]
    assert stems
    report = []
    for stem in sorted(stems, key=lambda s: (s.isdigit(), s)):       # pass-1 files first, then final maps
        got, want = np.load(wd / f"{stem}.npy"), z[f"{name}_{stem}"]
        if _same(got, want):
            continue
        g, w = got.astype(np.float64), want.astype(np.float64)
        diff = (g != w) & ~(np.isnan(g) & np.isnan(w))
        both = np.isfinite(g) & np.isfinite(w) & (w != 0)
        rel = np.abs(g[both] - w[both]) / np.abs(w[both])
        report.append((stem, int(diff.sum()), float(np.mean(rel <= 2e-3)) if rel.size else 1.0))
        assert diff.mean() <= 0.005 and (rel.size == 0 or np.mean(rel <= 2e-3) >= 0.99), report
    print("densify", name, "files differing from the reference:", report)
    if name == "pm":   # the repair pass changed pixels: final maps are not just the filtered pass-1 maps
        moved = sum(int((np.nan_to_num(z[f"pm_{k:05d}"]) != np.nan_to_num(z[f"pm_p{k:05d}"])).sum())
                    for k in range(9) if f"pm_{k:05d}" in z.files)
        assert moved > 100, moved
    want = z[f"{name}_xyz"]
    have = {row.tobytes() for row in np.ascontiguousarray(cloud.xyz)}
    found = np.mean([row.tobytes() in have for row in np.ascontiguousarray(want)])
    print("densify", name, "points", len(cloud.xyz), "reference", len(want), "found bit for bit", found)
    assert abs(len(cloud.xyz) - len(want)) <= 0.01 * len(want) and found >= 0.98
    assert cloud.spacing == float(z[f"{name}_spacing"])


def test_densify_keeps_no_files_by_default(tmp_path):
    from py_ortho_amd import densify, mvs
    z = np.load(GOLDEN)
    ar, rec, gains, biases, load_rgb, load_mask = _inputs(z)
    (tmp_path / "depthmaps").mkdir()
    (tmp_path / "depthmaps" / "stale.npy").write_bytes(b"x")      # a leftover of an interrupted run
    opt = mvs.DepthOptions(**SETTINGS["plain"])
    cloud = densify.densify(ar, rec, gains, biases, str(tmp_path), opt, load_rgb=load_rgb, load_mask=load_mask)
    assert len(cloud.xyz) == len(z["plain_xyz"]) and not (tmp_path / "depthmaps").exists()


@pytest.mark.parametrize("name", list(SETTINGS))
def test_densify_exact_matches_reference_bit_for_bit(name, tmp_path):
    """exact=True: every sweep in the reference's summation order -- the cloud and every depth-map
    file equal the reference's bit for bit."""
    from py_ortho_amd import densify, mvs
    z = np.load(GOLDEN)
    ar, rec, gains, biases, load_rgb, load_mask = _inputs(z)
    opt = mvs.DepthOptions(keep_depthmaps=True, workers=2, **SETTINGS[name])
    cloud = densify.densify(ar, rec, gains, biases, str(tmp_path), opt, load_rgb=load_rgb, load_mask=load_mask,
                            exact=True)
    wd = tmp_path / "depthmaps"
    stems = [k.split("_", 1)[1] for k in z.files if k.startswith(name + "_")]
    stems = [s for s in stems if s.isdigit() or (s[0] in "pst" and s[1:].isdigit())]
    for stem in stems:
        assert _same(np.load(wd / f"{stem}.npy"), z[f"{name}_{stem}"]), stem
    for field in ("xyz", "rgb", "views", "cam"):
        assert _same(getattr(cloud, field), z[f"{name}_{field}"]), field


@pytest.mark.parametrize("max_batch_pixels", [64 << 20, 20000])
def test_densify_pass2_batched_equals_per_view(max_batch_pixels):
    """Pass 2 with every view's PatchMatch in one om_batch (or, with a small pixel budget, in several
    batches) gives the per-view pm_refine calls' final maps and change counts bit for bit."""
    from py_ortho_amd import densify, mvs
    z = np.load(GOLDEN)
    ar, rec, gains, biases, load_rgb, load_mask = _inputs(z)
    opt = mvs.DepthOptions(**SETTINGS["pm"])
    run = densify.DensifyRun(ar, rec, gains, biases, opt, load_rgb, load_mask)
    run.pass1()
    raw = {k: v.clone() for k, v in run.raw.items()}
    run.pass2(batched=False)
    per_view, stats = {k: v.clone() for k, v in run.final.items()}, dict(run.stats)
    run.stats = {"repaired": 0, "active": 0, "pixels": 0}
    run.final = {}
    run.pass2(batched=True, max_batch_pixels=max_batch_pixels)
    torch.cuda.synchronize()
    assert sorted(run.final) == sorted(per_view)
    for k in per_view:
        assert torch.equal(run.final[k].view(torch.int32), per_view[k].view(torch.int32)), k
        assert torch.equal(run.raw[k].view(torch.int32), raw[k].view(torch.int32)), k   # pass-1 maps untouched
    assert run.stats == stats and stats["active"] > 0 and stats["repaired"] > 0, (run.stats, stats)
