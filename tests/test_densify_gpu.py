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
    from py_ortho_amd import densify, mvs
    z = np.load(GOLDEN)
    ar, rec, gains, biases, load_rgb, load_mask = _inputs(z)
    opt = mvs.DepthOptions(keep_depthmaps=True, workers=2, **SETTINGS[name])
    cloud = densify.densify(ar, rec, gains, biases, str(tmp_path), opt, load_rgb=load_rgb, load_mask=load_mask)
    torch.cuda.synchronize()
    wd = tmp_path / "depthmaps"
    files = sorted(k for k in z.files if k.startswith(name + "_") and k.split("_", 1)[1][0] in "pst0123456789")
    assert files
    for key in files:
        stem = key.split("_", 1)[1]
        got = np.load(wd / f"{stem}.npy")
        assert _same(got, z[key]), (key, np.nanmax(np.abs(got.astype(np.float64) - z[key])) if got.dtype.kind == "f"
                                    else int((got != z[key]).sum()))
    if name == "pm":   # the repair pass changed pixels: final maps are not just the filtered pass-1 maps
        moved = sum(int((np.nan_to_num(z[f"pm_{k:05d}"]) != np.nan_to_num(z[f"pm_p{k:05d}"])).sum())
                    for k in range(9) if f"pm_{k:05d}" in z.files)
        assert moved > 100, moved
    for field in ("xyz", "rgb", "views", "cam"):
        assert _same(getattr(cloud, field), z[f"{name}_{field}"]), field
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
