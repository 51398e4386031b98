"""Loader of the depth_map golden vectors (tests/golden/depthmap_*.npz)."""
import os
from types import SimpleNamespace

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name="grid3x3"):
    z = np.load(os.path.join(GOLDEN, f"depthmap_{name}.npz"))
    f, cx, cy, W, H = z["intr"]
    intr = [SimpleNamespace(f=float(f), cx=float(cx), cy=float(cy), k1=0.0, k2=0.0, k3=0.0, width=int(W),
                            height=int(H))]
    N = z["R"].shape[0]
    rec = SimpleNamespace(used=list(range(N)), R=z["R"], C=z["C"], X=z["X"], obs_cam=z["obs_cam"],
                          obs_pt=z["obs_pt"], intr=intr, cam_group=np.zeros(N, np.int32))
    nd, win, nb, rs, tk, mt = z["opt"]
    opt = dict(num_depths=int(nd), window=int(win), neighbors=int(nb), refine_steps=int(rs), top_k=int(tk),
               min_texture=float(mt))
    nbrs = [[int(j) for j in row if j >= 0] for row in z["nbrs"]]
    return z, rec, opt, nbrs, (int(W), int(H), float(f), float(cx), float(cy))
