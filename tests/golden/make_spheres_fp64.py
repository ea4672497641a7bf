"""Writes tests/golden/spheres_192x108_fp64_4096spp.npy: the FP64 oracle's
192x108 render of the 1M-triangle sphere scene (srt_amd.scenes.spheres:
camera, lights and materials of its generator), 4096 spp, max depth 8, rr
depth 4, NEE+MIS, seed 7, stored as float32.  The converged reference for the
radiance-MLP comparison on that scene."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from srt_amd import scenes  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

if __name__ == "__main__":
    W, H = 192, 108
    s = scenes.spheres()
    osc = O.Scene(s)
    img, secs = O.render(osc, s.camera, W, H, 4096, 8, 4, 1, 1, 7, fp64=True, bvh=O.Bvh(osc))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"spheres_{W}x{H}_fp64_4096spp.npy")
    np.save(out, img[:, :3].astype(np.float32))
    print(f"wrote {out} in {secs:.1f} s, mean {img[:, :3].mean():.6f}")
