"""Writes tests/golden/cornell_64x64_fp64_4096spp.npy: the FP64 oracle's
64x64 render of the 36-triangle Cornell box (camera and materials of
srt_amd.scenes.cornell), 4096 spp, max depth 5, rr depth 4, NEE+MIS, seed 7.
The image is the converged reference the float renders are compared with."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from srt_amd import scenes  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

if __name__ == "__main__":
    s = scenes.cornell()
    osc = O.Scene(s)
    img, secs = O.render(osc, s.camera, 64, 64, 4096, 5, 4, 1, 1, 7, fp64=True, bvh=O.Bvh(osc))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cornell_64x64_fp64_4096spp.npy")
    np.save(out, img[:, :3].astype(np.float64))
    print(f"wrote {out} in {secs:.1f} s, mean {img[:, :3].mean():.6f}")
