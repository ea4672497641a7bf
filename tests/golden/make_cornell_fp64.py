"""Writes tests/golden/cornell_<N>x<N>_fp64_4096spp.npy: the FP64 oracle's
NxN render (N = 64 by default, or the first argument) of the 36-triangle
Cornell box (camera and materials of srt_amd.scenes.cornell), 4096 spp, max
depth 5, rr depth 4, NEE+MIS, seed 7.  The image is the converged reference
the float renders are compared with (64: float64 values; larger sizes are
stored as float32 to keep the fixture small)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from srt_amd import scenes  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

if __name__ == "__main__":
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 64
    s = scenes.cornell()
    osc = O.Scene(s)
    img, secs = O.render(osc, s.camera, n, n, 4096, 5, 4, 1, 1, 7, fp64=True, bvh=O.Bvh(osc))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"cornell_{n}x{n}_fp64_4096spp.npy")
    np.save(out, img[:, :3].astype(np.float64 if n == 64 else np.float32))
    print(f"wrote {out} in {secs:.1f} s, mean {img[:, :3].mean():.6f}")
