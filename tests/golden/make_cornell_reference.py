"""Writes cornell256_4096spp.npz: the mean luminance Y (256x256 float32) of
the Cornell box workload (256x256, box filter, max depth 5, Russian roulette
after depth 3) rendered by the oracle at 4096 samples per pixel (8x8
stratified, 64 passes, seed 4242).  It is the reference image that image RMSE
is stated against.

    python tests/golden/make_cornell_reference.py [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from bling_amd import scenes  # noqa: E402
from bling_amd.runtime import JobSpec  # noqa: E402
from oracle import pyoracle  # noqa: E402

SPEC = dict(spp_nu=8, spp_nv=8, max_depth=5, sample_depth=3, rr_start_depth=3, passes=64, seed=4242)


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    flat = scenes.cornell(256)
    t0 = time.time()
    img, _, rays = pyoracle.render(flat, JobSpec(**SPEC), threads=threads)
    dt = time.time() - t0
    y = (img[..., 2] / img[..., 0]).astype(np.float32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cornell256_4096spp.npz")
    np.savez_compressed(out, y=y, spp=np.int32(4096), seconds=np.float64(dt), rays=np.uint64(rays))
    print("wrote %s in %.0fs (%d rays)" % (out, dt, rays))


if __name__ == "__main__":
    main()
