"""Writes the FP64 converged-image fixtures used by the RMSE gate (bench.py, tests).

For a benchmark scene it picks seeded 3x3 pixel blocks, runs the FP64 oracle (the same
estimator in double precision) at many samples per pixel on all nine pixels of each block,
and stores the 3x3-Gaussian-filtered expectation at the block centres - the expectation of
the running estimate (the per-pass filter of Source/filterframeestimate.cuh:28-64 is linear).

usage: python tests/golden/make_fp64_reference.py {small|ct} [blocks] [spp] [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "exposure-render_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from exposure_render import scenes  # noqa: E402
import oracle as O  # noqa: E402

G = np.exp(-np.array([[2, 1, 2], [1, 0, 1], [2, 1, 2]], dtype=np.float64) / 2.0)


def block_centres(W, H, n, seed):
    rng = np.random.default_rng(seed)
    xs = rng.integers(1, W - 1, size=n)
    ys = rng.integers(1, H - 1, size=n)
    return xs.astype(np.int32), ys.astype(np.int32)


def make(name, blocks, spp, threads):
    builder = {"small": scenes.config_small, "ct": scenes.config_ct}[name]
    sc = builder()
    orc = O.Oracle(sc, nthreads=threads)
    orc.compute_gm_max()
    W, H = sc.film
    cx, cy = block_centres(W, H, blocks, 2024)
    px = (cx[:, None] + np.array([-1, 0, 1] * 3)[None, :]).reshape(-1)
    py = (cy[:, None] + np.repeat([-1, 0, 1], 3)[None, :]).reshape(-1)
    t0 = time.time()
    e = orc.expect_f64(px, py, spp, seed=777).reshape(blocks, 3, 3, 4)
    filt = (e * G[None, :, :, None]).sum(axis=(1, 2)) / G.sum()
    filt[:, 3] = e[:, 1, 1, 3] / G.sum()    # the filter divides the centre alpha by the total weight
    out = os.path.join(HERE, f"fp64_{name}.npz")
    np.savez_compressed(out, cx=cx, cy=cy, expect=filt, spp=spp, exposure=sc.tracer.camera.exposure,
                        film=np.array([W, H]))
    print(f"{name}: {blocks} blocks x 9 px x {spp} spp in {time.time() - t0:.1f} s -> {out}")


if __name__ == "__main__":
    name = sys.argv[1] if len(sys.argv) > 1 else "small"
    blocks = int(sys.argv[2]) if len(sys.argv) > 2 else 256
    spp = int(sys.argv[3]) if len(sys.argv) > 3 else 4096
    threads = int(sys.argv[4]) if len(sys.argv) > 4 else 8
    make(name, blocks, spp, threads)
