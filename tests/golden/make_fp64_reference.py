"""Writes the FP64 converged-image fixtures used by the RMSE gate (bench.py, tests).

The converged image is the expectation of the running estimate: the mean of the per-pass
frames after the 3x3 Gaussian (Source/filterframeestimate.cuh:28-64), which is linear, so it
is the 3x3-Gaussian-filtered per-pixel expectation. The FP64 oracle (the same estimator in
double precision) estimates each needed pixel's expectation at many samples per pixel, and the
filter is applied to those.

  small        every pixel of the 256 x 256 film (the filter renormalised at the borders);
  ct, large,   4096 gated pixels: 16 seeded 16 x 16 tiles, one in each cell of a 4 x 4 grid
  ct_woodcock  over the volume's footprint on the film, each centred on a ray that passes
               through non-zero opacity and at least one pixel from the film's edge; each
               needs the expectations of its 18 x 18 neighbourhood.

Stored: cx, cy (gated pixels), expect (N x 4: filtered XYZ and the filtered alpha), spp,
exposure, film, seconds.

usage: python tests/golden/make_fp64_reference.py {small|ct|large|ct_woodcock} [spp] [threads]
"""
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "exposure-render_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from exposure_render import capi, scenes  # noqa: E402
import oracle as O  # noqa: E402

TILE = 16
GRID = 4


def scene(name):
    if name == "small":
        return scenes.config_small()
    if name == "ct":
        return scenes.config_ct()
    if name == "large":
        return scenes.config_large()
    if name == "ct_woodcock":
        sc = scenes.config_ct()
        sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
        return sc
    raise ValueError(name)


def _rays(orc, sc, xs, ys):
    """Camera rays through pixel centres (the film jitter's mean, Q12)."""
    W, H = sc.film
    cam = orc.camera().astype(np.float64)
    N, U, V = cam[0:3], cam[3:6], cam[6:9]
    sx = cam[9] + cam[11] * (xs + 0.5 / W)
    sy = cam[10] + cam[12] * (ys + 0.5 / H)
    D = N + sx[..., None] * U - sy[..., None] * V
    return D / np.linalg.norm(D, axis=-1)[..., None]


def _box(sc):
    res = np.array([sc.volume.resolution[i] for i in range(3)], dtype=np.float64)
    sp = np.array([sc.volume.spacing[i] for i in range(3)], dtype=np.float64)
    ext = res * sp / np.max(res * sp)
    return -0.5 * ext, 0.5 * ext


def _chord(O, D, lo, hi):
    t0 = np.zeros(D.shape[:-1])
    t1 = np.full(D.shape[:-1], np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(3):
            a, b = (lo[i] - O[i]) / D[..., i], (hi[i] - O[i]) / D[..., i]
            t0, t1 = np.maximum(t0, np.minimum(a, b)), np.minimum(t1, np.maximum(a, b))
    return t0, t1


def tiles(orc, sc, seed=2024):
    """Top-left corners of GRID x GRID seeded tiles over the volume's footprint on the film
    (the bounding rectangle of the pixels whose centre ray meets the volume's box), one per
    cell, 1 px inside the film; a candidate is taken when its centre ray passes through
    non-zero opacity (up to 200 candidates per cell)."""
    W, H = sc.film
    org = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    lo, hi = _box(sc)
    t0, t1 = _chord(org, _rays(orc, sc, xs, ys), lo, hi)
    hit = t1 > t0
    hx, hy = xs[hit], ys[hit]
    fx0, fx1 = max(1, int(hx.min())), min(W - 1 - TILE, int(hx.max()) - TILE)
    fy0, fy1 = max(1, int(hy.min())), min(H - 1 - TILE, int(hy.max()) - TILE)
    lib = O.load()

    def opaque(x, y):
        D = _rays(orc, sc, np.array(float(x)), np.array(float(y)))
        a, b = _chord(org, D, lo, hi)
        if not b > a:
            return False
        for t in np.linspace(a, b, 256):
            I = orc.intensity(org + t * D)
            if lib.orc_plf_eval(C.byref(sc.tracer.opacity), I) > 0.0:
                return True
        return False

    rng = np.random.default_rng(seed)
    out = []
    for gy in range(GRID):
        for gx in range(GRID):
            x0 = fx0 + gx * (fx1 - fx0) // GRID
            x1 = max(x0, fx0 + (gx + 1) * (fx1 - fx0) // GRID)
            y0 = fy0 + gy * (fy1 - fy0) // GRID
            y1 = max(y0, fy0 + (gy + 1) * (fy1 - fy0) // GRID)
            for _ in range(200):
                x, y = int(rng.integers(x0, x1 + 1)), int(rng.integers(y0, y1 + 1))
                if opaque(x + TILE // 2, y + TILE // 2):
                    break
            out.append((x, y))
    return out


def make(name, spp, threads):
    sc = scene(name)
    orc = O.Oracle(sc, nthreads=threads)
    if sc.tracer.settings.shading_type not in (capi.SHADING_BRDF_ONLY, capi.SHADING_PHASE_ONLY):
        orc.compute_gm_max()
    W, H = sc.film
    t0 = time.time()
    if name == "small":
        ys, xs = np.mgrid[0:H, 0:W]
        e = orc.expect_f64(xs.ravel(), ys.ravel(), spp, seed=777).reshape(H, W, 4)
        filt = O.gauss3x3_f64(e)                          # borders renormalise; alpha / total weight
        cy, cx = ys.ravel().astype(np.int32), xs.ravel().astype(np.int32)
        expect = filt.reshape(-1, 4).astype(np.float32)
    else:
        cxs, cys, exps = [], [], []
        for k, (x0, y0) in enumerate(tiles(orc, sc)):
            ys, xs = np.mgrid[y0 - 1:y0 + TILE + 1, x0 - 1:x0 + TILE + 1]
            e = orc.expect_f64(xs.ravel(), ys.ravel(), spp, seed=777 + k).reshape(TILE + 2, TILE + 2, 4)
            filt = O.gauss3x3_f64(e)[1:-1, 1:-1]          # interior pixels: all nine taps
            gy, gx = np.mgrid[y0:y0 + TILE, x0:x0 + TILE]
            cxs.append(gx.ravel())
            cys.append(gy.ravel())
            exps.append(filt.reshape(-1, 4))
            print(f"{name}: tile {k + 1}/{GRID * GRID} at ({x0},{y0}) {time.time() - t0:.0f} s", flush=True)
        cx = np.concatenate(cxs).astype(np.int32)
        cy = np.concatenate(cys).astype(np.int32)
        expect = np.concatenate(exps)
    secs = time.time() - t0
    out = os.path.join(HERE, f"fp64_{name}.npz")
    np.savez_compressed(out, cx=cx, cy=cy, expect=expect, spp=spp, exposure=sc.tracer.camera.exposure,
                        film=np.array([W, H]), seconds=secs)
    print(f"{name}: {len(cx)} gated pixels at {spp} spp in {secs:.1f} s -> {out}", flush=True)


if __name__ == "__main__":
    name = sys.argv[1] if len(sys.argv) > 1 else "small"
    spp = int(sys.argv[2]) if len(sys.argv) > 2 else (4096 if name == "small" else 16384)
    threads = int(sys.argv[3]) if len(sys.argv) > 3 else 8
    make(name, spp, threads)
