"""er_adjoint_lut_noalloc: the adjoint with every work buffer in the library's static device
scratch array instead of per-call device allocations.

It runs the same kernels as er_adjoint_lut, one pass at a time, so its gradient must equal
er_adjoint_lut's up to the FP32 summation order of the workgroup histograms. Two calls in a
row, the second on a smaller film, check that what the first call leaves in the shared array
(queue counters, histograms, the FP64 sum) does not leak into the second. Oversized films and
tracers whose device tables would need an allocation are rejected with an error.
"""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import scenes
import oracle as O

pytestmark = pytest.mark.gpu


def _positive_lut(seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, 256)
    lut = np.zeros((256, 4), dtype=np.float32)
    lut[:, 0] = 0.05 + 0.5 * x + 0.1 * np.sin(6.0 * x + rng.uniform(0, 3))
    for c in range(3):
        lut[:, 1 + c] = 0.5 + 0.3 * np.sin(4.0 * x + rng.uniform(0, 6))
    return lut


def _tracer(scale, seed):
    sc = scenes.config_adjoint(scale=scale)
    r = er.Renderer(sc)
    r.set_tf_lut(_positive_lut(seed), sc.tf_range)
    s1, s2 = O.seeds_for(r.W, r.H, 700 + seed)
    A = np.random.default_rng(seed).standard_normal((r.H, r.W, 3)).astype(np.float32)
    return r, s1, s2, A


def _close(got, want):
    scale = np.abs(want).max()
    assert scale > 0.0
    err = np.abs(got - want)
    print(f"[noalloc] max |diff| {err.max():.3e} of max |g| {scale:.3e}", flush=True)
    # FP32 histogram sums in a different order: 1e-4 relative per entry, plus 1e-5 of the
    # largest entry for entries that are sums of cancelling contributions
    assert np.all(err <= 1e-4 * np.abs(want) + 1e-5 * scale)


def test_noalloc_adjoint_matches_adjoint_lut_twice_in_a_row(fresh):
    big, b1, b2, bA = _tracer(0.5, 1)         # 256 x 256 film, 64^3 volume
    small, q1, q2, qA = _tracer(0.25, 2)      # 128 x 128 film, 32^3 volume
    print(f"[noalloc] films {big.W}x{big.H} then {small.W}x{small.H}", flush=True)
    want_big = big.adjoint_lut(b1, b2, 4, bA)
    want_small = small.adjoint_lut(q1, q2, 3, qA)
    got_big = big.adjoint_lut_noalloc(b1, b2, 4, bA)
    got_small = small.adjoint_lut_noalloc(q1, q2, 3, qA)
    _close(got_big, want_big)
    _close(got_small, want_small)
    # and once more on the larger film after the smaller call
    _close(big.adjoint_lut_noalloc(b1, b2, 4, bA), want_big)
    big.close()
    small.close()


def test_noalloc_adjoint_rejects_what_it_cannot_do_without_allocating(fresh):
    sc = scenes.config_adjoint(scale=0.125)
    scenes.set_camera(sc.tracer, (600, 600), (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), exposure=4.0)
    r = er.Renderer(sc)
    r.set_tf_lut(_positive_lut(3), sc.tf_range)
    s1, s2 = O.seeds_for(r.W, r.H, 5)
    A = np.ones((r.H, r.W, 3), np.float32)
    with pytest.raises(er.ExposureRenderError, match="scratch limit"):
        r.adjoint_lut_noalloc(s1, s2, 2, A)        # 600 x 600 > 512 x 512
    r.close()
    # a tracer that has never rendered has no device tables yet: building them would allocate
    t, s1, s2, A = _tracer(0.125, 4)
    with pytest.raises(er.ExposureRenderError, match="allocates no device memory"):
        t.adjoint_lut_noalloc(s1, s2, 2, A)
    t.render_estimate()
    g = t.adjoint_lut_noalloc(s1, s2, 2, A)
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0.0
    t.close()
