"""The RMSE gate on the GPU at full size: per-pixel RMSE of the tonemapped running estimate
against the FP64 oracle's converged image (tests/golden: every pixel of the small workload,
4096 pixels of the others, at 4096 / 16384 samples per pixel) must reach 1%. Passes are
rendered in blocks until it does; the workload's own pass count is the first stop, and the
test reports the passes used (a configuration whose spp does not reach the gate runs on)."""
import os
import sys

import pytest

import exposure_render as er
from exposure_render import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = {
    "small": (scenes.config_small, "fp64_small.npz", 512),
    "ct": (scenes.config_ct, "fp64_ct.npz", 512),
    "large": (scenes.config_large, "fp64_large.npz", 1024),
    "ct_woodcock": (scenes.config_ct, "fp64_ct_woodcock.npz", 512),
}


@pytest.mark.parametrize("case", list(CASES))
def test_rmse_gate(fresh, case):
    builder, fixture, cap = CASES[case]
    golden = os.path.join(ROOT, "tests", "golden", fixture)
    sc = builder()
    if case == "ct_woodcock":
        sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
    r = er.Renderer(sc)
    r.seed(4321)
    passes, block = 0, sc.passes
    while True:
        r.render_passes(block)
        passes += block
        rm = bench.rmse_vs_golden(r.running(), golden)
        print(f"[rmse] {case}: {passes} passes, display RMSE {rm['display']:.5f} over {rm['pixels']} pixels "
              f"(reference {rm['reference_spp']} spp)", flush=True)
        if rm["display"] <= 0.01 or passes >= cap:
            break
        block = passes
    assert rm["display"] <= 0.01, (case, passes, rm)
    r.close()
