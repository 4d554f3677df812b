"""An ingested volume on the GPU (Source/ervolume.h:63-69, BindVoxels): a CT-like phantom in
Hounsfield-style signed 16-bit values with anisotropic spacing is written as a big-endian
MetaImage file, read back through er_mhd_read (shifted by its minimum into uint16), bound and
rendered, and checked sample for sample against the oracle on the same voxels."""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import scenes
from test_gpu_parity import run_parity
from test_ingest import _write_mhd

pytestmark = pytest.mark.gpu


def test_ingested_ct_volume_renders_like_the_oracle(fresh, tmp_path):
    base = scenes.phantom_ellipsoids((36, 40, 48), 8, seed=12)          # [z, y, x], 0 .. 4095
    hu = base.astype(np.int32) - 1024
    hu[0, 0, 0] = -1024
    path = _write_mhd(tmp_path, "ct", hu.astype(np.int16), "MET_SHORT", spacing=(0.8, 0.8, 1.5), msb=True)
    vox, spacing, off, scale = er.load_mhd(path)
    assert off == -1024.0 and scale == 1.0
    assert np.array_equal(vox.astype(np.int32), hu + 1024)
    ref = scenes.config_ct(scale=0.125)          # transfer function, lights, camera, hybrid shading
    sc = scenes.Scene(vox, spacing=spacing)
    sc.tracer, sc.lights, sc.textures = ref.tracer, ref.lights, ref.textures
    sc.name = "ingested_mhd_ct"
    print(f"[ingest] {path}: {vox.shape[::-1]} spacing {spacing}", flush=True)
    assert run_parity(sc, passes=2) > 0.99
