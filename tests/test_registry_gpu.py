"""Registry and bind-time checks that need the device (binding uploads to it).

* A volume rebind without new voxels (voxels_dirty = 0) keeps the volume's own resolution, as
  Buffer3D::operator= skips a source that is not Dirty (Source/buffer3d.h:52-65); a changed
  resolution in such a descriptor must not reach the kernels, whose buffers keep their size.
* The cell records are indexed with 24-bit factors: a volume whose Rz + 1 reaches 2^24 is
  rejected at bind time instead of wrapping its z index.
* The autotuner leaves the displayed image as it was, including the bilateral filter's output
  that er_get_estimate returns when the filter is on.
"""
import ctypes as C

import numpy as np
import pytest

import exposure_render as er
from exposure_render import capi, scenes

pytestmark = pytest.mark.gpu


def test_volume_rebind_without_new_voxels_keeps_its_resolution(fresh, lib):
    sc = scenes.config_small(scale=0.25)
    r = er.Renderer(sc, seed=2)
    s1, s2 = r.get_seeds()
    r.render_estimate()
    f0 = r.frame().copy()
    v = r.volume
    print(f"[registry] rebind {v.resolution[0]}^3 volume claiming a larger resolution, voxels_dirty=0", flush=True)
    v.resolution[0] = 2 * v.resolution[0]
    v.resolution[2] = 3 * v.resolution[2]
    v.voxels_dirty = 0
    er._check(lib, lib.er_bind_volume(C.byref(v), 1))
    r.set_seeds(s1, s2)
    r.render_estimate()
    assert np.array_equal(r.frame(), f0)
    r.close()


def test_volume_beyond_24_bit_cell_coordinates_is_rejected(fresh, lib):
    v = capi.VolumeDesc()
    lib.er_volume_desc_init(C.byref(v))
    v.resolution[0], v.resolution[1], v.resolution[2] = 1, 1, 1 << 24
    v.spacing[0] = v.spacing[1] = v.spacing[2] = 1.0
    v.voxels_dirty = 1
    v.voxels = None          # rejected before any voxel is read or any device memory is taken
    st = lib.er_bind_volume(C.byref(v), 1)
    buf = C.create_string_buffer(512)
    lib.er_last_error(buf, 512)
    assert st == capi.ER_ERROR and b"2^24" in buf.value, buf.value
    assert v.id == -1


def test_autotune_restores_the_filtered_display(fresh):
    sc = scenes.config_ct(scale=0.125)
    assert sc.tracer.settings.bilateral and sc.tracer.settings.estimator == capi.ESTIMATOR_FULL
    print(f"[registry] autotune with the bilateral filter, film {sc.film}", flush=True)
    r = er.Renderer(sc, seed=4)
    for _ in range(2):
        r.render_estimate()
    before = r.get_estimate().copy()
    r.autotune(passes=2, force=True)
    assert np.array_equal(r.get_estimate(), before)
    r.close()


def test_host_seeds_equal_the_device_seeding(fresh):
    """exposure_render.seeds_for (what the benchmark and tests hand to er_set_seeds) equals
    er_seed_tracer's device-side seeding of the same tracer."""
    sc = scenes.config_small(scale=0.25)
    r = er.Renderer(sc)
    r.seed(123456789)
    s1, s2 = r.get_seeds()
    h1, h2 = er.seeds_for(r.W, r.H, 123456789)
    assert np.array_equal(s1, h1) and np.array_equal(s2, h2)
    r.close()
