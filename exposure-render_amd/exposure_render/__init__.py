"""Python host binding of the MI355X single-scattering volume tracer.

The product is the native library lib/libexposurerender_amd.so (gfx950 HIP kernels behind the
C-ABI of include/er_capi.h). This module only loads it and offers a thin object wrapper over
the reference's binding API (Source/exposurerender.h:32-42): bind descriptors, render
progressive passes, read the estimate back. Every call fails loudly when the native library is
missing or returns a non-zero status; there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

from . import capi
from .capi import load, check_abi  # noqa: F401

LEVELS = {capi.ER_INFO: "Info", capi.ER_WARNING: "Warning", capi.ER_ERROR: "Error", capi.ER_FATAL: "Fatal"}


class ExposureRenderError(RuntimeError):
    """ExposureRender::Exception (Source/exception.h:27-55): level + message."""

    def __init__(self, status, message):
        self.status = status
        self.level = LEVELS.get(status, "Error")
        super().__init__(f"[{self.level}] {message}")


def _check(lib, status):
    if status != capi.ER_OK:
        buf = C.create_string_buffer(512)
        lib.er_last_error(buf, 512)
        raise ExposureRenderError(status, buf.value.decode(errors="replace"))
    return status


class Renderer:
    """Binds one Scene (scenes.Scene) through the C-ABI and renders it."""

    def __init__(self, scene, seed=None):
        """Binds copies of the scene's descriptors. The scene refers to textures, lights and
        clipping objects by their position in its lists; the copies are remapped to the IDs the
        registry assigns."""
        self.lib = load()
        self.scene = scene
        lib = self.lib

        def cp(d):
            return type(d).from_buffer_copy(d)

        self.volume = cp(scene.volume)
        self.volume.voxels = scene.volume.voxels
        _check(lib, lib.er_bind_volume(C.byref(self.volume), 1))
        self._pixels = list(scene.bitmaps)
        self.bitmaps = []
        for a in self._pixels:
            B = capi.BitmapDesc()
            lib.er_bitmap_desc_init(C.byref(B))
            B.resolution[0], B.resolution[1] = a.shape[1], a.shape[0]
            B.pixels = a.ctypes.data_as(C.POINTER(C.c_uint8))
            B.pixels_dirty = 1
            _check(lib, lib.er_bind_bitmap(C.byref(B), 1))
            self.bitmaps.append(B)
        self.textures = [cp(T) for T in scene.textures]
        for T in self.textures:
            if T.type == capi.TEXTURE_BITMAP:
                T.bitmap_id = self.bitmaps[T.bitmap_id].id if 0 <= T.bitmap_id < len(self.bitmaps) else -1
            _check(lib, lib.er_bind_texture(C.byref(T), 1))

        def tex_id(j):
            return self.textures[j].id if 0 <= j < len(self.textures) else -1
        self.lights = [cp(L) for L in scene.lights]
        for L in self.lights:
            if 0 <= L.texture_id < len(self.textures):
                L.texture_id = self.textures[L.texture_id].id
            _check(lib, lib.er_bind_light(C.byref(L), 1))
        self.clips = [cp(Cb) for Cb in scene.clips]
        for Cb in self.clips:
            _check(lib, lib.er_bind_clipping_object(C.byref(Cb), 1))
        self.objects = [cp(O) for O in scene.objects]
        for O in self.objects:
            O.diffuse_texture_id = tex_id(O.diffuse_texture_id)
            O.specular_texture_id = tex_id(O.specular_texture_id)
            O.glossiness_texture_id = tex_id(O.glossiness_texture_id)
            _check(lib, lib.er_bind_object(C.byref(O), 1))
        t = cp(scene.tracer)
        t.volume_id = self.volume.id
        for i in range(t.light_count):
            j = scene.tracer.light_ids[i]
            t.light_ids[i] = self.lights[j].id if 0 <= j < len(self.lights) else -1
        for i in range(t.clipping_count):
            j = scene.tracer.clipping_ids[i]
            t.clipping_ids[i] = self.clips[j].id if 0 <= j < len(self.clips) else -1
        for i in range(t.object_count):
            j = scene.tracer.object_ids[i]
            t.object_ids[i] = self.objects[j].id if 0 <= j < len(self.objects) else -1
        self.tracer = t
        _check(lib, lib.er_bind_tracer(C.byref(t), 1))
        self.id = t.id
        self.W, self.H = scene.film
        if seed is not None:
            self.seed(seed)

    # -- reference entry points ------------------------------------------------------------
    def render_estimate(self):
        _check(self.lib, self.lib.er_render_estimate(self.id))

    def get_estimate(self):
        out = np.zeros((self.H, self.W, 4), dtype=np.uint8)
        _check(self.lib, self.lib.er_get_estimate(self.id, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.nbytes))
        return out

    def autofocus_distance(self, u, v):
        d = C.c_float(-1.0)
        _check(self.lib, self.lib.er_get_autofocus_distance(self.id, int(u), int(v), C.byref(d)))
        return d.value

    def autofocus_seeds(self, u, v):
        a, b = C.c_uint32(), C.c_uint32()
        _check(self.lib, self.lib.er_autofocus_seeds(self.id, int(u), int(v), C.byref(a), C.byref(b)))
        return a.value, b.value

    def no_iterations(self):
        n = C.c_int(0)
        _check(self.lib, self.lib.er_get_no_iterations(self.id, C.byref(n)))
        return n.value

    # -- extensions -------------------------------------------------------------------------
    def render_passes(self, passes):
        t = capi.Timing()
        _check(self.lib, self.lib.er_render_passes(self.id, passes, C.byref(t)))
        return t

    def autotune(self, passes=6, force=False):
        """Pick the fastest brick size, rays per wave (primary and shadow), tile shape, primary
        workgroups per CU, shadow waves per CU and heavy-tile count for this scene (er_autotune);
        the render state is left as it was. Returns the chosen er_tune."""
        t = capi.Tune()
        _check(self.lib, self.lib.er_autotune(self.id, int(passes), capi.TUNE_FORCE if force else 0, C.byref(t)))
        return t.as_dict()

    def set_tune(self, brick, tile_w, blocks_per_cu, shadow_waves_per_cu=4, heavy_tiles=0, rays_per_wave=64,
                 shadow_rays_per_wave=64, heavy_rays_per_wave=4):
        t = capi.Tune(brick=brick, tile_w=tile_w, tile_h=64 // tile_w, blocks_per_cu=blocks_per_cu,
                      shadow_waves_per_cu=shadow_waves_per_cu, heavy_tiles=heavy_tiles, rays_per_wave=rays_per_wave,
                      shadow_rays_per_wave=shadow_rays_per_wave, heavy_rays_per_wave=heavy_rays_per_wave)
        _check(self.lib, self.lib.er_set_tune(self.id, C.byref(t)))

    def get_tune(self):
        t = capi.Tune()
        _check(self.lib, self.lib.er_get_tune(self.id, C.byref(t)))
        return t.as_dict()

    def seed(self, seed):
        _check(self.lib, self.lib.er_seed_tracer(self.id, int(seed)))

    def set_seeds(self, s1, s2):
        s1 = np.ascontiguousarray(s1, dtype=np.uint32)
        s2 = np.ascontiguousarray(s2, dtype=np.uint32)
        _check(self.lib, self.lib.er_set_seeds(self.id, s1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               s2.ctypes.data_as(C.POINTER(C.c_uint32)), s1.size))

    def get_seeds(self):
        n = self.W * self.H
        s1 = np.zeros(n, dtype=np.uint32)
        s2 = np.zeros(n, dtype=np.uint32)
        _check(self.lib, self.lib.er_get_seeds(self.id, s1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               s2.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return s1, s2

    def frame(self):
        out = np.zeros((self.H, self.W, 4), dtype=np.float32)
        _check(self.lib, self.lib.er_get_frame_xyza(self.id, out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes))
        return out

    def running(self):
        out = np.zeros((self.H, self.W, 4), dtype=np.float32)
        _check(self.lib, self.lib.er_get_running_xyza(self.id, out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes))
        return out

    def set_counting(self, on):
        _check(self.lib, self.lib.er_set_counting(self.id, 1 if on else 0))

    def counters(self):
        c = capi.Counters()
        _check(self.lib, self.lib.er_get_counters(self.id, C.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        _check(self.lib, self.lib.er_reset_counters(self.id))

    def gm_max(self):
        v = C.c_float(0)
        _check(self.lib, self.lib.er_get_gradient_magnitude_max(self.id, C.byref(v)))
        return v.value

    # -- node-table transfer function and its adjoint --------------------------------------
    def set_tf_lut(self, lut, range_):
        """lut: (ER_TF_LUT_NODES, 4) float32 (opacity, diffuse X, Y, Z) or None for the PLFs."""
        if lut is None:
            _check(self.lib, self.lib.er_set_tf_lut(self.id, None, 0.0))
            return
        lut = np.ascontiguousarray(lut, dtype=np.float32).reshape(capi.TF_LUT_NODES, 4)
        _check(self.lib, self.lib.er_set_tf_lut(self.id, lut.ctypes.data_as(C.POINTER(C.c_float)), float(range_)))

    def adjoint_lut(self, s1, s2, passes, dl_drunning):
        """dL/d(node table) (ER_TF_LUT_NODES, 4) float64 for dL/d(running XYZ) (H, W, 3)."""
        s1 = np.ascontiguousarray(s1, dtype=np.uint32)
        s2 = np.ascontiguousarray(s2, dtype=np.uint32)
        g = np.ascontiguousarray(dl_drunning, dtype=np.float32).reshape(self.H, self.W, 3)
        out = np.zeros((capi.TF_LUT_NODES, 4), dtype=np.float64)
        _check(self.lib, self.lib.er_adjoint_lut(self.id, s1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                 s2.ctypes.data_as(C.POINTER(C.c_uint32)), s1.size, int(passes),
                                                 g.ctypes.data_as(C.POINTER(C.c_float)),
                                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def adjoint_lut_noalloc(self, s1, s2, passes, dl_drunning):
        """adjoint_lut through er_adjoint_lut_noalloc: no device allocation, films up to 512 x 512."""
        s1 = np.ascontiguousarray(s1, dtype=np.uint32)
        s2 = np.ascontiguousarray(s2, dtype=np.uint32)
        g = np.ascontiguousarray(dl_drunning, dtype=np.float32).reshape(self.H, self.W, 3)
        out = np.zeros((capi.TF_LUT_NODES, 4), dtype=np.float64)
        _check(self.lib, self.lib.er_adjoint_lut_noalloc(self.id, s1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                         s2.ctypes.data_as(C.POINTER(C.c_uint32)), s1.size,
                                                         int(passes), g.ctypes.data_as(C.POINTER(C.c_float)),
                                                         out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def adjoint_value(self, s1, s2, passes, eval_lut):
        """The smooth estimator's running XYZ (H, W, 3) with values from eval_lut and the
        decisions of the current node table."""
        s1 = np.ascontiguousarray(s1, dtype=np.uint32)
        s2 = np.ascontiguousarray(s2, dtype=np.uint32)
        lut = np.ascontiguousarray(eval_lut, dtype=np.float32).reshape(capi.TF_LUT_NODES, 4)
        out = np.zeros((self.H, self.W, 3), dtype=np.float32)
        _check(self.lib, self.lib.er_adjoint_value(self.id, s1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                   s2.ctypes.data_as(C.POINTER(C.c_uint32)), s1.size, int(passes),
                                                   lut.ctypes.data_as(C.POINTER(C.c_float)),
                                                   out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def rebind(self):
        """BindTracer with the existing ID: restarts accumulation (NoIterations from the desc)."""
        self.tracer.no_iterations = 0
        _check(self.lib, self.lib.er_bind_tracer(C.byref(self.tracer), 1))

    def close(self):
        lib = self.lib
        lib.er_bind_tracer(C.byref(self.tracer), 0)
        for Cb in self.clips:
            lib.er_bind_clipping_object(C.byref(Cb), 0)
        for O in self.objects:
            lib.er_bind_object(C.byref(O), 0)
        for B in self.bitmaps:
            lib.er_bind_bitmap(C.byref(B), 0)
        for L in self.lights:
            lib.er_bind_light(C.byref(L), 0)
        for T in self.textures:
            lib.er_bind_texture(C.byref(T), 0)
        lib.er_bind_volume(C.byref(self.volume), 0)


def load_mhd(path):
    """A MetaImage volume (er_mhd_read) as (voxels[z, y, x] uint16, spacing (x, y, z), offset,
    scale), with u16 = (file value - offset) * scale."""
    lib = load()
    res = (C.c_int * 3)()
    sp = (C.c_float * 3)()
    et = C.c_int()
    b = os.fsencode(path)
    _check(lib, lib.er_mhd_info(b, res, sp, C.byref(et)))
    vox = np.empty((res[2], res[1], res[0]), dtype=np.uint16)
    off, sc = C.c_float(), C.c_float()
    _check(lib, lib.er_mhd_read(b, vox.ctypes.data_as(C.POINTER(C.c_uint16)), vox.size, C.byref(off), C.byref(sc)))
    return vox, (sp[0], sp[1], sp[2]), off.value, sc.value


def load_raw(path, resolution, element_type, msb=False, header_bytes=0):
    """A headerless raw volume (er_raw_read) of resolution (x, y, z): (voxels, offset, scale)."""
    lib = load()
    res = (C.c_int * 3)(*resolution)
    vox = np.empty((resolution[2], resolution[1], resolution[0]), dtype=np.uint16)
    off, sc = C.c_float(), C.c_float()
    _check(lib, lib.er_raw_read(os.fsencode(path), res, int(element_type), 1 if msb else 0, int(header_bytes),
                                vox.ctypes.data_as(C.POINTER(C.c_uint16)), vox.size, C.byref(off), C.byref(sc)))
    return vox, off.value, sc.value


def autotune_batch(renderers, rounds=4):
    """er_autotune_batch: one brick size, tile shape and workgroups-per-CU setting for the
    whole batch, timed as `rounds` batch rounds; the jobs' render states are restored and the
    jobs keep the chosen configuration. Returns the er_tune as a dict."""
    lib = load()
    items = (capi.BatchItem * len(renderers))()
    for k, r in enumerate(renderers):
        items[k].tracer_id = r.id
        items[k].passes = rounds
    t = capi.Tune()
    _check(lib, lib.er_autotune_batch(items, len(renderers), int(rounds), C.byref(t)))
    return t.as_dict()


def render_batch(renderers, passes):
    """er_render_batch: passes[i] progressive passes of renderers[i], all in one graph-captured
    pass pipeline per round. Returns the capi.Timing of the whole batch."""
    lib = load()
    items = (capi.BatchItem * len(renderers))()
    for k, (r, n) in enumerate(zip(renderers, passes)):
        items[k].tracer_id = r.id
        items[k].passes = int(n)
    t = capi.Timing()
    _check(lib, lib.er_render_batch(items, len(renderers), C.byref(t)))
    return t


def shutdown():
    lib = load()
    _check(lib, lib.er_shutdown())


def seeds_for(width, height, seed):
    """The per-pixel RNG seeds er_seed_tracer writes, on the host: SplitMix64(seed ^ pixel)
    split into its low and high 32-bit halves, a zero half replaced by 1 (zero is a fixed point
    of the multiply-with-carry generators, Source/rng.h:35-51)."""
    n = width * height
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ idx) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    s1 = (z & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    s2 = (z >> np.uint64(32)).astype(np.uint32)
    s1[s1 == 0] = 1
    s2[s2 == 0] = 1
    return s1, s2
