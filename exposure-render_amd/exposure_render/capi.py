"""ctypes mirror of include/er_capi.h and the loader of the native library.

The struct layouts follow the C header field for field; `check_abi()` compares every
sizeof with the library's own (`er_abi_sizeof`).
"""
import ctypes as C
import os

MAX_NODES = 128
MAX_INDICES = 256

ER_OK, ER_INFO, ER_WARNING, ER_ERROR, ER_FATAL = 0, 1, 2, 3, 4

SHAPE_PLANE, SHAPE_DISK, SHAPE_RING, SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYLINDER = range(6)
(SHADING_BRDF_ONLY, SHADING_PHASE_ONLY, SHADING_HYBRID, SHADING_MODULATION,
 SHADING_THRESHOLD, SHADING_GRADIENT_MAGNITUDE) = range(6)
GRADIENT_FORWARD, GRADIENT_CENTRAL, GRADIENT_FILTERED = range(3)
UNIT_POWER, UNIT_LUX, UNIT_INTENSITY = range(3)
TEXTURE_PROCEDURAL, TEXTURE_BITMAP = range(2)
PROCEDURAL_UNIFORM, PROCEDURAL_CHECKER, PROCEDURAL_GRADIENT = range(3)
ESTIMATOR_COMPAT, ESTIMATOR_FULL = range(2)
SAMPLER_RAYMARCH, SAMPLER_WOODCOCK = range(2)

F3 = C.c_float * 3
F2 = C.c_float * 2
I2 = C.c_int * 2
I3 = C.c_int * 3
M44 = (C.c_float * 4) * 4


class Plf(C.Structure):
    _fields_ = [("node_range", F2), ("position", C.c_float * MAX_NODES),
                ("value", C.c_float * MAX_NODES), ("count", C.c_int)]

    def init(self):
        """PiecewiseFunction() (Source/pf.h:36-43): range {FLT_MAX, FLT_MIN}, no nodes."""
        self.node_range[0] = 3.4028234663852886e38
        self.node_range[1] = 1.1754943508222875e-38
        self.count = 0
        return self

    def add(self, position, value):
        """PiecewiseLinearFunction::AddNode (Source/plf.h:50-66)."""
        if self.count + 1 >= MAX_NODES:
            return self
        self.position[self.count] = position
        self.value[self.count] = value
        if position < self.node_range[0]:
            self.node_range[0] = position
        if position > self.node_range[1]:
            self.node_range[1] = position
        self.count += 1
        return self


class ColorTf(C.Structure):
    _fields_ = [("plf", Plf * 3)]

    def init(self):
        for i in range(3):
            self.plf[i].init()
        return self

    def add(self, position, rgb):
        for i in range(3):
            self.plf[i].add(position, rgb[i])
        return self


class Camera(C.Structure):
    _fields_ = [("film_size", I2), ("pos", F3), ("target", F3), ("up", F3),
                ("focal_distance", C.c_float), ("aperture_size", C.c_float),
                ("clip_near", C.c_float), ("clip_far", C.c_float), ("exposure", C.c_float),
                ("gamma", C.c_float), ("fov", C.c_float)]


class RenderSettings(C.Structure):
    _fields_ = [("step_factor_primary", C.c_float), ("step_factor_shadow", C.c_float),
                ("shadows", C.c_int), ("max_shadow_distance", C.c_float),
                ("shading_type", C.c_int), ("density_scale", C.c_float),
                ("opacity_modulated", C.c_int), ("gradient_computation", C.c_int),
                ("gradient_threshold", C.c_float), ("gradient_factor", C.c_float),
                ("estimator", C.c_int), ("sampler", C.c_int), ("bilateral", C.c_int),
                ("bilateral_radius", C.c_int), ("bilateral_sigma_spatial", C.c_float),
                ("bilateral_sigma_range", C.c_float)]


class TracerDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("opacity", Plf), ("diffuse", ColorTf),
                ("specular", ColorTf), ("glossiness", Plf), ("emission", ColorTf),
                ("camera", Camera), ("settings", RenderSettings), ("no_iterations", C.c_int),
                ("volume_id", C.c_int), ("light_ids", C.c_int * MAX_INDICES),
                ("light_count", C.c_int), ("object_ids", C.c_int * MAX_INDICES),
                ("object_count", C.c_int), ("clipping_ids", C.c_int * MAX_INDICES),
                ("clipping_count", C.c_int)]


class Shape(C.Structure):
    _fields_ = [("tm", M44), ("inv_tm", M44), ("one_sided", C.c_int), ("type", C.c_int),
                ("size", F3), ("area", C.c_float), ("inner_radius", C.c_float),
                ("outer_radius", C.c_float)]


class VolumeDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("resolution", I3), ("spacing", F3),
                ("normalize_size", C.c_int), ("voxels_dirty", C.c_int),
                ("voxels", C.POINTER(C.c_uint16))]


class LightDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("visible", C.c_int),
                ("texture_id", C.c_int), ("multiplier", C.c_float), ("unit", C.c_int),
                ("shape", Shape)]


class Procedural(C.Structure):
    _fields_ = [("type", C.c_int), ("uniform_color", F3), ("checker_color1", F3),
                ("checker_color2", F3), ("gradient", ColorTf)]


class TextureDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("type", C.c_int),
                ("output_level", C.c_float), ("bitmap_id", C.c_int), ("offset", F2),
                ("repeat", F2), ("flip", I2), ("procedural", Procedural)]


class ClippingDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("invert", C.c_int), ("shape", Shape)]


class ObjectDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("diffuse_texture_id", C.c_int),
                ("specular_texture_id", C.c_int), ("glossiness_texture_id", C.c_int),
                ("ior", C.c_float), ("shape", Shape)]


class BitmapDesc(C.Structure):
    _fields_ = [("id", C.c_int), ("enabled", C.c_int), ("resolution", I2),
                ("pixels_dirty", C.c_int), ("pixels", C.POINTER(C.c_uint8))]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("trace_ms", C.c_float), ("post_ms", C.c_float),
                ("passes", C.c_int), ("primary_ms", C.c_float)]


class Counters(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("primary_samples", C.c_uint64),
                ("shadow_samples", C.c_uint64), ("shading_samples", C.c_uint64),
                ("primary_fetches", C.c_uint64), ("shadow_fetches", C.c_uint64),
                ("scatter_events", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("march_iterations", C.c_uint64), ("skipped_samples", C.c_uint64),
                ("skips_undone", C.c_uint64), ("reserved", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}



PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libexposurerender_amd.so")
TF_LUT_NODES = 256   # er_capi.h ER_TF_LUT_NODES

# Every extern "C" symbol include/er_capi.h declares, with its ctypes signature.
_P = C.POINTER
class BatchItem(C.Structure):
    _fields_ = [("tracer_id", C.c_int), ("passes", C.c_int)]


class Tune(C.Structure):
    _fields_ = [("brick", C.c_int), ("tile_w", C.c_int), ("tile_h", C.c_int), ("blocks_per_cu", C.c_int),
                ("shadow_waves_per_cu", C.c_int), ("heavy_tiles", C.c_int), ("rays_per_wave", C.c_int), ("shadow_rays_per_wave", C.c_int), ("heavy_rays_per_wave", C.c_int), ("ms_per_pass", C.c_float), ("default_ms_per_pass", C.c_float), ("trials", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TUNE_FORCE = 1

ABI_STRUCTS = [TracerDesc, VolumeDesc, LightDesc, TextureDesc, ClippingDesc, ObjectDesc,
               BitmapDesc, Plf, Shape, Timing, Counters, BatchItem, Tune]

(ELEMENT_UINT8, ELEMENT_INT8, ELEMENT_UINT16, ELEMENT_INT16, ELEMENT_UINT32, ELEMENT_INT32, ELEMENT_FLOAT32,
 ELEMENT_FLOAT64) = range(8)


SIGNATURES = {
    "er_abi_version": (C.c_int, []),
    "er_abi_sizeof": (C.c_int, [_P(C.c_int), C.c_int]),
    "er_plf_init": (None, [_P(Plf)]),
    "er_plf_add_node": (None, [_P(Plf), C.c_float, C.c_float]),
    "er_tracer_desc_init": (None, [_P(TracerDesc)]),
    "er_volume_desc_init": (None, [_P(VolumeDesc)]),
    "er_light_desc_init": (None, [_P(LightDesc)]),
    "er_texture_desc_init": (None, [_P(TextureDesc)]),
    "er_clipping_desc_init": (None, [_P(ClippingDesc)]),
    "er_object_desc_init": (None, [_P(ObjectDesc)]),
    "er_bitmap_desc_init": (None, [_P(BitmapDesc)]),
    "er_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "er_bind_tracer": (C.c_int, [_P(TracerDesc), C.c_int]),
    "er_bind_volume": (C.c_int, [_P(VolumeDesc), C.c_int]),
    "er_bind_light": (C.c_int, [_P(LightDesc), C.c_int]),
    "er_bind_object": (C.c_int, [_P(ObjectDesc), C.c_int]),
    "er_bind_clipping_object": (C.c_int, [_P(ClippingDesc), C.c_int]),
    "er_bind_texture": (C.c_int, [_P(TextureDesc), C.c_int]),
    "er_bind_bitmap": (C.c_int, [_P(BitmapDesc), C.c_int]),
    "er_render_estimate": (C.c_int, [C.c_int]),
    "er_get_estimate": (C.c_int, [C.c_int, _P(C.c_uint8), C.c_size_t]),
    "er_get_autofocus_distance": (C.c_int, [C.c_int, C.c_int, C.c_int, _P(C.c_float)]),
    "er_get_no_iterations": (C.c_int, [C.c_int, _P(C.c_int)]),
    "er_init": (C.c_int, []),
    "er_shutdown": (C.c_int, []),
    "er_synchronize": (C.c_int, []),
    "er_get_seeds": (C.c_int, [C.c_int, _P(C.c_uint32), _P(C.c_uint32), C.c_size_t]),
    "er_set_seeds": (C.c_int, [C.c_int, _P(C.c_uint32), _P(C.c_uint32), C.c_size_t]),
    "er_seed_tracer": (C.c_int, [C.c_int, C.c_uint64]),
    "er_get_frame_xyza": (C.c_int, [C.c_int, _P(C.c_float), C.c_size_t]),
    "er_get_running_xyza": (C.c_int, [C.c_int, _P(C.c_float), C.c_size_t]),
    "er_render_passes": (C.c_int, [C.c_int, C.c_int, _P(Timing)]),
    "er_set_tf_lut": (C.c_int, [C.c_int, _P(C.c_float), C.c_float]),
    "er_render_batch": (C.c_int, [_P(BatchItem), C.c_int, _P(Timing)]),
    "er_set_list_capacity": (C.c_int, [C.c_int]),
    "er_autotune": (C.c_int, [C.c_int, C.c_int, C.c_int, _P(Tune)]),
    "er_set_tune": (C.c_int, [C.c_int, _P(Tune)]),
    "er_autotune_batch": (C.c_int, [_P(BatchItem), C.c_int, C.c_int, _P(Tune)]),
    "er_mhd_info": (C.c_int, [C.c_char_p, _P(C.c_int), _P(C.c_float), _P(C.c_int)]),
    "er_mhd_read": (C.c_int, [C.c_char_p, _P(C.c_uint16), C.c_size_t, _P(C.c_float), _P(C.c_float)]),
    "er_raw_read": (C.c_int, [C.c_char_p, _P(C.c_int), C.c_int, C.c_int, C.c_longlong, _P(C.c_uint16), C.c_size_t,
                              _P(C.c_float), _P(C.c_float)]),
    "er_get_tune": (C.c_int, [C.c_int, _P(Tune)]),
    "er_autofocus_seeds": (C.c_int, [C.c_int, C.c_int, C.c_int, _P(C.c_uint32), _P(C.c_uint32)]),
    "er_adjoint_lut": (C.c_int, [C.c_int, _P(C.c_uint32), _P(C.c_uint32), C.c_size_t, C.c_int, _P(C.c_float),
                                 _P(C.c_double)]),
    "er_adjoint_lut_noalloc": (C.c_int, [C.c_int, _P(C.c_uint32), _P(C.c_uint32), C.c_size_t, C.c_int,
                                         _P(C.c_float), _P(C.c_double)]),
    "er_adjoint_value": (C.c_int, [C.c_int, _P(C.c_uint32), _P(C.c_uint32), C.c_size_t, C.c_int, _P(C.c_float),
                                   _P(C.c_float)]),
    "er_set_counting": (C.c_int, [C.c_int, C.c_int]),
    "er_get_counters": (C.c_int, [C.c_int, _P(Counters)]),
    "er_reset_counters": (C.c_int, [C.c_int]),
    "er_get_gradient_magnitude_max": (C.c_int, [C.c_int, _P(C.c_float)]),
    "er_neglog_table_entry": (C.c_int, [C.c_uint32, _P(C.c_float)]),
    "er_phantom_blobs": (C.c_int, [_P(C.c_uint16), C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_uint64]),
    "er_phantom_ellipsoids": (C.c_int, [_P(C.c_uint16), C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                        C.c_float, C.c_uint64]),
}




_lib = None


def load(path=None):
    """Load the native library (built by __graft_entry__.build()). Raises if it is missing:
    there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"native library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check_abi(lib=None):
    lib = lib or load()
    sizes = (C.c_int * len(ABI_STRUCTS))()
    n = lib.er_abi_sizeof(sizes, len(ABI_STRUCTS))
    mine = [C.sizeof(s) for s in ABI_STRUCTS]
    theirs = list(sizes)[:n]
    if mine != theirs:
        raise RuntimeError(f"ABI mismatch: ctypes {mine} vs library {theirs}")
    return theirs
