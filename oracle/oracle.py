"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (oracle/_build/liborc.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use this module, and only
as the checker. The oracle restates the reference estimator (see er_oracle.cpp).
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_ROOT, "exposure-render_amd"))
from exposure_render import capi  # noqa: E402

LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None
_P = C.POINTER


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH}")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_create": (C.c_void_p, [_P(capi.TracerDesc), _P(capi.VolumeDesc), _P(capi.LightDesc), C.c_int,
                                    _P(capi.TextureDesc), C.c_int, _P(capi.ClippingDesc), C.c_int]),
        "orc_destroy": (None, [C.c_void_p]),
        "orc_set_objects": (None, [C.c_void_p, _P(capi.ObjectDesc), C.c_int]),
        "orc_set_bitmap": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, _P(C.c_uint8)]),
        "orc_eval_texture": (None, [C.c_void_p, C.c_int, C.c_float, C.c_float, _P(C.c_float)]),
        "orc_set_tf_lut": (None, [C.c_void_p, _P(C.c_float), C.c_float]),
        "orc_tf_lut_eval": (C.c_float, [_P(C.c_float), C.c_float, C.c_float, C.c_int]),
        "orc_autofocus": (C.c_float, [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32]),
        "orc_set_gm_max": (None, [C.c_void_p, C.c_float]),
        "orc_get_gm_max": (C.c_float, [C.c_void_p]),
        "orc_compute_gm_max": (C.c_float, [C.c_void_p, C.c_int]),
        "orc_trace_f32": (C.c_int, [C.c_void_p, _P(C.c_uint32), _P(C.c_uint32), _P(C.c_float), C.c_int, _P(C.c_uint64)]),
        "orc_trace_pixels_f32": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), C.c_int, _P(C.c_uint32),
                                           _P(C.c_uint32), _P(C.c_float), _P(C.c_uint32), _P(C.c_uint64)]),
        "orc_expect_f64": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), C.c_int, C.c_int, _P(C.c_uint32),
                                     _P(C.c_uint32), _P(C.c_double), C.c_int]),
        "orc_post": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_uint8), C.c_int, _P(C.c_float)]),
        "orc_bilateral": (C.c_int, [C.c_void_p, _P(C.c_uint8), _P(C.c_uint8)]),
        "orc_intensity": (C.c_float, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "orc_plf_eval": (C.c_float, [_P(capi.Plf), C.c_float]),
        "orc_plf_eval_f64": (C.c_double, [_P(capi.Plf), C.c_double]),
        "orc_rng_get1": (C.c_float, [_P(C.c_uint32), _P(C.c_uint32)]),
        "orc_neglog": (C.c_float, [C.c_float]),
        "orc_camera": (C.c_int, [C.c_void_p, _P(C.c_float)]),
        "orc_march_ray": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_float, C.c_float, C.c_int,
                                    C.c_float, C.c_float, _P(C.c_float), _P(C.c_float), _P(C.c_uint64)]),
        "orc_gradient": (None, [C.c_void_p, C.c_int, _P(C.c_float), _P(C.c_float), _P(C.c_float)]),
        "orc_gradient_magnitude": (C.c_float, [C.c_void_p, _P(C.c_float)]),
        "orc_fresnel": (C.c_float, [C.c_float, C.c_float]),
        "orc_blinn_d": (C.c_float, [C.c_float, C.c_float]),
        "orc_blinn_pdf": (C.c_float, [C.c_float, _P(C.c_float), _P(C.c_float)]),
        "orc_microfacet_f": (None, [_P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float, C.c_float,
                                    _P(C.c_float)]),
        "orc_build_info": (C.c_char_p, []),
        "orc_abi_sizeof": (C.c_int, [_P(C.c_int), C.c_int]),
        "orc_fmath": (None, [C.c_int, _P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(_P(t))


def check_abi():
    lib = load()
    sizes = (C.c_int * len(capi.ABI_STRUCTS))()
    n = lib.orc_abi_sizeof(sizes, len(capi.ABI_STRUCTS))
    mine = [C.sizeof(s) for s in capi.ABI_STRUCTS]
    if list(sizes)[:n] != mine:
        raise RuntimeError(f"oracle ABI mismatch {list(sizes)[:n]} vs {mine}")


def seeds_for(width, height, seed):
    """The per-pixel seeds the runtime's er_seed_tracer writes: SplitMix64(seed ^ pixel),
    low and high halves, zero halves replaced by 1 (zero is a fixed point of the MWC)."""
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


class Oracle:
    def __init__(self, scene, gm_max=None, nthreads=0):
        self.lib = load()
        self.scene = scene
        lights, nl = scene.array("lights")
        tex, nt = scene.array("textures")
        clips, nc = scene.array("clips")
        self._keep = (lights, tex, clips)
        self.h = self.lib.orc_create(C.byref(scene.tracer), C.byref(scene.volume), lights, nl, tex, nt, clips, nc)
        objs, no = scene.array("objects")
        self.lib.orc_set_objects(self.h, objs, no)
        for i, a in enumerate(scene.bitmaps):
            self.lib.orc_set_bitmap(self.h, i, a.shape[1], a.shape[0], _p(a, C.c_uint8))
        self.nthreads = nthreads
        self.W, self.H = scene.film
        if gm_max is not None:
            self.lib.orc_set_gm_max(self.h, gm_max)

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def set_tf_lut(self, lut, range_):
        if lut is None:
            self.lib.orc_set_tf_lut(self.h, None, 0.0)
            return
        self._lut = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
        self.lib.orc_set_tf_lut(self.h, _p(self._lut, C.c_float), float(range_))

    def autofocus(self, u, v, s0, s1):
        return self.lib.orc_autofocus(self.h, int(u), int(v), int(s0), int(s1))

    def compute_gm_max(self):
        return self.lib.orc_compute_gm_max(self.h, self.nthreads)

    def set_gm_max(self, v):
        self.lib.orc_set_gm_max(self.h, v)

    def trace(self, s1, s2, counters=None):
        """One FP32 pass over the frame; seeds (uint32 arrays of W*H) are advanced in place."""
        frame = np.zeros((self.H, self.W, 4), dtype=np.float32)
        cnt = np.zeros(8, dtype=np.uint64) if counters is None else counters
        self.lib.orc_trace_f32(self.h, _p(s1, C.c_uint32), _p(s2, C.c_uint32), _p(frame, C.c_float),
                               self.nthreads, _p(cnt, C.c_uint64))
        return frame, cnt

    def trace_pixels(self, px, py, s1, s2):
        px = np.ascontiguousarray(px, dtype=np.int32)
        py = np.ascontiguousarray(py, dtype=np.int32)
        n = len(px)
        out = np.zeros((n, 4), dtype=np.float32)
        draws = np.zeros(n, dtype=np.uint32)
        cnt = np.zeros(8, dtype=np.uint64)
        self.lib.orc_trace_pixels_f32(self.h, _p(px, C.c_int), _p(py, C.c_int), n, _p(s1, C.c_uint32),
                                      _p(s2, C.c_uint32), _p(out, C.c_float), _p(draws, C.c_uint32),
                                      _p(cnt, C.c_uint64))
        return out, draws, cnt

    def expect_f64(self, px, py, spp, seed=12345):
        px = np.ascontiguousarray(px, dtype=np.int32)
        py = np.ascontiguousarray(py, dtype=np.int32)
        n = len(px)
        s1, s2 = seeds_for(n, 1, seed)
        out = np.zeros((n, 4), dtype=np.float64)
        self.lib.orc_expect_f64(self.h, _p(px, C.c_int), _p(py, C.c_int), n, spp, _p(s1, C.c_uint32),
                                _p(s2, C.c_uint32), _p(out, C.c_double), self.nthreads)
        return out

    def eval_texture(self, tex, u, v):
        out = np.zeros(3, dtype=np.float32)
        self.lib.orc_eval_texture(self.h, int(tex), float(u), float(v), _p(out, C.c_float))
        return out

    def post(self, frame, running, no_iterations):
        display = np.zeros((self.H, self.W, 4), dtype=np.uint8)
        filtered = np.zeros((self.H, self.W, 4), dtype=np.float32)
        self.lib.orc_post(self.h, _p(frame, C.c_float), _p(running, C.c_float), _p(display, C.c_uint8),
                          no_iterations, _p(filtered, C.c_float))
        return display, filtered

    def bilateral(self, display):
        out = np.zeros_like(display)
        self.lib.orc_bilateral(self.h, _p(np.ascontiguousarray(display), C.c_uint8), _p(out, C.c_uint8))
        return out

    def render(self, passes, seed=None, s1=None, s2=None):
        """`passes` progressive passes (RenderEstimate semantics). Returns (display, running,
        last frame, seeds1, seeds2, counters)."""
        if s1 is None:
            s1, s2 = seeds_for(self.W, self.H, seed if seed is not None else 1)
        s1 = s1.copy()
        s2 = s2.copy()
        running = np.zeros((self.H, self.W, 4), dtype=np.float32)
        cnt = np.zeros(8, dtype=np.uint64)
        display = None
        frame = None
        for k in range(passes):
            frame, _ = self.trace(s1, s2, cnt)
            display, _ = self.post(frame, running, k)
        if self.scene.tracer.settings.bilateral and self.scene.tracer.settings.estimator == capi.ESTIMATOR_FULL:
            display = self.bilateral(display)
        return display, running, frame, s1, s2, cnt

    def gradient(self, mode, p):
        """(gradient, normalised gradient) at p for gradient mode `mode` (ER_GRADIENT_*)."""
        p = np.asarray(p, dtype=np.float32)
        g = np.zeros(3, dtype=np.float32)
        n = np.zeros(3, dtype=np.float32)
        self.lib.orc_gradient(self.h, int(mode), _p(p, C.c_float), _p(g, C.c_float), _p(n, C.c_float))
        return g, n

    def gradient_magnitude(self, p):
        p = np.asarray(p, dtype=np.float32)
        return self.lib.orc_gradient_magnitude(self.h, _p(p, C.c_float))

    def intensity(self, p):
        return self.lib.orc_intensity(self.h, float(p[0]), float(p[1]), float(p[2]))

    def camera(self):
        out = np.zeros(13, dtype=np.float32)
        self.lib.orc_camera(self.h, _p(out, C.c_float))
        return out

    def march_ray(self, o, d, mint, maxt, primary, u1, u2):
        o = np.asarray(o, dtype=np.float32)
        d = np.asarray(d, dtype=np.float32)
        T = C.c_float(0)
        P = np.zeros(3, dtype=np.float32)
        steps = C.c_uint64(0)
        hit = self.lib.orc_march_ray(self.h, _p(o, C.c_float), _p(d, C.c_float), mint, maxt, int(primary),
                                     u1, u2, C.byref(T), _p(P, C.c_float), C.byref(steps))
        return hit, T.value, P, steps.value


def tonemap_f64(xyza, exposure):
    """Tonemap (tonemap.cuh:30-47) in float64 before u8 quantisation, values in [0, 1]."""
    x, y, z = xyza[..., 0], xyza[..., 1], xyza[..., 2]
    rgb = np.stack([3.240479 * x - 1.537150 * y - 0.498535 * z,
                    -0.969256 * x + 1.875991 * y + 0.041556 * z,
                    0.055648 * x - 0.204043 * y + 1.057311 * z], axis=-1)
    return np.clip(1.0 - np.exp(-(rgb / exposure)), 0.0, 1.0)


def gauss3x3_f64(img):
    """The per-pass 3x3 Gaussian (sigma 1) with border renormalisation and alpha / total
    weight (filterframeestimate.cuh:28-64), in float64 over an H x W x 4 image."""
    H, W = img.shape[:2]
    out = np.zeros_like(img, dtype=np.float64)
    tw = np.zeros((H, W), dtype=np.float64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            w = np.exp(-(dx * dx + dy * dy) / 2.0)
            ys = slice(max(0, dy), H + min(0, dy))
            yd = slice(max(0, -dy), H + min(0, -dy))
            xs = slice(max(0, dx), W + min(0, dx))
            xd = slice(max(0, -dx), W + min(0, -dx))
            out[yd, xd, :] += w * img[ys, xs, :]
            tw[yd, xd] += w
    out[..., 3] = img[..., 3]
    return out / tw[..., None]


def tf_lut_eval(lut, range_, intensity, ch):
    lut = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
    return load().orc_tf_lut_eval(_p(lut, C.c_float), float(range_), float(intensity), int(ch))


def fmath(fn, x, y=None):
    """Evaluate the project's float elementary functions (0 exp, 1 log, 2 pow, 3 sin, 4 cos,
    5 atan2, 6 acos, 7 expm1) over float32 arrays."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float32)
    out = np.zeros_like(x)
    lib.orc_fmath(fn, _p(x, C.c_float), _p(y, C.c_float), _p(out, C.c_float), x.size)
    return out


def fresnel(cosi, ior):
    return load().orc_fresnel(float(cosi), float(ior))


def blinn_d(exponent, whz):
    return load().orc_blinn_d(float(exponent), float(whz))


def blinn_pdf(exponent, wo, wi):
    wo = np.asarray(wo, dtype=np.float32)
    wi = np.asarray(wi, dtype=np.float32)
    return load().orc_blinn_pdf(float(exponent), _p(wo, C.c_float), _p(wi, C.c_float))


def microfacet_f(wo, wi, ks, exponent, ior=15.0):
    wo = np.asarray(wo, dtype=np.float32)
    wi = np.asarray(wi, dtype=np.float32)
    ks = np.asarray(ks, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    load().orc_microfacet_f(_p(wo, C.c_float), _p(wi, C.c_float), _p(ks, C.c_float), float(exponent), float(ior),
                            _p(out, C.c_float))
    return out


def build_info():
    """Compiler and flags the oracle library was built with."""
    return load().orc_build_info().decode()
