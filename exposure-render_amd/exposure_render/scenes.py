"""Scene descriptors and the synthetic benchmark scenes.

`Scene` bundles the descriptors one tracer needs (tracer, volume, lights, textures, clipping
objects) as the C-ABI structs, plus the voxel array that backs the volume descriptor. The
`config_*` builders produce the five benchmark workloads (plain small, CT-like, large
anisotropic, differentiation, batching) from seeded phantoms; `scale` shrinks image and
volume for quick parity tests while keeping the scene content.
"""
import ctypes as C
import math

import numpy as np

from . import capi


# ---- transforms --------------------------------------------------------------------------

def _rigid(R, t):
    tm = np.eye(4)
    tm[:3, :3] = R
    tm[:3, 3] = t
    inv = np.eye(4)
    inv[:3, :3] = R.T
    inv[:3, 3] = -R.T @ np.asarray(t, dtype=np.float64)
    return tm, inv


def look_at(position, target, up=(0.0, 1.0, 0.0)):
    """TM / InvTM of a shape at `position` whose local +z axis points at `target`."""
    p = np.asarray(position, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - p
    z /= np.linalg.norm(z)
    u = np.asarray(up, dtype=np.float64)
    if abs(float(np.dot(u, z))) > 0.99:
        u = np.array([1.0, 0.0, 0.0])
    x = np.cross(u, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return _rigid(np.stack([x, y, z], axis=1), p)


def translate(t):
    return _rigid(np.eye(3), t)


def set_matrix(dst, m):
    for i in range(4):
        for j in range(4):
            dst[i][j] = float(m[i, j])


# ---- descriptor defaults (the reference constructors) ------------------------------------

def default_settings(s):
    """RenderSettings() defaults (Source/rendersettings.h:33-38,69-76)."""
    s.step_factor_primary = 0.1
    s.step_factor_shadow = 0.1
    s.shadows = 1
    s.max_shadow_distance = 1.0
    s.shading_type = capi.SHADING_BRDF_ONLY
    s.density_scale = 100.0
    s.opacity_modulated = 1
    s.gradient_computation = capi.GRADIENT_CENTRAL
    s.gradient_threshold = 0.5
    s.gradient_factor = 0.5
    s.estimator = capi.ESTIMATOR_COMPAT
    s.sampler = capi.SAMPLER_RAYMARCH
    s.bilateral = 0
    s.bilateral_radius = 3
    s.bilateral_sigma_spatial = 2.0
    s.bilateral_sigma_range = 25.0


def new_tracer():
    t = capi.TracerDesc()
    t.id = -1
    t.enabled = 1
    t.opacity.init()
    t.diffuse.init()
    t.specular.init()
    t.glossiness.init()
    t.emission.init()
    default_settings(t.settings)
    t.no_iterations = 0
    t.volume_id = 0
    for i in range(capi.MAX_INDICES):
        t.light_ids[i] = -1
        t.object_ids[i] = -1
        t.clipping_ids[i] = -1
    return t


def set_camera(t, film, pos, target, up=(0.0, 1.0, 0.0), fov=35.0, exposure=1.0,
               clip_near=0.0, clip_far=1000.0, aperture=0.0, focal=-1.0):
    c = t.camera
    c.film_size[0], c.film_size[1] = int(film[0]), int(film[1])
    for i in range(3):
        c.pos[i], c.target[i], c.up[i] = pos[i], target[i], up[i]
    c.focal_distance = focal
    c.aperture_size = aperture
    c.clip_near = clip_near
    c.clip_far = clip_far
    c.exposure = exposure
    c.gamma = 2.2
    c.fov = fov


def plane_light(pos, target, size, multiplier, texture_id=0, one_sided=0, visible=1, unit=capi.UNIT_POWER):
    L = capi.LightDesc()
    L.id = -1
    L.enabled = 1
    L.visible = visible
    L.texture_id = texture_id
    L.multiplier = multiplier
    L.unit = unit
    tm, inv = look_at(pos, target)
    set_matrix(L.shape.tm, tm)
    set_matrix(L.shape.inv_tm, inv)
    L.shape.one_sided = one_sided
    L.shape.type = capi.SHAPE_PLANE
    L.shape.size[0], L.shape.size[1], L.shape.size[2] = size[0], size[1], 0.0
    return L


def sphere_light(radius, multiplier, texture_id=0, center=(0.0, 0.0, 0.0), visible=1):
    """A uniform environment: a large two-sided sphere light around the scene."""
    L = capi.LightDesc()
    L.id = -1
    L.enabled = 1
    L.visible = visible
    L.texture_id = texture_id
    L.multiplier = multiplier
    L.unit = capi.UNIT_POWER
    tm, inv = translate(center)
    set_matrix(L.shape.tm, tm)
    set_matrix(L.shape.inv_tm, inv)
    L.shape.one_sided = 0
    L.shape.type = capi.SHAPE_SPHERE
    L.shape.outer_radius = radius
    return L


def shape_light(shape_type, pos, target, multiplier, texture_id=0, size=(0.0, 0.0, 0.0), inner=0.0, outer=0.0,
                one_sided=0, visible=1, unit=capi.UNIT_POWER):
    """A disk, ring or box light at `pos` facing `target` (Source/disk.h, ring.h, box.h)."""
    L = capi.LightDesc()
    L.id = -1
    L.enabled = 1
    L.visible = visible
    L.texture_id = texture_id
    L.multiplier = multiplier
    L.unit = unit
    tm, inv = look_at(pos, target)
    set_matrix(L.shape.tm, tm)
    set_matrix(L.shape.inv_tm, inv)
    L.shape.one_sided = one_sided
    L.shape.type = shape_type
    for i in range(3):
        L.shape.size[i] = size[i]
    L.shape.inner_radius = inner
    L.shape.outer_radius = outer
    return L


def object_desc(shape_type, tm, inv, size=(1.0, 1.0, 1.0), diffuse_tex=-1, specular_tex=-1,
                glossiness_tex=-1, inner_radius=0.0, outer_radius=0.5, one_sided=0):
    """A surface object (Source/erobject.h): a shape placed by tm / inv, shaded by the BRDF with
    diffuse, specular and glossiness read from textures at the hit's UV (transport.h:141-149)."""
    O = capi.ObjectDesc()
    capi.load().er_object_desc_init(C.byref(O))
    O.diffuse_texture_id, O.specular_texture_id, O.glossiness_texture_id = diffuse_tex, specular_tex, glossiness_tex
    O.shape.type = shape_type
    O.shape.one_sided = one_sided
    for i in range(3):
        O.shape.size[i] = size[i]
    O.shape.inner_radius = inner_radius
    O.shape.outer_radius = outer_radius
    set_matrix(O.shape.tm, tm)
    set_matrix(O.shape.inv_tm, inv)
    return O


def bitmap_texture(bitmap_index, repeat=(1.0, 1.0), offset=(0.0, 0.0), flip=(0, 0), output_level=1.0):
    """A bitmap texture (textures.h:99-105) over the scene's bitmap at bitmap_index."""
    T = capi.TextureDesc()
    capi.load().er_texture_desc_init(C.byref(T))
    T.type = capi.TEXTURE_BITMAP
    T.bitmap_id = bitmap_index
    T.output_level = output_level
    for i in range(2):
        T.repeat[i] = repeat[i]
        T.offset[i] = offset[i]
        T.flip[i] = flip[i]
    return T


def checker_texture(c1, c2, repeat=(4.0, 4.0), output_level=1.0):
    T = uniform_texture((1.0, 1.0, 1.0), output_level)
    T.procedural.type = capi.PROCEDURAL_CHECKER
    for i in range(3):
        T.procedural.checker_color1[i] = c1[i]
        T.procedural.checker_color2[i] = c2[i]
    T.repeat[0], T.repeat[1] = repeat
    return T


def gradient_texture(nodes, output_level=1.0):
    """A procedural gradient texture: nodes = [(position in [0, 1], (X, Y, Z)), ...]."""
    T = uniform_texture((1.0, 1.0, 1.0), output_level)
    T.procedural.type = capi.PROCEDURAL_GRADIENT
    for x, c in nodes:
        T.procedural.gradient.add(x, c)
    T.repeat[0], T.repeat[1] = 1.0, 1.0
    return T


def uniform_texture(rgb=(1.0, 1.0, 1.0), output_level=1.0):
    T = capi.TextureDesc()
    T.id = -1
    T.enabled = 1
    T.type = capi.TEXTURE_PROCEDURAL
    T.output_level = output_level
    T.bitmap_id = -1
    T.procedural.type = capi.PROCEDURAL_UNIFORM
    for i in range(3):
        T.procedural.uniform_color[i] = rgb[i]
    T.procedural.gradient.init()
    return T


def clipping_box(center, size, invert=0):
    return clipping_shape(capi.SHAPE_BOX, *translate(center), size=size, invert=invert)


def clipping_shape(shape_type, tm, inv, size=(1.0, 1.0, 1.0), outer=0.0, invert=0):
    """A clipping object (ErClippingObject): samples inside the shape (outside, if inverted)
    have zero opacity in the full estimator."""
    Cb = capi.ClippingDesc()
    Cb.id = -1
    Cb.enabled = 1
    Cb.invert = invert
    set_matrix(Cb.shape.tm, tm)
    set_matrix(Cb.shape.inv_tm, inv)
    Cb.shape.type = shape_type
    for i in range(3):
        Cb.shape.size[i] = size[i]
    Cb.shape.outer_radius = outer
    return Cb


class Scene:
    """One tracer with its volume, lights, textures and clipping objects. IDs follow the
    reference's registry: each kind counts from 0 in bind order, so list index == ID."""

    def __init__(self, voxels, spacing=(1.0, 1.0, 1.0)):
        self.voxels = np.ascontiguousarray(voxels, dtype=np.uint16)   # [z, y, x]
        self.volume = capi.VolumeDesc()
        self.volume.id = -1
        self.volume.enabled = 1
        rz, ry, rx = self.voxels.shape
        self.volume.resolution[0], self.volume.resolution[1], self.volume.resolution[2] = rx, ry, rz
        for i in range(3):
            self.volume.spacing[i] = spacing[i]
        self.volume.normalize_size = 1
        self.volume.voxels_dirty = 1
        self.volume.voxels = self.voxels.ctypes.data_as(C.POINTER(C.c_uint16))
        self.tracer = new_tracer()
        self.lights = []
        self.textures = []
        self.clips = []
        self.objects = []
        self.bitmaps = []       # (H, W, 4) uint8 arrays, RGBA
        self.name = "scene"

    @property
    def film(self):
        return self.tracer.camera.film_size[0], self.tracer.camera.film_size[1]

    def add_texture(self, T):
        self.textures.append(T)
        return len(self.textures) - 1

    def add_light(self, L):
        self.lights.append(L)
        i = len(self.lights) - 1
        self.tracer.light_ids[self.tracer.light_count] = i
        self.tracer.light_count += 1
        return i

    def add_clip(self, Cb):
        self.clips.append(Cb)
        i = len(self.clips) - 1
        self.tracer.clipping_ids[self.tracer.clipping_count] = i
        self.tracer.clipping_count += 1
        return i

    def add_object(self, O):
        self.objects.append(O)
        i = len(self.objects) - 1
        self.tracer.object_ids[self.tracer.object_count] = i
        self.tracer.object_count += 1
        return i

    def add_bitmap(self, rgba):
        """An (H, W, 4) uint8 RGBA image; returns the index a bitmap texture names."""
        a = np.ascontiguousarray(rgba, dtype=np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("bitmap must be (H, W, 4) uint8")
        self.bitmaps.append(a)
        return len(self.bitmaps) - 1

    def array(self, kind):
        lst = {"lights": self.lights, "textures": self.textures, "clips": self.clips, "objects": self.objects}[kind]
        typ = {"lights": capi.LightDesc, "textures": capi.TextureDesc, "clips": capi.ClippingDesc,
               "objects": capi.ObjectDesc}[kind]
        arr = (typ * max(1, len(lst)))()
        for i, d in enumerate(lst):
            arr[i] = d
        return arr, len(lst)


# ---- phantoms (native generators, seeded) -------------------------------------------------

def phantom_blobs(shape, n_blobs, sigma=(0.08, 0.2), peak=(1000.0, 4000.0), noise=50.0, seed=1):
    rz, ry, rx = shape
    out = np.empty((rz, ry, rx), dtype=np.uint16)
    lib = capi.load()
    st = lib.er_phantom_blobs(out.ctypes.data_as(C.POINTER(C.c_uint16)), rx, ry, rz, n_blobs,
                              sigma[0], sigma[1], peak[0], peak[1], noise, seed)
    if st != 0:
        raise RuntimeError("er_phantom_blobs failed")
    return out


def phantom_ellipsoids(shape, n_ellipsoids, n_tubes=0, axes=(0.05, 0.45), n_bands=5,
                       max_intensity=4095.0, noise_sigma=20.0, seed=2):
    rz, ry, rx = shape
    out = np.empty((rz, ry, rx), dtype=np.uint16)
    lib = capi.load()
    st = lib.er_phantom_ellipsoids(out.ctypes.data_as(C.POINTER(C.c_uint16)), rx, ry, rz,
                                   n_ellipsoids, n_tubes, axes[0], axes[1], n_bands,
                                   max_intensity, noise_sigma, seed)
    if st != 0:
        raise RuntimeError("er_phantom_ellipsoids failed")
    return out


# ---- the benchmark workloads ---------------------------------------------------------------

def _scaled(n, scale):
    return max(8, int(round(n * scale)))


def config_small(scale=1.0, estimator=capi.ESTIMATOR_FULL, seed=1):
    """Plain tracer, small: 64^3 blobs, 4-node TF, density scale 10, one 0.5 x 0.5 plane
    light of intensity 100, pinhole camera, 256 x 256, phase-only shading, 16 passes."""
    r = _scaled(64, scale)
    vox = phantom_blobs((r, r, r), 4, seed=seed)
    sc = Scene(vox)
    sc.name = "small_64c_256px"
    t = sc.tracer
    t.opacity.add(0.0, 0.0).add(400.0, 0.0).add(1500.0, 0.6).add(4000.0, 1.0)
    t.diffuse.add(0.0, (1.0, 1.0, 1.0)).add(400.0, (0.9, 0.55, 0.4)).add(1500.0, (0.85, 0.75, 0.6)).add(4000.0, (1.0, 1.0, 1.0))
    t.specular.add(0.0, (0.0, 0.0, 0.0)).add(4000.0, (0.0, 0.0, 0.0))
    t.glossiness.add(0.0, 0.5).add(4000.0, 0.5)
    t.settings.density_scale = 10.0
    t.settings.estimator = estimator
    t.settings.shading_type = capi.SHADING_PHASE_ONLY
    px = _scaled(256, scale)
    set_camera(t, (px, px), (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), exposure=4.0)
    tex = sc.add_texture(uniform_texture((1.0, 1.0, 1.0)))
    sc.add_light(plane_light((0.9, 1.0, 0.6), (0.0, 0.0, 0.0), (0.5, 0.5), 100.0, texture_id=tex))
    sc.passes = 16
    return sc


def _ct_tf(t):
    t.opacity.add(0.0, 0.0).add(300.0, 0.0).add(800.0, 0.02).add(1100.0, 0.1).add(1500.0, 0.3) \
        .add(2500.0, 0.5).add(3500.0, 0.9).add(4095.0, 1.0)
    t.diffuse.add(0.0, (0.8, 0.8, 0.8)).add(300.0, (0.9, 0.7, 0.6)).add(800.0, (0.85, 0.45, 0.35)) \
        .add(1100.0, (0.8, 0.3, 0.25)).add(1500.0, (0.9, 0.6, 0.5)).add(2500.0, (0.95, 0.9, 0.8)) \
        .add(3500.0, (1.0, 1.0, 0.95)).add(4095.0, (1.0, 1.0, 1.0))
    t.specular.add(0.0, (0.05, 0.05, 0.05)).add(1500.0, (0.2, 0.2, 0.2)).add(4095.0, (0.5, 0.5, 0.5))
    t.glossiness.add(0.0, 0.3).add(1500.0, 0.5).add(4095.0, 0.8)
    t.emission.add(0.0, (0.0, 0.0, 0.0)).add(4095.0, (0.0, 0.0, 0.0))


def config_ct(scale=1.0, estimator=capi.ESTIMATOR_FULL, seed=2, bilateral=True):
    """Plain tracer, CT-like phantom: 256^3, 12 nested ellipsoids in 5 bands + N(0, 20^2),
    1024 x 768, 2 plane lights + uniform environment sphere, hybrid shading, 8-node TF,
    64 passes, bilateral display filter, exponential tonemap."""
    r = _scaled(256, scale)
    vox = phantom_ellipsoids((r, r, r), 12, 0, (0.05, 0.45), 5, 4095.0, 20.0, seed)
    sc = Scene(vox)
    sc.name = "ct_256c_1024x768"
    t = sc.tracer
    _ct_tf(t)
    t.settings.density_scale = 100.0
    t.settings.estimator = estimator
    t.settings.shading_type = capi.SHADING_HYBRID
    t.settings.bilateral = 1 if bilateral else 0
    set_camera(t, (_scaled(1024, scale), _scaled(768, scale)), (0.9, 0.6, 1.9), (0.0, 0.0, 0.0), exposure=6.0)
    tex = sc.add_texture(uniform_texture((1.0, 1.0, 1.0)))
    sc.add_light(plane_light((1.0, 1.2, 0.8), (0.0, 0.0, 0.0), (0.6, 0.6), 80.0, texture_id=tex))
    sc.add_light(plane_light((-1.1, 0.4, 0.9), (0.0, 0.0, 0.0), (0.4, 0.4), 50.0, texture_id=tex))
    sc.add_light(sphere_light(10.0, 0.3, texture_id=tex))
    sc.passes = 64
    return sc


def config_large(scale=1.0, estimator=capi.ESTIMATOR_FULL, seed=3):
    """Plain tracer, large anisotropic volume: 512 x 512 x 460, 0.7 x 0.7 x 1.0 spacing,
    30 cylinders and ellipsoids + N(0, 30^2), clipping box removing 30% of the volume,
    1920 x 1080, 3 plane lights, hybrid shading, 256 passes."""
    rx, rz = _scaled(512, scale), _scaled(460, scale)
    vox = phantom_ellipsoids((rz, rx, rx), 15, 15, (0.05, 0.45), 5, 4095.0, 30.0, seed)
    sc = Scene(vox, spacing=(0.7, 0.7, 1.0))
    sc.name = "large_512x512x460_1080p"
    t = sc.tracer
    _ct_tf(t)
    t.settings.density_scale = 100.0
    t.settings.estimator = estimator
    t.settings.shading_type = capi.SHADING_HYBRID
    set_camera(t, (_scaled(1920, scale), _scaled(1080, scale)), (1.0, 0.7, 1.8), (0.0, 0.0, 0.0), exposure=6.0)
    tex = sc.add_texture(uniform_texture((1.0, 1.0, 1.0)))
    sc.add_light(plane_light((1.0, 1.2, 0.8), (0.0, 0.0, 0.0), (0.6, 0.6), 80.0, texture_id=tex))
    sc.add_light(plane_light((-1.1, 0.4, 0.9), (0.0, 0.0, 0.0), (0.4, 0.4), 50.0, texture_id=tex))
    sc.add_light(plane_light((0.2, -1.2, 0.9), (0.0, 0.0, 0.0), (0.5, 0.5), 40.0, texture_id=tex))
    # world extent of the volume after normalisation: 0.7*512 : 0.7*512 : 460 -> longest axis 1
    ext = np.array([0.7 * rx, 0.7 * rx, 1.0 * rz])
    ext = ext / ext.max()
    # remove 30% of the volume: a box over the +x side spanning 30% of the x extent
    cx = 0.5 * ext[0] - 0.15 * ext[0]
    sc.add_clip(clipping_box((cx, 0.0, 0.0), (0.3 * ext[0] + 1e-3, ext[1] + 0.01, ext[2] + 0.01)))
    sc.passes = 256
    return sc


# ---- differentiation workload ----------------------------------------------------------------

def tf_lut_from_nodes(positions, opacity, colors, range_=4095.0):
    """A node table (ER_TF_LUT_NODES x 4) sampling a piecewise-linear transfer function given
    by node positions (intensities), opacities and XYZ colours."""
    n = capi.TF_LUT_NODES
    x = np.arange(n, dtype=np.float64) * (range_ / (n - 1))
    lut = np.zeros((n, 4), dtype=np.float32)
    lut[:, 0] = np.interp(x, positions, opacity)
    for c in range(3):
        lut[:, 1 + c] = np.interp(x, positions, np.asarray(colors)[:, c])
    return lut


def random_tf_lut(seed, range_=4095.0):
    """A seeded random transfer function with 4-8 nodes: transparent below a random
    threshold in [200, 800], random opacities in [0.05, 1] and colours in [0.3, 1] above."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(4, 9))
    thr = float(rng.uniform(200.0, 800.0))
    pos = np.sort(rng.uniform(thr + 50.0, range_, size=k - 2))
    positions = np.concatenate([[0.0, thr], pos, [range_]])
    opacity = np.concatenate([[0.0, 0.0], rng.uniform(0.05, 1.0, size=k - 1)])
    colors = rng.uniform(0.3, 1.0, size=(k + 1, 3))
    return tf_lut_from_nodes(positions, opacity, colors, range_)


def initial_tf_lut(range_=4095.0):
    """The starting point of a transfer-function recovery: a grey ramp, transparent at 0."""
    return tf_lut_from_nodes([0.0, 300.0, range_], [0.0, 0.05, 0.5], [[0.7, 0.7, 0.7]] * 3, range_)


def config_adjoint(scale=1.0, seed=4):
    """Differentiation: 128^3 volume of 8 Gaussian blobs, 512 x 512, 32 passes, phase-function
    shading, full estimator, one plane light; the transfer function is a 256-node table over
    intensities [0, 4095] (er_set_tf_lut), set by the caller."""
    r = _scaled(128, scale)
    vox = phantom_blobs((r, r, r), 8, seed=seed)
    sc = Scene(vox)
    sc.name = "adjoint_128c_512px"
    t = sc.tracer
    t.opacity.add(0.0, 0.0).add(4095.0, 1.0)
    t.diffuse.add(0.0, (1.0, 1.0, 1.0)).add(4095.0, (1.0, 1.0, 1.0))
    t.specular.add(0.0, (0.0, 0.0, 0.0)).add(4095.0, (0.0, 0.0, 0.0))
    t.glossiness.add(0.0, 0.5).add(4095.0, 0.5)
    t.settings.density_scale = 10.0
    t.settings.estimator = capi.ESTIMATOR_FULL
    t.settings.shading_type = capi.SHADING_PHASE_ONLY
    px = _scaled(512, scale)
    set_camera(t, (px, px), (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), exposure=4.0)
    tex = sc.add_texture(uniform_texture((1.0, 1.0, 1.0)))
    sc.add_light(plane_light((0.9, 1.0, 0.6), (0.0, 0.0, 0.0), (0.5, 0.5), 100.0, texture_id=tex))
    sc.passes = 32
    sc.tf_range = 4095.0
    return sc


# ---- batching workload -----------------------------------------------------------------------

def _random_plf_tf(t, rng, range_=4095.0):
    """A seeded 4-8-node transfer function on the tracer's PLFs: transparent below a random
    threshold, random opacities and colours above."""
    k = int(rng.integers(4, 9))
    thr = float(rng.uniform(150.0, 700.0))
    pos = np.sort(rng.uniform(thr + 20.0, range_, size=k - 2))
    positions = np.concatenate([[0.0, thr], pos])
    op = np.concatenate([[0.0, 0.0], rng.uniform(0.05, 1.0, size=k - 2)])
    for x, o in zip(positions, op):
        t.opacity.add(float(x), float(o))
        c = rng.uniform(0.3, 1.0, size=3)
        t.diffuse.add(float(x), tuple(float(v) for v in c))
    t.specular.add(0.0, (0.1, 0.1, 0.1)).add(range_, (0.3, 0.3, 0.3))
    t.glossiness.add(0.0, 0.4).add(range_, 0.6)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def batch_job(seed, scale=1.0):
    """One job of the batching workload: a blob or ellipsoid phantom with an edge length from
    {32, 64, 128, 192, 256}, an image side uniform in [128, 1024], a 4-8-node transfer function,
    1-3 plane lights and 8-128 passes (sc.passes)."""
    rng = np.random.default_rng(seed)
    edge = _scaled(int(rng.choice([32, 64, 128, 192, 256])), scale)
    if rng.random() < 0.5:
        vox = phantom_blobs((edge, edge, edge), int(rng.integers(2, 9)), seed=int(rng.integers(1 << 30)))
    else:
        vox = phantom_ellipsoids((edge, edge, edge), int(rng.integers(4, 13)), 0, (0.05, 0.45), 5, 4095.0, 20.0,
                                 int(rng.integers(1 << 30)))
    sc = Scene(vox)
    sc.name = f"batch_job_{seed}"
    t = sc.tracer
    _random_plf_tf(t, rng)
    t.settings.density_scale = float(rng.uniform(10.0, 100.0))
    t.settings.estimator = capi.ESTIMATOR_FULL
    t.settings.shading_type = capi.SHADING_PHASE_ONLY if rng.random() < 0.5 else capi.SHADING_HYBRID
    side = _scaled(int(rng.integers(128, 1025)), scale)
    cam = 2.2 * _unit(rng.normal(size=3))
    set_camera(t, (side, side), tuple(cam), (0.0, 0.0, 0.0), exposure=float(rng.uniform(2.0, 8.0)))
    tex = sc.add_texture(uniform_texture((1.0, 1.0, 1.0)))
    for _ in range(int(rng.integers(1, 4))):
        pos = 1.5 * _unit(rng.normal(size=3))
        s = float(rng.uniform(0.3, 0.6))
        sc.add_light(plane_light(tuple(pos), (0.0, 0.0, 0.0), (s, s), float(rng.uniform(50.0, 150.0)), texture_id=tex))
    sc.passes = int(rng.integers(8, 129))
    return sc


def config_batch(n_jobs=64, scale=1.0, seed=5):
    """Batching: `n_jobs` independent jobs (batch_job) from one seed."""
    return [batch_job(seed * 1000 + k, scale) for k in range(n_jobs)]
