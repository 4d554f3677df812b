"""Per-sample parity of the gfx950 tracer with the CPU oracle (through the C-ABI).

With identical per-pixel seeds the RNG streams, the scatter / no-scatter and light-index
decisions (visible in the alpha channel, the sample counters and the RNG state after each pass)
must match bit for bit, and the FP32 radiance within 1e-4 relative per sample. Most cases are
in fact bit-identical; the tests report the exact-match fraction.
"""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import capi, scenes
import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-4      # FP32 radiance per sample, relative
ATOL = 1e-6

COUNTER_KEYS = ["paths", "primary_samples", "shadow_samples", "shading_samples", "scatter_events", "shadow_rays"]
ORC_INDEX = {"paths": 0, "primary_samples": 1, "shadow_samples": 2, "shading_samples": 3, "scatter_events": 6,
             "shadow_rays": 7}


def run_parity(sc, passes=3, seed=5, counting=True, check_display=True, lut=None, lut_range=4095.0):
    print(f"[parity] scene={sc.name} film={sc.film} vol={sc.voxels.shape} estimator={sc.tracer.settings.estimator} "
          f"shading={sc.tracer.settings.shading_type} sampler={sc.tracer.settings.sampler} lut={lut is not None}",
          flush=True)
    r = er.Renderer(sc)
    orc = O.Oracle(sc)
    if lut is not None:
        r.set_tf_lut(lut, lut_range)
        orc.set_tf_lut(lut, lut_range)
    if sc.tracer.settings.estimator == capi.ESTIMATOR_FULL and sc.tracer.settings.shading_type not in (0, 1):
        gm = r.gm_max()
        gm_ref = orc.compute_gm_max()
        assert gm == gm_ref, (gm, gm_ref)
    s1, s2 = O.seeds_for(r.W, r.H, seed)
    r.set_seeds(s1, s2)
    r.set_counting(counting)
    r.reset_counters()
    running = np.zeros((r.H, r.W, 4), dtype=np.float32)
    cnt_ref = np.zeros(8, dtype=np.uint64)
    exact = 0
    total = 0
    disp_ref = None
    for k in range(passes):
        r.render_estimate()
        f_ref, _ = orc.trace(s1, s2, cnt_ref)
        disp_ref, _ = orc.post(f_ref, running, k)
        f = r.frame()
        g1, g2 = r.get_seeds()
        assert np.array_equal(g1, s1) and np.array_equal(g2, s2), f"RNG state differs after pass {k}"
        assert np.array_equal(f[..., 3], f_ref[..., 3]), f"alpha (valid event) differs in pass {k}"
        np.testing.assert_allclose(f, f_ref, rtol=RTOL, atol=ATOL, err_msg=f"frame pass {k}")
        exact += int(np.sum(f == f_ref))
        total += f.size
    if counting:
        c = r.counters()
        for key in COUNTER_KEYS:
            assert c[key] == int(cnt_ref[ORC_INDEX[key]]), (key, c[key], int(cnt_ref[ORC_INDEX[key]]))
    np.testing.assert_allclose(r.running(), running, rtol=RTOL, atol=ATOL)
    if check_display:
        d = r.get_estimate()
        if sc.tracer.settings.estimator == capi.ESTIMATOR_FULL and sc.tracer.settings.bilateral:
            disp_ref = orc.bilateral(disp_ref)
        diff = np.abs(d.astype(np.int32) - disp_ref.astype(np.int32))
        assert diff.max() <= 1, f"display differs by {diff.max()} LSB"
    assert r.no_iterations() == passes
    print(f"[parity] exact-match fraction {exact / max(total, 1):.6f}", flush=True)
    r.close()
    return exact / max(total, 1)


def test_small_compat(fresh):
    sc = scenes.config_small(scale=0.5, estimator=capi.ESTIMATOR_COMPAT)
    assert run_parity(sc) > 0.999


def test_small_full_phase(fresh):
    sc = scenes.config_small(scale=0.5, estimator=capi.ESTIMATOR_FULL)
    run_parity(sc)


def test_ct_full_hybrid_env_bilateral(fresh):
    sc = scenes.config_ct(scale=0.125)
    run_parity(sc, passes=2)


def test_ct_compat_brdf(fresh):
    sc = scenes.config_ct(scale=0.125, estimator=capi.ESTIMATOR_COMPAT)
    run_parity(sc, passes=2)


def test_large_clipping_anisotropic(fresh):
    sc = scenes.config_large(scale=0.0625)
    run_parity(sc, passes=2)


def _rot(axis, angle):
    c, s = np.cos(angle), np.sin(angle)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    R = np.eye(3)
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


@pytest.mark.parametrize("case", ["sphere", "cylinder_rotated", "box_inverted", "sphere_inverted",
                                  "cylinder_inverted", "plane", "plane_inverted", "two_boxes"])
def test_clipping_shapes(fresh, case):
    """Every clipping shape, plain and inverted, rotated: macrocells wholly clipped away are
    skipped like empty ones, and the samples, decisions and radiance stay those of the oracle,
    which tests every sample."""
    sc = scenes.config_ct(scale=0.125, bilateral=False)
    sc.name = f"ct_clip_{case}"
    R = _rot(1, 0.6) @ _rot(0, 0.3)
    if case == "sphere":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_SPHERE, *scenes.translate((0.2, 0.1, 0.0)), outer=0.35))
    elif case == "cylinder_rotated":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_CYLINDER, *scenes._rigid(R, (0.1, 0.0, -0.1)),
                                          size=(0.0, 0.8, 0.0), outer=0.3))
    elif case == "box_inverted":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_BOX, *scenes._rigid(R, (0.05, 0.0, 0.0)),
                                          size=(0.5, 0.6, 0.55), invert=1))
    elif case == "sphere_inverted":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_SPHERE, *scenes.translate((0.0, -0.1, 0.1)), outer=0.3, invert=1))
    elif case == "cylinder_inverted":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_CYLINDER, *scenes._rigid(R, (0.0, 0.0, 0.0)),
                                          size=(0.0, 0.5, 0.0), outer=0.25, invert=1))
    elif case == "plane":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_PLANE, *scenes._rigid(R, (0.0, 0.05, 0.0))))
    elif case == "plane_inverted":
        sc.add_clip(scenes.clipping_shape(capi.SHAPE_PLANE, *scenes._rigid(R, (0.0, 0.05, 0.0)), invert=1))
    else:
        sc.add_clip(scenes.clipping_box((0.3, 0.0, 0.0), (0.4, 1.1, 1.1)))
        sc.add_clip(scenes.clipping_box((-0.3, 0.0, 0.0), (0.3, 0.4, 1.1)))
    run_parity(sc, passes=2)


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_clipping_random_affine(fresh, seed):
    """Random clipping objects with scaled and sheared (non-rigid) transforms, plain and
    inverted, one to three at once: the clip-aware occupancy bits never skip a sample the
    oracle keeps."""
    rng = np.random.default_rng(seed)
    sc = scenes.config_ct(scale=0.125, bilateral=False)
    sc.name = f"ct_clip_random_{seed}"
    for _ in range(int(rng.integers(1, 4))):
        A = np.eye(3) + rng.uniform(-0.3, 0.3, size=(3, 3))
        A = A @ _rot(int(rng.integers(0, 3)), float(rng.uniform(0, np.pi)))
        tm = np.eye(4)
        tm[:3, :3] = A
        tm[:3, 3] = rng.uniform(-0.3, 0.3, size=3)
        inv = np.linalg.inv(tm)
        shape = [capi.SHAPE_BOX, capi.SHAPE_SPHERE, capi.SHAPE_CYLINDER, capi.SHAPE_PLANE][int(rng.integers(0, 4))]
        size = tuple(rng.uniform(0.2, 0.9, size=3))
        sc.add_clip(scenes.clipping_shape(shape, tm, inv, size=size, outer=float(rng.uniform(0.15, 0.45)),
                                          invert=int(rng.integers(0, 2))))
    run_parity(sc, passes=1)


def test_fully_clipped_volume_is_skipped(fresh):
    """A clipping box around the whole volume: every macrocell is clipped away, so the march
    skips through it in a few iterations per ray while counting every sample the reference
    takes, and nothing scatters."""
    sc = scenes.config_ct(scale=0.125, bilateral=False)
    sc.add_clip(scenes.clipping_box((0.0, 0.0, 0.0), (1.2, 1.2, 1.2)))
    run_parity(sc, passes=1)
    r = er.Renderer(sc)
    r.set_counting(True)
    r.reset_counters()
    r.render_estimate()
    c = r.counters()
    print("[fully clipped]", c, flush=True)
    assert c["scatter_events"] == 0 and c["primary_fetches"] == 0
    assert c["primary_samples"] > 50 * c["paths"]
    assert c["march_iterations"] <= 8 * c["paths"]
    r.close()


def test_shading_modes(fresh):
    for mode in (capi.SHADING_BRDF_ONLY, capi.SHADING_MODULATION, capi.SHADING_THRESHOLD,
                 capi.SHADING_GRADIENT_MAGNITUDE):
        sc = scenes.config_ct(scale=0.0625, bilateral=False)
        sc.tracer.settings.shading_type = mode
        run_parity(sc, passes=1)
        er.shutdown()


def test_gradient_modes(fresh):
    for mode in (capi.GRADIENT_FORWARD, capi.GRADIENT_FILTERED):
        sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
        sc.tracer.settings.gradient_computation = mode
        run_parity(sc, passes=1)
        er.shutdown()


def _woodcock_scene(case):
    if case == "small_phase":
        sc = scenes.config_small(scale=0.5, estimator=capi.ESTIMATOR_FULL)
    elif case == "ct_hybrid_env_bilateral":
        sc = scenes.config_ct(scale=0.125)
    else:
        sc = scenes.config_ct(scale=0.125, bilateral=False)
        sc.add_clip(scenes.clipping_box((0.2, 0.0, 0.0), (0.5, 1.1, 1.1)))
    sc.tracer.settings.sampler = capi.SAMPLER_WOODCOCK
    sc.name = f"woodcock_{case}"
    return sc


@pytest.mark.parametrize("case", ["small_phase", "ct_hybrid_env_bilateral", "ct_clip_box"])
def test_woodcock(fresh, case):
    """Delta tracking with per-macrocell majorants (full estimator): the same tentative
    collisions, acceptances, RNG streams and radiance as the oracle's walk."""
    run_parity(_woodcock_scene(case), passes=2)


def test_woodcock_image_does_not_depend_on_the_brick_size(fresh):
    """The majorants live on their own fixed grid of 8-cell macrocells, so the tuned brick
    size (occupancy bits) changes no decision of the Woodcock walk."""
    sc = _woodcock_scene("ct_hybrid_env_bilateral")
    frames = []
    for brick in (2, 8, 16):
        r = er.Renderer(sc, seed=9)
        r.set_tune(brick, 8, 7)
        r.render_estimate()
        r.render_estimate()
        frames.append((r.frame().copy(), r.get_seeds()))
        r.close()
    for f, (g1, g2) in frames[1:]:
        assert np.array_equal(f, frames[0][0])
        assert np.array_equal(g1, frames[0][1][0]) and np.array_equal(g2, frames[0][1][1])


def test_no_lights_emission_only(fresh):
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    sc.tracer.light_count = 0
    sc.tracer.emission.init()
    sc.tracer.emission.add(0.0, (0.1, 0.2, 0.3)).add(4000.0, (0.5, 0.4, 0.3))
    run_parity(sc, passes=2)


def test_odd_film_sizes(fresh):
    for film in ((1, 1), (17, 5), (33, 70)):
        sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
        sc.tracer.camera.film_size[0], sc.tracer.camera.film_size[1] = film
        run_parity(sc, passes=2)
        er.shutdown()


def test_camera_inside_volume_and_thin_lens(fresh):
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    scenes.set_camera(sc.tracer, sc.film, (0.1, 0.05, 0.2), (0.0, 0.0, 0.0), clip_near=0.0, exposure=4.0)
    run_parity(sc, passes=1)
    er.shutdown()
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    scenes.set_camera(sc.tracer, sc.film, (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), exposure=4.0, aperture=0.05, focal=-1.0)
    run_parity(sc, passes=1)


def test_node_table_transfer_function(fresh):
    """er_set_tf_lut: opacity and diffuse colour from a 256-node table, bit-for-bit decisions."""
    sc = scenes.config_adjoint(scale=0.25)
    run_parity(sc, passes=3, lut=scenes.random_tf_lut(7))


def test_autofocus_distance(fresh):
    """GetAutoFocusDistance (autofocus.cuh:26-75): the same 100 jittered rays and events as the
    oracle from the same seeds, at the film centre and at a corner that misses the volume."""
    sc = scenes.config_small(scale=0.5)
    r = er.Renderer(sc)
    orc = O.Oracle(sc)
    for (u, v) in [(r.W // 2, r.H // 2), (r.W // 3, r.H // 2 + 5), (0, 0)]:
        s0, s1 = r.autofocus_seeds(u, v)
        d = r.autofocus_distance(u, v)
        d_ref = orc.autofocus(u, v, s0, s1)
        print(f"[autofocus] ({u},{v}) gpu {d:.7f} oracle {d_ref:.7f}", flush=True)
        assert d == pytest.approx(d_ref, rel=1e-5, abs=1e-6)
        assert d > 0.0
    r.close()


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
def test_light_shapes_and_textures(fresh, estimator):
    """Disk (Lux, one-sided), ring and box lights (disk.h, ring.h, box.h:122-191) with checker
    and gradient procedural textures (textures.h:64-111), next to the scene's plane light."""
    sc = scenes.config_small(scale=0.5, estimator=estimator)
    chk = sc.add_texture(scenes.checker_texture((1.0, 0.9, 0.8), (0.2, 0.3, 0.4)))
    grd = sc.add_texture(scenes.gradient_texture([(0.0, (0.2, 0.2, 1.0)), (1.0, (1.0, 0.6, 0.2))]))
    sc.add_light(scenes.shape_light(capi.SHAPE_DISK, (-0.9, 0.8, 0.7), (0.0, 0.0, 0.0), 40.0, texture_id=chk,
                                    outer=0.3, one_sided=1, unit=capi.UNIT_LUX))
    sc.add_light(scenes.shape_light(capi.SHAPE_RING, (0.2, -1.0, 0.8), (0.0, 0.0, 0.0), 60.0, texture_id=grd,
                                    inner=0.1, outer=0.35))
    sc.add_light(scenes.shape_light(capi.SHAPE_BOX, (0.0, 1.1, -0.9), (0.0, 0.0, 0.0), 30.0, texture_id=0,
                                    size=(0.3, 0.2, 0.25)))
    assert run_parity(sc, passes=3) > 0.99


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
def test_surface_objects_and_bitmaps(fresh, estimator):
    """Surface objects (objects.h:26-82) as camera-ray candidates after the volume and the
    lights, as shadow-ray occluders after the lights (transport.h:30-44), shaded by the BRDF
    from textures at the hit (transport.h:141-149): a checker floor plane, a sphere with a
    bitmap texture (textures.h:99-105) and a box with a glossiness bitmap."""
    sc = scenes.config_small(scale=0.5, estimator=estimator)
    rng = np.random.default_rng(21)
    bmp = sc.add_bitmap(rng.integers(0, 256, size=(8, 16, 4), dtype=np.uint8))
    gls = sc.add_bitmap(rng.integers(0, 256, size=(4, 4, 4), dtype=np.uint8))
    chk = sc.add_texture(scenes.checker_texture((0.9, 0.8, 0.7), (0.2, 0.25, 0.3)))
    btx = sc.add_texture(scenes.bitmap_texture(bmp, repeat=(2.0, 1.0)))
    gtx = sc.add_texture(scenes.bitmap_texture(gls, flip=(1, 0)))
    spc = sc.add_texture(scenes.uniform_texture((0.2, 0.2, 0.2)))
    glu = sc.add_texture(scenes.uniform_texture((0.5, 0.5, 0.5)))
    tm, inv = scenes.look_at((0.0, -0.6, 0.0), (0.0, 1.0, 0.0))
    sc.add_object(scenes.object_desc(capi.SHAPE_PLANE, tm, inv, size=(3.0, 3.0, 0.0), diffuse_tex=chk,
                                     specular_tex=spc, glossiness_tex=glu))
    tm, inv = scenes.translate((0.55, 0.25, 0.1))
    sc.add_object(scenes.object_desc(capi.SHAPE_SPHERE, tm, inv, outer_radius=0.2, diffuse_tex=btx,
                                     specular_tex=spc, glossiness_tex=glu))
    tm, inv = scenes.translate((-0.45, 0.1, 0.45))
    sc.add_object(scenes.object_desc(capi.SHAPE_BOX, tm, inv, size=(0.25, 0.3, 0.2), diffuse_tex=chk,
                                     specular_tex=spc, glossiness_tex=gtx))
    sc.name = "small_objects"
    assert run_parity(sc, passes=3) > 0.99
