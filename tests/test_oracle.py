"""Pins the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, and its hot-path sources cannot be
built here without stand-in headers, so every check below pins an oracle function against
what the mathematics fixes: closed forms, invariants, special cases, exhaustive sweeps of
finite domains and brute force on tiny inputs. Each check is chosen so that a plausible slip
(a transposed index, a dropped term, a wrong power of the density scale, a wrong draw order)
fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from exposure_render import capi, scenes


def _u_from(s0, s1):
    return ((((s0 << 16) + s1) & 0xFFFFFFFF) & 0x7FFFFF) / 8388608.0


# ---- RNG (rng.h:35-51) -----------------------------------------------------------------

def test_rng_mwc_closed_form(orc_lib):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = (int(v) for v in rng.integers(1, 2**32, size=2))
        s0 = C.c_uint32(a)
        s1 = C.c_uint32(b)
        for _ in range(20):
            u = orc_lib.orc_rng_get1(C.byref(s0), C.byref(s1))
            # Marsaglia's two 16-bit multiply-with-carry generators
            a = (36969 * (a & 0xFFFF) + (a >> 16)) & 0xFFFFFFFF
            b = (18000 * (b & 0xFFFF) + (b >> 16)) & 0xFFFFFFFF
            assert s0.value == a and s1.value == b
            assert u == _u_from(a, b)
            assert 0.0 <= u < 1.0


def test_rng_zero_seed_is_fixed_point(orc_lib):
    s0, s1 = C.c_uint32(0), C.c_uint32(0)
    for _ in range(5):
        assert orc_lib.orc_rng_get1(C.byref(s0), C.byref(s1)) == 0.0
    assert s0.value == 0 and s1.value == 0


def test_seed_hash_never_zero():
    s1, s2 = O.seeds_for(4096, 4096, 0)
    assert s1.min() > 0 and s2.min() > 0
    assert len(np.unique(s1[:100000])) > 99000


# ---- -log(u) (raymarching.h:45,89) -----------------------------------------------------

def test_neglog_correctly_rounded_exhaustive(orc_lib):
    """Over all 2^23 values u = m 2^-23 the oracle's -log(u) must be the float nearest the exact
    value. float64 log has < 1 ulp(double) error; the check asserts that no exact value lies
    within 4 ulp(double) of a float rounding boundary, so rounding the float64 result is the
    correctly rounded float for every m."""
    m = np.arange(1, 2**23, dtype=np.float64)
    v = -np.log(m * 2.0 ** -23)
    f = v.astype(np.float32)
    # distance of v to the nearest float midpoint, in units of ulp(double)
    lo = np.nextafter(f, np.float32(-np.inf)).astype(np.float64)
    hi = np.nextafter(f, np.float32(np.inf)).astype(np.float64)
    mid_lo = (f.astype(np.float64) + lo) / 2
    mid_hi = (f.astype(np.float64) + hi) / 2
    gap = np.minimum(np.abs(v - mid_lo), np.abs(v - mid_hi)) / np.spacing(v)
    assert gap.min() > 4.0
    # the oracle on a dense sample, including the extremes
    idx = np.unique(np.concatenate([np.arange(1, 5000), np.arange(2**23 - 5000, 2**23),
                                    np.random.default_rng(1).integers(1, 2**23, 20000)]))
    got = np.array([orc_lib.orc_neglog(float(k) * 2.0 ** -23) for k in idx], dtype=np.float32)
    assert np.array_equal(got, f[idx - 1])
    assert math.isinf(orc_lib.orc_neglog(0.0))


# ---- float elementary functions (er_fmath.h) ----------------------------------------------

def _ulp_err(got, want):
    want32 = want.astype(np.float32)
    ulp = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / np.maximum(ulp, 1e-45)


def test_fmath_accuracy():
    x = np.linspace(-103.0, 88.0, 400001, dtype=np.float32)
    assert _ulp_err(O.fmath(0, x), np.exp(x.astype(np.float64))).max() < 2.0
    x = np.concatenate([np.geomspace(1e-38, 1e38, 400001), np.linspace(0.5, 2.0, 100001)]).astype(np.float32)
    assert _ulp_err(O.fmath(1, x), np.log(x.astype(np.float64))).max() < 2.5
    x = np.linspace(0.0, 2 * np.pi, 400001, dtype=np.float32)
    s = O.fmath(3, x)
    c = O.fmath(4, x)
    assert np.abs(s - np.sin(x.astype(np.float64))).max() < 2e-7
    assert np.abs(c - np.cos(x.astype(np.float64))).max() < 2e-7
    # pow on the shader's domain: cosines in [0, 1], exponents up to 1e6
    rng = np.random.default_rng(2)
    b = rng.uniform(0.0, 1.0, 200000).astype(np.float32)
    e = np.concatenate([rng.uniform(0, 8, 100000), 10 ** rng.uniform(0, 6, 100000)]).astype(np.float32)
    want = np.power(b.astype(np.float64), e.astype(np.float64))
    got = O.fmath(2, b, e).astype(np.float64)
    keep = want > 1e-30
    rel = np.abs(got[keep] - want[keep]) / want[keep]
    assert rel.max() < 2.5e-7 * np.maximum(np.abs(e[keep] * np.log(b[keep])), 1.0).max() + 1e-6
    x = np.linspace(-1.0, 1.0, 200001, dtype=np.float32)
    assert np.abs(O.fmath(6, x) - np.arccos(x.astype(np.float64))).max() < 5e-7
    # expm1 (the adjoint's free-flight probabilities): relative to the value, also near 0
    x = np.concatenate([-np.geomspace(1e-30, 40.0, 200001), np.geomspace(1e-30, 20.0, 200001)]).astype(np.float32)
    want = np.expm1(x.astype(np.float64))
    assert _ulp_err(O.fmath(7, x), want).max() < 4.0
    yy, xx = np.meshgrid(np.linspace(-1, 1, 401), np.linspace(-1, 1, 401))
    yy = yy.astype(np.float32).ravel()
    xx = xx.astype(np.float32).ravel()
    assert np.abs(O.fmath(5, yy, xx) - np.arctan2(yy.astype(np.float64), xx.astype(np.float64))).max() < 5e-7


# ---- transfer functions (plf.h:68-91) ---------------------------------------------------

def test_plf_closed_form_and_last_node_zero(orc_lib):
    p = capi.Plf().init()
    p.add(0.0, 0.0).add(400.0, 0.0).add(1500.0, 0.6).add(4000.0, 1.0)
    ev = lambda x: orc_lib.orc_plf_eval(C.byref(p), x)
    assert ev(-5.0) == 0.0                 # below the first node -> first value
    assert ev(5000.0) == 1.0               # above the last node -> last value
    assert ev(4000.0) == 0.0               # exactly the last node -> 0 (Q5)
    assert ev(1500.0) == np.float32(0.6)   # a node is the start of its segment
    assert abs(ev(950.0) - 0.3) < 1e-7     # midpoint of a segment
    for x in np.linspace(400.0, 1499.0, 50):
        assert abs(ev(float(x)) - 0.6 * (x - 400.0) / 1100.0) < 1e-6
    e = capi.Plf().init()
    assert orc_lib.orc_plf_eval(C.byref(e), 10.0) == 0.0   # empty function
    # the node capacity: 127 nodes (plf.h:53)
    big = capi.Plf().init()
    for k in range(200):
        big.add(float(k), float(k))
    assert big.count == 127


def test_plf_brute_force_random(orc_lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = int(rng.integers(2, 12))
        pos = np.sort(rng.uniform(0, 4000, n)).astype(np.float32)
        val = rng.uniform(0, 1, n).astype(np.float32)
        p = capi.Plf().init()
        for a, b in zip(pos, val):
            p.add(float(a), float(b))
        xs = rng.uniform(-100, 4100, 200).astype(np.float32)
        for x in xs:
            got = orc_lib.orc_plf_eval(C.byref(p), float(x))
            if x < pos[0]:
                want = val[0]
            elif x > pos[-1]:
                want = val[-1]
            elif x == pos[-1]:
                want = 0.0
            else:
                k = np.searchsorted(pos, x, side="right") - 1
                want = val[k] + (x - pos[k]) / (pos[k + 1] - pos[k]) * (val[k + 1] - val[k])
            assert abs(got - want) < 2e-6


# ---- trilinear sampling (buffer3d.h:210-242, volume.h:131-138) ----------------------------

def _tiny_scene(vox):
    sc = scenes.Scene(vox)
    sc.tracer.opacity.add(0.0, 0.5).add(70000.0, 0.5)
    scenes.set_camera(sc.tracer, (4, 4), (0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    return sc


def test_trilinear_lattice_points_and_layout():
    """At lattice points the truncating trilinear returns the voxel itself; a distinct value
    per voxel pins the x-fastest, z-slowest layout and the corner-aligned lattice."""
    rz, ry, rx = 2, 4, 8          # powers of two: the lattice coordinates of these points are exact
    vox = (np.arange(rx * ry * rz, dtype=np.uint16) * 7 + 3).reshape(rz, ry, rx)
    orc = O.Oracle(_tiny_scene(vox))
    ext = np.array([rx, ry, rz], dtype=np.float64)
    size = ext / ext.max()
    for z in range(rz):
        for y in range(ry):
            for x in range(rx):
                p = (np.array([x, y, z]) / ext - 0.5) * size
                assert orc.intensity(p) == vox[z, y, x], (x, y, z)
    # off the exact lattice the chain truncates: a point rounding to just below a lattice
    # coordinate loses one unit (the reference's uint16 lerps, Q4)
    assert orc.intensity(np.array([1.0 / 8 - 0.5 - 1e-7, -0.25, -0.125])) in (vox[0, 0, 1] - 1, vox[0, 0, 1])


def test_trilinear_bounds_and_clamp():
    rng = np.random.default_rng(4)
    vox = rng.integers(0, 5000, size=(6, 6, 6)).astype(np.uint16)
    orc = O.Oracle(_tiny_scene(vox))
    for _ in range(2000):
        L = rng.uniform(-1.5, 7.5, 3)          # lattice coordinates, also outside the volume
        p = (L / 6.0 - 0.5) * 1.0
        got = orc.intensity(p)
        assert got == int(got)                  # every lerp truncates to uint16 (Q4)
        v = np.floor(L).astype(int)
        xs = np.clip([v[0], v[0] + 1], 0, 5)
        ys = np.clip([v[1], v[1] + 1], 0, 5)
        zs = np.clip([v[2], v[2] + 1], 0, 5)
        corners = vox[np.ix_(zs, ys, xs)]
        d = L - v
        w = [[(1 - d[2]) if k == 0 else d[2] for k in range(2)], [(1 - d[1]) if k == 0 else d[1] for k in range(2)],
             [(1 - d[0]) if k == 0 else d[0] for k in range(2)]]
        exact = sum(corners[a, b, c] * w[0][a] * w[1][b] * w[2][c] for a in range(2) for b in range(2) for c in range(2))
        assert corners.min() - 3 <= got <= corners.max()
        assert exact - 3.0 - 1e-3 <= got <= exact + 1e-3   # three truncation levels lose < 1 each


# ---- free flight (raymarching.h:30-113) ---------------------------------------------------

def _slab(intensity, opacity, ds, res=16):
    vox = np.full((res, res, res), intensity, dtype=np.uint16)
    sc = scenes.Scene(vox)
    sc.tracer.opacity.add(0.0, opacity).add(60000.0, opacity)
    sc.tracer.settings.density_scale = ds
    scenes.set_camera(sc.tracer, (4, 4), (0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    return sc


def test_march_homogeneous_step_count_and_T_offset():
    """Constant sigma: the march scatters after n = ceil(S / (sigma dt)) samples, reports P at
    the n-th sample and T one step further (Q13)."""
    ds, op = 10.0, 0.25
    orc = O.Oracle(_slab(1000, op, ds))
    dt = np.float32(0.1) * np.float32(1.0 / 16)
    sig = np.float32(ds) * np.float32(op)
    for u1 in (0.9, 0.5, 0.1, 0.01):
        hit, T, P, steps = orc.march_ray((0.0, 0.0, -2.0), (0.0, 0.0, 1.0), 0.0, 1000.0, 1, u1, 0.5)
        S = np.float32(-np.log(np.float32(u1))) / np.float32(ds)
        acc, n = np.float32(0), 0
        while acc < S:
            acc = np.float32(acc + np.float32(sig * dt))
            n += 1
        assert hit == 1 and steps == n
        t_first = np.float32(np.float32(1.5) + np.float32(0.5) * dt)
        assert abs(P[2] - (-2.0 + t_first + (n - 1) * dt)) < 1e-5
        assert abs(T - (t_first + n * dt)) < 1e-5


def test_march_transmittance_density_scale_squared():
    """P(no scatter through a slab of length L) = exp(-DS^2 * opacity * L) (Q3: the threshold is
    divided by DS and the accumulation multiplied by DS)."""
    ds, op = 4.0, 0.2
    orc = O.Oracle(_slab(1000, op, ds))
    us = (np.arange(1, 2001) - 0.5) / 2000.0
    passed = 0
    for u1 in us:
        hit, T, P, steps = orc.march_ray((0.1, 0.05, -2.0), (0.0, 0.0, 1.0), 0.0, 1000.0, 1, float(u1), 0.3)
        passed += hit == 0
    L = 1.0
    assert abs(passed / len(us) - math.exp(-ds * ds * op * L)) < 0.01


def test_march_misses_and_shadow_tmax():
    orc = O.Oracle(_slab(1000, 0.5, 10.0))
    hit, *_ = orc.march_ray((2.0, 2.0, -2.0), (0.0, 0.0, 1.0), 0.0, 1000.0, 1, 0.5, 0.5)
    assert hit == -1                                   # misses the box
    hit, *_ = orc.march_ray((0.0, 0.0, -2.0), (0.0, 0.0, 1.0), 0.0, 1.0, 0, 0.5, 0.5)
    assert hit == -1                                   # box starts after MaxT
    empty = O.Oracle(_slab(1000, 0.0, 10.0))
    hit, T, P, steps = empty.march_ray((0.0, 0.0, -2.0), (0.0, 0.0, 1.0), 0.0, 1000.0, 1, 0.5, 0.5)
    assert hit == 0 and steps == 160                    # 1.0 / (0.1 / 16) samples, no event


# ---- camera (camera.h:61-96) ----------------------------------------------------------------

def test_camera_frame():
    sc = scenes.config_small(scale=0.25)
    scenes.set_camera(sc.tracer, (64, 48), (1.0, 2.0, 3.0), (0.0, 0.5, 0.0), fov=40.0)
    cam = O.Oracle(sc).camera()
    N, U, V = cam[0:3], cam[3:6], cam[6:9]
    d = -np.array([1.0, 1.5, 3.0])
    assert np.allclose(N, d / np.linalg.norm(d), atol=1e-6)
    assert abs(np.dot(N, U)) < 1e-6 and abs(np.dot(N, V)) < 1e-6 and abs(np.dot(U, V)) < 1e-6
    assert np.allclose(U, np.cross(N, [0, 1, 0]) / np.linalg.norm(np.cross(N, [0, 1, 0])), atol=1e-6)
    scale = math.tan(math.radians(20.0))
    aspect = 48 / 64
    assert abs(cam[9] + scale / aspect) < 1e-6 and abs(cam[10] + scale) < 1e-6
    assert abs(cam[11] - 2 * scale / aspect / 64) < 1e-7 and abs(cam[12] - 2 * scale / 48) < 1e-7


# ---- post-processing (filterframeestimate.cuh, estimate.cuh, tonemap.cuh) ---------------------

def test_post_filter_cma_tonemap_closed_forms():
    sc = scenes.config_small(scale=0.125)
    orc = O.Oracle(sc)
    W, H = sc.film
    running = np.zeros((H, W, 4), dtype=np.float32)
    vals = [0.7, 0.2, 0.5, 0.9]
    for k, v in enumerate(vals):
        frame = np.zeros((H, W, 4), dtype=np.float32)
        frame[..., :3] = v
        frame[..., 3] = 1.0
        disp, filt = orc.post(frame, running, k)
        # constant image: the renormalised Gaussian returns the constant; alpha / total weight
        assert np.allclose(filt[..., :3], v, atol=1e-6)
        tw_inner = 1 + 4 * math.exp(-0.5) + 4 * math.exp(-1.0)
        assert abs(filt[H // 2, W // 2, 3] - 1.0 / tw_inner) < 1e-6
        assert abs(filt[0, 0, 3] - 1.0 / (1 + 2 * math.exp(-0.5) + math.exp(-1.0))) < 1e-6
    # running estimate = mean of passes 2..n (pass 1 discarded, Q1)
    assert np.allclose(running[..., 0], np.mean(vals[1:]), atol=1e-6)
    # tonemap: XYZ -> linear sRGB, 1 - exp(-c / exposure), truncation to u8, A = a * 255
    m = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556], [0.055648, -0.204043, 1.057311]])
    rgb = m @ np.full(3, np.mean(vals[1:]))
    want = np.floor(255 * np.clip(1 - np.exp(-rgb / sc.tracer.camera.exposure), 0, 1))
    assert np.all(np.abs(disp[H // 2, W // 2, :3].astype(float) - want) <= 1)


def test_bilateral_constant_and_range_limits():
    sc = scenes.config_ct(scale=0.0625)
    orc = O.Oracle(sc)
    W, H = sc.film
    img = np.full((H, W, 4), 120, dtype=np.uint8)
    out = orc.bilateral(img)
    # normalisation by the summed weights, then truncation to u8: at most one unit per pass
    assert np.all(out[..., :3] >= 118) and np.all(out[..., :3] <= 120) and np.all(out[..., 3] == 255)
    # a step edge far above the range sigma survives (edge-preserving)
    img[:, : W // 2, :3] = 10
    img[:, W // 2:, :3] = 250
    out = orc.bilateral(img)
    assert out[H // 2, W // 2 - 2, 0] <= 12 and out[H // 2, W // 2 + 2, 0] >= 246


# ---- estimator special cases ------------------------------------------------------------------

def test_no_lights_gives_unscaled_emission():
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    sc.tracer.light_count = 0
    sc.tracer.emission.init()
    sc.tracer.emission.add(0.0, (0.25, 0.5, 0.75)).add(70000.0, (0.25, 0.5, 0.75))
    orc = O.Oracle(sc)
    s1, s2 = O.seeds_for(*sc.film, 3)
    f, cnt = orc.trace(s1, s2)
    hit = f[..., 3] == 1.0
    assert hit.any()
    assert np.allclose(f[hit][:, :3], [0.25, 0.5, 0.75])     # Q10: no light-count scaling without lights
    assert np.all(f[~hit][:, :3] == 0.0)


def test_transparent_volume_is_invisible():
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    sc.tracer.opacity.init()
    sc.tracer.opacity.add(0.0, 0.0).add(70000.0, 0.0)
    sc.lights[0].visible = 0
    orc = O.Oracle(sc)
    s1, s2 = O.seeds_for(*sc.film, 3)
    f, cnt = orc.trace(s1, s2)
    assert np.all(f == 0.0) and cnt[6] == 0


def test_draw_schedule():
    """11 draws per path, +2 when the camera ray meets the volume box, +2 per shadow march."""
    sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_COMPAT)
    sc.tracer.opacity.init()
    sc.tracer.opacity.add(0.0, 0.0).add(70000.0, 0.0)     # no scatter -> no shadow rays
    orc = O.Oracle(sc)
    W, H = sc.film
    px = np.array([0, W // 2], dtype=np.int32)
    py = np.array([0, H // 2], dtype=np.int32)
    s1, s2 = O.seeds_for(2, 1, 9)
    out, draws, cnt = orc.trace_pixels(px, py, s1, s2)
    assert list(draws) == [11, 13]                      # corner misses the box, centre hits it


def test_fp32_mean_matches_fp64_expectation():
    sc = scenes.config_small(scale=0.125, estimator=capi.ESTIMATOR_FULL)
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    px = np.repeat(np.arange(W), 1)[: W].astype(np.int32)
    py = np.full(W, H // 2, dtype=np.int32)
    e64 = orc.expect_f64(px, py, 4000, seed=1)
    acc = np.zeros((W, 4))
    s1, s2 = O.seeds_for(W, 1, 2)
    n = 4000
    for _ in range(n):
        out, _, _ = orc.trace_pixels(px, py, s1, s2)
        acc += out
    m32 = acc / n
    sel = e64[:, 3] > 0.5
    err = np.abs(m32[sel, :3] - e64[sel, :3]).mean()
    assert err < 0.05 * np.abs(e64[sel, :3]).mean()


def test_node_table_transfer_function():
    """Node n at intensity n * range / 255; linear between nodes; last node's value beyond the
    range; negative interpolated values clamp to 0 (closed form, against numpy.interp)."""
    rng = np.random.default_rng(3)
    lut = rng.uniform(-0.2, 1.0, size=(256, 4)).astype(np.float32)
    rng_i = 4095.0
    nodes = np.arange(256) * (rng_i / 255.0)
    for ch in range(4):
        for I in [0.0, 1.0, 16.0, 17.0, 2047.0, 4094.0, 4095.0, 5000.0, 65535.0]:
            got = O.tf_lut_eval(lut, rng_i, I, ch)
            want = max(float(np.interp(min(I, rng_i), nodes, lut[:, ch].astype(np.float64))), 0.0)
            assert abs(got - want) <= 2e-6 * max(1.0, abs(want)), (ch, I, got, want)
    # exactly at a node: the node's value (when non-negative)
    lut2 = np.abs(lut)
    for n in [0, 1, 100, 254, 255]:
        I = n * rng_i / 255.0
        assert abs(O.tf_lut_eval(lut2, rng_i, I, 0) - lut2[n, 0]) <= 1e-6 * max(1.0, lut2[n, 0])


# ---- textures (textures.h:26-111, buffer2d.h:227-254, color.h:195-211) ---------------------

# sRGB (D65) to CIE XYZ; a white texel maps to the D65 white point, the row sums
_RGB2XYZ = np.array([[0.412453, 0.357580, 0.180423],
                     [0.212671, 0.715160, 0.072169],
                     [0.019334, 0.119193, 0.950227]])


def _texture_oracle(textures, bitmaps, estimator):
    sc = scenes.config_small(scale=0.125, estimator=estimator)
    sc.textures = []
    for T in textures:
        sc.add_texture(T)
    for b in bitmaps:
        sc.add_bitmap(b)
    return O.Oracle(sc)


def test_procedural_textures_closed_form():
    uni = scenes.uniform_texture((0.3, 0.6, 0.9), output_level=2.0)
    chk = scenes.checker_texture((1.0, 0.0, 0.0), (0.0, 0.0, 1.0), repeat=(4.0, 4.0))
    orc = _texture_oracle([uni, chk], [], capi.ESTIMATOR_COMPAT)
    assert np.allclose(orc.eval_texture(0, 0.37, 0.81), [0.6, 1.2, 1.8])
    rng = np.random.default_rng(7)
    for u, v in rng.uniform(0.0, 1.0, size=(64, 2)):
        # repeat, then V is offset by 1 - 0 before the wrap (textures.h:71-78)
        a = math.floor(2.0 * ((4.0 * u) % 1.0))
        b = math.floor(2.0 * ((4.0 * v + 1.0) % 1.0))
        want = [1.0, 0.0, 0.0] if (a + b) % 2 == 0 else [0.0, 0.0, 1.0]
        assert np.array_equal(orc.eval_texture(1, u, v), np.float32(want)), (u, v)
    assert np.array_equal(orc.eval_texture(-1, 0.5, 0.5), np.zeros(3, np.float32))   # ID < 0: black


def test_bitmap_texture_full_is_bilinear():
    rng = np.random.default_rng(11)
    bmp = rng.integers(0, 256, size=(4, 8, 4), dtype=np.uint8)     # H = 4, W = 8
    T = scenes.bitmap_texture(0)
    orc = _texture_oracle([T], [bmp], capi.ESTIMATOR_FULL)
    H, W = bmp.shape[:2]
    white = _texture_oracle([T], [np.full((2, 2, 4), 255, np.uint8)], capi.ESTIMATOR_FULL)
    assert np.allclose(white.eval_texture(0, 0.3, 0.6), _RGB2XYZ.sum(axis=1), rtol=1e-5)
    for u, v in rng.uniform(0.0, 0.999, size=(64, 2)):
        X, Y = u * W, v * H
        x0, y0 = int(math.floor(X)), int(math.floor(Y))
        x1, y1 = min(x0 + 1, W - 1), min(y0 + 1, H - 1)
        du, dv = X - x0, Y - y0
        c = bmp[..., :3].astype(np.float64)
        rgb = (1 - dv) * ((1 - du) * c[y0, x0] + du * c[y0, x1]) + dv * ((1 - du) * c[y1, x0] + du * c[y1, x1])
        want = _RGB2XYZ @ (rgb / 255.0)
        np.testing.assert_allclose(orc.eval_texture(0, u, v), want, rtol=2e-5, atol=2e-6)


def test_bitmap_texture_compat_reads_exact_texels_only():
    """The reference blends texels with `float * ColorRGBAuc`, which converts each bilinear
    weight to unsigned char: exact texel coordinates read that texel, anything between texels
    has weights 0 and reads black."""
    rng = np.random.default_rng(12)
    bmp = rng.integers(0, 256, size=(4, 8, 4), dtype=np.uint8)
    orc = _texture_oracle([scenes.bitmap_texture(0)], [bmp], capi.ESTIMATOR_COMPAT)
    H, W = bmp.shape[:2]
    for j in range(H):
        for i in range(W):
            want = _RGB2XYZ @ (bmp[j, i, :3].astype(np.float64) / 255.0)
            np.testing.assert_allclose(orc.eval_texture(0, i / W, j / H), want, rtol=2e-5, atol=2e-6)
    for u, v in [(0.5 / W, 0.0), (0.0, 0.5 / H), (1.3 / W, 2.6 / H)]:
        assert np.array_equal(orc.eval_texture(0, u, v), np.zeros(3, np.float32)), (u, v)


# ---- surface objects (objects.h:26-82, transport.h:115-156) ---------------------------------

LIGHT_POS, LIGHT_SIZE = (0.3, 0.1, 0.45), 0.8
OCC_CENTER, OCC_RADIUS = (0.35, 0.15, 0.2), 0.15


def _floor_scene(estimator, occluder):
    """A transparent volume, a 2 x 2 floor object in z = 0 (diffuse 0.8 / 0.6 / 0.4, no
    specular, glossiness 0), an invisible 0.8 x 0.8 plane light 0.45 above it (large and close,
    so that the MIS weight matters) and, optionally, a disk object of radius 0.15 at z = 0.2
    that shadows part of the light."""
    sc = scenes.config_small(scale=0.25, estimator=estimator)
    t = sc.tracer
    t.opacity.init()
    t.opacity.add(0.0, 0.0).add(70000.0, 0.0)
    sc.textures, sc.lights = [], []
    t.light_count = 0
    white = sc.add_texture(scenes.uniform_texture((1.0, 1.0, 1.0)))
    kd = sc.add_texture(scenes.uniform_texture((0.8, 0.6, 0.4)))
    black = sc.add_texture(scenes.uniform_texture((0.0, 0.0, 0.0)))
    sc.add_light(scenes.plane_light(LIGHT_POS, (0.25, 0.05, 0.0), (LIGHT_SIZE, LIGHT_SIZE), 5.0, texture_id=white,
                                    visible=0))
    tm, inv = scenes.translate((0.0, 0.0, 0.0))
    sc.add_object(scenes.object_desc(capi.SHAPE_PLANE, tm, inv, size=(2.0, 2.0, 0.0), diffuse_tex=kd,
                                     specular_tex=black, glossiness_tex=black))
    if occluder:
        tm, inv = scenes.translate(OCC_CENTER)
        sc.add_object(scenes.object_desc(capi.SHAPE_DISK, tm, inv, outer_radius=OCC_RADIUS, diffuse_tex=kd,
                                         specular_tex=black, glossiness_tex=black))
    scenes.set_camera(t, (32, 32), (-0.6, -0.4, 1.5), (0.0, 0.0, 0.0), fov=30.0)
    return sc


def _floor_quadrature(sc, P, Wo, weighted, occluder, n=400):
    """Direct light at floor point P by midpoint quadrature over the light's area, written from
    the rendering equation: (Kd / pi) Le |cos_i| |cos_l| / d^2 dA, times the light-sampling MIS
    weight LightPdf^2 / (LightPdf^2 + BsdfPdf^2) when `weighted` (the light-sampling half
    alone), with the BSDF pdf Lambert + Blinn (exponent 0)."""
    L = sc.lights[0]
    tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
    s = (np.arange(n) + 0.5) / n
    uu, vv = np.meshgrid(s, s, indexing="ij")
    local = np.stack([(uu - 0.5) * LIGHT_SIZE, (vv - 0.5) * LIGHT_SIZE, np.zeros_like(uu), np.ones_like(uu)], axis=-1)
    X = (local @ tm.T)[..., :3]
    Nl = tm[:3, :3] @ np.array([0.0, 0.0, 1.0])
    D = X - P
    d2 = np.sum(D * D, axis=-1)
    Wi = D / np.sqrt(d2)[..., None]
    N = np.array([0.0, 0.0, 1.0])
    cos_i = np.abs(Wi @ N)
    cos_l = np.abs(Wi @ Nl)
    A = LIGHT_SIZE * LIGHT_SIZE
    f = np.array([0.8, 0.6, 0.4]) / math.pi
    Le = 5.0     # multiplier x the uniform XYZ (1, 1, 1) texture
    vis = np.ones_like(d2)
    if occluder:
        t = (OCC_CENTER[2] - P[2]) / Wi[..., 2]
        q = P + t[..., None] * Wi
        vis = (np.hypot(q[..., 0] - OCC_CENTER[0], q[..., 1] - OCC_CENTER[1]) > OCC_RADIUS).astype(np.float64)
    integrand = cos_i * cos_l / d2 * vis
    if weighted:
        light_pdf = d2 / (cos_l * A)
        # shading frame: Nn = N, Nu = norm(N x Wo), Nv = norm(N x Nu)
        Nu = np.cross(N, Wo); Nu /= np.linalg.norm(Nu)
        Nv = np.cross(N, Nu); Nv /= np.linalg.norm(Nv)
        wo = np.array([Wo @ Nu, Wo @ Nv, Wo @ N])
        wi = np.stack([Wi @ Nu, Wi @ Nv, Wi @ N], axis=-1)
        lam = np.where(wi[..., 2] * wo[2] > 0, np.abs(wi[..., 2]) / math.pi, 0.0)
        wh = wi + wo
        wh /= np.linalg.norm(wh, axis=-1)[..., None]
        wowh = wh @ wo
        blinn = np.where((wowh > 0) & (wi[..., 2] * wo[2] > 0), 1.0 / (8.0 * math.pi * wowh), 0.0)
        bsdf_pdf = lam + blinn
        integrand = integrand * light_pdf ** 2 / (light_pdf ** 2 + bsdf_pdf ** 2)
    return f * Le * A * integrand.mean()


def _camera_hits(orc, W, H, px, py):
    cam = orc.camera().astype(np.float64)
    Nc, U, V = cam[0:3], cam[3:6], cam[6:9]
    sx = cam[9] + cam[11] * (px + 0.5 / W)      # the mean of the film jitter
    sy = cam[10] + cam[12] * (py + 0.5 / H)
    d = Nc + sx * U - sy * V
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
@pytest.mark.parametrize("occluder", [False, True])
def test_object_direct_light_quadrature(estimator, occluder):
    """An object event is shaded by the BRDF from its textures at the hit (transport.h:141-149).
    COMPAT (light sampling with its MIS weight) must equal the weighted quadrature, FULL (both
    MIS halves) the unweighted rendering-equation integral; a shadowing object
    (transport.h:36-37) removes the light it covers."""
    sc = _floor_scene(estimator, occluder)
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([12, 16, 20, 16], dtype=np.int32)
    py = np.array([16, 12, 16, 20], dtype=np.int32)
    e = orc.expect_f64(px, py, 20000, seed=4)
    for k in range(len(px)):
        D = _camera_hits(orc, W, H, px[k], py[k])
        P = O_cam + (-O_cam[2] / D[2]) * D
        assert abs(P[0]) < 1 and abs(P[1]) < 1
        if occluder:
            q = O_cam + ((OCC_CENTER[2] - O_cam[2]) / D[2]) * D
            assert np.hypot(q[0] - OCC_CENTER[0], q[1] - OCC_CENTER[1]) > OCC_RADIUS + 0.01   # the camera sees the floor
        want = _floor_quadrature(sc, P, -D, estimator == capi.ESTIMATOR_COMPAT, occluder)
        assert e[k, 3] == 1.0
        np.testing.assert_allclose(e[k, :3], want, rtol=0.03, err_msg=f"pixel {px[k]},{py[k]}")


def _round_light_scene(estimator, shape, inner, outer):
    """The floor scene of _floor_scene with a disk or ring light (Source/disk.h, ring.h) of the
    given radii in place of the plane light."""
    sc = _floor_scene(estimator, False)
    sc.lights = []
    sc.tracer.light_count = 0
    sc.add_light(scenes.shape_light(shape, LIGHT_POS, (0.25, 0.05, 0.0), 5.0, texture_id=0, inner=inner, outer=outer,
                                    visible=0))
    return sc


def _round_light_quadrature(sc, P, Wo, full, inner, outer, n=600):
    """Direct light at floor point P from a disk (inner = 0) or ring light, by midpoint
    quadrature in polar coordinates over the light. The light sample is not area-uniform on a
    ring: SampleUnitRing (ring.h:59-67) takes r = a + (1 - a) sqrt(u) with a = inner / outer,
    an area density of (rho - a) / (pi rho (1 - a)^2) on the unit ring, while the estimator
    divides by the uniform density 1 / Area. The light-sampling half therefore carries the
    factor Area * density; the BSDF half (full mode) does not. Compat: the light half with its
    MIS weight; full: both halves with the mixture pdf 0.5 (Lambert + Blinn)."""
    L = sc.lights[0]
    tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
    a = inner / outer
    rr = a + (1.0 - a) * (np.arange(n) + 0.5) / n            # rho in [a, 1]
    th = 2.0 * math.pi * (np.arange(n) + 0.5) / n
    R, TH = np.meshgrid(rr, th, indexing="ij")
    dA = (1.0 - a) / n * (2.0 * math.pi / n) * R * outer * outer
    local = np.stack([R * outer * np.cos(TH), R * outer * np.sin(TH), np.zeros_like(R), np.ones_like(R)], axis=-1)
    X = (local @ tm.T)[..., :3]
    Nl = tm[:3, :3] @ np.array([0.0, 0.0, 1.0])
    D = X - P
    d2 = np.sum(D * D, axis=-1)
    Wi = D / np.sqrt(d2)[..., None]
    N = np.array([0.0, 0.0, 1.0])
    cos_i, cos_l = np.abs(Wi @ N), np.abs(Wi @ Nl)
    area = math.pi * (outer * outer - inner * inner)
    Nu = np.cross(N, Wo); Nu /= np.linalg.norm(Nu)
    Nv = np.cross(N, Nu); Nv /= np.linalg.norm(Nv)
    wo = np.array([Wo @ Nu, Wo @ Nv, Wo @ N])
    wi = np.stack([Wi @ Nu, Wi @ Nv, Wi @ N], axis=-1)
    same = wi[..., 2] * wo[2] > 0
    lam = np.where(same, np.abs(wi[..., 2]) / math.pi, 0.0)
    wh = wi + wo
    wh /= np.linalg.norm(wh, axis=-1)[..., None]
    wowh = wh @ wo
    blinn = np.where((wowh > 0) & same, 1.0 / (8.0 * math.pi * wowh), 0.0)
    bsdf = 0.5 * (lam + blinn) if full else lam + blinn
    light_pdf = d2 / (cos_l * area)
    w_light = light_pdf ** 2 / (light_pdf ** 2 + bsdf ** 2)
    sample_weight = (1.0 + a) * (R - a) / (R * (1.0 - a)) if a > 0 else np.ones_like(R)   # Area x density
    weight = sample_weight * w_light + (bsdf ** 2 / (bsdf ** 2 + light_pdf ** 2) if full else 0.0)
    g = cos_i * cos_l / d2 * weight
    f = np.array([0.8, 0.6, 0.4]) / math.pi
    return f * 5.0 * float(np.sum(g * dA))


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
@pytest.mark.parametrize("shape,inner,outer", [(capi.SHAPE_DISK, 0.0, 0.45), (capi.SHAPE_RING, 0.2, 0.5)])
def test_round_light_direct_light_quadrature(estimator, shape, inner, outer):
    """Disk and ring lights (disk.h, ring.h): their sampling, area and intersection, against a
    quadrature that states the ring's non-uniform sample density explicitly. A ring sampled as
    if it were area-uniform would be off by several per cent here."""
    sc = _round_light_scene(estimator, shape, inner, outer)
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([12, 16, 20, 16], dtype=np.int32)
    py = np.array([16, 12, 16, 20], dtype=np.int32)
    e = orc.expect_f64(px, py, 20000, seed=6)
    for k in range(len(px)):
        D = _camera_hits(orc, W, H, px[k], py[k])
        P = O_cam + (-O_cam[2] / D[2]) * D
        want = _round_light_quadrature(sc, P, -D, estimator == capi.ESTIMATOR_FULL, inner, outer)
        assert e[k, 3] == 1.0
        np.testing.assert_allclose(e[k, :3], want, rtol=0.03, err_msg=f"pixel {px[k]},{py[k]}")


def test_environment_sphere_light():
    """A uniform two-sided sphere light of radius 10 around the scene (sphere.h:152-170) over a
    Lambertian floor object. Full mode, both MIS halves: the radiance is exactly Kd * Le (the
    cosine integral of a uniform environment, pi, times Kd / pi). Compat, light sampling with
    its MIS weight: a quadrature over the sphere."""
    Le = 2.0
    for estimator in (capi.ESTIMATOR_FULL, capi.ESTIMATOR_COMPAT):
        sc = _floor_scene(estimator, False)
        sc.lights = []
        sc.tracer.light_count = 0
        sc.add_light(scenes.sphere_light(10.0, Le, texture_id=0, visible=0))
        orc = O.Oracle(sc, nthreads=4)
        W, H = sc.film
        O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
        px = np.array([14, 18], dtype=np.int32)
        py = np.array([16, 15], dtype=np.int32)
        e = orc.expect_f64(px, py, 40000, seed=8)
        kd = np.array([0.8, 0.6, 0.4])
        for k in range(len(px)):
            assert e[k, 3] == 1.0
            if estimator == capi.ESTIMATOR_FULL:
                np.testing.assert_allclose(e[k, :3], kd * Le, rtol=0.03)
                continue
            D = _camera_hits(orc, W, H, px[k], py[k])
            P = O_cam + (-O_cam[2] / D[2]) * D
            Wo = -D
            n = 800
            th = math.pi * (np.arange(n) + 0.5) / n
            ph = 2.0 * math.pi * (np.arange(n) + 0.5) / n
            TH, PH = np.meshgrid(th, ph, indexing="ij")
            Rs = 10.0
            X = Rs * np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], axis=-1)
            dA = Rs * Rs * np.sin(TH) * (math.pi / n) * (2.0 * math.pi / n)
            Dv = X - P
            d2 = np.sum(Dv * Dv, axis=-1)
            Wi = Dv / np.sqrt(d2)[..., None]
            Nl = X / Rs
            N = np.array([0.0, 0.0, 1.0])
            cos_i, cos_l = np.abs(Wi @ N), np.abs(np.sum(Wi * Nl, axis=-1))
            Nu = np.cross(N, Wo); Nu /= np.linalg.norm(Nu)
            Nv = np.cross(N, Nu); Nv /= np.linalg.norm(Nv)
            wo = np.array([Wo @ Nu, Wo @ Nv, Wo @ N])
            wi = np.stack([Wi @ Nu, Wi @ Nv, Wi @ N], axis=-1)
            same = wi[..., 2] * wo[2] > 0
            lam = np.where(same, np.abs(wi[..., 2]) / math.pi, 0.0)
            wh = wi + wo
            wh /= np.linalg.norm(wh, axis=-1)[..., None]
            wowh = wh @ wo
            blinn = np.where((wowh > 0) & same, 1.0 / (8.0 * math.pi * wowh), 0.0)
            bsdf = lam + blinn
            light_pdf = d2 / (cos_l * 4.0 * math.pi * Rs * Rs)
            w = np.where(bsdf > 0, light_pdf ** 2 / (light_pdf ** 2 + bsdf ** 2), 0.0)
            want = kd / math.pi * Le * float(np.sum(cos_i * cos_l / d2 * w * dA))
            np.testing.assert_allclose(e[k, :3], want, rtol=0.03)


def test_box_light_direct_light_quadrature():
    """A box light (box.h:122-191): SampleUnitBox picks each of the six faces with probability
    1/6 whatever its area, so the light-sampling half carries Area / (6 * face area) per face;
    faces turned away from the shaded point are hidden by the box itself (the shadow test meets
    the box first). Compat and full, against a quadrature over the faces."""
    size = np.array([0.5, 0.3, 0.4])
    for estimator in (capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL):
        full = estimator == capi.ESTIMATOR_FULL
        sc = _floor_scene(estimator, False)
        sc.lights = []
        sc.tracer.light_count = 0
        sc.add_light(scenes.shape_light(capi.SHAPE_BOX, LIGHT_POS, (0.25, 0.05, 0.0), 5.0, texture_id=0,
                                        size=tuple(size), visible=0))
        orc = O.Oracle(sc, nthreads=4)
        W, H = sc.film
        O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
        L = sc.lights[0]
        tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
        px = np.array([12, 16, 20], dtype=np.int32)
        py = np.array([16, 12, 16], dtype=np.int32)
        e = orc.expect_f64(px, py, 20000, seed=9)
        area = 2.0 * (size[0] * size[1] + size[0] * size[2] + size[1] * size[2])
        n = 300
        s = (np.arange(n) + 0.5) / n - 0.5
        A_, B_ = np.meshgrid(s, s, indexing="ij")
        for k in range(len(px)):
            D = _camera_hits(orc, W, H, px[k], py[k])
            P = O_cam + (-O_cam[2] / D[2]) * D
            Wo = -D
            N = np.array([0.0, 0.0, 1.0])
            Nu = np.cross(N, Wo); Nu /= np.linalg.norm(Nu)
            Nv = np.cross(N, Nu); Nv /= np.linalg.norm(Nv)
            wo = np.array([Wo @ Nu, Wo @ Nv, Wo @ N])
            total = 0.0
            for axis in range(3):
                for sign in (1.0, -1.0):
                    u, v = [a for a in range(3) if a != axis]
                    local = np.zeros(A_.shape + (4,))
                    local[..., axis] = sign * 0.5 * size[axis]
                    local[..., u] = A_ * size[u]
                    local[..., v] = B_ * size[v]
                    local[..., 3] = 1.0
                    X = (local @ tm.T)[..., :3]
                    nl = np.zeros(3); nl[axis] = sign
                    Nl = tm[:3, :3] @ nl
                    face_area = size[u] * size[v]
                    Dv = X - P
                    d2 = np.sum(Dv * Dv, axis=-1)
                    Wi = Dv / np.sqrt(d2)[..., None]
                    facing = (Wi @ Nl) < 0            # the face looks at P
                    cos_i, cos_l = np.abs(Wi @ N), np.abs(Wi @ Nl)
                    wi = np.stack([Wi @ Nu, Wi @ Nv, Wi @ N], axis=-1)
                    same = wi[..., 2] * wo[2] > 0
                    lam = np.where(same, np.abs(wi[..., 2]) / math.pi, 0.0)
                    wh = wi + wo
                    wh /= np.linalg.norm(wh, axis=-1)[..., None]
                    wowh = wh @ wo
                    blinn = np.where((wowh > 0) & same, 1.0 / (8.0 * math.pi * wowh), 0.0)
                    bsdf = 0.5 * (lam + blinn) if full else lam + blinn
                    lp = d2 / (cos_l * area)
                    wl = np.where(bsdf > 0, lp ** 2 / (lp ** 2 + bsdf ** 2), 0.0)
                    weight = area / (6.0 * face_area) * wl
                    if full:
                        weight = weight + np.where(bsdf > 0, bsdf ** 2 / (bsdf ** 2 + lp ** 2), 0.0)
                    g = np.where(facing, cos_i * cos_l / d2 * weight, 0.0)
                    total += float(np.sum(g)) * face_area / (n * n)
            want = np.array([0.8, 0.6, 0.4]) / math.pi * 5.0 * total
            assert e[k, 3] == 1.0
            np.testing.assert_allclose(e[k, :3], want, rtol=0.03, err_msg=f"estimator {estimator} pixel {px[k]},{py[k]}")


def _slab_chord(o, d, lo, hi):
    """Length of the ray o + t d (t >= 0) inside the box [lo, hi]."""
    t0, t1 = 0.0, np.inf
    for i in range(3):
        if abs(d[i]) < 1e-12:
            if o[i] < lo[i] or o[i] > hi[i]:
                return 0.0
            continue
        a, b = (lo[i] - o[i]) / d[i], (hi[i] - o[i]) / d[i]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0)


@pytest.mark.parametrize("clip", [False, True])
def test_woodcock_free_flight_and_clipping_closed_form(clip):
    """Delta tracking (full mode) in a homogeneous cube with the majorant equal to the
    extinction sigma = DS^2 * opacity (Q3) samples exact exponential free flights: a camera
    ray's scatter probability is 1 - exp(-sigma L), L its chord through the cube. A clipping
    box (full mode) removes the density inside it: L loses the chord through the clip box."""
    vox = np.full((16, 16, 16), 1000, dtype=np.uint16)
    sc = scenes.Scene(vox)
    t = sc.tracer
    t.opacity.add(0.0, 0.5).add(70000.0, 0.5)
    t.settings.density_scale = 2.0
    t.settings.estimator = capi.ESTIMATOR_FULL
    t.settings.shading_type = capi.SHADING_PHASE_ONLY
    t.settings.sampler = capi.SAMPLER_WOODCOCK
    scenes.set_camera(t, (24, 24), (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), fov=30.0)
    clip_lo, clip_hi = np.array([0.0, -1.0, -1.0]), np.array([0.4, 1.0, 1.0])
    if clip:
        sc.add_clip(scenes.clipping_box(tuple((clip_lo + clip_hi) / 2), tuple(clip_hi - clip_lo)))
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    sigma = 2.0 * 2.0 * 0.5
    O_cam = np.array([t.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([4, 8, 12, 16, 20, 12], dtype=np.int32)
    py = np.array([12, 6, 12, 18, 12, 2], dtype=np.int32)
    e = orc.expect_f64(px, py, 20000, seed=10)
    lo, hi = np.full(3, -0.5), np.full(3, 0.5)
    for k in range(len(px)):
        D = _camera_hits(orc, W, H, px[k], py[k])
        L = _slab_chord(O_cam, D, lo, hi)
        if clip:
            L -= _slab_chord(O_cam, D, np.maximum(lo, clip_lo), np.minimum(hi, clip_hi))
        want = 1.0 - math.exp(-sigma * L)
        assert abs(e[k, 3] - want) < 0.012, (px[k], py[k], e[k, 3], want, L)


def test_shading_type_selection_limits():
    """Full-mode shader selection (design: scatterevent.h:155-212) at its limits: hybrid with
    gradient factor 0 and threshold above every normalised gradient always choose the phase
    function; threshold below 0 always the BRDF. Each must give the expectation of the
    corresponding fixed shading type (the selection's extra draw changes the streams, not the
    expectation)."""
    def expect(shading, factor=1.0, threshold=0.5):
        sc = scenes.config_small(scale=0.25, estimator=capi.ESTIMATOR_FULL)
        sc.tracer.settings.shading_type = shading
        sc.tracer.settings.gradient_factor = factor
        sc.tracer.settings.gradient_threshold = threshold
        orc = O.Oracle(sc, nthreads=4)
        orc.compute_gm_max()
        W, H = sc.film
        px = np.array([W // 2 - 4, W // 2, W // 2 + 3, W // 2], dtype=np.int32)
        py = np.array([H // 2, H // 2 - 3, H // 2 + 2, H // 2 + 5], dtype=np.int32)
        return orc.expect_f64(px, py, 20000, seed=11)

    phase = expect(capi.SHADING_PHASE_ONLY)
    brdf = expect(capi.SHADING_BRDF_ONLY)
    assert np.abs(phase[:, :3] - brdf[:, :3]).max() > 0.1 * np.abs(phase[:, :3]).max()   # the two differ
    for got, want in ((expect(capi.SHADING_HYBRID, factor=0.0), phase),
                      (expect(capi.SHADING_THRESHOLD, threshold=2.0), phase),
                      (expect(capi.SHADING_THRESHOLD, threshold=-1.0), brdf)):
        np.testing.assert_allclose(got[:, :3], want[:, :3], rtol=0.04, atol=0.02 * np.abs(want[:, :3]).max())
        assert np.array_equal(got[:, 3] > 0, want[:, 3] > 0)


def test_woodcock_heterogeneous_volume_matches_the_optical_depth():
    """Delta tracking with per-macrocell majorants on a blob volume, where the majorant exceeds
    the extinction almost everywhere: a camera ray's scatter probability must still be
    1 - exp(-tau), tau the integral of DS^2 * opacity(I(P(t))) along its chord, integrated here
    by quadrature over the oracle's own trilinear samples. A majorant below the extinction
    anywhere on the ray, or a tau spent wrongly across macrocells, biases it."""
    vox = scenes.phantom_blobs((32, 32, 32), 4, seed=7)
    sc = scenes.Scene(vox)
    t = sc.tracer
    t.opacity.add(0.0, 0.0).add(600.0, 0.0).add(2000.0, 0.5).add(4000.0, 1.0)
    t.settings.density_scale = 3.0
    t.settings.estimator = capi.ESTIMATOR_FULL
    t.settings.shading_type = capi.SHADING_PHASE_ONLY
    t.settings.sampler = capi.SAMPLER_WOODCOCK
    scenes.set_camera(t, (24, 24), (1.2, 0.9, 1.6), (0.0, 0.0, 0.0), fov=30.0)
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    O_cam = np.array([t.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([5, 2, 11, 14, 17, 2, 5], dtype=np.int32)
    py = np.array([2, 17, 14, 20, 17, 20, 20], dtype=np.int32)
    e = orc.expect_f64(px, py, 20000, seed=31)
    lo, hi = np.full(3, -0.5), np.full(3, 0.5)
    n = 3000
    for k in range(len(px)):
        D = _camera_hits(orc, W, H, px[k], py[k])
        t0, t1 = 0.0, np.inf
        for i in range(3):
            a, b = (lo[i] - O_cam[i]) / D[i], (hi[i] - O_cam[i]) / D[i]
            t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
        ts = t0 + (np.arange(n) + 0.5) / n * (t1 - t0)
        tau = 0.0
        for tt in ts:
            I = orc.intensity(O_cam + tt * D)
            tau += 9.0 * O.load().orc_plf_eval(C.byref(t.opacity), I)
        tau *= (t1 - t0) / n
        want = 1.0 - math.exp(-tau)
        assert 0.05 < want < 0.98, want
        assert abs(e[k, 3] - want) < 0.012, (px[k], py[k], e[k, 3], want)
