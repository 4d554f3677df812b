"""Pins of the oracle's gradients, specular shading, phase-function scattering and light choice
(CPU only).

Each check states what the mathematics fixes for a case it can compute independently, and is
chosen so that a plausible slip in the oracle fails it:

* gradients (Source/volumes.h:31-114): on a linear ramp volume the central, forward and
  filtered differences, their normalisation and the gradient magnitude have closed forms, at
  lattice points and (anisotropic spacing, delta = MinStep on every axis, Q23) at half-lattice
  points; the maximum over voxel centres is the interior value; a constant volume has a NaN
  normal (Q20);
* Fresnel (Source/shader.h:69-123): normal incidence ((n-1)/(n+1))^2, total internal
  reflection, Brewster's angle, and the angle form of the Fresnel equations
  (sin^2 / tan^2 of the angle difference and sum), which shares no code path with the
  cosine form the oracle uses;
* Blinn (shader.h:166-185): D is normalised, the integral of D cos(theta_h) over the hemisphere
  is 1; the pdf over outgoing directions integrates to 1 - 2^(-(e+1)/2) above the horizon for
  normal viewing, and in general to the probability mass of half vectors facing the viewer,
  computed by a quadrature over half vectors (this checks the 1 / (4 wo.wh) Jacobian);
* Torrance-Sparrow (Microfacet::F, shader.h:216-263): its value at normal incidence and its
  reciprocity;
* a floor object with a specular, glossy BRDF under a plane light: the full estimator (both
  MIS halves) must give the rendering-equation integral, the compat estimator its
  light-sampling half with the power-heuristic weight (transport.h:58-84);
* phase-function single scattering in a homogeneous cube (transport.h:115-156,
  raymarching.h:30-71): the exact expectation of the jittered discrete march (scatter at sample
  k with probability exp(-(k-1) s dt) - exp(-k s dt), s = DS^2 opacity, Q3) times the light
  integral (Kd / pi) * integral of Le cos_l / d^2 over the light (Q7);
* light choice (transport.h:126-155): with three lights the expectation is the sum of the three
  one-light expectations (the choice floor(u * Count) times the Count scaling, Q10), and the
  chosen index is uniform (chi-square over the events of many paths).

The draw order of `Get2`/`Get3` (Q22) is the one thing the mathematics cannot fix: any order
gives the same distributions. It follows clang's left-to-right evaluation and stays unpinned.
"""
import math

import numpy as np
import pytest

import oracle as O
from exposure_render import capi, scenes
from test_oracle import _camera_hits, _slab_chord


# ---- gradients (volumes.h:31-114) --------------------------------------------------------------

A, B, Cz = 6, -10, 14        # even ramp coefficients: half-lattice values stay integers


def _ramp(res, spacing=(1.0, 1.0, 1.0)):
    z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    vox = (2000 + A * x + B * y + Cz * z).astype(np.uint16)
    sc = scenes.Scene(vox, spacing=spacing)
    sc.tracer.opacity.add(0.0, 0.5).add(70000.0, 0.5)
    scenes.set_camera(sc.tracer, (4, 4), (0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    return sc


def test_gradients_on_a_ramp_closed_form():
    res = 16
    orc = O.Oracle(_ramp(res))
    g_len = math.sqrt(A * A + B * B + Cz * Cz)
    for x, y, z in [(2, 2, 2), (7, 9, 4), (13, 13, 13), (5, 11, 8)]:
        p = np.array([x, y, z], dtype=np.float64) / res - 0.5      # lattice point, exact in float
        cd, ncd = orc.gradient(capi.GRADIENT_CENTRAL, p)
        fd, nfd = orc.gradient(capi.GRADIENT_FORWARD, p)
        fl, nfl = orc.gradient(capi.GRADIENT_FILTERED, p)
        # I(P - d) - I(P + d) with d one lattice step; I(P) - I(P + d)
        assert np.array_equal(cd, [-2 * A, -2 * B, -2 * Cz]), cd
        assert np.array_equal(fd, [-A, -B, -Cz]), fd
        assert np.array_equal(fl, [-2 * A, -2 * B, -2 * Cz]), fl
        for n in (ncd, nfd, nfl):
            np.testing.assert_allclose(n, -np.array([A, B, Cz]) / g_len, rtol=0, atol=1e-6)
        # magnitude: the differences times 0.5 / spacing, spacing 1/16
        assert abs(orc.gradient_magnitude(p) - 16.0 * g_len) <= 1e-5 * 16.0 * g_len
    # the maximum over voxel centres is the interior value (edges clamp, differences shrink)
    assert abs(orc.compute_gm_max() - 16.0 * g_len) <= 1e-5 * 16.0 * g_len


def test_gradient_delta_is_the_minimum_spacing_on_every_axis():
    """Spacing (1, 2, 1) on 16^3 normalises to (1/32, 1/16, 1/32): delta = 1/32 is one lattice
    step along x and z but half a step along y (Q23), where the truncating trilinear gives
    v +- B/2, so the y difference is -B and the magnitude's y term (-B) * 0.5 * 16."""
    orc = O.Oracle(_ramp(16, spacing=(1.0, 2.0, 1.0)))
    for x, y, z in [(3, 4, 5), (8, 8, 8), (12, 3, 10)]:
        p = np.array([x / 32.0 - 0.25, y / 16.0 - 0.5, z / 32.0 - 0.25])
        cd, _ = orc.gradient(capi.GRADIENT_CENTRAL, p)
        assert np.array_equal(cd, [-2 * A, -B, -2 * Cz]), cd
        want = math.sqrt((2 * A * 0.5 * 32) ** 2 + (B * 0.5 * 16) ** 2 + (2 * Cz * 0.5 * 32) ** 2)
        assert abs(orc.gradient_magnitude(p) - want) <= 1e-5 * want


def test_constant_volume_has_nan_normal():
    sc = scenes.Scene(np.full((8, 8, 8), 1234, np.uint16))
    sc.tracer.opacity.add(0.0, 0.5).add(70000.0, 0.5)
    scenes.set_camera(sc.tracer, (4, 4), (0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    orc = O.Oracle(sc)
    for mode in (capi.GRADIENT_FORWARD, capi.GRADIENT_CENTRAL, capi.GRADIENT_FILTERED):
        g, n = orc.gradient(mode, (0.05, -0.1, 0.2))
        assert np.all(g == 0.0) and np.all(np.isnan(n))
    assert orc.gradient_magnitude((0.05, -0.1, 0.2)) == 0.0


# ---- Fresnel, Blinn, Torrance-Sparrow (shader.h:69-263) ------------------------------------------

def _fresnel_angles(cosi, n):
    """Unpolarised dielectric reflectance from the angle form of the Fresnel equations:
    R_s = sin^2(ti - tt) / sin^2(ti + tt), R_p = tan^2(ti - tt) / tan^2(ti + tt)."""
    eta_i, eta_t = (1.0, n) if cosi > 0 else (n, 1.0)
    ti = math.acos(min(1.0, abs(cosi)))
    s = eta_i / eta_t * math.sin(ti)
    if s >= 1.0:
        return 1.0
    tt = math.asin(s)
    if ti == 0.0:
        return ((eta_t - eta_i) / (eta_t + eta_i)) ** 2
    rs = math.sin(ti - tt) ** 2 / math.sin(ti + tt) ** 2
    rp = math.tan(ti - tt) ** 2 / math.tan(ti + tt) ** 2
    return 0.5 * (rs + rp)


def test_fresnel_closed_forms():
    assert abs(O.fresnel(1.0, 15.0) - 0.765625) < 1e-7          # ((15 - 1) / (15 + 1))^2
    assert abs(O.fresnel(-1.0, 15.0) - 0.765625) < 1e-7         # leaving at normal incidence
    assert O.fresnel(-0.9, 15.0) == 1.0                          # sin t = 15 sin i > 1: total reflection
    assert O.fresnel(-math.sqrt(1 - (1 / 15.0) ** 2) - 1e-4, 15.0) < 1.0   # just inside the critical angle
    # Brewster's angle: R_p = 0, so R = R_s / 2
    cb = 1.0 / math.sqrt(1.0 + 15.0 * 15.0)
    st = math.sqrt(1 - cb * cb) / 15.0
    ct = math.sqrt(1 - st * st)
    rs = ((cb - 15.0 * ct) / (cb + 15.0 * ct)) ** 2
    assert abs(O.fresnel(cb, 15.0) - 0.5 * rs) < 2e-6
    for n in (1.5, 15.0):
        for c in np.concatenate([np.linspace(0.02, 1.0, 60), -np.linspace(0.02, 1.0, 60)]):
            assert abs(O.fresnel(float(c), n) - _fresnel_angles(float(c), n)) < 3e-6, (n, c)


@pytest.mark.parametrize("e", [0.5, 4.0, 64.0, 1000.0])
def test_blinn_d_is_normalised(e):
    # 2 pi int_0^1 D(mu) mu dmu = 1, by quadrature of the oracle's D on a fine grid
    mus = (np.arange(20000) + 0.5) / 20000
    vals = np.array([O.blinn_d(e, float(m)) for m in mus], dtype=np.float64)
    assert abs(2 * math.pi * np.mean(vals * mus) - 1.0) < (2e-3 if e >= 1000 else 2e-4)


def _blinn_pdf_mass(e, wo, n=600):
    """Integral over the sphere of outgoing directions wi of the oracle's Blinn pdf, by midpoint
    quadrature in (cos theta, phi)."""
    mu = -1.0 + 2.0 * (np.arange(n) + 0.5) / n
    ph = 2 * math.pi * (np.arange(n) + 0.5) / n
    tot = 0.0
    for m in mu:
        s = math.sqrt(max(0.0, 1 - m * m))
        for p in ph[::1]:
            wi = (s * math.cos(p), s * math.sin(p), m)
            tot += O.blinn_pdf(e, wo, wi)
    return tot * (2.0 / n) * (2 * math.pi / n)


@pytest.mark.parametrize("e", [1.0, 3.0, 10.0])
def test_blinn_pdf_integrates_to_the_half_vector_mass(e):
    # normal viewing: wi = reflect(wo, wh) lies above the horizon iff theta_h < 45 degrees
    wo = (0.0, 0.0, 1.0)
    n = 300
    mu = (np.arange(n) + 0.5) / n
    ph = 2 * math.pi * (np.arange(n) + 0.5) / n
    mass_upper = sum(O.blinn_pdf(e, wo, (math.sqrt(1 - m * m) * math.cos(p), math.sqrt(1 - m * m) * math.sin(p), m))
                     for m in mu[::3] for p in ph[::30]) * (1.0 / (n / 3)) * (2 * math.pi / (n / 30))
    assert abs(mass_upper - (1.0 - 2.0 ** (-(e + 1) / 2))) < 0.02
    # oblique viewing: wi -> wh = normalize(wo + wi) maps the sphere onto the half vectors facing
    # wo with Jacobian 1 / (4 wo.wh), so the pdf's integral over all wi is the mass of
    # (e + 1) / (2 pi) |cos theta_h|^e over those half vectors, integrated over wh instead
    th = math.radians(50.0)
    wo = (math.sin(th), 0.0, math.cos(th))
    got = _blinn_pdf_mass(e, wo, n=160)
    m = 4000
    muh = -1.0 + 2.0 * (np.arange(m) + 0.5) / m
    phh = 2 * math.pi * (np.arange(256) + 0.5) / 256
    MU, PH = np.meshgrid(muh, phh, indexing="ij")
    wh = np.stack([np.sqrt(1 - MU ** 2) * np.cos(PH), np.sqrt(1 - MU ** 2) * np.sin(PH), MU], axis=-1)
    facing = (wh @ np.array(wo)) > 0
    want = float(np.sum(np.where(facing, (e + 1) / (2 * math.pi) * np.abs(MU) ** e, 0.0))) * (2.0 / m) * \
        (2 * math.pi / 256)
    assert abs(got - want) < 0.01 * want, (got, want)


def test_microfacet_normal_incidence_and_reciprocity():
    ks = np.array([0.5, 0.25, 1.0])
    for e in (2.0, 12.8, 200.0):
        f = O.microfacet_f((0.0, 0.0, 1.0), (0.0, 0.0, 1.0), ks, e)
        # D = (e + 2) / 2 pi, G = 1, F_r = 0.765625, over 4 cos_i cos_o = 4
        np.testing.assert_allclose(f, ks * (e + 2) / (2 * math.pi) * 0.765625 / 4.0, rtol=2e-6)
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        a[2], b[2] = abs(a[2]) + 0.05, abs(b[2]) + 0.05
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        f1 = O.microfacet_f(a, b, ks, 12.8)
        f2 = O.microfacet_f(b, a, ks, 12.8)
        np.testing.assert_allclose(f1, f2, rtol=2e-5, atol=1e-9)
        assert np.all(f1 >= 0)


# ---- a specular floor under a plane light (transport.h:58-84,141-149) ----------------------------

LIGHT_POS, LIGHT_SIZE, LE = (0.3, 0.1, 0.45), 0.8, 5.0
KD = np.array([0.8, 0.6, 0.4])
KS = np.array([0.5, 0.5, 0.5])
GLOSS = 0.2                       # exponent 1e6 * 0.2^7 = 12.8 (utilities.h:28-31)


def _specular_floor(estimator):
    sc = scenes.config_small(scale=0.25, estimator=estimator)
    t = sc.tracer
    t.opacity.init()
    t.opacity.add(0.0, 0.0).add(70000.0, 0.0)
    sc.textures, sc.lights = [], []
    t.light_count = 0
    white = sc.add_texture(scenes.uniform_texture((1.0, 1.0, 1.0)))
    kd = sc.add_texture(scenes.uniform_texture(tuple(KD)))
    ks = sc.add_texture(scenes.uniform_texture(tuple(KS)))
    gl = sc.add_texture(scenes.uniform_texture((GLOSS, GLOSS, GLOSS)))
    sc.add_light(scenes.plane_light(LIGHT_POS, (0.25, 0.05, 0.0), (LIGHT_SIZE, LIGHT_SIZE), LE, texture_id=white,
                                    visible=0))
    tm, inv = scenes.translate((0.0, 0.0, 0.0))
    sc.add_object(scenes.object_desc(capi.SHAPE_PLANE, tm, inv, size=(2.0, 2.0, 0.0), diffuse_tex=kd,
                                     specular_tex=ks, glossiness_tex=gl))
    scenes.set_camera(t, (32, 32), (-0.6, -0.4, 1.5), (0.0, 0.0, 0.0), fov=30.0)
    return sc


def _textbook_brdf(wi, wo, e, n_ior=15.0):
    """Kd / pi + Torrance-Sparrow with Blinn D, the V-groove G and the dielectric Fresnel term
    (angle form) at the half vector, in the shading frame; returns (f (..., 3), Lambert pdf,
    Blinn pdf)."""
    same = wi[..., 2] * wo[2] > 0
    ci, co = np.abs(wi[..., 2]), abs(wo[2])
    wh = wi + wo
    wh = wh / np.linalg.norm(wh, axis=-1)[..., None]
    ch = np.abs(wh[..., 2])
    wowh = wh @ wo
    D = (e + 2) / (2 * math.pi) * ch ** e
    G = np.minimum(1.0, np.minimum(2 * ch * co / np.abs(wowh), 2 * ch * ci / np.abs(wowh)))
    cos_h = np.sum(wi * wh, axis=-1)
    Fr = np.vectorize(lambda c: _fresnel_angles(float(c), n_ior))(cos_h)
    spec = (D * G * Fr / (4 * ci * co))[..., None] * KS
    f = KD / math.pi + spec
    lam = np.where(same, ci / math.pi, 0.0)
    blinn = np.where((wowh > 0) & same, (e + 1) * ch ** e / (2 * math.pi * 4 * wowh), 0.0)
    return f, lam, blinn


@pytest.mark.parametrize("estimator", [capi.ESTIMATOR_COMPAT, capi.ESTIMATOR_FULL])
def test_specular_floor_direct_light_quadrature(estimator):
    sc = _specular_floor(estimator)
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([12, 16, 20, 18], dtype=np.int32)
    py = np.array([16, 12, 16, 19], dtype=np.int32)
    e_spp = orc.expect_f64(px, py, 40000, seed=14)
    L = sc.lights[0]
    tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
    n = 240
    s = (np.arange(n) + 0.5) / n
    uu, vv = np.meshgrid(s, s, indexing="ij")
    local = np.stack([(uu - 0.5) * LIGHT_SIZE, (vv - 0.5) * LIGHT_SIZE, np.zeros_like(uu), np.ones_like(uu)], axis=-1)
    X = (local @ tm.T)[..., :3]
    Nl = tm[:3, :3] @ np.array([0.0, 0.0, 1.0])
    area = LIGHT_SIZE * LIGHT_SIZE
    e = 1e6 * GLOSS ** 7
    for k in range(len(px)):
        D = _camera_hits(orc, W, H, px[k], py[k])
        P = O_cam + (-O_cam[2] / D[2]) * D
        Wo = -D
        N = np.array([0.0, 0.0, 1.0])
        Nu = np.cross(N, Wo); Nu /= np.linalg.norm(Nu)
        Nv = np.cross(N, Nu); Nv /= np.linalg.norm(Nv)
        wo = np.array([Wo @ Nu, Wo @ Nv, Wo @ N])
        Dv = X - P
        d2 = np.sum(Dv * Dv, axis=-1)
        Wi = Dv / np.sqrt(d2)[..., None]
        wi = np.stack([Wi @ Nu, Wi @ Nv, Wi @ N], axis=-1)
        cos_i, cos_l = np.abs(Wi @ N), np.abs(Wi @ Nl)
        f, lam, blinn = _textbook_brdf(wi, wo, e)
        g = cos_i * cos_l / d2
        if estimator == capi.ESTIMATOR_COMPAT:
            lp = d2 / (cos_l * area)
            bsdf = lam + blinn                                     # Q32: the sum of both pdfs
            g = g * lp ** 2 / (lp ** 2 + bsdf ** 2)
        want = LE * area * np.mean(f * g[..., None], axis=(0, 1))
        assert e_spp[k, 3] == 1.0
        np.testing.assert_allclose(e_spp[k, :3], want, rtol=0.025, err_msg=f"estimator {estimator} pixel {px[k]},{py[k]}")
    # the specular term matters here: without it the compat result would be this much lower
    assert np.max(np.abs(want - LE * area * np.mean((KD / math.pi)[None, None, :] * g[..., None], axis=(0, 1)))) > \
        0.05 * np.max(want)


# ---- phase-function single scattering in a homogeneous cube (transport.h:115-156) -----------------

DS, OPACITY, KD_PHASE = 2.0, 0.5, (0.8, 0.6, 0.4)


def _phase_cube(lights):
    """16^3 cube of one intensity, opacity 0.5 everywhere, density scale 2 (extinction
    DS^2 * opacity = 2), diffuse (0.8, 0.6, 0.4), no emission, shadows off, full estimator with
    phase-function shading. `lights`: (position, target, size, multiplier, XYZ colour)."""
    vox = np.full((16, 16, 16), 1000, dtype=np.uint16)
    sc = scenes.Scene(vox)
    t = sc.tracer
    t.opacity.add(0.0, OPACITY).add(70000.0, OPACITY)
    t.diffuse.add(0.0, KD_PHASE).add(70000.0, KD_PHASE)
    t.specular.add(0.0, (0.0, 0.0, 0.0)).add(70000.0, (0.0, 0.0, 0.0))
    t.glossiness.add(0.0, 0.5).add(70000.0, 0.5)
    t.settings.density_scale = DS
    t.settings.estimator = capi.ESTIMATOR_FULL
    t.settings.shading_type = capi.SHADING_PHASE_ONLY
    t.settings.shadows = 0
    scenes.set_camera(t, (24, 24), (1.3, 0.8, 1.7), (0.0, 0.0, 0.0), fov=30.0)
    for pos, target, size, mult, col in lights:
        tex = sc.add_texture(scenes.uniform_texture(col))
        sc.add_light(scenes.plane_light(pos, target, size, mult, texture_id=tex, visible=0))
    return sc


def _light_integral(L, P, col, n=32):
    """int over the plane light of Le cos_l / d^2 dA at each point of P (..., 3)."""
    tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
    sx, sy = L.shape.size[0], L.shape.size[1]
    s = (np.arange(n) + 0.5) / n
    uu, vv = np.meshgrid(s, s, indexing="ij")
    local = np.stack([(uu - 0.5) * sx, (vv - 0.5) * sy, np.zeros_like(uu), np.ones_like(uu)], axis=-1)
    X = (local @ tm.T)[..., :3].reshape(-1, 3)
    Nl = tm[:3, :3] @ np.array([0.0, 0.0, 1.0])
    Dv = X[None, :, :] - P.reshape(-1, 3)[:, None, :]
    d2 = np.sum(Dv * Dv, axis=-1)
    cos_l = np.abs(Dv @ Nl) / np.sqrt(d2)
    irr = np.mean(cos_l / d2, axis=1) * sx * sy * L.multiplier
    return irr.reshape(P.shape[:-1])[..., None] * np.asarray(col, dtype=np.float64)


def _phase_expectation(sc, orc, px, py, colours, jitter=8):
    """Exact expectation of the jittered march's scatter position (raymarching.h:30-71: the
    first sample at MinT + u2 dt, scatter at sample k with probability e^{-(k-1) s dt} -
    e^{-k s dt}, no event once the sample reaches MaxT) times the phase estimate's expectation
    (Kd / pi) * sum over lights of int Le cos_l / d^2 dA (F = Kd / pi, Q7; unbiased MIS)."""
    W, H = sc.film
    O_cam = np.array([sc.tracer.camera.pos[i] for i in range(3)], dtype=np.float64)
    D = _camera_hits(orc, W, H, px, py)
    lo, hi = np.full(3, -0.5), np.full(3, 0.5)
    t0, t1 = 0.0, np.inf
    for i in range(3):
        a, b = (lo[i] - O_cam[i]) / D[i], (hi[i] - O_cam[i]) / D[i]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    if t1 <= t0:
        return np.zeros(3)
    s = DS * DS * OPACITY
    dt = 0.1 * (1.0 / 16.0)
    kmax = int(math.ceil((t1 - t0) / dt)) + 1
    k = np.arange(1, kmax + 1)
    pk = np.exp(-(k - 1) * s * dt) - np.exp(-k * s * dt)
    u2 = (np.arange(jitter) + 0.5) / jitter
    T = t0 + (u2[None, :] + (k[:, None] - 1)) * dt                     # (k, jitter)
    P = O_cam + T[..., None] * D
    inside = T < t1
    Ld = np.zeros(T.shape + (3,))
    for L, col in zip(sc.lights, colours):
        Ld += _light_integral(L, P, col)
    Ld *= np.asarray(KD_PHASE) / math.pi
    Ld[~inside] = 0.0
    return np.einsum("k,kjc->c", pk, Ld) / jitter


LIGHTS = [((0.9, 1.0, 0.6), (0.0, 0.0, 0.0), (0.6, 0.5), 40.0, (1.0, 0.0, 0.0)),
          ((-1.0, 0.7, -0.3), (0.0, 0.0, 0.0), (0.5, 0.7), 25.0, (0.0, 1.0, 0.0)),
          ((0.2, -1.1, 0.5), (0.0, 0.0, 0.0), (0.4, 0.4), 60.0, (0.0, 0.0, 1.0))]


def test_phase_single_scatter_matches_the_discrete_march_integral():
    sc = _phase_cube([LIGHTS[0][:4] + ((1.0, 1.0, 1.0),)])
    orc = O.Oracle(sc, nthreads=4)
    px = np.array([8, 12, 15, 11], dtype=np.int32)
    py = np.array([12, 12, 9, 15], dtype=np.int32)
    got = orc.expect_f64(px, py, 40000, seed=21)
    for i in range(len(px)):
        want = _phase_expectation(sc, orc, px[i], py[i], [(1.0, 1.0, 1.0)])
        assert want.min() > 0.0
        np.testing.assert_allclose(got[i, :3], want, rtol=0.02, err_msg=f"pixel {px[i]},{py[i]}")


def test_three_lights_expectation_is_the_sum_and_choice_is_uniform():
    three = _phase_cube(LIGHTS)
    orc3 = O.Oracle(three, nthreads=4)
    px = np.array([8, 12, 15], dtype=np.int32)
    py = np.array([12, 12, 9], dtype=np.int32)
    got = orc3.expect_f64(px, py, 40000, seed=22)
    singles = np.zeros((len(px), 3))
    for L in LIGHTS:
        one = O.Oracle(_phase_cube([L]), nthreads=4)
        singles += one.expect_f64(px, py, 40000, seed=23)[:, :3]
    for i in range(len(px)):
        # each light's colour is one XYZ channel: the channels separate the three expectations
        want = _phase_expectation(three, orc3, px[i], py[i], [L[4] for L in LIGHTS])
        np.testing.assert_allclose(got[i, :3], singles[i], rtol=0.03)
        np.testing.assert_allclose(got[i, :3], want, rtol=0.03)
    # chi-square of the chosen light: the non-zero channel of each scatter event's radiance
    W, H = three.film
    ys, xs = np.mgrid[4:20, 4:20]
    qx = xs.ravel().astype(np.int32)
    qy = ys.ravel().astype(np.int32)
    s1, s2 = O.seeds_for(len(qx), 1, 24)
    counts = np.zeros(3)
    for _ in range(60):
        out, _, _ = orc3.trace_pixels(qx, qy, s1, s2)
        ev = out[:, 3] == 1.0
        rad = out[ev, :3]
        assert np.all(np.count_nonzero(rad, axis=1) <= 1)     # one light per path
        lit = rad.max(axis=1) > 0
        counts += np.bincount(np.argmax(rad[lit], axis=1), minlength=3)
    n = counts.sum()
    assert n > 5000
    chi2 = float(np.sum((counts - n / 3) ** 2 / (n / 3)))
    assert chi2 < 13.8, (counts, chi2)       # 2 degrees of freedom, p = 0.001


def test_woodcock_with_shadows_matches_the_transmittance_integral():
    """Woodcock delta tracking for both the camera ray and the shadow ray (full estimator; the
    shadow march on its own stream): in the homogeneous cube the scatter position has density
    s e^{-s (t - t0)} on the camera ray's chord [t0, t1], s = DS^2 opacity, and the light
    arrives through the shadow segment's chord c inside the cube with transmittance e^{-s c}:
    the expectation is the integral of (Kd / pi) Le cos_l / d^2 e^{-s c} over the chord and the
    light. A shadow stream that reused the path's draws, or a biased shadow tracker, would
    miss it."""
    L0 = LIGHTS[0]
    sc = _phase_cube([L0[:4] + ((1.0, 1.0, 1.0),)])
    t = sc.tracer
    t.settings.sampler = capi.SAMPLER_WOODCOCK
    t.settings.shadows = 1
    t.settings.max_shadow_distance = 100.0
    orc = O.Oracle(sc, nthreads=4)
    W, H = sc.film
    s = DS * DS * OPACITY
    O_cam = np.array([t.camera.pos[i] for i in range(3)], dtype=np.float64)
    px = np.array([8, 12, 15, 11], dtype=np.int32)
    py = np.array([12, 12, 9, 15], dtype=np.int32)
    got = orc.expect_f64(px, py, 40000, seed=25)
    L = sc.lights[0]
    tm = np.array([[L.shape.tm[i][j] for j in range(4)] for i in range(4)], dtype=np.float64)
    n = 32
    u = (np.arange(n) + 0.5) / n
    uu, vv = np.meshgrid(u, u, indexing="ij")
    local = np.stack([(uu - 0.5) * L.shape.size[0], (vv - 0.5) * L.shape.size[1], np.zeros_like(uu),
                      np.ones_like(uu)], axis=-1)
    X = (local @ tm.T)[..., :3].reshape(-1, 3)
    Nl = tm[:3, :3] @ np.array([0.0, 0.0, 1.0])
    lo, hi = np.full(3, -0.5), np.full(3, 0.5)
    for i in range(len(px)):
        D = _camera_hits(orc, W, H, px[i], py[i])
        t0, t1 = 0.0, np.inf
        for a in range(3):
            p0, p1 = (lo[a] - O_cam[a]) / D[a], (hi[a] - O_cam[a]) / D[a]
            t0, t1 = max(t0, min(p0, p1)), min(t1, max(p0, p1))
        m = 400
        ts = t0 + (np.arange(m) + 0.5) / m * (t1 - t0)
        dens = s * np.exp(-s * (ts - t0))
        total = 0.0
        for tt, w in zip(ts, dens):
            P = O_cam + tt * D
            Dv = X - P
            d2 = np.sum(Dv * Dv, axis=1)
            dist = np.sqrt(d2)
            cos_l = np.abs(Dv @ Nl) / dist
            chord = np.array([_slab_chord(P, Dv[k] / dist[k], lo, hi) for k in range(len(X))])
            chord = np.minimum(chord, dist)
            irr = np.mean(cos_l / d2 * np.exp(-s * chord)) * L.shape.size[0] * L.shape.size[1] * L.multiplier
            total += w * irr
        total *= (t1 - t0) / m
        want = total * np.asarray(KD_PHASE) / math.pi
        np.testing.assert_allclose(got[i, :3], want, rtol=0.03, err_msg=f"pixel {px[i]},{py[i]}")
