/*
 * er_oracle.cpp - TEST INFRASTRUCTURE. CPU restatement of Exposure Render's single-scattering
 * estimator and per-pass post-processing, used only as the checker of the HIP path (by tests/,
 * by __graft_entry__.smoke() and as bench.py's cpu_baseline). It is never linked into the
 * product library.
 *
 * One template, two instantiations:
 *   Real = float  : follows the reference's FP32 operation order exactly (build with
 *                   -ffp-contract=off). Used for per-sample parity: RNG streams, scatter /
 *                   no-scatter and light-index decisions, sample counts, XYZA radiance.
 *   Real = double : the FP64 oracle. Same estimator, double arithmetic; its high-spp mean is
 *                   the converged image the RMSE gate compares against.
 *
 * Every function cites the reference file:line it restates (paths relative to Source/).
 * Elementary functions: -log(u) is the correctly rounded value (-logl in long double, rounded
 * to float); sin/cos/exp/pow/acos/atan2 in float are the project's er_fmath.h functions, whose
 * accuracy the tests pin against double libm; Real = double uses <cmath>.
 *
 * Parity status per function. Each line names the tests that pin it (tests/test_oracle.py,
 * tests/test_oracle_pins.py); a function marked "unpinned" has no such test.
 *   rng (rng.h:35-61)                     test_rng_mwc_closed_form, test_rng_zero_seed_is_fixed_point
 *   -log(u) (raymarching.h:45,89)         test_neglog_correctly_rounded_exhaustive
 *   float elementary functions            test_fmath_accuracy
 *   plf_eval (plf.h:68-91)                test_plf_closed_form_and_last_node_zero, test_plf_brute_force_random
 *   get_intensity (buffer3d.h:222-242)    test_trilinear_lattice_points_and_layout, test_trilinear_bounds_and_clamp
 *   march (raymarching.h:30-113)          test_march_homogeneous_step_count_and_T_offset,
 *                                         test_march_transmittance_density_scale_squared,
 *                                         test_march_misses_and_shadow_tmax
 *   intersect_box (box.h:51-83)           the march tests above (box entry, misses, MaxT)
 *   camera (camera.h:61-96)               test_camera_frame
 *   gradients FD/CD/Filtered, normalised  test_gradients_on_a_ramp_closed_form,
 *     (volumes.h:31-91)                   test_gradient_delta_is_the_minimum_spacing_on_every_axis,
 *                                         test_constant_volume_has_nan_normal
 *   gradient_magnitude, gm max            test_gradients_on_a_ramp_closed_form,
 *     (volumes.h:93-114, gradientmagnitude.cuh:28-71)  test_gradient_delta_is_the_minimum_spacing_on_every_axis
 *   fresnel (shader.h:69-123)             test_fresnel_closed_forms
 *   blinn_d / blinn_pdf (shader.h:166-185) test_blinn_d_is_normalised,
 *                                         test_blinn_pdf_integrates_to_the_half_vector_mass
 *   microfacet_f (shader.h:216-263)       test_microfacet_normal_incidence_and_reciprocity,
 *                                         test_specular_floor_direct_light_quadrature
 *   Lambert BRDF, MIS, estimate_direct    test_object_direct_light_quadrature, test_specular_floor_direct_light_quadrature
 *     (transport.h:58-113)
 *   phase scattering (shader.h:29-60,     test_phase_single_scatter_matches_the_discrete_march_integral
 *     transport.h:137-139)
 *   light choice, Count scaling           test_three_lights_expectation_is_the_sum_and_choice_is_uniform
 *     (transport.h:115-156)
 *   lights: plane, disk, ring, box,       test_object_direct_light_quadrature, test_round_light_direct_light_quadrature,
 *     sphere (lights.h, plane.h, disk.h,  test_box_light_direct_light_quadrature, test_environment_sphere_light
 *     ring.h, box.h, sphere.h)
 *   textures, bitmaps (textures.h,        test_procedural_textures_closed_form, test_bitmap_texture_full_is_bilinear,
 *     buffer2d.h:227-254)                 test_bitmap_texture_compat_reads_exact_texels_only
 *   composition special cases             test_no_lights_gives_unscaled_emission, test_transparent_volume_is_invisible,
 *     (singlescattering.h:52-102)         test_draw_schedule, test_fp32_mean_matches_fp64_expectation
 *   post (filterframeestimate.cuh, estimate.cuh, tonemap.cuh)  test_post_filter_cma_tonemap_closed_forms
 *   bilateral (filterrunningestimate.cuh) test_bilateral_constant_and_range_limits
 *   node-table transfer function (no counterpart in the reference)  test_node_table_transfer_function
 *   Woodcock and clipping (full mode)     test_woodcock_free_flight_and_clipping_closed_form
 *   shader selection (full mode; design   test_shading_type_selection_limits at the limits; between them the
 *     scatterevent.h:155-212)             hybrid and modulation probabilities are unpinned
 *   draw order within Get2/Get3 (Q22)     unpinned: every order gives the same distributions
 */
#include "../include/er_capi.h"
#include "../exposure-render_amd/csrc/er_fmath.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

/* ---- constants (defines.h:58-77) ------------------------------------------------------ */
static const float PI_F          = 3.141592654f;
static const float INV_PI_F      = 0.31830988618379067154f;
static const float INV_TWO_PI_F  = 0.15915494309189533577f;
static const float INV_FOUR_PI_F = 1.0f / 4.0f * 3.141592654f;   /* Q6: expands to pi/4 */
static const float RAY_EPS       = 0.0001f;
static const float RAY_EPS_2     = 2.0f * 0.0001f;
static const float TRUE_INV_FOUR_PI_F = 0.0795774715459476678844f; /* full mode phase pdf */
static const int   WC_SHIFT = 3;              /* Woodcock majorant macrocells: 8 cells a side */
static const int   WC_MAX_SEGMENTS = 65536;   /* segments per tracking walk (then: no event) */

/* ---- elementary functions ------------------------------------------------------------- */
template<class R> struct M;
template<> struct M<float> {
    static float neglog(float u) { return (float)(-logl((long double)u)); }
    static float log(float x) { return erm::logf_(x); }
    static float exp(float x) { return erm::expf_(x); }
    static float pow(float x, float y) { return erm::powf_(x, y); }
    static float sin(float x) { return erm::sinf_(x); }
    static float cos(float x) { return erm::cosf_(x); }
    static float acos(float x) { return erm::acosf_(x); }
    static float atan2(float y, float x) { return erm::atan2f_(y, x); }
    static float sqrt(float x) { return std::sqrt(x); }
};
template<> struct M<double> {
    static double neglog(double u) { return -std::log(u); }
    static double log(double x) { return std::log(x); }
    static double exp(double x) { return std::exp(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double acos(double x) { return std::acos(x); }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double sqrt(double x) { return std::sqrt(x); }
};

/* ---- vectors (vector.h:47-395; component-wise, left-to-right) --------------------------- */
template<class R> struct V3 {
    R v[3];
    V3() : v{R(0), R(0), R(0)} {}
    V3(R a, R b, R c) : v{a, b, c} {}
    R  operator[](int i) const { return v[i]; }
    R& operator[](int i) { return v[i]; }
};
template<class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return V3<R>(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template<class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return V3<R>(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
template<class R> inline V3<R> operator*(const V3<R>& a, const V3<R>& b) { return V3<R>(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
template<class R> inline V3<R> operator-(const V3<R>& a) { return V3<R>(-a[0], -a[1], -a[2]); }
/* float * Vec3f returns V[i] * F (vector.h:529) */
template<class R> inline V3<R> scale(R f, const V3<R>& a) { return V3<R>(a[0] * f, a[1] * f, a[2] * f); }
template<class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template<class R> inline R length(const V3<R>& a) { return M<R>::sqrt(dot(a, a)); }
template<class R> inline V3<R> normalize(const V3<R>& a) { const R L = length(a); return V3<R>(a[0] / L, a[1] / L, a[2] / L); }
template<class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b)
{
    return V3<R>((a[1] * b[2]) - (a[2] * b[1]), (a[2] * b[0]) - (a[0] * b[2]), (a[0] * b[1]) - (a[1] * b[0]));
}
template<class R> inline R abs_dot(const V3<R>& a, const V3<R>& b) { return std::fabs(dot(a, b)); }
template<class R> inline R dist2(const V3<R>& a, const V3<R>& b) { const V3<R> d = a - b; return dot(d, d); }
template<class R> inline bool is_black(const V3<R>& c) { return c[0] == R(0) && c[1] == R(0) && c[2] == R(0); }
template<class R> inline V3<R> from(const float* f) { return V3<R>(R(f[0]), R(f[1]), R(f[2])); }
template<class R> inline R clampf(R v, R lo, R hi) { return std::max(lo, std::min(v, hi)); }

/* matrix transforms (geometry.h:30-52) */
template<class R> inline V3<R> transform_point(const float tm[4][4], const V3<R>& p)
{
    const R x = p[0], y = p[1], z = p[2];
    return V3<R>(R(tm[0][0]) * x + R(tm[0][1]) * y + R(tm[0][2]) * z + R(tm[0][3]),
                 R(tm[1][0]) * x + R(tm[1][1]) * y + R(tm[1][2]) * z + R(tm[1][3]),
                 R(tm[2][0]) * x + R(tm[2][1]) * y + R(tm[2][2]) * z + R(tm[2][3]));
}
template<class R> inline V3<R> transform_vector(const float tm[4][4], const V3<R>& v)
{
    const R x = v[0], y = v[1], z = v[2];
    return V3<R>(R(tm[0][0]) * x + R(tm[0][1]) * y + R(tm[0][2]) * z,
                 R(tm[1][0]) * x + R(tm[1][1]) * y + R(tm[1][2]) * z,
                 R(tm[2][0]) * x + R(tm[2][1]) * y + R(tm[2][2]) * z);
}

/* Ray (ray.h:27-55); operator() re-normalises D (Q34) */
template<class R> struct Ray {
    V3<R> O, D; R MinT, MaxT;
    Ray() : O(), D(R(0), R(0), R(1)), MinT(R(0)), MaxT(R(1000000.0f)) {}
    Ray(const V3<R>& o, const V3<R>& d, R mn, R mx) : O(o), D(d), MinT(mn), MaxT(mx) {}
    V3<R> at(R t) const { return O + scale(t, normalize(D)); }
};

/* TransformRay (geometry.h:54-68) */
template<class R> inline Ray<R> transform_ray(const float tm[4][4], const Ray<R>& r)
{
    const V3<R> P = transform_point(tm, r.O);
    const V3<R> MinP = transform_point(tm, r.at(r.MinT));
    const V3<R> MaxP = transform_point(tm, r.at(r.MaxT));
    Ray<R> t;
    t.O = P;
    t.D = normalize(MaxP - t.O);
    t.MinT = length(MinP - t.O);
    t.MaxT = length(MaxP - t.O);
    return t;
}

/* Intersection (intersection.h:27-77) */
template<class R> struct Isect {
    bool Valid = false, Front = true; R NearT = R(0), FarT = R(FLT_MAX);
    V3<R> P, N; R UV[2] = {R(0), R(0)};
};

/* ---- RNG (rng.h:26-66). Q22: Get2/Get3 draw left to right. ---------------------------- */
struct Rng;
static Rng shadow_stream(const Rng& r);
struct Rng {
    uint32_t s0, s1;
    uint32_t draws = 0;
    float get1()
    {
        s0 = 36969u * (s0 & 65535u) + (s0 >> 16);
        s1 = 18000u * (s1 & 65535u) + (s1 >> 16);
        const uint32_t ires = (s0 << 16) + s1;
        const uint32_t b = (ires & 0x007fffffu) | 0x40000000u;
        float f; std::memcpy(&f, &b, 4);
        ++draws;
        return (f - 2.f) / 2.f;
    }
};

/* The stream of a Woodcock shadow march (full mode; the march's draw count is not known before
 * it runs): SplitMix64's finaliser of the path's state (s0 high, s1 low, xor a constant), in two
 * non-zero halves. The path's own stream advances by the two draws of a ray-march shadow test. */
static Rng shadow_stream(const Rng& r)
{
    uint64_t z = (((uint64_t)r.s0 << 32) | (uint64_t)r.s1) ^ 0x5368616457636b21ull;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    Rng s{(uint32_t)z ? (uint32_t)z : 1u, (uint32_t)(z >> 32) ? (uint32_t)(z >> 32) : 1u};
    return s;
}

/* ---- scene --------------------------------------------------------------------------- */
struct Scene {
    er_tracer_desc t;
    /* derived camera (camera.h:61-96) */
    float camN[3], camU[3], camV[3], screen[2][2], invscreen[2], focal;
    /* volume (volume.h:100-129) */
    int res[3];
    std::vector<uint16_t> vox;
    float spacing[3], size[3], invsize[3], minp[3], maxp[3], minstep;
    std::vector<er_light_desc> lights;     /* index = light ID */
    std::vector<er_texture_desc> textures; /* index = texture ID */
    std::vector<er_clipping_desc> clips;   /* index = clipping ID */
    std::vector<er_object_desc> objects;   /* index = object ID, the reference's device list order */
    struct Bitmap { int w = 0, h = 0; std::vector<uint8_t> rgba; };
    std::vector<Bitmap> bitmaps;           /* index = bitmap ID */
    float gm_max = 0.0f;
    /* Woodcock majorants: max sigma over the intensities a sample can take in each macrocell
     * of 2^WC_SHIFT cells (the GPU's wc_sigma grid) */
    std::vector<float> wc_sigma;
    int wc_dim[3] = {0, 0, 0};
    std::vector<float> tf_lut;             /* node table (er_set_tf_lut), empty: the PLFs */
    float tf_scale = 0.0f;
};

static float shape_area(const er_shape& s)   /* shape.h:28-56,91-103 */
{
    switch (s.type) {
        case ER_SHAPE_PLANE:    return s.size[0] * s.size[1];
        case ER_SHAPE_DISK:     return PI_F * (s.outer_radius * s.outer_radius);
        case ER_SHAPE_RING:     return PI_F * (s.outer_radius * s.outer_radius) - PI_F * (s.inner_radius * s.inner_radius);
        case ER_SHAPE_BOX:      return (2.0f * s.size[0] * s.size[1]) + (2.0f * s.size[0] * s.size[2]) + (2.0f * s.size[1] * s.size[2]);
        case ER_SHAPE_SPHERE:   return 4.0f * PI_F * (s.outer_radius * s.outer_radius);
        case ER_SHAPE_CYLINDER: return (2.0f * PI_F * (s.outer_radius * s.outer_radius)) + (2.0f * PI_F * s.outer_radius * s.size[2]);
    }
    return s.area;
}

static void camera_update(Scene& S)   /* camera.h:61-96 */
{
    const er_camera& c = S.t.camera;
    V3<float> pos = from<float>(c.pos), target = from<float>(c.target), up = from<float>(c.up);
    V3<float> N = normalize(target - pos);
    V3<float> U = normalize(cross(N, up));
    V3<float> V = normalize(cross(N, U));
    for (int i = 0; i < 3; ++i) { S.camN[i] = N[i]; S.camU[i] = U[i]; S.camV[i] = V[i]; }
    S.focal = c.focal_distance;
    if (S.focal == -1.0f) S.focal = length(target - pos);
    const float Scale = tanf((0.5f * c.fov / 57.29577951308232f));
    const float Aspect = (float)c.film_size[1] / (float)c.film_size[0];
    if (Aspect > 1.0f) {
        S.screen[0][0] = -Scale; S.screen[0][1] = Scale;
        S.screen[1][0] = -Scale * Aspect; S.screen[1][1] = Scale * Aspect;
    } else {
        S.screen[0][0] = -Scale / Aspect; S.screen[0][1] = Scale / Aspect;
        S.screen[1][0] = -Scale; S.screen[1][1] = Scale;
    }
    S.invscreen[0] = (S.screen[0][1] - S.screen[0][0]) / (float)c.film_size[0];
    S.invscreen[1] = (S.screen[1][1] - S.screen[1][0]) / (float)c.film_size[1];
}

static void volume_update(Scene& S, const er_volume_desc& v)   /* volume.h:100-129 */
{
    for (int i = 0; i < 3; ++i) S.res[i] = v.resolution[i];
    float scale = 0.0f;
    if (v.normalize_size) {
        const float ps[3] = {(float)S.res[0] * v.spacing[0], (float)S.res[1] * v.spacing[1], (float)S.res[2] * v.spacing[2]};
        scale = 1.0f / std::max(ps[0], std::max(ps[1], ps[2]));
    }
    for (int i = 0; i < 3; ++i) {
        S.spacing[i] = v.spacing[i] * scale;
        S.size[i] = (float)S.res[i] * S.spacing[i];
        S.invsize[i] = 1.0f / S.size[i];
        S.minp[i] = S.size[i] * -0.5f;
        S.maxp[i] = S.size[i] * 0.5f;
    }
    S.minstep = std::min(S.spacing[0], std::min(S.spacing[1], S.spacing[2]));
}

/* ---- transfer functions (plf.h:68-91, transferfunction.h:110-113,148-151) ----------- */
template<class R> inline R plf_eval(const er_plf& p, R x)
{
    if (p.count <= 0) return R(0);
    if (x < R(p.node_range[0])) return R(p.value[0]);
    if (x > R(p.node_range[1])) return R(p.value[p.count - 1]);
    for (int i = 1; i < p.count; ++i) {
        const R P1 = R(p.position[i - 1]);
        const R P2 = R(p.position[i]);
        const R DeltaP = P2 - P1;
        const R LerpT = (x - P1) / DeltaP;
        if (x >= P1 && x < P2)
            return R(p.value[i - 1]) + LerpT * (R(p.value[i]) - R(p.value[i - 1]));
    }
    return R(0);   /* Q5: x == last node position */
}
template<class R> inline V3<R> ctf_eval(const er_color_tf& c, R x)
{
    return V3<R>(plf_eval(c.plf[0], x), plf_eval(c.plf[1], x), plf_eval(c.plf[2], x));
}

/* Node-table transfer function (er_capi.h er_set_tf_lut): node n at intensity n / scale,
 * linear between nodes, the last node's value beyond it, negative values clamped to 0. */
template<class R> inline R tf_lut_eval(const std::vector<float>& lut, float scale, R I, int ch)
{
    const int last = ER_TF_LUT_NODES - 1;
    const R x = std::min(I * R(scale), R(last));
    int b = (int)x;
    if (b > last - 1) b = last - 1;
    const R f = x - R(b);
    const R v = (R(1) - f) * R(lut[b * 4 + ch]) + f * R(lut[(b + 1) * 4 + ch]);
    return std::max(v, R(0));
}
template<class R> inline R opacity_at(const Scene& S, R I)
{
    return S.tf_lut.empty() ? plf_eval(S.t.opacity, I) : tf_lut_eval(S.tf_lut, S.tf_scale, I, 0);
}
template<class R> inline V3<R> diffuse_at(const Scene& S, R I)
{
    if (S.tf_lut.empty()) return ctf_eval(S.t.diffuse, I);
    return V3<R>(tf_lut_eval(S.tf_lut, S.tf_scale, I, 1), tf_lut_eval(S.tf_lut, S.tf_scale, I, 2),
                 tf_lut_eval(S.tf_lut, S.tf_scale, I, 3));
}

/* ---- estimator ----------------------------------------------------------------------- */
struct Counters { uint64_t c[8]; };   /* er_counters field order */

template<class R> struct Est {
    const Scene& S;
    Rng& rng;
    uint64_t* cnt;   /* paths, primary, shadow, shading, primary_fetch, shadow_fetch, scatters, shadow_rays */
    bool full;

    Est(const Scene& s, Rng& r, uint64_t* c) : S(s), rng(r), cnt(c), full(s.t.settings.estimator == ER_ESTIMATOR_FULL) {}

    R u() { return R(rng.get1()); }

    /* Buffer3D<ushort>::operator()(Vec3f) (buffer3d.h:222-242) via Volume::operator()
     * (volume.h:131-138): corner-aligned lattice, clamped fetches, 7 lerps each truncated
     * to uint16 (Q4, geometry.h:87-90). */
    R get_intensity(const V3<R>& P) const
    {
        const R L[3] = {((P[0] - R(S.minp[0])) * R(S.invsize[0])) * R((float)S.res[0]),
                        ((P[1] - R(S.minp[1])) * R(S.invsize[1])) * R((float)S.res[1]),
                        ((P[2] - R(S.minp[2])) * R(S.invsize[2])) * R((float)S.res[2])};
        const int vx = (int)std::floor(L[0]), vy = (int)std::floor(L[1]), vz = (int)std::floor(L[2]);
        const R dx = L[0] - R(vx), dy = L[1] - R(vy), dz = L[2] - R(vz);
        auto at = [&](int x, int y, int z) -> R {
            x = std::max(0, std::min(x, S.res[0] - 1));
            y = std::max(0, std::min(y, S.res[1] - 1));
            z = std::max(0, std::min(z, S.res[2] - 1));
            return R(S.vox[(size_t)z * S.res[0] * S.res[1] + (size_t)y * S.res[0] + x]);
        };
        auto lerp = [](R t, R a, R b) -> R { return R((uint16_t)((R(1) - t) * a + t * b)); };
        const R d00 = lerp(dx, at(vx, vy, vz), at(vx + 1, vy, vz));
        const R d10 = lerp(dx, at(vx, vy + 1, vz), at(vx + 1, vy + 1, vz));
        const R d01 = lerp(dx, at(vx, vy, vz + 1), at(vx + 1, vy, vz + 1));
        const R d11 = lerp(dx, at(vx, vy + 1, vz + 1), at(vx + 1, vy + 1, vz + 1));
        const R d0 = lerp(dy, d00, d10);
        const R d1 = lerp(dy, d01, d11);
        return lerp(dz, d0, d1);
    }

    /* GradientCD / FD / Filtered (volumes.h:31-74) */
    V3<R> grad_cd(const V3<R>& P) const
    {
        const R d = R(S.minstep), z = R(0);
        const V3<R> dX(d, z, z), dY(z, d, z), dZ(z, z, d);
        const R i00 = get_intensity(P + dX), i01 = get_intensity(P - dX);
        const R i10 = get_intensity(P + dY), i11 = get_intensity(P - dY);
        const R i20 = get_intensity(P + dZ), i21 = get_intensity(P - dZ);
        return V3<R>(i01 - i00, i11 - i10, i21 - i20);
    }
    V3<R> grad_fd(const V3<R>& P) const
    {
        const R d = R(S.minstep), z = R(0);
        const R i0 = get_intensity(P);
        const R i1 = get_intensity(P + V3<R>(d, z, z));
        const R i2 = get_intensity(P + V3<R>(z, d, z));
        const R i3 = get_intensity(P + V3<R>(z, z, d));
        return V3<R>(i0 - i1, i0 - i2, i0 - i3);
    }
    static V3<R> vlerp(const V3<R>& a, const V3<R>& b, R c) { return a + scale(c, b - a); }   /* vector.h:548 */
    V3<R> grad_filtered(const V3<R>& P) const
    {
        const R o = R(S.minstep);
        const V3<R> G0 = grad_cd(P);
        const V3<R> G1 = grad_cd(P + V3<R>(-o, -o, -o));
        const V3<R> G2 = grad_cd(P + V3<R>(o, o, o));
        const V3<R> G3 = grad_cd(P + V3<R>(-o, o, -o));
        const V3<R> G4 = grad_cd(P + V3<R>(o, -o, o));
        const V3<R> G5 = grad_cd(P + V3<R>(-o, -o, o));
        const V3<R> G6 = grad_cd(P + V3<R>(o, o, -o));
        const V3<R> G7 = grad_cd(P + V3<R>(-o, o, o));
        const V3<R> G8 = grad_cd(P + V3<R>(o, -o, -o));
        const R h = R(0.5f);
        const V3<R> L0 = vlerp(vlerp(G1, G2, h), vlerp(G3, G4, h), h);
        const V3<R> L1 = vlerp(vlerp(G5, G6, h), vlerp(G7, G8, h), h);
        return vlerp(G0, vlerp(L0, L1, h), R(0.75f));
    }
    int gradient_samples() const
    {
        switch (S.t.settings.gradient_computation) {
            case ER_GRADIENT_FORWARD: return 4;
            case ER_GRADIENT_CENTRAL: return 6;
            case ER_GRADIENT_FILTERED: return 54;
        }
        return 4;
    }
    V3<R> normalized_gradient(const V3<R>& P) const   /* volumes.h:76-91 (Q20: zero -> NaN) */
    {
        switch (S.t.settings.gradient_computation) {
            case ER_GRADIENT_FORWARD: return normalize(grad_fd(P));
            case ER_GRADIENT_CENTRAL: return normalize(grad_cd(P));
            case ER_GRADIENT_FILTERED: return normalize(grad_filtered(P));
        }
        return normalize(grad_fd(P));
    }
    /* GradientMagnitude (volumes.h:93-114) */
    R gradient_magnitude(const V3<R>& P) const
    {
        const R d = R(S.minstep), z = R(0);
        const V3<R> dl[3] = {V3<R>(d, z, z), V3<R>(z, d, z), V3<R>(z, z, d)};
        R Sum = R(0);
        for (int i = 0; i < 3; ++i) {
            R D = get_intensity(P - dl[i]) - get_intensity(P + dl[i]);
            D *= R(0.5f) / R(S.spacing[i]);
            Sum += D * D;
        }
        return M<R>::sqrt(Sum);
    }

    /* ---- clipping (full mode; box.h:116-120, sphere.h:147-150, plane.h:90-93, cylinder.h:96-99) */
    bool clipped(const V3<R>& P) const
    {
        for (int k = 0; k < S.t.clipping_count; ++k) {
            const int id = S.t.clipping_ids[k];
            if (id < 0 || id >= (int)S.clips.size()) continue;
            const er_clipping_desc& c = S.clips[id];
            const V3<R> Pl = transform_point(c.shape.inv_tm, P);
            bool inside = false;
            switch (c.shape.type) {
                case ER_SHAPE_BOX: {
                    const R h0 = R(0.5f * c.shape.size[0]), h1 = R(0.5f * c.shape.size[1]), h2 = R(0.5f * c.shape.size[2]);
                    inside = Pl[0] > -h0 && Pl[0] < h0 && Pl[1] > -h1 && Pl[1] < h1 && Pl[2] > -h2 && Pl[2] < h2;
                    break;
                }
                case ER_SHAPE_SPHERE: inside = length(Pl) < R(c.shape.outer_radius); break;
                case ER_SHAPE_CYLINDER:
                    inside = M<R>::sqrt((Pl[0] * Pl[0]) + (Pl[2] * Pl[2])) < R(c.shape.outer_radius) &&
                             std::fabs(Pl[1]) < R(0.5f * c.shape.size[1]);
                    break;
                default: inside = Pl[2] > R(0); break;   /* plane, disk, ring: half-space */
            }
            if (c.invert ? !inside : inside) return true;
        }
        return false;
    }

    /* IntersectBox (box.h:51-83): slab test; fminf/fmaxf semantics for the per-axis min/max */
    static bool intersect_box(const Ray<R>& r, const V3<R>& mn, const V3<R>& mx, R& nearT, R& farT)
    {
        const V3<R> InvR(R(1) / r.D[0], R(1) / r.D[1], R(1) / r.D[2]);
        const V3<R> B = InvR * (mn - r.O);
        const V3<R> T = InvR * (mx - r.O);
        R mnT[3], mxT[3];
        for (int i = 0; i < 3; ++i) { mnT[i] = std::fmin(T[i], B[i]); mxT[i] = std::fmax(T[i], B[i]); }
        const R LargestMinT = std::fmax(std::fmax(mnT[0], mnT[1]), std::fmax(mnT[0], mnT[2]));
        const R LargestMaxT = std::fmin(std::fmin(mxT[0], mxT[1]), std::fmin(mxT[0], mxT[2]));
        if (LargestMaxT < LargestMinT) return false;
        nearT = LargestMinT > R(0) ? LargestMinT : R(0);
        farT = LargestMaxT;
        if (nearT < r.MinT || nearT > r.MaxT) return false;
        return true;
    }

    R sigma_at(const V3<R>& Ps, bool primary)
    {
        if (full && clipped(Ps)) return R(0);
        const R I = get_intensity(Ps);
        if (cnt) cnt[primary ? 4 : 5]++;
        return R(S.t.settings.density_scale) * opacity_at(S, I);
    }

    /* SampleVolume (raymarching.h:30-71) / ScatterEventInVolume (raymarching.h:73-113):
     * jittered fixed-step tracking. primary: exit when t >= tmax and report the event with
     * T one step ahead of P (Q13); shadow: exit when t > tmax (Q14). */
    bool march(const Ray<R>& r, bool primary, R& Tout, V3<R>& Pout)
    {
        const V3<R> mn = from<R>(S.minp), mx = from<R>(S.maxp);
        R nearT, farT;
        if (!intersect_box(r, mn, mx, nearT, farT)) return false;
        R MinT = std::max(nearT, r.MinT);
        const R MaxT = std::min(farT, r.MaxT);
        if (full && S.t.settings.sampler == ER_SAMPLER_WOODCOCK)
            return woodcock(r, primary, MinT, MaxT, Tout, Pout);
        const R Sthr = M<R>::neglog(u()) / R(S.t.settings.density_scale);
        R Sum = R(0);
        const R StepSize = R(primary ? S.t.settings.step_factor_primary : S.t.settings.step_factor_shadow) * R(S.minstep);
        MinT += u() * StepSize;
        /* a march that could not advance (zero-size volume, Q2, or a step below the spacing of
         * floats at MaxT) would not terminate; it is reported as no event */
        if (!(StepSize > R(0)) || !(MaxT + StepSize > MaxT)) return false;
        V3<R> Ps;
        while (Sum < Sthr) {
            Ps = r.O + scale(MinT, r.D);
            if (primary ? (MinT >= MaxT) : (MinT > MaxT)) return false;
            if (cnt) cnt[primary ? 1 : 2]++;
            const R SigmaT = sigma_at(Ps, primary);
            Sum += SigmaT * StepSize;
            MinT += StepSize;
        }
        Tout = MinT;
        Pout = Ps;
        return true;
    }

    /* Woodcock delta tracking with a piecewise constant majorant (full mode only; design from
     * the commented sampler at raymarching.h:115-273: macrocells :155, majorant texture :205,
     * acceptance sigma / sigma_max > u :261). The effective extinction of the ray-marcher is
     * DS^2 * opacity (Q3); its majorant in a macrocell of 2^WC_SHIFT cells is DS * the largest
     * sigma_lut value the macrocell's samples can take. One optical depth -log(u) per tentative
     * collision is spent across macrocells: the collision is at t + tau / mu inside the current
     * segment [t, t_exit), else tau loses mu (t_exit - t) and the walk moves on. Segments end
     * at the nearest macrocell face in lattice coordinates (edge macrocells extend to the end
     * of the ray, as the clamped lookups do; lattice distance times the reciprocal of the ray's
     * lattice rate) and advance t by at least a millionth of it. */
    static int wc_axis(R L, int dim) { return std::min(std::max(((int)std::floor(L)) >> WC_SHIFT, 0), dim - 1); }
    bool woodcock(const Ray<R>& r, bool primary, R MinT, R MaxT, R& Tout, V3<R>& Pout)
    {
        const R DS = R(S.t.settings.density_scale);
        const R res[3] = {R((float)S.res[0]), R((float)S.res[1]), R((float)S.res[2])};
        const R g[3] = {(r.D[0] * R(S.invsize[0])) * res[0], (r.D[1] * R(S.invsize[1])) * res[1],
                        (r.D[2] * R(S.invsize[2])) * res[2]};
        const R ig[3] = {g[0] != R(0) ? R(1) / g[0] : R(0), g[1] != R(0) ? R(1) / g[1] : R(0),
                         g[2] != R(0) ? R(1) / g[2] : R(0)};
        const int* dim = S.wc_dim;
        const R big = R(3.402823466e+38f);
        R t = MinT;
        int segments = 0;
        for (;;) {
            R tau = -M<R>::log(u());   /* the project's logf_ in FP32, as the GPU walk */
            R mu = R(0);
            for (;;) {
                if (++segments > WC_MAX_SEGMENTS || !(t < MaxT)) return false;
                const V3<R> P = r.O + scale(t, r.D);
                const R L[3] = {((P[0] - R(S.minp[0])) * R(S.invsize[0])) * res[0],
                                ((P[1] - R(S.minp[1])) * R(S.invsize[1])) * res[1],
                                ((P[2] - R(S.minp[2])) * R(S.invsize[2])) * res[2]};
                int m[3];
                R e[3];
                for (int a = 0; a < 3; ++a) {
                    m[a] = wc_axis(L[a], dim[a]);
                    if (g[a] > R(0)) e[a] = m[a] == dim[a] - 1 ? big : (R((float)((m[a] + 1) << WC_SHIFT)) - L[a]) * ig[a];
                    else if (g[a] < R(0)) e[a] = m[a] == 0 ? big : (R((float)(m[a] << WC_SHIFT)) - L[a]) * ig[a];
                    else e[a] = big;
                }
                mu = DS * R(S.wc_sigma[((size_t)m[2] * dim[1] + m[1]) * dim[0] + m[0]]);
                const R t_exit = std::fmax(std::fmin(t + std::fmin(e[0], std::fmin(e[1], e[2])), MaxT),
                                           t + std::fmax(t * R(1.0e-6f), R(1.0e-7f)));
                if (mu > R(0)) {
                    const R tc = t + tau / mu;
                    if (tc < t_exit) { t = tc; break; }
                    tau = std::fmax(tau - mu * (t_exit - t), R(0));
                }
                t = t_exit;
            }
            if (primary ? (t >= MaxT) : (t > MaxT)) return false;
            const V3<R> Ps = r.O + scale(t, r.D);
            if (cnt) cnt[primary ? 1 : 2]++;
            const R sigma = sigma_at(Ps, primary) * DS;
            if (u() * mu < sigma) { Tout = t; Pout = Ps; return true; }
        }
    }

    /* ---- shapes (plane.h:41-88, sphere.h:84-140, box.h:51-83, disk.h, ring.h) --------- */
    static void intersect_plane(const Ray<R>& r, bool one_sided, Isect<R>& I)
    {
        if (std::fabs(r.O[2] - r.D[2]) < R(RAY_EPS)) return;          /* Q16 */
        I.NearT = (R(0) - r.O[2]) / r.D[2];
        if (I.NearT < r.MinT || I.NearT > r.MaxT) return;
        I.P = r.at(I.NearT);
        I.UV[0] = I.P[0]; I.UV[1] = I.P[1];
        I.N = V3<R>(R(0), R(0), R(1));
        if (one_sided && r.D[2] >= R(0)) { I.Front = false; I.N = V3<R>(R(0), R(0), R(-1)); }
        I.Valid = true;
    }
    static void intersect_sphere(const Ray<R>& r, R radius, Isect<R>& I)
    {
        const R a = dot(r.D, r.D);
        const R b = R(2) * dot(r.D, r.O);
        const R c = dot(r.O, r.O) - (radius * radius);
        const R disc = b * b - R(4) * a * c;
        if (disc < R(0)) return;
        const R distSqrt = M<R>::sqrt(disc);
        const R q = (b < R(0)) ? (-b - distSqrt) / R(2) : (-b + distSqrt) / R(2);
        R t0 = q / a, t1 = c / q;
        if (t0 > t1) std::swap(t0, t1);
        if (t0 >= r.MinT && t0 < r.MaxT) I.NearT = t0;
        else if (t1 >= r.MinT && t1 < r.MaxT) I.NearT = t1;
        else return;
        if (I.NearT < r.MinT || I.NearT > r.MaxT) return;
        I.Valid = true;
        I.P = r.at(I.NearT);
        I.N = normalize(I.P);
        /* SphericalToUV (geometry.h:70-83) */
        const V3<R> V = normalize(I.P);
        R phi = M<R>::atan2(V[2], V[0]);
        phi = (phi < R(0)) ? phi + R(2.0f * PI_F) : phi;
        const R theta = M<R>::acos(clampf(V[1], R(-1), R(1)));
        I.UV[0] = R(INV_TWO_PI_F) * phi;
        I.UV[1] = R(1) - (R(INV_PI_F) * theta);
    }
    static void intersect_shape(const er_shape& s, const Ray<R>& r, Isect<R>& I)
    {
        switch (s.type) {
            case ER_SHAPE_PLANE: {
                intersect_plane(r, s.one_sided != 0, I);
                const R hx = R(0.5f * s.size[0]), hy = R(0.5f * s.size[1]);
                if (I.Valid && (I.UV[0] < -hx || I.UV[0] > hx || I.UV[1] < -hy || I.UV[1] > hy)) I.Valid = false;
                I.UV[0] /= R(s.size[0]); I.UV[1] /= R(s.size[1]);
                I.UV[0] += R(0.5f); I.UV[1] += R(0.5f);
                I.UV[0] = R(1) - I.UV[0];
                break;
            }
            case ER_SHAPE_DISK: {
                intersect_plane(r, s.one_sided != 0, I);
                const R l = M<R>::sqrt(I.UV[0] * I.UV[0] + I.UV[1] * I.UV[1]);
                if (I.Valid && l > R(s.outer_radius)) I.Valid = false;
                const R dia = R(2.0f * s.outer_radius);
                I.UV[0] /= dia; I.UV[1] /= dia;
                I.UV[0] += R(0.5f); I.UV[1] += R(0.5f);
                I.UV[0] = R(1) - I.UV[0];
                break;
            }
            case ER_SHAPE_RING: {
                intersect_plane(r, s.one_sided != 0, I);
                const R l = M<R>::sqrt(I.UV[0] * I.UV[0] + I.UV[1] * I.UV[1]);
                if (I.Valid && (l < R(s.inner_radius) || l > R(s.outer_radius))) I.Valid = false;
                const R dia = R(2.0f * s.outer_radius);
                I.UV[0] /= dia; I.UV[1] /= dia;
                I.UV[0] += R(0.5f); I.UV[1] += R(0.5f);
                I.UV[0] = R(1) - I.UV[0];
                break;
            }
            case ER_SHAPE_BOX: {
                const V3<R> mx(R(0.5f * s.size[0]), R(0.5f * s.size[1]), R(0.5f * s.size[2]));
                const V3<R> mn(R(-0.5f * s.size[0]), R(-0.5f * s.size[1]), R(-0.5f * s.size[2]));
                R nT, fT;
                if (intersect_box(r, mn, mx, nT, fT)) {
                    I.NearT = nT; I.FarT = fT; I.Valid = true;
                    I.P = r.at(nT);
                    I.N = V3<R>();
                    for (int i = 0; i < 3; ++i) {
                        if (I.P[i] <= mn[i] + R(0.0001f)) I.N[i] = R(-1);
                        if (I.P[i] >= mx[i] - R(0.0001f)) I.N[i] = R(1);
                    }
                }
                break;
            }
            case ER_SHAPE_SPHERE: intersect_sphere(r, R(s.outer_radius), I); break;
            default: break;
        }
    }

    /* surface samples (plane.h:95-107, sphere.h:152-170, box.h:122-191, disk.h, ring.h) */
    struct SS { V3<R> P, N; R UV[2]; };
    void sample_shape(const er_shape& s, const V3<R>& uvw, SS& ss) const
    {
        ss.N = V3<R>(R(0), R(0), R(1));
        ss.UV[0] = ss.UV[1] = R(0);
        switch (s.type) {
            case ER_SHAPE_PLANE:
                ss.P = V3<R>(R(-0.5f) + uvw[0], R(-0.5f) + uvw[1], R(0));
                ss.UV[0] = uvw[0]; ss.UV[1] = uvw[1];
                ss.P = ss.P * V3<R>(R(s.size[0]), R(s.size[1]), R(0));
                break;
            case ER_SHAPE_DISK: {
                const R rr = M<R>::sqrt(uvw[0]);
                const R th = R(2.0f * PI_F) * uvw[1];
                ss.P = V3<R>(rr * M<R>::cos(th), rr * M<R>::sin(th), R(0));
                ss.UV[0] = uvw[0]; ss.UV[1] = uvw[1];
                ss.P = scale(R(s.outer_radius), ss.P);
                break;
            }
            case ER_SHAPE_RING: {
                const R inner = R(s.inner_radius) / R(s.outer_radius);
                const R rr = inner + (R(1) - inner) * M<R>::sqrt(uvw[0]);
                const R th = R(2.0f * PI_F) * uvw[1];
                ss.P = V3<R>(rr * M<R>::cos(th), rr * M<R>::sin(th), R(0));
                ss.P = scale(R(s.outer_radius), ss.P);
                ss.UV[0] = ss.P[0]; ss.UV[1] = ss.P[1];
                break;
            }
            case ER_SHAPE_BOX: {
                const int side = (int)std::floor(uvw[0] * R(6));
                const R h = R(0.5f);
                switch (side) {
                    case 0: ss.P = V3<R>(h, -h + uvw[2], -h + uvw[1]); ss.N = V3<R>(R(1), R(0), R(0)); break;
                    case 1: ss.P = V3<R>(-h, -h + uvw[2], -h + uvw[1]); ss.N = V3<R>(R(-1), R(0), R(0)); break;
                    case 2: ss.P = V3<R>(-h + uvw[1], h, -h + uvw[2]); ss.N = V3<R>(R(0), R(1), R(0)); break;
                    case 3: ss.P = V3<R>(-h + uvw[1], -h, -h + uvw[2]); ss.N = V3<R>(R(0), R(-1), R(0)); break;
                    case 4: ss.P = V3<R>(-h + uvw[1], -h + uvw[2], h); ss.N = V3<R>(R(0), R(0), R(1)); break;
                    case 5: ss.P = V3<R>(-h + uvw[1], -h + uvw[2], -h); ss.N = V3<R>(R(0), R(0), R(-1)); break;
                }
                ss.UV[0] = uvw[1]; ss.UV[1] = uvw[2];
                ss.P = ss.P * from<R>(s.size);
                break;
            }
            case ER_SHAPE_SPHERE: {
                const R z = R(1) - R(2) * uvw[0];
                const R rr = M<R>::sqrt(std::max(R(0), R(1) - z * z));
                const R phi = R(2.0f * PI_F) * uvw[1];
                const R x = rr * M<R>::cos(phi);
                const R y = rr * M<R>::sin(phi);
                ss.P = V3<R>(x, y, z);
                ss.N = ss.P;
                /* UV = (SphericalTheta, SphericalPhi) (sphere.h:162, geometry.h:70-78) */
                R ph = M<R>::atan2(ss.P[2], ss.P[0]);
                ph = (ph < R(0)) ? ph + R(2.0f * PI_F) : ph;
                ss.UV[0] = M<R>::acos(clampf(ss.P[1], R(-1), R(1)));
                ss.UV[1] = ph;
                ss.P = scale(R(s.outer_radius), ss.P);
                break;
            }
            default: break;
        }
        ss.P = transform_point(s.tm, ss.P);
        ss.N = transform_vector(s.tm, ss.N);
    }

    /* Buffer2D<ColorRGBAuc>::operator()(Vec2f, true) (buffer2d.h:227-254) then
     * ColorXYZf::FromRGBAuc (color.h:195-211). COMPAT keeps the colour type's arithmetic:
     * `float * ColorRGBAuc` binds the float weight to `const unsigned char&` (truncation, so
     * a weight in (0, 1) is 0), the products and sums are stored back into unsigned char
     * components (mod 256). FULL blends the four texels in float. */
    V3<R> bitmap_lookup(const Scene::Bitmap& B, const R uv[2]) const
    {
        if (B.w <= 0 || B.h <= 0 || B.rgba.size() < (size_t)B.w * B.h * 4) return V3<R>();
        const R X = uv[0] * R((float)B.w), Y = uv[1] * R((float)B.h);
        const int cx0 = (int)std::floor(X), cx1 = (int)std::ceil(X), cy0 = (int)std::floor(Y), cy1 = (int)std::ceil(Y);
        const R du = X - R((float)cx0), dv = Y - R((float)cy0);
        const int x0 = std::min(std::max(cx0, 0), B.w - 1), x1 = std::min(std::max(cx1, 0), B.w - 1);
        const int y0 = std::min(std::max(cy0, 0), B.h - 1), y1 = std::min(std::max(cy1, 0), B.h - 1);
        const uint8_t* t00 = &B.rgba[((size_t)y0 * B.w + x0) * 4];
        const uint8_t* t10 = &B.rgba[((size_t)y0 * B.w + x1) * 4];
        const uint8_t* t01 = &B.rgba[((size_t)y1 * B.w + x0) * 4];
        const uint8_t* t11 = &B.rgba[((size_t)y1 * B.w + x1) * 4];
        R rgb[3];
        for (int k = 0; k < 3; ++k) {
            if (full) {
                rgb[k] = (R(1) - dv) * ((R(1) - du) * R(t00[k]) + du * R(t10[k])) + dv * ((R(1) - du) * R(t01[k]) + du * R(t11[k]));
            } else {
                const uint8_t w_a = (uint8_t)(R(1) - dv), w_b = (uint8_t)(R(1) - du), w_c = (uint8_t)du, w_d = (uint8_t)dv;
                const uint8_t top = (uint8_t)((uint8_t)(w_b * t00[k]) + (uint8_t)(w_c * t10[k]));
                const uint8_t bot = (uint8_t)((uint8_t)(w_b * t01[k]) + (uint8_t)(w_c * t11[k]));
                rgb[k] = R((float)(uint8_t)((uint8_t)(w_a * top) + (uint8_t)(w_d * bot)));
            }
        }
        const R s255 = R(1.0f / 255.0f);
        const R r = s255 * rgb[0], g = s255 * rgb[1], b = s255 * rgb[2];
        return V3<R>(R(0.412453f) * r + R(0.357580f) * g + R(0.180423f) * b,
                     R(0.212671f) * r + R(0.715160f) * g + R(0.072169f) * b,
                     R(0.019334f) * r + R(0.119193f) * g + R(0.950227f) * b);
    }

    /* EvaluateProcedural / EvaluateTexture (textures.h:26-111) */
    V3<R> eval_texture(int id, const R UVin[2]) const
    {
        if (id < 0 || id >= (int)S.textures.size()) return V3<R>();
        const er_texture_desc& T = S.textures[id];
        R uv[2] = {UVin[0], UVin[1]};
        uv[0] *= R(T.repeat[0]); uv[1] *= R(T.repeat[1]);
        uv[0] += R(T.offset[0]); uv[1] += R(1.0f - T.offset[1]);
        uv[0] = uv[0] - std::floor(uv[0]); uv[1] = uv[1] - std::floor(uv[1]);
        uv[0] = clampf(uv[0], R(0), R(1)); uv[1] = clampf(uv[1], R(0), R(1));
        if (T.flip[0]) uv[0] = R(1) - uv[0];
        if (T.flip[1]) uv[1] = R(1) - uv[1];
        V3<R> L;
        if (T.type == ER_TEXTURE_BITMAP) {
            if (T.bitmap_id >= 0 && T.bitmap_id < (int)S.bitmaps.size()) L = bitmap_lookup(S.bitmaps[T.bitmap_id], uv);
        } else if (T.type == ER_TEXTURE_PROCEDURAL) {
            const er_procedural& p = T.procedural;
            switch (p.type) {
                case ER_PROCEDURAL_UNIFORM: L = from<R>(p.uniform_color); break;
                case ER_PROCEDURAL_CHECKER: {
                    const int a = (int)(uv[0] * R(2)), b = (int)(uv[1] * R(2));
                    const bool c1 = (a % 2 == 0) ? (b % 2 == 0) : (b % 2 != 0);
                    L = c1 ? from<R>(p.checker_color1) : from<R>(p.checker_color2);
                    break;
                }
                case ER_PROCEDURAL_GRADIENT: L = ctf_eval(p.gradient, uv[1]); break;
                default: break;
            }
        }
        return scale(R(T.output_level), L);
    }

    /* ---- shading (shader.h:29-483) ----------------------------------------------------- */
    struct Shader {
        bool brdf;
        V3<R> Nn, Nu, Nv, Kd, Ks;
        R exponent, ior;
        bool full;
    };
    Shader make_shader(bool brdf, const V3<R>& N, const V3<R>& Wo, const V3<R>& Kd, const V3<R>& Ks, R ior, R exponent) const
    {
        Shader sh;
        sh.brdf = brdf; sh.Kd = Kd; sh.Ks = Ks; sh.exponent = exponent; sh.ior = ior; sh.full = full;
        sh.Nn = normalize(N);
        sh.Nu = normalize(cross(N, Wo));
        sh.Nv = normalize(cross(N, sh.Nu));
        return sh;
    }
    static V3<R> to_local(const Shader& s, const V3<R>& W) { return V3<R>(dot(W, s.Nu), dot(W, s.Nv), dot(W, s.Nn)); }
    static V3<R> to_world(const Shader& s, const V3<R>& W)
    {
        return V3<R>(s.Nu[0] * W[0] + s.Nv[0] * W[1] + s.Nn[0] * W[2],
                     s.Nu[1] * W[0] + s.Nv[1] * W[1] + s.Nn[1] * W[2],
                     s.Nu[2] * W[0] + s.Nv[2] * W[1] + s.Nn[2] * W[2]);
    }
    static bool same_hemi(const V3<R>& a, const V3<R>& b) { return (a[2] * b[2]) > R(0); }
    /* Fresnel::Evaluate + FrDiel (shader.h:69-123), EtaI = 1, EtaT = ior */
    static R fresnel(R cosi, R ior)
    {
        cosi = clampf(cosi, R(-1), R(1));
        const bool entering = cosi > R(0);
        R ei = R(1), et = ior;
        if (!entering) std::swap(ei, et);
        const R sint = ei / et * M<R>::sqrt(std::max(R(0), R(1) - cosi * cosi));
        if (sint >= R(1)) return R(1);
        const R cost = M<R>::sqrt(std::max(R(0), R(1) - sint * sint));
        const R ci = std::fabs(cosi);
        const R Rparl = ((et * ci) - (ei * cost)) / ((et * ci) + (ei * cost));
        const R Rperp = ((ei * ci) - (et * cost)) / ((ei * ci) + (et * cost));
        return (Rparl * Rparl + Rperp * Rperp) / R(2);
    }
    static R blinn_d(R exponent, const V3<R>& wh)   /* shader.h:181-185 */
    {
        const R c = std::fabs(wh[2]);
        return (exponent + R(2)) * R(INV_TWO_PI_F) * M<R>::pow(c, exponent);
    }
    static R blinn_pdf(R exponent, const V3<R>& wo, const V3<R>& wi)   /* shader.h:166-179 */
    {
        const V3<R> wh = normalize(wo + wi);
        const R c = std::fabs(wh[2]);
        R pdf = ((exponent + R(1)) * M<R>::pow(c, exponent)) / (R(2.0f) * R(PI_F) * R(4.0f) * dot(wo, wh));
        if (dot(wo, wh) <= R(0)) pdf = R(0);
        return pdf;
    }
    /* Microfacet::F (shader.h:216-236) */
    static V3<R> microfacet_f(const Shader& s, const V3<R>& wo, const V3<R>& wi)
    {
        const R cO = std::fabs(wo[2]), cI = std::fabs(wi[2]);
        if (cI == R(0) || cO == R(0)) return V3<R>();
        V3<R> wh = wi + wo;
        if (wh[0] == R(0) && wh[1] == R(0) && wh[2] == R(0)) return V3<R>();
        wh = normalize(wh);
        const R cosH = dot(wi, wh);
        const R Fr = fresnel(cosH, s.ior);
        const R D = blinn_d(s.exponent, wh);
        /* G (shader.h:255-263) */
        const R NdotWh = std::fabs(wh[2]), NdotWo = std::fabs(wo[2]), NdotWi = std::fabs(wi[2]);
        const R WOdotWh = abs_dot(wo, wh);
        const R G = std::min(R(1), std::min((R(2) * NdotWh * NdotWo / WOdotWh), (R(2) * NdotWh * NdotWi / WOdotWh)));
        const R den = R(4) * cI * cO;
        return V3<R>((((s.Ks[0] * D) * G) * Fr) / den, (((s.Ks[1] * D) * G) * Fr) / den, (((s.Ks[2] * D) * G) * Fr) / den);
    }
    static R lambert_pdf(const V3<R>& wo, const V3<R>& wi) { return same_hemi(wo, wi) ? std::fabs(wi[2]) * R(INV_PI_F) : R(0); }
    static R microfacet_pdf(const Shader& s, const V3<R>& wo, const V3<R>& wi) { return same_hemi(wo, wi) ? blinn_pdf(s.exponent, wo, wi) : R(0); }

    /* Shader::F (shader.h:429-441) */
    static V3<R> shader_f(const Shader& s, const V3<R>& Wo, const V3<R>& Wi)
    {
        if (!s.brdf) return scale(R(INV_PI_F), s.Kd);   /* IsotropicPhase::F (Q7) */
        const V3<R> wo = to_local(s, Wo), wi = to_local(s, Wi);
        V3<R> Rr;
        Rr = Rr + scale(R(INV_PI_F), s.Kd);
        Rr = Rr + microfacet_f(s, wo, wi);
        return Rr;
    }
    /* Shader::Pdf (shader.h:457-469). COMPAT: Lambert + Blinn (Q32), phase pi/4 (Q6).
     * FULL: the density of the one-sample mixture the BSDF sampler draws from. */
    static R shader_pdf(const Shader& s, const V3<R>& Wo, const V3<R>& Wi)
    {
        if (!s.brdf) return s.full ? R(TRUE_INV_FOUR_PI_F) : R(INV_FOUR_PI_F);
        const V3<R> wo = to_local(s, Wo), wi = to_local(s, Wi);
        R pdf = R(0);
        pdf += lambert_pdf(wo, wi);
        pdf += microfacet_pdf(s, wo, wi);
        return s.full ? R(0.5f) * pdf : pdf;
    }

    /* montecarlo.h:86-137,187-195 */
    static void concentric_disk(R u0, R u1, R& ox, R& oy)
    {
        R r, theta;
        const R sx = R(2) * u0 - R(1), sy = R(2) * u1 - R(1);
        if (sx == R(0) && sy == R(0)) { ox = oy = R(0); return; }
        if (sx >= -sy) {
            if (sx > sy) { r = sx; theta = (sy > R(0)) ? sy / r : R(8) + sy / r; }
            else { r = sy; theta = R(2) - sx / r; }
        } else {
            if (sx <= sy) { r = -sx; theta = R(4) - sy / r; }
            else { r = -sy; theta = R(6) + sx / r; }
        }
        theta *= R(PI_F / 4.f);
        ox = r * M<R>::cos(theta);
        oy = r * M<R>::sin(theta);
    }
    /* BRDF::SampleF / IsotropicPhase::SampleF (shader.h:297-303,350-374), full mode only */
    V3<R> shader_sample_f(const Shader& s, const V3<R>& Wo, V3<R>& Wi, R& pdf, R comp, R d0, R d1) const
    {
        if (!s.brdf) {
            const R z = R(1) - R(2) * d0;
            const R rr = M<R>::sqrt(std::max(R(0), R(1) - z * z));
            const R phi = R(2.f * PI_F) * d1;
            Wi = V3<R>(rr * M<R>::cos(phi), rr * M<R>::sin(phi), z);
            pdf = R(TRUE_INV_FOUR_PI_F);
            return scale(R(INV_PI_F), s.Kd);
        }
        const V3<R> wo = to_local(s, Wo);
        V3<R> wi;
        if (comp <= R(0.5f)) {
            R dx, dy;
            concentric_disk(d0, d1, dx, dy);
            wi = V3<R>(dx, dy, M<R>::sqrt(std::max(R(0), R(1) - dx * dx - dy * dy)));
            if (wo[2] < R(0)) wi[2] *= R(-1);
        } else {
            const R costheta = M<R>::pow(d0, R(1) / (s.exponent + R(1)));
            const R sintheta = M<R>::sqrt(std::max(R(0), R(1) - costheta * costheta));
            const R phi = d1 * R(2) * R(PI_F);
            V3<R> wh(sintheta * M<R>::cos(phi), sintheta * M<R>::sin(phi), costheta);
            if (!same_hemi(wo, wh)) wh = -wh;
            const R two_d = R(2) * dot(wo, wh);
            wi = (-wo) + scale(two_d, wh);
        }
        pdf = R(0.5f) * (lambert_pdf(wo, wi) + microfacet_pdf(s, wo, wi));
        V3<R> Rr;
        Rr = Rr + scale(R(INV_PI_F), s.Kd);
        Rr = Rr + microfacet_f(s, wo, wi);
        Wi = to_world(s, wi);
        return Rr;
    }

    static R power_heuristic(R f, R g) { return (f * f) / (f * f + g * g); }   /* montecarlo.h:223-227 */

    /* ---- lights (lights.h:28-131) ------------------------------------------------------- */
    const er_light_desc* light(int id) const { return (id >= 0 && id < (int)S.lights.size()) ? &S.lights[id] : nullptr; }

    /* IntersectLight (lights.h:59-88) */
    bool intersect_light(const er_light_desc& L, const Ray<R>& r, R& T, V3<R>& P, V3<R>& N, V3<R>& Le) const
    {
        const Ray<R> rt = transform_ray(L.shape.inv_tm, r);
        Isect<R> I;
        intersect_shape(L.shape, rt, I);
        if (!I.Valid) return false;
        P = transform_point(L.shape.tm, I.P);
        N = transform_vector(L.shape.tm, I.N);
        T = length(P - r.O);
        Le = I.Front ? scale(R(L.multiplier), eval_texture(L.texture_id, I.UV)) : V3<R>();
        if (L.unit == ER_UNIT_LUX) Le = V3<R>(Le[0] / R(L.shape.area), Le[1] / R(L.shape.area), Le[2] / R(L.shape.area));
        return true;
    }
    /* IntersectLights (lights.h:90-113): nearest hit; SE.LightID = list position */
    bool intersect_lights(const Ray<R>& r, bool respect_visibility, R& T, V3<R>& P, V3<R>& N, V3<R>& Le, int& idx) const
    {
        R best = R(FLT_MAX);
        bool any = false;
        for (int i = 0; i < S.t.light_count; ++i) {
            const er_light_desc* L = light(S.t.light_ids[i]);
            if (!L) continue;
            if (respect_visibility && !L->visible) continue;
            R t; V3<R> p, n, le;
            if (intersect_light(*L, r, t, p, n, le) && t < best) {
                best = t; T = t; P = p; N = n; Le = le; idx = i; any = true;
            }
        }
        return any;
    }
    /* IntersectsLight (lights.h:115-131) */
    bool intersects_any_light(const Ray<R>& r) const
    {
        for (int i = 0; i < S.t.light_count; ++i) {
            const er_light_desc* L = light(S.t.light_ids[i]);
            if (!L) continue;
            Isect<R> I;
            intersect_shape(L->shape, transform_ray(L->shape.inv_tm, r), I);
            if (I.Valid) return true;
        }
        return false;
    }

    /* The objects SampleRay and Visible test. COMPAT: gpObjects[i] for i < ObjectIDs.Count,
     * the live objects in ID order (objects.h:50-52,75-77), clamped to the live count; FULL:
     * the objects ObjectIDs names. */
    int object_count() const
    {
        return full ? std::min(S.t.object_count, ER_MAX_INDICES) : std::min(S.t.object_count, (int)S.objects.size());
    }
    const er_object_desc* object(int i) const
    {
        const int id = full ? S.t.object_ids[i] : i;
        return (id >= 0 && id < (int)S.objects.size()) ? &S.objects[id] : nullptr;
    }
    /* IntersectObjects (objects.h:26-66): nearest hit; SE.ObjectID = list position */
    bool intersect_objects(const Ray<R>& r, R& T, V3<R>& P, V3<R>& N, R UV[2], int& idx) const
    {
        R best = R(FLT_MAX);
        bool any = false;
        for (int i = 0; i < object_count(); ++i) {
            const er_object_desc* O = object(i);
            if (!O) continue;
            Isect<R> I;
            intersect_shape(O->shape, transform_ray(O->shape.inv_tm, r), I);
            if (!I.Valid) continue;
            const V3<R> Pw = transform_point(O->shape.tm, I.P);
            const R t = length(Pw - r.O);
            if (t < best) {
                best = t; T = t; P = Pw; N = transform_vector(O->shape.tm, I.N);
                UV[0] = I.UV[0]; UV[1] = I.UV[1]; idx = i; any = true;
            }
        }
        return any;
    }
    /* IntersectsObject (objects.h:68-82) */
    bool intersects_any_object(const Ray<R>& r) const
    {
        for (int i = 0; i < object_count(); ++i) {
            const er_object_desc* O = object(i);
            if (!O) continue;
            Isect<R> I;
            intersect_shape(O->shape, transform_ray(O->shape.inv_tm, r), I);
            if (I.Valid) return true;
        }
        return false;
    }

    /* Visible (transport.h:46-56) -> Intersect (transport.h:30-44) */
    bool visible(const V3<R>& P1, const V3<R>& P2)
    {
        if (!S.t.settings.shadows) return true;
        const V3<R> W = normalize(P2 - P1);
        const Ray<R> r(P1 + scale(R(RAY_EPS), W), W, R(0), std::min(length(P2 - P1) - R(RAY_EPS_2), R(S.t.settings.max_shadow_distance)));
        if (cnt) cnt[7]++;
        if (intersects_any_light(r)) return false;
        if (intersects_any_object(r)) return false;
        R T; V3<R> P;
        if (full && S.t.settings.sampler == ER_SAMPLER_WOODCOCK) {
            /* own stream for the march; the path's stream takes the two draws of a ray-march
             * test when the ray meets the volume (Q18) */
            R nearT, farT;
            if (!intersect_box(r, from<R>(S.minp), from<R>(S.maxp), nearT, farT)) return true;
            const Rng path = rng;
            rng = shadow_stream(path);
            const bool hit = march(r, false, T, P);
            rng = path;
            (void)rng.get1();
            (void)rng.get1();
            return !hit;
        }
        if (march(r, false, T, P)) return false;
        return true;
    }

    /* SampleLight (lights.h:43-57) */
    void sample_light(const er_light_desc& L, const V3<R>& uvw, const V3<R>& SEP, SS& ss, V3<R>& Wi, V3<R>& Le) const
    {
        sample_shape(L.shape, uvw, ss);
        Wi = normalize(ss.P - SEP);
        Le = scale(R(L.multiplier), eval_texture(L.texture_id, ss.UV));
        if (L.shape.one_sided && dot(SEP - ss.P, ss.N) < R(0)) Le = V3<R>();
        if (L.unit == ER_UNIT_LUX) Le = V3<R>(Le[0] / R(L.shape.area), Le[1] / R(L.shape.area), Le[2] / R(L.shape.area));
    }

    struct Sample { R comp, dir0, dir1, uvw[3], lightnum, film0, film1, lens0, lens1; };

    /* EstimateDirectLight (transport.h:58-113). COMPAT stops after the light-sampling half
     * (the early `return Ld;` at transport.h:84); FULL runs both halves with the power
     * heuristic on both. */
    V3<R> estimate_direct(const er_light_desc& L, int light_pos, const Sample& smp, const V3<R>& SEP, const V3<R>& SEN, const V3<R>& Wo, const Shader& sh)
    {
        V3<R> Ld;
        SS ss; V3<R> Wi, Li;
        sample_light(L, V3<R>(smp.uvw[0], smp.uvw[1], smp.uvw[2]), SEP, ss, Wi, Li);
        V3<R> F = shader_f(sh, Wo, Wi);
        R BsdfPdf = shader_pdf(sh, Wo, Wi);
        const R area = R(L.shape.area);
        if (!is_black(Li) && !is_black(F) && BsdfPdf > R(0) && visible(SEP, ss.P)) {
            const R LightPdf = dist2(SEP, ss.P) / (abs_dot(ss.N, -Wi) * area);
            const R Weight = power_heuristic(LightPdf, BsdfPdf);
            const V3<R> FL = F * Li;
            if (sh.brdf) Ld = Ld + scale(abs_dot(Wi, SEN) * Weight / LightPdf, FL);
            else if (full) Ld = Ld + scale(Weight / LightPdf, FL);
            else Ld = Ld + V3<R>(FL[0] / LightPdf, FL[1] / LightPdf, FL[2] / LightPdf);
        }
        if (!full) return Ld;

        F = shader_sample_f(sh, Wo, Wi, BsdfPdf, smp.comp, smp.dir0, smp.dir1);
        if (is_black(F) || BsdfPdf <= R(0)) return Ld;
        R T2; V3<R> P2, N2, Le2; int idx2 = -1;
        const Ray<R> r2(SEP, Wi, R(0), R(1000000.0f));
        if (!intersect_lights(r2, false, T2, P2, N2, Le2, idx2) || idx2 != light_pos) return Ld;
        if (!is_black(Le2) && visible(SEP, P2)) {
            const R LightPdf = dist2(SEP, P2) / (abs_dot(N2, -Wi) * area);
            const R Weight = power_heuristic(BsdfPdf, LightPdf);
            const V3<R> FL = F * Le2;
            if (sh.brdf) Ld = Ld + scale(abs_dot(Wi, SEN) * Weight / BsdfPdf, FL);
            else Ld = Ld + scale(Weight / BsdfPdf, FL);
        }
        return Ld;
    }

    /* UniformSampleOneLight (transport.h:115-156) with the full-mode shader selection
     * (design: scatterevent.h:155-212) */
    V3<R> uniform_sample_one_light(const Sample& smp, const V3<R>& SEP, const V3<R>& SEN, const V3<R>& Wo,
                                   const er_object_desc* obj = nullptr, const R* UV = nullptr)
    {
        V3<R> Ld;
        const R I = get_intensity(SEP);
        if (cnt) cnt[3] += 1 + gradient_samples();
        Ld = Ld + ctf_eval(S.t.emission, I);
        const int count = S.t.light_count;
        if (count <= 0) return Ld;
        const int pos = (int)std::floor(smp.lightnum * R((float)count));
        const int LightID = S.t.light_ids[pos];
        const er_light_desc* L = light(LightID);
        if (LightID < 0 || !L) return Ld;
        if (obj) {   /* object event (transport.h:141-149) */
            const V3<R> Kd = eval_texture(obj->diffuse_texture_id, UV);
            const V3<R> Ks = eval_texture(obj->specular_texture_id, UV);
            const V3<R> G = eval_texture(obj->glossiness_texture_id, UV);
            const Shader sh = make_shader(true, SEN, Wo, Kd, Ks, R(15.0f), R(1000000.0f) * M<R>::pow(G[1], R(7)));
            Ld = Ld + estimate_direct(*L, pos, smp, SEP, SEN, Wo, sh);
            return scale(R((float)count), Ld);
        }
        bool brdf = true;
        if (full) {
            const int type = S.t.settings.shading_type;
            if (type == ER_SHADING_PHASE_ONLY) brdf = false;
            else if (type == ER_SHADING_HYBRID || type == ER_SHADING_GRADIENT_MAGNITUDE || type == ER_SHADING_MODULATION || type == ER_SHADING_THRESHOLD) {
                const R gm = gradient_magnitude(SEP);
                if (cnt) cnt[3] += 6;
                const R NGM = S.gm_max > 0.0f ? gm * (R(1) / R(S.gm_max)) : R(0);
                if (type == ER_SHADING_THRESHOLD) {
                    brdf = NGM > R(S.t.settings.gradient_threshold);
                } else if (type == ER_SHADING_MODULATION) {
                    const R PdfBrdf = R(1) - M<R>::pow(R(1) - NGM, R(2));
                    brdf = u() < PdfBrdf;
                } else {
                    const R e = R(25) * M<R>::pow(R(S.t.settings.gradient_factor), R(3)) * NGM;
                    const R base = R(1) - M<R>::exp(-e);
                    const R PdfBrdf = S.t.settings.opacity_modulated ? opacity_at(S, I) * base : base;
                    brdf = u() <= PdfBrdf;
                }
            }
        }
        const V3<R> Kd = diffuse_at(S, I);
        const V3<R> Ks = ctf_eval(S.t.specular, I);
        const R gloss = R(1000000.0f) * M<R>::pow(plf_eval(S.t.glossiness, I), R(7));   /* utilities.h:28-31 */
        const Shader sh = make_shader(brdf, SEN, Wo, Kd, Ks, R(15.0f), gloss);
        Ld = Ld + estimate_direct(*L, pos, smp, SEP, SEN, Wo, sh);
        return scale(R((float)count), Ld);
    }

    /* SingleScattering (singlescattering.h:29-102) */
    void single_scattering(int px, int py, R out[4])
    {
        if (cnt) cnt[0]++;
        Sample smp;
        smp.comp = u();                                       /* sample.h:142-146 */
        smp.dir0 = u(); smp.dir1 = u();
        smp.uvw[0] = u(); smp.uvw[1] = u(); smp.uvw[2] = u(); /* sample.h:101-104 */
        smp.lightnum = u();                                   /* sample.h:184 */
        smp.film0 = u(); smp.film1 = u();                     /* sample.h:221-225 */
        smp.lens0 = u(); smp.lens1 = u();

        /* SampleCamera (singlescattering.h:29-50) */
        const er_camera& c = S.t.camera;
        const R sx = R(S.screen[0][0]) + (R(S.invscreen[0]) * R((float)px + smp.film0 * R(1.0f) / R((float)c.film_size[0])));
        const R sy = R(S.screen[1][0]) + (R(S.invscreen[1]) * R((float)py + smp.film1 * R(1.0f) / R((float)c.film_size[1])));
        const V3<R> N = from<R>(S.camN), U = from<R>(S.camU), V = from<R>(S.camV);
        Ray<R> r;
        r.O = from<R>(c.pos);
        r.D = normalize((N + scale(sx, U)) - scale(sy, V));
        r.MinT = R(c.clip_near);
        r.MaxT = R(c.clip_far);
        if (c.aperture_size != 0.0f) {
            R lx, ly;
            concentric_disk(smp.lens0, smp.lens1, lx, ly);
            lx = lx * R(c.aperture_size); ly = ly * R(c.aperture_size);
            const V3<R> LI = scale(lx, U) + scale(ly, V);
            r.O = r.O + LI;
            r.D = normalize(scale(R(S.focal), r.D) - LI);
        }

        /* SampleRay (singlescattering.h:52-74): volume, lights, objects; nearest by T */
        R Tv = R(0); V3<R> Pv;
        const bool vol = march(r, true, Tv, Pv);
        R Tl = R(0); V3<R> Pl, Nl, Lel; int il = -1;
        const bool lig = intersect_lights(r, true, Tl, Pl, Nl, Lel, il);
        int type = -1;
        R T = R(FLT_MAX);
        if (vol && Tv < T) { type = 0; T = Tv; }
        if (lig && Tl < T) { type = 1; T = Tl; }
        R To = R(0), UVo[2] = {R(0), R(0)}; V3<R> Po, No; int io = -1;
        if (intersect_objects(r, To, Po, No, UVo, io) && To < T) { type = 2; T = To; }

        V3<R> Lv;
        if (type == 0) {
            if (cnt) cnt[6]++;
            const V3<R> Nv = normalized_gradient(Pv);   /* raymarching.h:70 */
            Lv = Lv + uniform_sample_one_light(smp, Pv, Nv, -r.D);
        }
        if (type == 2) {
            if (cnt) cnt[6]++;
            Lv = Lv + uniform_sample_one_light(smp, Po, No, -r.D, object(io), UVo);
        }
        if (type == 1) Lv = Lv + Lel;
        out[0] = Lv[0]; out[1] = Lv[1]; out[2] = Lv[2];
        out[3] = type >= 0 ? R(1) : R(0);
    }
};

/* ---- post (filterframeestimate.cuh:28-64, estimate.cuh:27-32, tonemap.cuh:30-59) ------- */
static void post_pass(const Scene& S, const float* frame, float* running, uint8_t* display, int no_iterations, float* filtered)
{
    const int W = S.t.camera.film_size[0], H = S.t.camera.film_size[1];
    /* Gauss2D(1, X, Y) = expf(-(X*X + Y*Y) / 2) for X, Y in {-1, 0, 1} */
    const float g[3] = {erm::expf_(-0.0f), erm::expf_(-0.5f), erm::expf_(-1.0f)};
    const float exposure = S.t.camera.exposure;
    const float nf = std::max((float)no_iterations, 1.0f);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const int x0 = std::max(x - 1, 0), x1 = std::min(x + 1, W - 1);
            const int y0 = std::max(y - 1, 0), y1 = std::min(y + 1, H - 1);
            float sum[4] = {0, 0, 0, 0}, tw = 0.0f;
            for (int yy = y0; yy <= y1; ++yy)
                for (int xx = x0; xx <= x1; ++xx) {
                    const int d2 = (xx - x) * (xx - x) + (yy - y) * (yy - y);
                    const float w = g[d2];
                    const float* f = frame + 4 * ((size_t)yy * W + xx);
                    for (int c = 0; c < 4; ++c) sum[c] += f[c] * w;
                    tw += w;
                }
            sum[3] = frame[4 * ((size_t)y * W + x) + 3];
            float fe[4];
            for (int c = 0; c < 4; ++c) fe[c] = tw > 0.0f ? sum[c] / tw : 0.0f;   /* alpha / tw too */
            if (filtered) for (int c = 0; c < 4; ++c) filtered[4 * ((size_t)y * W + x) + c] = fe[c];
            float* r = running + 4 * ((size_t)y * W + x);
            for (int c = 0; c < 4; ++c) r[c] = r[c] + (fe[c] - r[c]) / nf;
            /* ColorRGBf::FromXYZAf (color.h:155-164) + tonemap */
            float rgb[3];
            rgb[0] = 3.240479f * r[0] - 1.537150f * r[1] - 0.498535f * r[2];
            rgb[1] = -0.969256f * r[0] + 1.875991f * r[1] + 0.041556f * r[2];
            rgb[2] = 0.055648f * r[0] - 0.204043f * r[1] + 1.057311f * r[2];
            uint8_t* d = display + 4 * ((size_t)y * W + x);
            for (int c = 0; c < 3; ++c) {
                float v = 1.0f - erm::expf_(-(rgb[c] / exposure));
                v = std::fmax(0.0f, std::fmin(v, 1.0f));   /* Clamp via device min/max */
                d[c] = (uint8_t)(255.0f * v);
            }
            d[3] = (uint8_t)(r[3] * 255.0f);
        }
    }
}

/* Separable bilateral filter on RGBA8 (design: filterrunningestimate.cuh:28-200,
 * filter.h:34-40). Taps k in [max(0, i-R), min(i+R, W-1)) - the upper bound is exclusive
 * as written (Q27). Luminance weights of color.h:37-51 on 0..255 values. */
static void bilateral_pass(const Scene& S, const uint8_t* in, uint8_t* tmp, uint8_t* out)
{
    const int W = S.t.camera.film_size[0], H = S.t.camera.film_size[1];
    const int R = S.t.settings.bilateral_radius;
    const float ss = S.t.settings.bilateral_sigma_spatial, sr = S.t.settings.bilateral_sigma_range;
    std::vector<float> kd(2 * R + 1), gs(256);
    for (int k = -R; k <= R; ++k) kd[k + R] = erm::expf_(-((float)(k * k)) / (2.0f * ss * ss));
    for (int d = 0; d < 256; ++d) gs[d] = erm::expf_(-((float)(d * d)) / (2.0f * sr * sr));
    auto lum = [](const uint8_t* p) { return 0.212671f * (float)p[0] + 0.715160f * (float)p[1] + 0.072169f * (float)p[2]; };
    for (int pass = 0; pass < 2; ++pass) {
        const uint8_t* src = pass == 0 ? in : tmp;
        uint8_t* dst = pass == 0 ? tmp : out;
        #pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int i = pass == 0 ? x : y;
                const int n = pass == 0 ? W : H;
                const int lo = std::max(0, i - R), hi = std::min(i + R, n - 1);
                const uint8_t* c = src + 4 * ((size_t)y * W + x);
                const float yc = lum(c);
                float sum[3] = {0, 0, 0}, tw = 0.0f;
                for (int k = lo; k < hi; ++k) {
                    const uint8_t* p = pass == 0 ? src + 4 * ((size_t)y * W + k) : src + 4 * ((size_t)k * W + x);
                    const float w = kd[R + k - i] * gs[(int)std::fabs(lum(p) - yc)];
                    tw += w;
                    for (int ch = 0; ch < 3; ++ch) sum[ch] += (float)p[ch] * w;
                }
                uint8_t* d = dst + 4 * ((size_t)y * W + x);
                for (int ch = 0; ch < 3; ++ch)
                    d[ch] = tw > 0.0f ? (uint8_t)std::max(0.0f, std::min(sum[ch] / tw, 255.0f)) : 0;
                d[3] = 255;
            }
    }
}

} /* namespace orc */

using namespace orc;

/* ---- C API (ctypes) -------------------------------------------------------------------- */
/* Woodcock majorant grid: for each macrocell of 2^WC_SHIFT cells, the voxel range
 * [b E, min(b E + E, R - 1)] per axis that its (clamped) trilinear samples read, the
 * intensities [min - 3, max] a truncating sample can take (three truncating lerp levels), and
 * the largest sigma = DS * opacity over them. */
static void build_woodcock(Scene& S)
{
    const int E = 1 << WC_SHIFT;
    for (int i = 0; i < 3; ++i) S.wc_dim[i] = (S.res[i] + E - 1) / E;
    uint16_t vmax = 0;
    for (uint16_t v : S.vox) vmax = std::max(vmax, v);
    std::vector<float> sig((size_t)vmax + 1);
    for (int i = 0; i <= (int)vmax; ++i) sig[i] = S.t.settings.density_scale * opacity_at(S, (float)i);
    const int mx = S.wc_dim[0], my = S.wc_dim[1], mz = S.wc_dim[2];
    S.wc_sigma.assign((size_t)mx * my * mz, 0.0f);
    for (int bz = 0; bz < mz; ++bz)
        for (int by = 0; by < my; ++by)
            for (int bx = 0; bx < mx; ++bx) {
                int lo = 0xffff, hi = 0;
                for (int z = bz * E; z <= std::min(bz * E + E, S.res[2] - 1); ++z)
                    for (int y = by * E; y <= std::min(by * E + E, S.res[1] - 1); ++y)
                        for (int x = bx * E; x <= std::min(bx * E + E, S.res[0] - 1); ++x) {
                            const int v = S.vox[((size_t)z * S.res[1] + y) * S.res[0] + x];
                            lo = std::min(lo, v);
                            hi = std::max(hi, v);
                        }
                float m = 0.0f;
                for (int i = std::max(lo - 3, 0); i <= std::min(hi, (int)vmax); ++i) m = std::fmax(m, sig[i]);
                S.wc_sigma[((size_t)bz * my + by) * mx + bx] = m;
            }
}

extern "C" {

void* orc_create(const er_tracer_desc* t, const er_volume_desc* v, const er_light_desc* lights, int nlights,
                 const er_texture_desc* textures, int ntextures, const er_clipping_desc* clips, int nclips)
{
    Scene* S = new Scene();
    S->t = *t;
    camera_update(*S);
    volume_update(*S, *v);
    const size_t n = (size_t)v->resolution[0] * v->resolution[1] * v->resolution[2];
    S->vox.assign(v->voxels, v->voxels + n);
    for (int i = 0; i < nlights; ++i) { er_light_desc L = lights[i]; L.shape.area = shape_area(L.shape); S->lights.push_back(L); }
    for (int i = 0; i < ntextures; ++i) S->textures.push_back(textures[i]);
    for (int i = 0; i < nclips; ++i) S->clips.push_back(clips[i]);
    build_woodcock(*S);
    return S;
}

void orc_destroy(void* h) { delete (Scene*)h; }

/* EvaluateTexture (textures.h:62-111) of texture `id` at (u, v), in FP32 */
void orc_eval_texture(void* h, int id, float u, float v, float* xyz)
{
    Scene& S = *(Scene*)h;
    Rng rng{1u, 1u};
    Est<float> e(S, rng, nullptr);
    const float uv[2] = {u, v};
    const V3<float> L = e.eval_texture(id, uv);
    for (int k = 0; k < 3; ++k) xyz[k] = L[k];
}

/* registered surface objects, index = object ID */
void orc_set_objects(void* h, const er_object_desc* objs, int n)
{
    Scene* S = (Scene*)h;
    S->objects.clear();
    for (int i = 0; i < n; ++i) { er_object_desc O = objs[i]; O.shape.area = shape_area(O.shape); S->objects.push_back(O); }
}

/* a registered bitmap, index = bitmap ID: w * h RGBA8 texels, row-major */
void orc_set_bitmap(void* h, int id, int w, int hgt, const uint8_t* rgba)
{
    Scene* S = (Scene*)h;
    if (id < 0) return;
    if ((int)S->bitmaps.size() <= id) S->bitmaps.resize(id + 1);
    Scene::Bitmap& B = S->bitmaps[id];
    B.w = w; B.h = hgt;
    B.rgba.assign(rgba, rgba + (size_t)std::max(0, w) * std::max(0, hgt) * 4);
}

/* ComputeAutoFocusDistance (autofocus.cuh:26-63): 100 rays through film point (u, v), each
 * jittered by 0.01 x a concentric-disk sample of the screen (Get2, Q22 order: first draw ->
 * first component), marched with SampleVolume; the mean distance of the rays' volume events,
 * or the distance to the last event point when none (the reference's stale-event quirk is not
 * kept: a ray without an event adds nothing). */
float orc_autofocus(void* h, int u, int v, uint32_t s0, uint32_t s1)
{
    Scene& S = *(Scene*)h;
    Rng r{s0, s1};
    Est<float> est(S, r, nullptr);
    const V3<float> O = from<float>(S.t.camera.pos), N = from<float>(S.camN), U = from<float>(S.camU),
                    V = from<float>(S.camV);
    float sum = 0.0f, w = 0.0f;
    V3<float> lastP(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 100; ++i) {
        float sx = S.screen[0][0] + (S.invscreen[0] * (float)u);
        float sy = S.screen[1][0] + (S.invscreen[1] * (float)v);
        const float d0 = est.u(), d1 = est.u();
        float dx, dy;
        Est<float>::concentric_disk(d0, d1, dx, dy);
        sx += 0.01f * dx;
        sy += 0.01f * dy;
        const Ray<float> ray(O, normalize((N + scale(sx, U)) - scale(sy, V)), S.t.camera.clip_near, S.t.camera.clip_far);
        float T; V3<float> P;
        if (est.march(ray, true, T, P)) {
            sum += length(P - O);
            w += 1.0f;
            lastP = P;
        }
    }
    return sum <= 0.0f ? length(lastP - O) : sum / w;
}

/* the node-table evaluation at intensity I, channel ch (float and double) */
float orc_tf_lut_eval(const float* lut, float range, float I, int ch)
{
    std::vector<float> v(lut, lut + ER_TF_LUT_NODES * 4);
    return tf_lut_eval(v, (float)(ER_TF_LUT_NODES - 1) / range, I, ch);
}

/* node-table transfer function; lut == NULL returns to the PLFs */
void orc_set_tf_lut(void* h, const float* lut, float range)
{
    Scene& S = *(Scene*)h;
    if (!lut) { S.tf_lut.clear(); S.tf_scale = 0.0f; }
    else { S.tf_lut.assign(lut, lut + ER_TF_LUT_NODES * 4); S.tf_scale = (float)(ER_TF_LUT_NODES - 1) / range; }
    build_woodcock(S);
}

void orc_set_gm_max(void* h, float gm) { ((Scene*)h)->gm_max = gm; }
float orc_get_gm_max(void* h) { return ((Scene*)h)->gm_max; }

/* Maximum gradient magnitude over voxel centres (design: gradientmagnitude.cuh:28-70);
 * P = MinAABB + Size * ((X + 0.5, Y + 0.5, Z + 0.5) / Resolution). */
float orc_compute_gm_max(void* h, int nthreads)
{
    Scene& S = *(Scene*)h;
    float best = 0.0f;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(static) reduction(max:best)
    for (int z = 0; z < S.res[2]; ++z) {
        Rng dummy{1, 1};
        Est<float> e(S, dummy, nullptr);
        for (int y = 0; y < S.res[1]; ++y)
            for (int x = 0; x < S.res[0]; ++x) {
                const V3<float> P(S.minp[0] + S.size[0] * (((float)x + 0.5f) / (float)S.res[0]),
                                  S.minp[1] + S.size[1] * (((float)y + 0.5f) / (float)S.res[1]),
                                  S.minp[2] + S.size[2] * (((float)z + 0.5f) / (float)S.res[2]));
                best = std::max(best, e.gradient_magnitude(P));
            }
    }
    S.gm_max = best;
    return best;
}

/* One FP32 pass over every pixel: frame[W*H*4] = SingleScattering; seeds updated in place.
 * counters (may be NULL): 8 x uint64 in er_counters order, accumulated. */
int orc_trace_f32(void* h, uint32_t* seeds1, uint32_t* seeds2, float* frame, int nthreads, uint64_t* counters)
{
    Scene& S = *(Scene*)h;
    const int W = S.t.camera.film_size[0], H = S.t.camera.film_size[1];
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    std::vector<uint64_t> acc(8, 0);
    #pragma omp parallel
    {
        uint64_t local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        #pragma omp for schedule(dynamic, 1)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)y * W + x;
                Rng rng{seeds1[i], seeds2[i]};
                Est<float> e(S, rng, counters ? local : nullptr);
                float out[4];
                e.single_scattering(x, y, out);
                for (int c = 0; c < 4; ++c) frame[4 * i + c] = out[c];
                seeds1[i] = rng.s0; seeds2[i] = rng.s1;
            }
        #pragma omp critical
        for (int k = 0; k < 8; ++k) acc[k] += local[k];
    }
    if (counters) for (int k = 0; k < 8; ++k) counters[k] += acc[k];
    return 0;
}

/* FP32 pass over a pixel subset (px, py, n); seeds are per listed pixel. Also returns the
 * number of RNG draws per pixel (draws[n], may be NULL). */
int orc_trace_pixels_f32(void* h, const int* px, const int* py, int n, uint32_t* seeds1, uint32_t* seeds2,
                         float* out, uint32_t* draws, uint64_t* counters)
{
    Scene& S = *(Scene*)h;
    for (int i = 0; i < n; ++i) {
        Rng rng{seeds1[i], seeds2[i]};
        Est<float> e(S, rng, counters);
        e.single_scattering(px[i], py[i], out + 4 * i);
        seeds1[i] = rng.s0; seeds2[i] = rng.s1;
        if (draws) draws[i] = rng.draws;
    }
    return 0;
}

/* FP64 oracle: per listed pixel, the mean of `spp` samples of the Real=double estimator
 * (unfiltered XYZA), with per-pixel streams from (seed1[i], seed2[i]). */
int orc_expect_f64(void* h, const int* px, const int* py, int n, int spp, const uint32_t* seeds1,
                   const uint32_t* seeds2, double* out, int nthreads)
{
    Scene& S = *(Scene*)h;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        Rng rng{seeds1[i], seeds2[i]};
        Est<double> e(S, rng, nullptr);
        double acc[4] = {0, 0, 0, 0};
        for (int s = 0; s < spp; ++s) {
            double o[4];
            e.single_scattering(px[i], py[i], o);
            for (int c = 0; c < 4; ++c) acc[c] += o[c];
        }
        for (int c = 0; c < 4; ++c) out[4 * i + c] = acc[c] / spp;
    }
    return 0;
}

/* Filter + running average + tonemap for one pass (no_iterations = passes before this one). */
int orc_post(void* h, const float* frame, float* running, uint8_t* display, int no_iterations, float* filtered)
{
    post_pass(*(Scene*)h, frame, running, display, no_iterations, filtered);
    return 0;
}

int orc_bilateral(void* h, const uint8_t* in, uint8_t* out)
{
    Scene& S = *(Scene*)h;
    std::vector<uint8_t> tmp((size_t)S.t.camera.film_size[0] * S.t.camera.film_size[1] * 4);
    bilateral_pass(S, in, tmp.data(), out);
    return 0;
}

/* compiler and flags of this build, for the CPU baseline's record */
const char* orc_build_info(void)
{
#ifdef __clang__
    return "clang " __clang_version__ ", -O3 -march=x86-64-v2 -ffp-contract=off, OpenMP";
#else
    return "g++ " __VERSION__ ", -O3 -march=x86-64-v2 -ffp-contract=off, OpenMP";
#endif
}

/* Single functions for pinning tests. */
float orc_intensity(void* h, float x, float y, float z)
{
    Rng r{1, 1};
    Est<float> e(*(Scene*)h, r, nullptr);
    return e.get_intensity(V3<float>(x, y, z));
}
/* GradientFD / CD / Filtered at P (volumes.h:31-74, mode = ER_GRADIENT_*), as computed
 * (grad3) and normalised (norm3, volumes.h:76-91) */
void orc_gradient(void* h, int mode, const float* p, float* grad3, float* norm3)
{
    Scene& S = *(Scene*)h;
    const int keep = S.t.settings.gradient_computation;
    S.t.settings.gradient_computation = mode;
    Rng r{1, 1};
    Est<float> e(S, r, nullptr);
    const V3<float> P(p[0], p[1], p[2]);
    const V3<float> g = mode == ER_GRADIENT_FORWARD ? e.grad_fd(P) : (mode == ER_GRADIENT_CENTRAL ? e.grad_cd(P) : e.grad_filtered(P));
    const V3<float> n = e.normalized_gradient(P);
    S.t.settings.gradient_computation = keep;
    for (int i = 0; i < 3; ++i) { grad3[i] = g[i]; norm3[i] = n[i]; }
}
/* GradientMagnitude at P (volumes.h:93-114) */
float orc_gradient_magnitude(void* h, const float* p)
{
    Rng r{1, 1};
    Est<float> e(*(Scene*)h, r, nullptr);
    return e.gradient_magnitude(V3<float>(p[0], p[1], p[2]));
}
/* shader pieces (shader.h:69-263), in the shading frame (z = normal) */
float orc_fresnel(float cosi, float ior) { return Est<float>::fresnel(cosi, ior); }
float orc_blinn_d(float exponent, float whz) { return Est<float>::blinn_d(exponent, V3<float>(0.0f, 0.0f, whz)); }
float orc_blinn_pdf(float exponent, const float* wo, const float* wi)
{
    return Est<float>::blinn_pdf(exponent, V3<float>(wo[0], wo[1], wo[2]), V3<float>(wi[0], wi[1], wi[2]));
}
void orc_microfacet_f(const float* wo, const float* wi, const float* ks, float exponent, float ior, float* out3)
{
    typename Est<float>::Shader s{};
    s.brdf = true;
    s.Ks = V3<float>(ks[0], ks[1], ks[2]);
    s.exponent = exponent;
    s.ior = ior;
    const V3<float> f = Est<float>::microfacet_f(s, V3<float>(wo[0], wo[1], wo[2]), V3<float>(wi[0], wi[1], wi[2]));
    for (int i = 0; i < 3; ++i) out3[i] = f[i];
}
float orc_plf_eval(const er_plf* p, float x) { return plf_eval(*p, x); }
double orc_plf_eval_f64(const er_plf* p, double x) { return plf_eval(*p, x); }
float orc_rng_get1(uint32_t* s0, uint32_t* s1) { Rng r{*s0, *s1}; const float v = r.get1(); *s0 = r.s0; *s1 = r.s1; return v; }
float orc_neglog(float u) { return M<float>::neglog(u); }
int orc_camera(void* h, float* out13)
{
    Scene& S = *(Scene*)h;
    for (int i = 0; i < 3; ++i) { out13[i] = S.camN[i]; out13[3 + i] = S.camU[i]; out13[6 + i] = S.camV[i]; }
    out13[9] = S.screen[0][0]; out13[10] = S.screen[1][0]; out13[11] = S.invscreen[0]; out13[12] = S.invscreen[1];
    return 0;
}
/* Primary march of one explicit ray (FP32): returns 1 on a scatter event, writes T and P and
 * the number of steps taken; u1, u2 are the two draws. */
int orc_march_ray(void* h, const float* o, const float* d, float mint, float maxt, int primary, float u1, float u2,
                  float* T, float* P, uint64_t* steps)
{
    /* Restates march() (raymarching.h:30-113) for one explicit ray with given draws u1
     * (free-path threshold) and u2 (jitter). */
    Scene& S = *(Scene*)h;
    Rng r{1, 1};
    Est<float> est(S, r, nullptr);
    const V3<float> mn = from<float>(S.minp), mx = from<float>(S.maxp);
    const Ray<float> ray(V3<float>(o[0], o[1], o[2]), V3<float>(d[0], d[1], d[2]), mint, maxt);
    float nearT, farT;
    *steps = 0;
    if (!Est<float>::intersect_box(ray, mn, mx, nearT, farT)) return -1;
    float MinT = std::max(nearT, ray.MinT);
    const float MaxT = std::min(farT, ray.MaxT);
    const float Sthr = M<float>::neglog(u1) / S.t.settings.density_scale;
    float Sum = 0.0f;
    const float StepSize = (primary ? S.t.settings.step_factor_primary : S.t.settings.step_factor_shadow) * S.minstep;
    MinT += u2 * StepSize;
    if (!(StepSize > 0.0f) || !(MaxT + StepSize > MaxT)) return 0;
    V3<float> Ps;
    uint64_t n = 0;
    while (Sum < Sthr) {
        Ps = ray.O + scale(MinT, ray.D);
        if (primary ? (MinT >= MaxT) : (MinT > MaxT)) { *steps = n; return 0; }
        ++n;
        const float SigmaT = est.sigma_at(Ps, primary != 0);
        Sum += SigmaT * StepSize;
        MinT += StepSize;
    }
    *steps = n;
    *T = MinT;
    P[0] = Ps[0]; P[1] = Ps[1]; P[2] = Ps[2];
    return 1;
}

/* the project's float elementary functions (er_fmath.h), vectorised for the accuracy tests:
 * fn 0 exp, 1 log, 2 pow, 3 sin, 4 cos, 5 atan2(x, y), 6 acos */
void orc_fmath(int fn, const float* x, const float* y, float* out, int n)
{
    for (int i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = erm::expf_(x[i]); break;
            case 1: out[i] = erm::logf_(x[i]); break;
            case 2: out[i] = erm::powf_(x[i], y[i]); break;
            case 3: out[i] = erm::sinf_(x[i]); break;
            case 4: out[i] = erm::cosf_(x[i]); break;
            case 5: out[i] = erm::atan2f_(x[i], y[i]); break;
            case 7: out[i] = erm::expm1f_(x[i]); break;
            default: out[i] = erm::acosf_(x[i]); break;
        }
    }
}

int orc_abi_sizeof(int* sizes, int n)
{
    const int s[11] = {(int)sizeof(er_tracer_desc), (int)sizeof(er_volume_desc), (int)sizeof(er_light_desc),
                       (int)sizeof(er_texture_desc), (int)sizeof(er_clipping_desc), (int)sizeof(er_object_desc),
                       (int)sizeof(er_bitmap_desc), (int)sizeof(er_plf), (int)sizeof(er_shape),
                       (int)sizeof(er_timing), (int)sizeof(er_counters)};
    int k = 0;
    for (; k < n && k < 11; ++k) sizes[k] = s[k];
    return k;
}

} /* extern "C" */
