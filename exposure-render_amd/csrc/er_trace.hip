/*
 * er_trace.hip - the per-pixel single-scattering estimator kernel for gfx950.
 *
 * One thread per pixel and pass (KrnlSingleScattering, Source/singlescattering.cuh:27-38).
 * A 256-thread workgroup covers a 16 x 16 pixel tile; each wave64 an 8 x 8 sub-tile, so the
 * 64 rays of a wave start from neighbouring pixels and march through neighbouring cells.
 * Workgroups are remapped so that each of the 8 XCDs (which receive workgroups round-robin)
 * renders a contiguous band of tiles and keeps that band's volume cells in its own L2.
 * RNG state lives in registers for the whole path and is written back once (the reference
 * read-modify-writes it in global memory on every draw, Source/rng.h:37-38).
 *
 * Template parameters: FULL selects the full estimator (shading-mode selection, both halves
 * of MIS, clipping objects, Woodcock sampler); COUNT adds the sample counters.
 */
#include "er_device.h"

namespace erd {

struct Cnt {
    uint32_t primary, shadow, shading, pfetch, sfetch, scatter, shadow_rays;
};

struct Isect {
    bool Valid, Front;
    float NearT;
    f3 P, N;
    float U, V;
};

__device__ __forceinline__ void isect_init(Isect& I)
{
    I.Valid = false; I.Front = true; I.NearT = 0.0f;
    I.P = mk(0.0f, 0.0f, 0.0f); I.N = mk(0.0f, 0.0f, 0.0f); I.U = 0.0f; I.V = 0.0f;
}

/* ---- clipping objects (full mode; box.h:116-120, sphere.h:147-150, plane.h:90-93,
 *      cylinder.h:96-99) -------------------------------------------------------------- */
__device__ __forceinline__ bool clipped(const DevParams& p, f3 P)
{
    for (int k = 0; k < p.clip_count; ++k) {
        const DevClip& c = p.clips[k];
        const f3 Pl = xform_point(c.inv_tm, P);
        bool inside;
        if (c.type == ER_SHAPE_BOX) {
            const float h0 = 0.5f * c.size[0], h1 = 0.5f * c.size[1], h2 = 0.5f * c.size[2];
            inside = Pl.x > -h0 && Pl.x < h0 && Pl.y > -h1 && Pl.y < h1 && Pl.z > -h2 && Pl.z < h2;
        } else if (c.type == ER_SHAPE_SPHERE) {
            inside = len(Pl) < c.outer_radius;
        } else if (c.type == ER_SHAPE_CYLINDER) {
            inside = sqrtf((Pl.x * Pl.x) + (Pl.z * Pl.z)) < c.outer_radius && fabsf(Pl.y) < (0.5f * c.size[1]);
        } else {
            inside = Pl.z > 0.0f;
        }
        if (c.invert ? !inside : inside) return true;
    }
    return false;
}

/* ---- free-flight sampling --------------------------------------------------------------- */

/* sigma of a sample at Ps with lattice coordinates l; 0 without a fetch when its macrocell
 * holds no opacity or (full mode) the point is clipped */
template<int FULL, int COUNT>
__device__ __forceinline__ float sample_sigma(const DevParams& p, f3 Ps, bool primary, Cnt& cnt)
{
    const Lattice l = to_lattice(p, Ps);
    if (!(p.mc_sigma[macrocell(p, l)] > 0.0f)) return 0.0f;
    if (FULL && p.clip_count > 0 && clipped(p, Ps)) return 0.0f;
    if (COUNT) { if (primary) cnt.pfetch++; else cnt.sfetch++; }
    return sigma_lut(p, trilinear_cell(p, l));
}

/* SampleVolume (raymarching.h:30-71) for primary rays, ScatterEventInVolume
 * (raymarching.h:73-113) for shadow rays: jittered fixed-step tracking with threshold
 * S = -log(u) / DensityScale; T is reported one step beyond P (Q13); the primary march ends
 * at t >= tmax, the shadow march at t > tmax (Q14). */
template<int FULL, int COUNT>
__device__ bool march(const DevParams& p, const Ray& r, bool primary, Rng& rng, float& Tout, f3& Pout, Cnt& cnt)
{
    float nearT, farT;
    if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return false;
    float MinT = fmaxf(nearT, r.MinT);
    const float MaxT = fminf(farT, r.MaxT);

    if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) {
        /* delta tracking against the global majorant DS^2 * max opacity */
        const float mu = p.woodcock_majorant;
        if (!(mu > 0.0f)) return false;
        float t = MinT;
        for (;;) {
            t += p.neglog[rng.next_bits()] / mu;
            if (primary ? (t >= MaxT) : (t > MaxT)) return false;
            const f3 Ps = add(r.O, scl(t, r.D));
            if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; }
            const float sigma = sample_sigma<FULL, COUNT>(p, Ps, primary, cnt) * p.density_scale;
            if (rng.u() * mu < sigma) { Tout = t; Pout = Ps; return true; }
        }
    }

    const float S = p.neglog[rng.next_bits()] / p.density_scale;
    const float StepSize = (primary ? p.step_primary : p.step_shadow) * p.minstep;
    MinT += rng.u() * StepSize;
    if (!(StepSize > 0.0f)) return false;   /* zero-size volume (Q2): the march would not advance */
    float Sum = 0.0f;
    f3 Ps = mk(0.0f, 0.0f, 0.0f);
    while (Sum < S) {
        Ps = add(r.O, scl(MinT, r.D));
        if (primary ? (MinT >= MaxT) : (MinT > MaxT)) return false;
        if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; }
        const float SigmaT = sample_sigma<FULL, COUNT>(p, Ps, primary, cnt);
        Sum += SigmaT * StepSize;
        MinT += StepSize;
    }
    Tout = MinT;
    Pout = Ps;
    return true;
}

/* ---- shapes (plane.h:41-88, disk.h, ring.h, box.h:51-83, sphere.h:84-140) ------------ */
__device__ __forceinline__ void intersect_plane(const Ray& r, bool one_sided, Isect& I)
{
    if (fabsf(r.O.z - r.D.z) < ER_RAY_EPS) return;   /* Q16 */
    I.NearT = (0.0f - r.O.z) / r.D.z;
    if (I.NearT < r.MinT || I.NearT > r.MaxT) return;
    I.P = ray_at(r, I.NearT);
    I.U = I.P.x; I.V = I.P.y;
    I.N = mk(0.0f, 0.0f, 1.0f);
    if (one_sided && r.D.z >= 0.0f) { I.Front = false; I.N = mk(0.0f, 0.0f, -1.0f); }
    I.Valid = true;
}

__device__ __forceinline__ void intersect_sphere(const Ray& r, float radius, bool need_uv, Isect& I)
{
    const float a = dot(r.D, r.D);
    const float b = 2.0f * dot(r.D, r.O);
    const float c = dot(r.O, r.O) - (radius * radius);
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return;
    const float distSqrt = sqrtf(disc);
    const float q = (b < 0.0f) ? (-b - distSqrt) / 2.0f : (-b + distSqrt) / 2.0f;
    float t0 = q / a, t1 = c / q;
    if (t0 > t1) { const float tmp = t0; t0 = t1; t1 = tmp; }
    if (t0 >= r.MinT && t0 < r.MaxT) I.NearT = t0;
    else if (t1 >= r.MinT && t1 < r.MaxT) I.NearT = t1;
    else return;
    if (I.NearT < r.MinT || I.NearT > r.MaxT) return;
    I.Valid = true;
    I.P = ray_at(r, I.NearT);
    I.N = nrm(I.P);
    if (need_uv) {   /* SphericalToUV (geometry.h:70-83) */
        const f3 V = nrm(I.P);
        float phi = erm::atan2f_(V.z, V.x);
        phi = (phi < 0.0f) ? phi + (2.0f * ER_PI_F) : phi;
        const float theta = erm::acosf_(clampf_(V.y, -1.0f, 1.0f));
        I.U = ER_INV_TWO_PI_F * phi;
        I.V = 1.0f - (ER_INV_PI_F * theta);
    }
}

__device__ __forceinline__ void intersect_shape(const DevLight& s, const Ray& r, bool need_uv, Isect& I)
{
    if (s.type == ER_SHAPE_PLANE) {
        intersect_plane(r, s.one_sided != 0, I);
        const float hx = 0.5f * s.size[0], hy = 0.5f * s.size[1];
        if (I.Valid && (I.U < -hx || I.U > hx || I.V < -hy || I.V > hy)) I.Valid = false;
        I.U /= s.size[0]; I.V /= s.size[1];
        I.U += 0.5f; I.V += 0.5f;
        I.U = 1.0f - I.U;
    } else if (s.type == ER_SHAPE_DISK || s.type == ER_SHAPE_RING) {
        intersect_plane(r, s.one_sided != 0, I);
        const float l = sqrtf(I.U * I.U + I.V * I.V);
        if (s.type == ER_SHAPE_DISK) { if (I.Valid && l > s.outer_radius) I.Valid = false; }
        else if (I.Valid && (l < s.inner_radius || l > s.outer_radius)) I.Valid = false;
        const float dia = 2.0f * s.outer_radius;
        I.U /= dia; I.V /= dia;
        I.U += 0.5f; I.V += 0.5f;
        I.U = 1.0f - I.U;
    } else if (s.type == ER_SHAPE_BOX) {
        const f3 mx = mk(0.5f * s.size[0], 0.5f * s.size[1], 0.5f * s.size[2]);
        const f3 mn = mk(-0.5f * s.size[0], -0.5f * s.size[1], -0.5f * s.size[2]);
        float nT, fT;
        if (intersect_box(r, mn, mx, nT, fT)) {
            I.NearT = nT; I.Valid = true;
            I.P = ray_at(r, nT);
            float n[3] = {0.0f, 0.0f, 0.0f};
            const float pp[3] = {I.P.x, I.P.y, I.P.z}, lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
            for (int i = 0; i < 3; ++i) {
                if (pp[i] <= lo[i] + 0.0001f) n[i] = -1.0f;
                if (pp[i] >= hi[i] - 0.0001f) n[i] = 1.0f;
            }
            I.N = mk(n[0], n[1], n[2]);
        }
    } else if (s.type == ER_SHAPE_SPHERE) {
        intersect_sphere(r, s.outer_radius, need_uv, I);
    }
}

/* EvaluateTexture / EvaluateProcedural (textures.h:26-111) */
__device__ __forceinline__ bool texture_needs_uv(const DevParams& p, int id)
{
    if (id < 0 || id >= ER_LIST_CAPACITY) return false;
    const DevTexture& T = p.textures[id];
    return T.valid && !(T.type == ER_TEXTURE_PROCEDURAL && T.proc_type == ER_PROCEDURAL_UNIFORM);
}

__device__ f3 eval_texture(const DevParams& p, int id, float u, float v)
{
    if (id < 0 || id >= ER_LIST_CAPACITY || !p.textures[id].valid) return mk(0.0f, 0.0f, 0.0f);
    const DevTexture& T = p.textures[id];
    u *= T.repeat[0]; v *= T.repeat[1];
    u += T.offset[0]; v += 1.0f - T.offset[1];
    u = u - floorf(u); v = v - floorf(v);
    u = clampf_(u, 0.0f, 1.0f); v = clampf_(v, 0.0f, 1.0f);
    if (T.flip[0]) u = 1.0f - u;
    if (T.flip[1]) v = 1.0f - v;
    f3 L = mk(0.0f, 0.0f, 0.0f);
    if (T.type == ER_TEXTURE_PROCEDURAL) {
        if (T.proc_type == ER_PROCEDURAL_UNIFORM) {
            L = ld3(T.uniform);
        } else if (T.proc_type == ER_PROCEDURAL_CHECKER) {
            const int a = (int)(u * 2.0f), b = (int)(v * 2.0f);
            const bool c1 = (a % 2 == 0) ? (b % 2 == 0) : (b % 2 != 0);
            L = c1 ? ld3(T.checker1) : ld3(T.checker2);
        } else if (T.proc_type == ER_PROCEDURAL_GRADIENT && T.gradient_index >= 0) {
            L = ctf_eval(p.tex_gradients[T.gradient_index], v);
        }
    }
    return scl(T.output_level, L);
}

/* ---- shading (shader.h:29-483) ------------------------------------------------------- */
struct Shader {
    bool brdf;
    f3 Nn, Nu, Nv, Kd, Ks;
    float exponent;
};

__device__ __forceinline__ f3 to_local(const Shader& s, f3 W) { return mk(dot(W, s.Nu), dot(W, s.Nv), dot(W, s.Nn)); }
__device__ __forceinline__ f3 to_world(const Shader& s, f3 W)
{
    return mk(s.Nu.x * W.x + s.Nv.x * W.y + s.Nn.x * W.z,
              s.Nu.y * W.x + s.Nv.y * W.y + s.Nn.y * W.z,
              s.Nu.z * W.x + s.Nv.z * W.y + s.Nn.z * W.z);
}
__device__ __forceinline__ bool same_hemi(f3 a, f3 b) { return (a.z * b.z) > 0.0f; }

/* Fresnel::Evaluate + FrDiel with EtaI = 1, EtaT = 15 (shader.h:69-123, transport.h:138) */
__device__ __forceinline__ float fresnel(float cosi)
{
    cosi = clampf_(cosi, -1.0f, 1.0f);
    const bool entering = cosi > 0.0f;
    float ei = 1.0f, et = 15.0f;
    if (!entering) { ei = 15.0f; et = 1.0f; }
    const float sint = ei / et * sqrtf(fmaxf(0.0f, 1.0f - cosi * cosi));
    if (sint >= 1.0f) return 1.0f;
    const float cost = sqrtf(fmaxf(0.0f, 1.0f - sint * sint));
    const float ci = fabsf(cosi);
    const float Rparl = ((et * ci) - (ei * cost)) / ((et * ci) + (ei * cost));
    const float Rperp = ((ei * ci) - (et * cost)) / ((ei * ci) + (et * cost));
    return (Rparl * Rparl + Rperp * Rperp) / 2.0f;
}

__device__ __forceinline__ float blinn_pdf(float exponent, f3 wo, f3 wi)   /* shader.h:166-179 */
{
    const f3 wh = nrm(add(wo, wi));
    const float c = fabsf(wh.z);
    float pdf = ((exponent + 1.0f) * erm::powf_(c, exponent)) / (2.0f * ER_PI_F * 4.0f * dot(wo, wh));
    if (dot(wo, wh) <= 0.0f) pdf = 0.0f;
    return pdf;
}

__device__ f3 microfacet_f(const Shader& s, f3 wo, f3 wi)   /* shader.h:216-263 */
{
    const float cO = fabsf(wo.z), cI = fabsf(wi.z);
    if (cI == 0.0f || cO == 0.0f) return mk(0.0f, 0.0f, 0.0f);
    f3 wh = add(wi, wo);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return mk(0.0f, 0.0f, 0.0f);
    wh = nrm(wh);
    const float cosH = dot(wi, wh);
    const float Fr = fresnel(cosH);
    const float D = (s.exponent + 2.0f) * ER_INV_TWO_PI_F * erm::powf_(fabsf(wh.z), s.exponent);
    const float NdotWh = fabsf(wh.z), NdotWo = fabsf(wo.z), NdotWi = fabsf(wi.z);
    const float WOdotWh = absdot(wo, wh);
    const float G = fminf(1.0f, fminf((2.0f * NdotWh * NdotWo / WOdotWh), (2.0f * NdotWh * NdotWi / WOdotWh)));
    const float den = 4.0f * cI * cO;
    return mk((((s.Ks.x * D) * G) * Fr) / den, (((s.Ks.y * D) * G) * Fr) / den, (((s.Ks.z * D) * G) * Fr) / den);
}

__device__ __forceinline__ float lambert_pdf(f3 wo, f3 wi) { return same_hemi(wo, wi) ? fabsf(wi.z) * ER_INV_PI_F : 0.0f; }
__device__ __forceinline__ float microfacet_pdf(const Shader& s, f3 wo, f3 wi) { return same_hemi(wo, wi) ? blinn_pdf(s.exponent, wo, wi) : 0.0f; }

template<int FULL>
__device__ f3 shader_f(const Shader& s, f3 Wo, f3 Wi)
{
    if (!s.brdf) return scl(ER_INV_PI_F, s.Kd);   /* IsotropicPhase::F (Q7) */
    const f3 wo = to_local(s, Wo), wi = to_local(s, Wi);
    f3 R = mk(0.0f, 0.0f, 0.0f);
    R = add(R, scl(ER_INV_PI_F, s.Kd));
    R = add(R, microfacet_f(s, wo, wi));
    return R;
}

template<int FULL>
__device__ float shader_pdf(const Shader& s, f3 Wo, f3 Wi)
{
    if (!s.brdf) return FULL ? 0.0795774715459476678844f : (1.0f / 4.0f * 3.141592654f);   /* Q6 in compat */
    const f3 wo = to_local(s, Wo), wi = to_local(s, Wi);
    float pdf = 0.0f;
    pdf += lambert_pdf(wo, wi);
    pdf += microfacet_pdf(s, wo, wi);
    return FULL ? 0.5f * pdf : pdf;
}

/* ConcentricSampleDisk (montecarlo.h:86-137) */
__device__ __forceinline__ void concentric_disk(float u0, float u1, float& ox, float& oy)
{
    float r, theta;
    const float sx = 2.0f * u0 - 1.0f, sy = 2.0f * u1 - 1.0f;
    if (sx == 0.0f && sy == 0.0f) { ox = 0.0f; oy = 0.0f; return; }
    if (sx >= -sy) {
        if (sx > sy) { r = sx; theta = (sy > 0.0f) ? sy / r : 8.0f + sy / r; }
        else { r = sy; theta = 2.0f - sx / r; }
    } else {
        if (sx <= sy) { r = -sx; theta = 4.0f - sy / r; }
        else { r = -sy; theta = 6.0f + sx / r; }
    }
    theta *= ER_PI_F / 4.0f;
    float sn, cs;
    erm::sincosf_(theta, &sn, &cs);
    ox = r * cs;
    oy = r * sn;
}

/* BRDF::SampleF / IsotropicPhase::SampleF (shader.h:297-303,350-374), full mode: the pdf is
 * the one-sample mixture density 0.5 (pdf_L + pdf_M); the phase pdf is 1 / (4 pi). */
__device__ f3 shader_sample_f(const Shader& s, f3 Wo, f3& Wi, float& pdf, float comp, float d0, float d1)
{
    if (!s.brdf) {
        const float z = 1.0f - 2.0f * d0;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = (2.0f * ER_PI_F) * d1;
        float sn, cs;
        erm::sincosf_(phi, &sn, &cs);
        Wi = mk(rr * cs, rr * sn, z);
        pdf = 0.0795774715459476678844f;
        return scl(ER_INV_PI_F, s.Kd);
    }
    const f3 wo = to_local(s, Wo);
    f3 wi;
    if (comp <= 0.5f) {
        float dx, dy;
        concentric_disk(d0, d1, dx, dy);
        wi = mk(dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy)));
        if (wo.z < 0.0f) wi.z *= -1.0f;
    } else {
        const float costheta = erm::powf_(d0, 1.0f / (s.exponent + 1.0f));
        const float sintheta = sqrtf(fmaxf(0.0f, 1.0f - costheta * costheta));
        const float phi = d1 * 2.0f * ER_PI_F;
        float sn, cs;
        erm::sincosf_(phi, &sn, &cs);
        f3 wh = mk(sintheta * cs, sintheta * sn, costheta);
        if (!same_hemi(wo, wh)) wh = neg(wh);
        const float two_d = 2.0f * dot(wo, wh);
        wi = add(neg(wo), scl(two_d, wh));
    }
    pdf = 0.5f * (lambert_pdf(wo, wi) + microfacet_pdf(s, wo, wi));
    f3 R = mk(0.0f, 0.0f, 0.0f);
    R = add(R, scl(ER_INV_PI_F, s.Kd));
    R = add(R, microfacet_f(s, wo, wi));
    Wi = to_world(s, wi);
    return R;
}

__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

/* ---- lights (lights.h:28-131) -------------------------------------------------------- */
__device__ __forceinline__ const DevLight* light_by_id(const DevParams& p, int id)
{
    if (id < 0 || id >= ER_LIST_CAPACITY || !p.lights[id].valid) return nullptr;
    return &p.lights[id];
}

/* IntersectLights (lights.h:90-113); returns the list position of the nearest hit, or -1 */
__device__ int intersect_lights(const DevParams& p, const Ray& r, bool respect_visibility, float& T, f3& P, f3& N, f3& Le)
{
    float best = 3.402823466e+38f;
    int hit = -1;
    for (int i = 0; i < p.light_count; ++i) {
        const DevLight* L = light_by_id(p, p.light_ids[i]);
        if (!L) continue;
        if (respect_visibility && !L->visible) continue;
        const Ray rt = xform_ray(L->inv_tm, r);
        Isect I;
        isect_init(I);
        intersect_shape(*L, rt, texture_needs_uv(p, L->texture_id), I);
        if (!I.Valid) continue;
        const f3 Pw = xform_point(L->tm, I.P);
        const float t = len(sub(Pw, r.O));
        if (t < best) {
            best = t; hit = i; T = t; P = Pw;
            N = xform_vector(L->tm, I.N);
            f3 le = I.Front ? scl(L->multiplier, eval_texture(p, L->texture_id, I.U, I.V)) : mk(0.0f, 0.0f, 0.0f);
            if (L->unit == ER_UNIT_LUX) le = mk(le.x / L->area, le.y / L->area, le.z / L->area);
            Le = le;
        }
    }
    return hit;
}

/* IntersectsLight (lights.h:115-131) */
__device__ bool intersects_any_light(const DevParams& p, const Ray& r)
{
    for (int i = 0; i < p.light_count; ++i) {
        const DevLight* L = light_by_id(p, p.light_ids[i]);
        if (!L) continue;
        Isect I;
        isect_init(I);
        intersect_shape(*L, xform_ray(L->inv_tm, r), false, I);
        if (I.Valid) return true;
    }
    return false;
}

/* Visible (transport.h:46-56) -> Intersect (transport.h:30-44) */
template<int FULL, int COUNT>
__device__ bool visible(const DevParams& p, f3 P1, f3 P2, Rng& rng, Cnt& cnt)
{
    if (!p.shadows) return true;
    const f3 W = nrm(sub(P2, P1));
    Ray r;
    r.O = add(P1, scl(ER_RAY_EPS, W));
    r.D = W;
    r.MinT = 0.0f;
    r.MaxT = fminf(len(sub(P2, P1)) - (2.0f * ER_RAY_EPS), p.max_shadow);
    if (COUNT) cnt.shadow_rays++;
    if (intersects_any_light(p, r)) return false;
    float T; f3 P;
    if (march<FULL, COUNT>(p, r, false, rng, T, P, cnt)) return false;
    return true;
}

/* SampleLightSurface (lights.h:28-41) */
__device__ void sample_light_surface(const DevParams& p, const DevLight& L, float u0, float u1, float u2, f3& P, f3& N, float& U, float& V)
{
    N = mk(0.0f, 0.0f, 1.0f);
    U = 0.0f; V = 0.0f;
    if (L.type == ER_SHAPE_PLANE) {
        P = mk(-0.5f + u0, -0.5f + u1, 0.0f);
        U = u0; V = u1;
        P = mul(P, mk(L.size[0], L.size[1], 0.0f));
    } else if (L.type == ER_SHAPE_DISK) {
        const float rr = sqrtf(u0);
        const float th = (2.0f * ER_PI_F) * u1;
        float sn, cs;
        erm::sincosf_(th, &sn, &cs);
        P = mk(rr * cs, rr * sn, 0.0f);
        U = u0; V = u1;
        P = scl(L.outer_radius, P);
    } else if (L.type == ER_SHAPE_RING) {
        const float inner = L.inner_radius / L.outer_radius;
        const float rr = inner + (1.0f - inner) * sqrtf(u0);
        const float th = (2.0f * ER_PI_F) * u1;
        float sn, cs;
        erm::sincosf_(th, &sn, &cs);
        P = mk(rr * cs, rr * sn, 0.0f);
        P = scl(L.outer_radius, P);
        U = P.x; V = P.y;
    } else if (L.type == ER_SHAPE_BOX) {
        const int side = (int)floorf(u0 * 6.0f);
        const float h = 0.5f;
        switch (side) {
            case 0: P = mk(h, -h + u2, -h + u1); N = mk(1.0f, 0.0f, 0.0f); break;
            case 1: P = mk(-h, -h + u2, -h + u1); N = mk(-1.0f, 0.0f, 0.0f); break;
            case 2: P = mk(-h + u1, h, -h + u2); N = mk(0.0f, 1.0f, 0.0f); break;
            case 3: P = mk(-h + u1, -h, -h + u2); N = mk(0.0f, -1.0f, 0.0f); break;
            case 4: P = mk(-h + u1, -h + u2, h); N = mk(0.0f, 0.0f, 1.0f); break;
            default: P = mk(-h + u1, -h + u2, -h); N = mk(0.0f, 0.0f, -1.0f); break;
        }
        U = u1; V = u2;
        P = mul(P, ld3(L.size));
    } else if (L.type == ER_SHAPE_SPHERE) {
        const float z = 1.0f - 2.0f * u0;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = (2.0f * ER_PI_F) * u1;
        float sn, cs;
        erm::sincosf_(phi, &sn, &cs);
        P = mk(rr * cs, rr * sn, z);
        N = P;
        if (texture_needs_uv(p, L.texture_id)) {
            float ph = erm::atan2f_(P.z, P.x);
            ph = (ph < 0.0f) ? ph + (2.0f * ER_PI_F) : ph;
            U = erm::acosf_(clampf_(P.y, -1.0f, 1.0f));
            V = ph;
        }
        P = scl(L.outer_radius, P);
    }
    P = xform_point(L.tm, P);
    N = xform_vector(L.tm, N);
}

struct Sample { float comp, dir0, dir1, uvw0, uvw1, uvw2, lightnum, film0, film1, lens0, lens1; };

/* EstimateDirectLight (transport.h:58-113) */
template<int FULL, int COUNT>
__device__ f3 estimate_direct(const DevParams& p, const DevLight& L, int light_pos, const Sample& smp, f3 SEP, f3 SEN, f3 Wo,
                              const Shader& sh, Rng& rng, Cnt& cnt)
{
    f3 Ld = mk(0.0f, 0.0f, 0.0f);
    f3 SP, SN; float su, sv;
    sample_light_surface(p, L, smp.uvw0, smp.uvw1, smp.uvw2, SP, SN, su, sv);
    f3 Wi = nrm(sub(SP, SEP));
    f3 Li = scl(L.multiplier, eval_texture(p, L.texture_id, su, sv));
    if (L.one_sided && dot(sub(SEP, SP), SN) < 0.0f) Li = mk(0.0f, 0.0f, 0.0f);
    if (L.unit == ER_UNIT_LUX) Li = mk(Li.x / L.area, Li.y / L.area, Li.z / L.area);
    f3 F = shader_f<FULL>(sh, Wo, Wi);
    float BsdfPdf = shader_pdf<FULL>(sh, Wo, Wi);
    if (!black(Li) && !black(F) && BsdfPdf > 0.0f && visible<FULL, COUNT>(p, SEP, SP, rng, cnt)) {
        const f3 d = sub(SEP, SP);
        const float LightPdf = dot(d, d) / (absdot(SN, neg(Wi)) * L.area);
        const float Weight = power_heuristic(LightPdf, BsdfPdf);
        const f3 FL = mul(F, Li);
        if (sh.brdf) Ld = add(Ld, scl(absdot(Wi, SEN) * Weight / LightPdf, FL));
        else if (FULL) Ld = add(Ld, scl(Weight / LightPdf, FL));
        else Ld = add(Ld, mk(FL.x / LightPdf, FL.y / LightPdf, FL.z / LightPdf));
    }
    if (!FULL) return Ld;

    F = shader_sample_f(sh, Wo, Wi, BsdfPdf, smp.comp, smp.dir0, smp.dir1);
    if (black(F) || BsdfPdf <= 0.0f) return Ld;
    Ray r2;
    r2.O = SEP; r2.D = Wi; r2.MinT = 0.0f; r2.MaxT = 1000000.0f;
    float T2; f3 P2, N2, Le2;
    const int idx2 = intersect_lights(p, r2, false, T2, P2, N2, Le2);
    if (idx2 < 0 || idx2 != light_pos) return Ld;
    if (!black(Le2) && visible<FULL, COUNT>(p, SEP, P2, rng, cnt)) {
        const f3 d = sub(SEP, P2);
        const float LightPdf = dot(d, d) / (absdot(N2, neg(Wi)) * L.area);
        const float Weight = power_heuristic(BsdfPdf, LightPdf);
        const f3 FL = mul(F, Le2);
        if (sh.brdf) Ld = add(Ld, scl(absdot(Wi, SEN) * Weight / BsdfPdf, FL));
        else Ld = add(Ld, scl(Weight / BsdfPdf, FL));
    }
    return Ld;
}

/* UniformSampleOneLight (transport.h:115-156) with the full-mode shader selection
 * (design at scatterevent.h:155-212) */
template<int FULL, int COUNT>
__device__ f3 uniform_sample_one_light(const DevParams& p, const Sample& smp, f3 SEP, f3 SEN, f3 Wo, Rng& rng, Cnt& cnt)
{
    f3 Ld = mk(0.0f, 0.0f, 0.0f);
    const float I = intensity(p, SEP);
    if (COUNT) cnt.shading += 1 + gradient_samples(p);
    Ld = add(Ld, ctf_eval(p.emission, I));
    const int count = p.light_count;
    if (count <= 0) return Ld;
    const int pos = (int)floorf(smp.lightnum * (float)count);
    const int LightID = p.light_ids[pos];
    const DevLight* L = light_by_id(p, LightID);
    if (LightID < 0 || !L) return Ld;
    bool brdf = true;
    if (FULL) {
        const int type = p.shading_type;
        if (type == ER_SHADING_PHASE_ONLY) {
            brdf = false;
        } else if (type != ER_SHADING_BRDF_ONLY) {
            const float gm = gradient_magnitude(p, SEP);
            if (COUNT) cnt.shading += 6;
            const float NGM = gm * p.gm_inv;
            if (type == ER_SHADING_THRESHOLD) {
                brdf = NGM > p.grad_threshold;
            } else if (type == ER_SHADING_MODULATION) {
                const float PdfBrdf = 1.0f - erm::powf_(1.0f - NGM, 2.0f);
                brdf = rng.u() < PdfBrdf;
            } else {
                const float e = 25.0f * erm::powf_(p.grad_factor, 3.0f) * NGM;
                const float base = 1.0f - erm::expf_(-e);
                const float PdfBrdf = p.opacity_modulated ? plf_eval(p.opacity, I) * base : base;
                brdf = rng.u() <= PdfBrdf;
            }
        }
    }
    Shader sh;
    sh.brdf = brdf;
    sh.Kd = ctf_eval(p.diffuse, I);
    sh.Ks = ctf_eval(p.specular, I);
    sh.exponent = 1000000.0f * erm::powf_(plf_eval(p.glossiness, I), 7.0f);   /* utilities.h:28-31 */
    sh.Nn = nrm(SEN);
    sh.Nu = nrm(crs(SEN, Wo));
    sh.Nv = nrm(crs(SEN, sh.Nu));
    Ld = add(Ld, estimate_direct<FULL, COUNT>(p, *L, pos, smp, SEP, SEN, Wo, sh, rng, cnt));
    return scl((float)count, Ld);
}

/* SingleScattering (singlescattering.h:29-102) */
template<int FULL, int COUNT>
__device__ float4 single_scattering(const DevParams& p, int px, int py, Rng& rng, Cnt& cnt)
{
    Sample smp;
    smp.comp = rng.u();
    smp.dir0 = rng.u(); smp.dir1 = rng.u();
    smp.uvw0 = rng.u(); smp.uvw1 = rng.u(); smp.uvw2 = rng.u();
    smp.lightnum = rng.u();
    smp.film0 = rng.u(); smp.film1 = rng.u();
    smp.lens0 = rng.u(); smp.lens1 = rng.u();

    /* SampleCamera (singlescattering.h:29-50) */
    const float sx = p.screen00 + (p.invscreen0 * ((float)px + smp.film0 * 1.0f / (float)p.film_w));
    const float sy = p.screen10 + (p.invscreen1 * ((float)py + smp.film1 * 1.0f / (float)p.film_h));
    const f3 N = ld3(p.cam_N), U = ld3(p.cam_U), V = ld3(p.cam_V);
    Ray r;
    r.O = ld3(p.cam_pos);
    r.D = nrm(sub(add(N, scl(sx, U)), scl(sy, V)));
    r.MinT = p.clip_near;
    r.MaxT = p.clip_far;
    if (p.aperture != 0.0f) {
        float lx, ly;
        concentric_disk(smp.lens0, smp.lens1, lx, ly);
        lx = lx * p.aperture; ly = ly * p.aperture;
        const f3 LI = add(scl(lx, U), scl(ly, V));
        r.O = add(r.O, LI);
        r.D = nrm(sub(scl(p.focal, r.D), LI));
    }

    /* SampleRay (singlescattering.h:52-74) */
    float Tv = 0.0f; f3 Pv = mk(0.0f, 0.0f, 0.0f);
    const bool vol = march<FULL, COUNT>(p, r, true, rng, Tv, Pv, cnt);
    float Tl = 0.0f; f3 Pl, Nl, Lel = mk(0.0f, 0.0f, 0.0f);
    const int il = intersect_lights(p, r, true, Tl, Pl, Nl, Lel);
    int type = -1;
    float T = 3.402823466e+38f;
    if (vol && Tv < T) { type = 0; T = Tv; }
    if (il >= 0 && Tl < T) { type = 1; T = Tl; }

    f3 Lv = mk(0.0f, 0.0f, 0.0f);
    if (type == 0) {
        if (COUNT) cnt.scatter++;
        const f3 Nv = normalized_gradient(p, Pv);   /* raymarching.h:70 */
        Lv = add(Lv, uniform_sample_one_light<FULL, COUNT>(p, smp, Pv, Nv, neg(r.D), rng, cnt));
    }
    if (type == 1) Lv = add(Lv, Lel);
    return make_float4(Lv.x, Lv.y, Lv.z, type >= 0 ? 1.0f : 0.0f);
}

/* XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a contiguous band */
__device__ __forceinline__ int tile_of_block(int b, int n)
{
    const int per = (n + 7) >> 3;
    const int xcd = b & 7, local = b >> 3;
    const int t = xcd * per + local;
    return t;
}

template<int FULL, int COUNT>
__global__ __launch_bounds__(256) void k_trace(const DevParams* __restrict__ pp, float4* __restrict__ frame,
                                              uint32_t* __restrict__ seeds1, uint32_t* __restrict__ seeds2,
                                              unsigned long long* __restrict__ counters, int tiles_x, int ntiles)
{
    const DevParams& p = *pp;
    /* blocks beyond the band layout of the last XCD get remapped tiles >= ntiles: skip */
    const int tile = tile_of_block(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0};
    bool active = tile < ntiles;
    int px = 0, py = 0;
    if (active) {
        const int tx = tile % tiles_x, ty = tile / tiles_x;
        px = tx * 16 + (wave & 1) * 8 + (lane & 7);
        py = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
        active = px < p.film_w && py < p.film_h;
    }
    if (active) {
        const size_t i = (size_t)py * p.film_w + px;
        Rng rng;
        rng.s0 = seeds1[i];
        rng.s1 = seeds2[i];
        frame[i] = single_scattering<FULL, COUNT>(p, px, py, rng, cnt);
        seeds1[i] = rng.s0;
        seeds2[i] = rng.s1;
    }
    if (COUNT) {
        __shared__ unsigned long long acc[8];
        if (threadIdx.x < 8) acc[threadIdx.x] = 0ull;
        __syncthreads();
        if (active) {
            atomicAdd(&acc[0], 1ull);
            atomicAdd(&acc[1], (unsigned long long)cnt.primary);
            atomicAdd(&acc[2], (unsigned long long)cnt.shadow);
            atomicAdd(&acc[3], (unsigned long long)cnt.shading);
            atomicAdd(&acc[4], (unsigned long long)cnt.pfetch);
            atomicAdd(&acc[5], (unsigned long long)cnt.sfetch);
            atomicAdd(&acc[6], (unsigned long long)cnt.scatter);
            atomicAdd(&acc[7], (unsigned long long)cnt.shadow_rays);
        }
        __syncthreads();
        if (threadIdx.x < 8) atomicAdd(&counters[threadIdx.x], acc[threadIdx.x]);
    }
}

} /* namespace erd */

/* launcher used by the runtime */
extern "C" hipError_t erd_launch_trace(const erd::DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2,
                                       unsigned long long* counters, int W, int H, int full, int count, hipStream_t st)
{
    const int tiles_x = (W + 15) / 16, tiles_y = (H + 15) / 16;
    const int ntiles = tiles_x * tiles_y;
    /* round the grid up to a multiple of 8 so every XCD band has the same length */
    const int per = (ntiles + 7) / 8;
    const int nblocks = per * 8;
    dim3 grid(nblocks), block(256);
    if (full) {
        if (count) hipLaunchKernelGGL((erd::k_trace<1, 1>), grid, block, 0, st, p, frame, s1, s2, counters, tiles_x, ntiles);
        else hipLaunchKernelGGL((erd::k_trace<1, 0>), grid, block, 0, st, p, frame, s1, s2, counters, tiles_x, ntiles);
    } else {
        if (count) hipLaunchKernelGGL((erd::k_trace<0, 1>), grid, block, 0, st, p, frame, s1, s2, counters, tiles_x, ntiles);
        else hipLaunchKernelGGL((erd::k_trace<0, 0>), grid, block, 0, st, p, frame, s1, s2, counters, tiles_x, ntiles);
    }
    return hipGetLastError();
}
