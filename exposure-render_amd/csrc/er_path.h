/*
 * er_path.h - device functions of one single-scattering path (Source/singlescattering.h:29-102):
 * free-flight sampling (ray march with exact empty-space skipping, Woodcock), light and shape
 * intersection, textures, shaders, direct lighting. Shared by the forward kernels (er_trace.hip)
 * and the path-replay adjoint (er_adjoint.hip).
 */
#ifndef ER_PATH_H
#define ER_PATH_H

#include "er_device.h"

#include <algorithm>

namespace erd {


struct Cnt {
    uint32_t primary, shadow, shading, pfetch, sfetch, scatter, shadow_rays, iters, skipped, undone;
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

/* dense samples per march iteration (loads in flight together) in the primary and shadow kernels */
#define ER_BATCH_CONST_JOBS 128   /* jobs per batch graph (parameter blocks in constant memory) */

#ifndef ER_BATCH_PRIMARY
#define ER_BATCH_PRIMARY 6
#endif
/* a primary wave finishes its marches together (coop_march) once this few lanes still march */
#ifndef ER_COOP_LANES
#define ER_COOP_LANES 4
#endif
#ifndef ER_BATCH_SHADOW
#define ER_BATCH_SHADOW 8
#endif
/* idle lanes of a shadow-march wave that claim new jobs together */
#ifndef ER_SHADOW_REFILL
#define ER_SHADOW_REFILL 16
#endif

/* ---- free-flight sampling --------------------------------------------------------------- */

/* Per-block copies of the macrocell occupancy bits and (when it fits) the opacity LUT. */
struct Tables {
    /* bit m set <=> macrocell m holds a sample with sigma > 0: in LDS, or (bits_glb set) in
     * global memory */
    const uint32_t* bits;
    const __attribute__((address_space(1))) uint32_t* bits_glb;
    /* DensityScale * Opacity(i): in LDS when it fits (lut_lds), else in global memory */
    const float* lut_lds;
    const __attribute__((address_space(1))) float* lut_glb;
    __device__ __forceinline__ float lut(int i) const { return lut_glb ? lut_glb[i] : lut_lds[i]; }
};

__device__ __forceinline__ bool mc_occupied(const Tables& T, int m)
{
    const uint32_t w = T.bits_glb ? T.bits_glb[m >> 5] : T.bits[m >> 5];
    return (w >> (m & 31)) & 1u;
}

/* the empty-space radius of group g (4 bits, eight groups to a word from word `base`): 0 for
 * an occupied group, else R >= 1 with every group within Chebyshev distance R - 1 of g empty */
__device__ __forceinline__ int group_radius(const Tables& T, int base, int g)
{
    const int wi = base + (g >> 3);
    const uint32_t w = T.bits_glb ? T.bits_glb[wi] : T.bits[wi];
    return (int)((w >> ((g & 7) << 2)) & 15u);
}

__device__ __forceinline__ int mc_axis(float L, int shift, int dim)
{
    return min(max(((int)floorf(L)) >> shift, 0), dim - 1);
}

/* macrocell index of a lattice point (x fastest); the grid has fewer than 2^24 macrocells a
 * side, so the products use the full-rate 24-bit multiply */
__device__ __forceinline__ int mc_index(const DevParams& p, const Lattice& l, int sh)
{
    return (int)(__umul24(__umul24((unsigned)mc_axis(l.Lz, sh, p.mc_dim[2]), (unsigned)p.mc_dim[1]) +
                          (unsigned)mc_axis(l.Ly, sh, p.mc_dim[1]), (unsigned)p.mc_dim[0]) +
                 (unsigned)mc_axis(l.Lx, sh, p.mc_dim[0]));
}

/* whether the clipping objects can cut the sample at lattice point l: false when the group of
 * macrocells holding it has its clip-free bit (every occupied macrocell of the group wholly
 * unclipped), so clipped() need not run for it */
__device__ __forceinline__ bool clip_possible(const DevParams& p, const Tables& T, const Lattice& l)
{
    const int sh = p.mc_shift + ER_GROUP_SHIFT;
    const int g = (int)(__umul24(__umul24((unsigned)mc_axis(l.Lz, sh, p.mc2_dim[2]), (unsigned)p.mc2_dim[1]) +
                                 (unsigned)mc_axis(l.Ly, sh, p.mc2_dim[1]), (unsigned)p.mc2_dim[0]) +
                        (unsigned)mc_axis(l.Lx, sh, p.mc2_dim[0]));
    return !mc_occupied(T, (p.mc3_base << 5) + g);
}

/* sigma of one sample (Woodcock path) */
template<int FULL, int COUNT>
__device__ __forceinline__ float sample_sigma(const DevParams& p, const Tables& T, f3 Ps, bool primary, Cnt& cnt)
{
    const Lattice l = to_lattice(p, Ps);
    const int m = mc_index(p, l, p.mc_shift);
    if (!mc_occupied(T, m)) return 0.0f;
    if (FULL && p.clip_count > 0 && clip_possible(p, T, l) && clipped(p, Ps)) return 0.0f;
    if (COUNT) { if (primary) cnt.pfetch++; else cnt.sfetch++; }
    return T.lut(min((int)trilinear_cell(p, l), p.lut_max));
}

/* floor(x / y) for integers x < 2^24, 1 <= y < 2^24: the hardware reciprocal (1 ulp) makes the
 * float quotient at most 2 too small or 1 too large (exact for y a power of 2), which the
 * remainder then corrects without a loop */
__device__ __forceinline__ uint32_t udiv24(uint32_t x, uint32_t y)
{
    uint32_t q = (uint32_t)((float)x * __builtin_amdgcn_rcpf((float)y));
    int r = (int)x - (int)(q * y);
    if (r < 0) { q -= 1u; r += (int)y; }
    if (r >= (int)y) { q += 1u; r -= (int)y; }
    if (r >= (int)y) { q += 1u; }
    return q;
}

/* Advance t by repeated float additions of b while t < t_lim, without performing them one by
 * one: returns the number of additions n and sets t to the value after the last one and t_prev
 * to the value before it - bit-identical to `while (t < t_lim) { t_prev = t; t += b; ++n; }`.
 * Within one binade [2^e, 2^(e+1)) of t every value is an integer multiple A of the spacing
 * u = 2^(e-23), so t + b rounds to (A + D) u with the constant D = round(b / u) unless b / u
 * has fractional part exactly 1/2 (then the rounding depends on the parity of A and the
 * addition is done plainly). The additions are counted with integer arithmetic: reaching
 * t_lim takes J = ceil((ceil(t_lim / u) - A) / D) of them (t_lim / u is exact, a power-of-2
 * scaling, and A + j D is an integer); if A + J D stays below 2^24 the binade holds them all,
 * else the binade's last addition, which crosses into the next binade, is done as a float
 * addition and the count goes on from there. Requires b > 0 and t_lim + b > t_lim (the march
 * checks it), so D >= 1 wherever t < t_lim. */
__device__ __forceinline__ uint32_t skip_additions(float& t, float b, float t_lim, float& t_prev)
{
    uint32_t n = 0;
    while (t < t_lim) {
        const uint32_t tb = __float_as_uint(t);
        const int e = (int)((tb >> 23) & 0xffu);
        if (e <= 24 || e >= 254 || (tb >> 31)) { t_prev = t; t += b; ++n; continue; }
        const float u = __uint_as_float((uint32_t)(e - 23) << 23);
        const float inv_u = __uint_as_float((uint32_t)(254 - (e - 23)) << 23);   /* 1 / u, exact */
        const float q = b * inv_u;                                                 /* exact: power-of-2 scaling */
        const float fq = floorf(q);
        const float fr = q - fq;
        if (fr == 0.5f || q >= 8388608.0f) { t_prev = t; t += b; ++n; continue; }
        const uint32_t D = (uint32_t)fq + (fr > 0.5f ? 1u : 0u);
        const uint32_t A = (tb & 0x007fffffu) | 0x00800000u;                      /* t / u */
        const float T = t_lim * inv_u;                                             /* exact */
        if (T < 16777216.0f) {
            /* T > A (t < t_lim), so J >= 1; A + j D >= ceil(T) <=> (A + j D) u >= t_lim */
            const uint32_t J = udiv24((uint32_t)ceilf(T) - A + D - 1u, D);
            if (A + J * D < 16777216u) {
                n += J;
                t_prev = (float)(A + (J - 1u) * D) * u;
                t = (float)(A + J * D) * u;
                break;
            }
        }
        /* every addition inside the binade is taken; the next one crosses into the next binade */
        const uint32_t kbin = udiv24(0x00ffffffu - A, D);
        n += kbin + 1u;
        t_prev = (float)(A + kbin * D) * u;
        t = t_prev + b;
    }
    return n;
}

/* SampleVolume (raymarching.h:30-71) for primary rays, ScatterEventInVolume
 * (raymarching.h:73-113) for shadow rays: jittered fixed-step tracking with threshold
 * S = -log(u) / DensityScale; T is reported one step beyond P (Q13); the primary march ends
 * at t >= tmax, the shadow march at t > tmax (Q14). */
/* Observer of the samples a march reads voxels for (the path-replay adjoint): begin() gets
 * the step size, sample() each fetched sample's integer intensity and sigma, in march order.
 * Samples in empty macrocells or clipped away have sigma 0 and are not reported. */
struct NoVisit {
    __device__ __forceinline__ void begin(float) {}
    __device__ __forceinline__ void sample(int, float) {}
    /* group marches (group_step_v): the event sample, reported to every lane of the group */
    __device__ __forceinline__ void event(int, float) {}
};
template<class V> struct IsNoVisit { static constexpr bool value = false; };
template<> struct IsNoVisit<NoVisit> { static constexpr bool value = true; };

/* march_v with no observer: the forward march */
template<int FULL, int COUNT, int K = 1, class Visit = NoVisit>
__device__ __forceinline__ bool march_v(const DevParams& p, const Tables& T, const Ray& r, bool primary, Rng& rng, float& Tout, f3& Pout,
                        Cnt& cnt, Visit& vis, float S_fixed);

template<int FULL, int COUNT, int K = 1>
__device__ __forceinline__ bool march(const DevParams& p, const Tables& T, const Ray& r, bool primary, Rng& rng, float& Tout,
                                      f3& Pout, Cnt& cnt)
{
    NoVisit nv;
    return march_v<FULL, COUNT, K, NoVisit>(p, T, r, primary, rng, Tout, Pout, cnt, nv, -1.0f);
}

/* The state of a ray-march between iterations of its loop (raymarching.h:55-68). */
struct MarchState {
    f3 O, D;          /* the ray */
    float MinT, MaxT; /* next sample position, end of the volume segment */
    float Sum, S;     /* optical depth so far / DensityScale, threshold */
};

/* Set-up of the fixed-step march (raymarching.h:38-53) after the box test: the threshold
 * S = -log(u) / DensityScale, then the jitter of the first sample (two draws). S_fixed >= 0
 * replaces the threshold after it is drawn. False when the march cannot advance (zero-size
 * volume, Q2, or a step below the spacing of floats at MaxT): the reference would not
 * terminate; it is reported as no event. */
__device__ __forceinline__ bool march_begin(const DevParams& p, const Ray& r, float nearT, float farT, float StepSize, Rng& rng,
                                            float S_fixed, MarchState& m)
{
    m.O = r.O;
    m.D = r.D;
    m.MinT = fmaxf(nearT, r.MinT);
    m.MaxT = fminf(farT, r.MaxT);
    m.S = gmem(p.neglog)[rng.next_bits()] / p.density_scale;
    if (S_fixed >= 0.0f) m.S = S_fixed;
    m.MinT += rng.u() * StepSize;
    m.Sum = 0.0f;
    return StepSize > 0.0f && m.MaxT + StepSize > m.MaxT;
}

/* One iteration of the march loop: up to K consecutive samples in an occupied macrocell, or
 * the skip over the rest of an empty one. Returns 1 on an event (Pout = the sample, m.MinT one
 * step beyond it, Q13), -1 when the ray leaves the segment first (the primary march ends at
 * t >= tmax, the shadow march at t > tmax, Q14), 0 to go on. */
template<int FULL, int COUNT, int K, class Visit>
__device__ __forceinline__ int march_step(const DevParams& p, const Tables& T, MarchState& m, bool primary, float StepSize,
                                          Cnt& cnt, Visit& vis, f3& Pout)
{
    const int sh = p.mc_shift;
    const f3 Ps = add(m.O, scl(m.MinT, m.D));
    if (primary ? (m.MinT >= m.MaxT) : (m.MinT > m.MaxT)) return -1;
    if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; cnt.iters++; }
    const Lattice l = to_lattice(p, Ps);
    const int mxi = mc_axis(l.Lx, sh, p.mc_dim[0]);
    const int myi = mc_axis(l.Ly, sh, p.mc_dim[1]);
    const int mzi = mc_axis(l.Lz, sh, p.mc_dim[2]);
    const int mc = (int)(__umul24(__umul24((unsigned)mzi, (unsigned)p.mc_dim[1]) + (unsigned)myi, (unsigned)p.mc_dim[0]) +
                         (unsigned)mxi);
    if (K == 1 && mc_occupied(T, mc)) {
        float SigmaT = 0.0f;
        if (!(FULL && p.clip_count > 0 && clip_possible(p, T, l) && clipped(p, Ps))) {
            if (COUNT) { if (primary) cnt.pfetch++; else cnt.sfetch++; }
            const int iv = (int)trilinear_cell(p, l);
            SigmaT = T.lut(min(iv, p.lut_max));
            vis.sample(iv, SigmaT);
        }
        m.Sum += SigmaT * StepSize;
        m.MinT += StepSize;
        if (!(m.Sum < m.S)) { Pout = Ps; return 1; }
        return 0;
    }
    if (K > 1 && mc_occupied(T, mc)) {
        /* K consecutive samples at once: their positions follow from the same sequence of
         * MinT additions and do not depend on the values, so all K cell loads are issued
         * before the first value is used. The values are then consumed in order, with the
         * loop's own end and threshold tests between them; loads past the point where the
         * loop stops are discarded (their addresses are clamped into the cell array). A
         * sample in an empty macrocell has sigma exactly 0 and is consumed without a fetch. */
        CellFetch f[K];
        bool live[K];
        float tk[K];
        int kn = K;
        tk[0] = m.MinT;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            f3 Pk = Ps;
            Lattice lk = l;
            bool occ = true;
            if (k > 0) {
                tk[k] = tk[k - 1] + StepSize;
                if (kn == K && (primary ? (tk[k] >= m.MaxT) : (tk[k] > m.MaxT))) kn = k;
                Pk = add(m.O, scl(tk[k], m.D));
                lk = to_lattice(p, Pk);
                occ = mc_occupied(T, mc_index(p, lk, sh));
            }
            /* no branch on the occupancy bit (the clip test branches on the uniform clip count
             * only): its LDS read completes while the following positions and loads issue */
            bool cut = false;
            if (FULL && p.clip_count > 0) cut = clip_possible(p, T, lk) && clipped(p, Pk);
            live[k] = occ & !cut;
            f[k] = cell_fetch(p, lk);
        }
        /* every sample's intensity and sigma first, so that the K table reads are in flight
         * together; then the samples are consumed in order */
        int iv[K];
        float sig[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            iv[k] = (int)cell_value(f[k]);
            sig[k] = live[k] ? T.lut(min(iv[k], p.lut_max)) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k > 0) {
                if (k >= kn) return -1;
                if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; cnt.iters++; }
            }
            const float SigmaT = sig[k];
            if (live[k]) {
                if (COUNT) { if (primary) cnt.pfetch++; else cnt.sfetch++; }
                vis.sample(iv[k], SigmaT);
            }
            m.Sum += SigmaT * StepSize;
            m.MinT += StepSize;
            if (!(m.Sum < m.S)) { Pout = add(m.O, scl(tk[k], m.D)); return 1; }
        }
        return 0;
    }
    /* Empty macrocell: this sample adds exactly 0. Skip the following samples that stay in
     * the same macrocell without computing them, keeping the exact sequence of MinT
     * additions. Each operation of P = O + t D, (P - Min) * InvSize * Res is monotone in t,
     * so if the first and the last skipped sample lie in this macrocell's lattice box,
     * every sample between them does too; the last one is checked with the same
     * expression as above, and the skip is undone if it left the box. The exit estimate
     * uses the lattice-space direction and hardware reciprocals (approximate: the skip is
     * verified). */
    const float Ldx = (m.D.x * p.invsize[0]) * p.resf[0];
    const float Ldy = (m.D.y * p.invsize[1]) * p.resf[1];
    const float Ldz = (m.D.z * p.invsize[2]) * p.resf[2];
    const float iLdx = __builtin_amdgcn_rcpf(Ldx), iLdy = __builtin_amdgcn_rcpf(Ldy), iLdz = __builtin_amdgcn_rcpf(Ldz);
    const float margin = 1.0e-3f * __builtin_amdgcn_rcpf(fmaxf(fabsf(Ldx), fmaxf(fabsf(Ldy), fabsf(Ldz))));
    m.MinT += StepSize;
    /* the box to skip: the macrocell, or, when its group of macrocells is empty, the cube of
     * groups [g - (R - 1), g + (R - 1)] on each axis that the group's radius R marks empty (the
     * same argument holds for any lattice-aligned box; a box reaching the grid's edge extends
     * beyond it, where the clamped lookups land in its edge groups) */
    int bsh = sh, nx = p.mc_dim[0], ny = p.mc_dim[1], nz = p.mc_dim[2];
    int x0 = mxi, x1 = mxi, y0 = myi, y1 = myi, z0 = mzi, z1 = mzi;
    {
        const int gx = mxi >> ER_GROUP_SHIFT, gy = myi >> ER_GROUP_SHIFT, gz = mzi >> ER_GROUP_SHIFT;
        const int R = group_radius(T, p.mc2_base, (gz * p.mc2_dim[1] + gy) * p.mc2_dim[0] + gx);
        if (R > 0) {
            bsh = sh + ER_GROUP_SHIFT; nx = p.mc2_dim[0]; ny = p.mc2_dim[1]; nz = p.mc2_dim[2];
            x0 = gx - (R - 1); x1 = gx + (R - 1);
            y0 = gy - (R - 1); y1 = gy + (R - 1);
            z0 = gz - (R - 1); z1 = gz + (R - 1);
        }
    }
    const float big = 3.0e38f;
    const float lo_x = x0 <= 0 ? -big : (float)(x0 << bsh), hi_x = x1 >= nx - 1 ? big : (float)((x1 + 1) << bsh);
    const float lo_y = y0 <= 0 ? -big : (float)(y0 << bsh), hi_y = y1 >= ny - 1 ? big : (float)((y1 + 1) << bsh);
    const float lo_z = z0 <= 0 ? -big : (float)(z0 << bsh), hi_z = z1 >= nz - 1 ? big : (float)((z1 + 1) << bsh);
    const float tx = Ldx > 0.0f ? (hi_x - l.Lx) * iLdx : (Ldx < 0.0f ? (lo_x - l.Lx) * iLdx : big);
    const float ty = Ldy > 0.0f ? (hi_y - l.Ly) * iLdy : (Ldy < 0.0f ? (lo_y - l.Ly) * iLdy : big);
    const float tz = Ldz > 0.0f ? (hi_z - l.Lz) * iLdz : (Ldz < 0.0f ? (lo_z - l.Lz) * iLdz : big);
    const float dt_exit = fminf(tx, fminf(ty, tz)) - margin;
    const float t_lim = fminf((m.MinT - StepSize) + dt_exit, m.MaxT);
    const float t_first = m.MinT;
    float t_prev = m.MinT;
    uint32_t n = skip_additions(m.MinT, StepSize, t_lim, t_prev);
    if (n > 0) {
        const Lattice lp = to_lattice(p, add(m.O, scl(t_prev, m.D)));
        const int ax = mc_axis(lp.Lx, bsh, nx), ay = mc_axis(lp.Ly, bsh, ny), az = mc_axis(lp.Lz, bsh, nz);
        if (ax < x0 || ax > x1 || ay < y0 || ay > y1 || az < z0 || az > z1) {
            m.MinT = t_first;
            n = 0;
            if (COUNT) cnt.undone++;
        }
    }
    if (COUNT) { if (primary) cnt.primary += n; else cnt.shadow += n; cnt.skipped += n; }
    return 0;
}

/* Woodcock delta tracking (full estimator; design from the commented sampler of
 * raymarching.h:115-273, its macrocells at :155 and majorant texture at :205) with a piecewise
 * constant majorant: mu = DS * wc_sigma of the 2^ER_WC_SHIFT-cell macrocell the ray is in. Each
 * tentative collision draws one optical depth tau = -log(u) (the project's logf_, computed:
 * a table lookup per collision would be a dependent random load) and spends it across macrocells:
 * in a segment [t, t_exit) the collision is at t + tau / mu if that is before t_exit, else tau
 * loses mu (t_exit - t) and the walk moves to the next macrocell. A collision is real with
 * probability sigma(P) / mu (one more draw), else tracking continues from it. The segment ends
 * at the nearest macrocell face in lattice coordinates (a macrocell at the grid's edge extends
 * to the end of the ray, as the clamped lookups do; the distance to a face is its lattice
 * distance times the reciprocal of the ray's lattice rate on that axis) and advances t by at
 * least a millionth of it, so every segment makes progress. The oracle runs the same float
 * operations. */
#define ER_WC_SHIFT 3
#define ER_WC_MAX_SEGMENTS 65536
template<int FULL, int COUNT>
__device__ __forceinline__ bool woodcock_track(const DevParams& p, const Tables& T, const Ray& r, float nearT, float farT,
                                               bool primary, Rng& rng, float& Tout, f3& Pout, Cnt& cnt)
{
    float t = fmaxf(nearT, r.MinT);
    const float MaxT = fminf(farT, r.MaxT);
    /* lattice units per unit of t along each axis, and their reciprocals */
    const float gx = (r.D.x * p.invsize[0]) * p.resf[0];
    const float gy = (r.D.y * p.invsize[1]) * p.resf[1];
    const float gz = (r.D.z * p.invsize[2]) * p.resf[2];
    const float igx = gx != 0.0f ? 1.0f / gx : 0.0f, igy = gy != 0.0f ? 1.0f / gy : 0.0f;
    const float igz = gz != 0.0f ? 1.0f / gz : 0.0f;
    const int dx = p.wc_dim[0], dy = p.wc_dim[1], dz = p.wc_dim[2];
    const float big = 3.402823466e+38f;
    int segments = 0;
    for (;;) {
        float tau = -erm::logf_(rng.u());   /* u = 0 gives +inf: no collision on this ray */
        float mu = 0.0f;
        for (;;) {
            if (++segments > ER_WC_MAX_SEGMENTS || !(t < MaxT)) return false;
            if (COUNT) cnt.iters++;
            const Lattice l = to_lattice(p, add(r.O, scl(t, r.D)));
            const int mx = mc_axis(l.Lx, ER_WC_SHIFT, dx), my = mc_axis(l.Ly, ER_WC_SHIFT, dy), mz = mc_axis(l.Lz, ER_WC_SHIFT, dz);
            mu = p.density_scale * gmem(p.wc_sigma)[(mz * dy + my) * dx + mx];
            const float ex = gx > 0.0f ? (mx == dx - 1 ? big : ((float)((mx + 1) << ER_WC_SHIFT) - l.Lx) * igx)
                                       : (gx < 0.0f ? (mx == 0 ? big : ((float)(mx << ER_WC_SHIFT) - l.Lx) * igx) : big);
            const float ey = gy > 0.0f ? (my == dy - 1 ? big : ((float)((my + 1) << ER_WC_SHIFT) - l.Ly) * igy)
                                       : (gy < 0.0f ? (my == 0 ? big : ((float)(my << ER_WC_SHIFT) - l.Ly) * igy) : big);
            const float ez = gz > 0.0f ? (mz == dz - 1 ? big : ((float)((mz + 1) << ER_WC_SHIFT) - l.Lz) * igz)
                                       : (gz < 0.0f ? (mz == 0 ? big : ((float)(mz << ER_WC_SHIFT) - l.Lz) * igz) : big);
            const float t_exit = fmaxf(fminf(t + fminf(ex, fminf(ey, ez)), MaxT), t + fmaxf(t * 1.0e-6f, 1.0e-7f));
            if (mu > 0.0f) {
                const float tc = t + tau / mu;
                if (tc < t_exit) { t = tc; break; }
                tau = fmaxf(tau - mu * (t_exit - t), 0.0f);
            }
            t = t_exit;
        }
        if (primary ? (t >= MaxT) : (t > MaxT)) return false;
        const f3 Ps = add(r.O, scl(t, r.D));
        if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; }
        const float sigma = sample_sigma<FULL, COUNT>(p, T, Ps, primary, cnt) * p.density_scale;
        if (rng.u() * mu < sigma) { Tout = t; Pout = Ps; return true; }
    }
}

/* woodcock_track one segment at a time, for loops that step many walks together (the shadow
 * marches with lane refill): the same operations and draws in the same order. */
struct WcWalk {
    f3 O, D;
    float t, MaxT, tau;
    float gx, gy, gz, igx, igy, igz;
    int segments;
    bool need_tau;
    Rng rng;
};

__device__ __forceinline__ void wc_begin(const DevParams& p, const Ray& r, float nearT, float farT, const Rng& rng, WcWalk& w)
{
    w.O = r.O;
    w.D = r.D;
    w.t = fmaxf(nearT, r.MinT);
    w.MaxT = fminf(farT, r.MaxT);
    w.gx = (r.D.x * p.invsize[0]) * p.resf[0];
    w.gy = (r.D.y * p.invsize[1]) * p.resf[1];
    w.gz = (r.D.z * p.invsize[2]) * p.resf[2];
    w.igx = w.gx != 0.0f ? 1.0f / w.gx : 0.0f;
    w.igy = w.gy != 0.0f ? 1.0f / w.gy : 0.0f;
    w.igz = w.gz != 0.0f ? 1.0f / w.gz : 0.0f;
    w.segments = 0;
    w.need_tau = true;
    w.rng = rng;
}

/* one step: draws the walk's next optical depth if it needs one, then walks one majorant
 * segment and, if the tentative collision falls in it, tests it. 1: a real collision (w.t its
 * distance), -1: the walk left the segment [t, MaxT] first, 0: go on. */
template<int FULL, int COUNT>
__device__ __forceinline__ int wc_step(const DevParams& p, const Tables& T, WcWalk& w, bool primary, Cnt& cnt)
{
    if (w.need_tau) {
        w.tau = -erm::logf_(w.rng.u());
        w.need_tau = false;
    }
    if (++w.segments > ER_WC_MAX_SEGMENTS || !(w.t < w.MaxT)) return -1;
    if (COUNT) cnt.iters++;
    const int dx = p.wc_dim[0], dy = p.wc_dim[1], dz = p.wc_dim[2];
    const float big = 3.402823466e+38f;
    const Lattice l = to_lattice(p, add(w.O, scl(w.t, w.D)));
    const int mx = mc_axis(l.Lx, ER_WC_SHIFT, dx), my = mc_axis(l.Ly, ER_WC_SHIFT, dy), mz = mc_axis(l.Lz, ER_WC_SHIFT, dz);
    const float mu = p.density_scale * gmem(p.wc_sigma)[(mz * dy + my) * dx + mx];
    const float ex = w.gx > 0.0f ? (mx == dx - 1 ? big : ((float)((mx + 1) << ER_WC_SHIFT) - l.Lx) * w.igx)
                                 : (w.gx < 0.0f ? (mx == 0 ? big : ((float)(mx << ER_WC_SHIFT) - l.Lx) * w.igx) : big);
    const float ey = w.gy > 0.0f ? (my == dy - 1 ? big : ((float)((my + 1) << ER_WC_SHIFT) - l.Ly) * w.igy)
                                 : (w.gy < 0.0f ? (my == 0 ? big : ((float)(my << ER_WC_SHIFT) - l.Ly) * w.igy) : big);
    const float ez = w.gz > 0.0f ? (mz == dz - 1 ? big : ((float)((mz + 1) << ER_WC_SHIFT) - l.Lz) * w.igz)
                                 : (w.gz < 0.0f ? (mz == 0 ? big : ((float)(mz << ER_WC_SHIFT) - l.Lz) * w.igz) : big);
    const float t_exit = fmaxf(fminf(w.t + fminf(ex, fminf(ey, ez)), w.MaxT), w.t + fmaxf(w.t * 1.0e-6f, 1.0e-7f));
    if (mu > 0.0f) {
        const float tc = w.t + w.tau / mu;
        if (tc < t_exit) {
            w.t = tc;
            if (primary ? (w.t >= w.MaxT) : (w.t > w.MaxT)) return -1;
            const f3 Ps = add(w.O, scl(w.t, w.D));
            if (COUNT) { if (primary) cnt.primary++; else cnt.shadow++; }
            const float sigma = sample_sigma<FULL, COUNT>(p, T, Ps, primary, cnt) * p.density_scale;
            if (w.rng.u() * mu < sigma) return 1;
            w.need_tau = true;
            return 0;
        }
        w.tau = fmaxf(w.tau - mu * (t_exit - w.t), 0.0f);
    }
    w.t = t_exit;
    return 0;
}

/* S_fixed >= 0: the threshold is still drawn (same draws) but S_fixed replaces it, so the march
 * reads every sample until that optical depth / DensityScale or the end of the ray (the
 * expected-transmittance form of a shadow test). */
template<int FULL, int COUNT, int K, class Visit>
__device__ __forceinline__ bool march_v(const DevParams& p, const Tables& T, const Ray& r, bool primary, Rng& rng, float& Tout, f3& Pout,
                        Cnt& cnt, Visit& vis, float S_fixed)
{
    float nearT, farT;
    if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return false;
    if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) return woodcock_track<FULL, COUNT>(p, T, r, nearT, farT, primary, rng, Tout, Pout, cnt);
    const float StepSize = (primary ? p.step_primary : p.step_shadow) * p.minstep;
    vis.begin(StepSize);
    MarchState m;
    if (!march_begin(p, r, nearT, farT, StepSize, rng, S_fixed, m)) return false;
    if (!(m.Sum < m.S)) { Tout = m.MinT; Pout = mk(0.0f, 0.0f, 0.0f); return true; }
    for (;;) {
        const int st = march_step<FULL, COUNT, K, Visit>(p, T, m, primary, StepSize, cnt, vis, Pout);
        if (st > 0) { Tout = m.MinT; return true; }
        if (st < 0) return false;
    }
}

/* The rest of one primary march (SampleVolume's loop, raymarching.h:55-68) run by the whole
 * wave on one ray: m is the same in every lane. A march whose wave has only a few lanes still
 * marching would otherwise spend one memory round trip per ER_BATCH_PRIMARY samples; here each
 * round trip reads 64 consecutive samples, lane k taking sample k. Lane k finds t_k with the
 * same k float additions of StepSize as the sequential loop, and the optical depth is then
 * accumulated in sample order (one addition per sample, broadcast from its lane), with the
 * loop's own end test before each sample and threshold test after it, so every decision,
 * position and sum is the sequential march's. Empty macrocells at the start of a block are
 * skipped by march_step's exact skip. Returns 1 on an event (Pout = the sample, m.MinT one step
 * beyond it), -1 when the ray leaves the volume first; the same in every lane. Counts go to
 * lane 0's cnt; steps counts the iterations the per-lane loop would have taken (a skip, or
 * ER_BATCH_PRIMARY samples). */
template<int FULL, int COUNT>
__device__ int coop_march(const DevParams& p, const Tables& T, MarchState& m, float StepSize, int lane, Cnt& cnt, f3& Pout,
                          int& steps)
{
    NoVisit nv;
    for (;;) {
        if (m.MinT >= m.MaxT) return -1;
        const f3 P0 = add(m.O, scl(m.MinT, m.D));
        const Lattice l0 = to_lattice(p, P0);
        const int sh = p.mc_shift;
        const int mc0 = mc_index(p, l0, sh);
        if (!mc_occupied(T, mc0)) {
            /* the sequential march would skip: the same skip in every lane */
            Cnt c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            f3 P;
            (void)march_step<FULL, COUNT, 1, NoVisit>(p, T, m, true, StepSize, c, nv, P);
            if (COUNT && lane == 0) { cnt.primary += c.primary; cnt.iters += c.iters; cnt.skipped += c.skipped; cnt.undone += c.undone; }
            ++steps;
            continue;
        }
        float tk = m.MinT;
        for (int s = 0; s < lane; ++s) tk += StepSize;
        const bool inseg = tk < m.MaxT;
        const f3 Pk = add(m.O, scl(tk, m.D));
        const Lattice lk = to_lattice(p, Pk);
        const bool occ = mc_occupied(T, mc_index(p, lk, sh));
        const bool live = inseg && occ && !(FULL && p.clip_count > 0 && clip_possible(p, T, lk) && clipped(p, Pk));
        float sigma = 0.0f;
        if (live) sigma = T.lut(min((int)trilinear_cell(p, lk), p.lut_max));
        const float c = sigma * StepSize;
        /* samples 0 .. kexit-1 lie before the end of the segment (t_k grows with k) */
        const unsigned long long segm = __ballot(inseg);
        const int kexit = segm == ~0ull ? 64 : __ffsll((long long)~segm) - 1;
        float Sum = m.Sum;
        int kend = 64, term = 0;
        for (int k = 0; k < 64; ++k) {
            if (k >= kexit) { term = -1; kend = k; break; }
            Sum = Sum + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), k));
            if (!(Sum < m.S)) { term = 1; kend = k; break; }
        }
        const unsigned long long lm = __ballot(live);
        const int ns = term == 1 ? kend + 1 : kend;
        steps += (ns + ER_BATCH_PRIMARY - 1) / ER_BATCH_PRIMARY;   /* the sequential loop's iterations */
        if (COUNT && lane == 0) {
            const unsigned long long within = ns >= 64 ? ~0ull : ((1ull << ns) - 1ull);
            cnt.primary += (uint32_t)ns;
            cnt.iters += (uint32_t)ns;
            cnt.pfetch += (uint32_t)__popcll(lm & within);
        }
        if (term == 1) {
            Pout = mk(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(Pk.x), kend)),
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Pk.y), kend)),
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Pk.z), kend)));
            m.MinT = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tk), kend)) + StepSize;
            m.Sum = Sum;
            return 1;
        }
        if (term == -1) return -1;
        m.Sum = Sum;
        m.MinT = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tk), 63)) + StepSize;
    }
}

/* samples each lane of a G-lane group loads per round trip in group_march */
#ifndef ER_GROUP4_BATCH
#define ER_GROUP4_BATCH 6
#endif
template<int G> struct GroupBatch { static constexpr int value = G >= 16 ? 2 : ER_GROUP4_BATCH; };

/* lane q's v within each aligned group of G lanes (q a constant): a quad permutation (DPP, no
 * LDS round trip) for G = 4, ds_bpermute otherwise */
template<int G, int Q>
__device__ __forceinline__ float group_bcast(float v)
{
    if constexpr (G == 4)
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Q | (Q << 2) | (Q << 4) | (Q << 6), 0xf, 0xf,
                                                          false));
    else
        return __shfl(v, Q, G);
}

template<int G, int KB, int J = 0, int Q = 0>
__device__ __forceinline__ void group_gather(const float (&c)[KB], float (&v)[KB * G])
{
    if constexpr (J < KB) {
        v[J * G + Q] = group_bcast<G, Q>(c[J]);
        if constexpr (Q + 1 < G) group_gather<G, KB, J, Q + 1>(c, v);
        else group_gather<G, KB, J + 1, 0>(c, v);
    }
}

/* The optical depth of samples S, S + 1, ... of a group_march block added in order, sample s
 * loaded by lane s % G as c[s / G]: before each sample the end test (s >= nin: term -1), after
 * it the threshold test (term 1); send is the sample that ended the block. The addends are
 * broadcast one at a time (for G = 4 a DPP quad permutation, as cheap as a move). */
template<int G, int KB, int S = 0>
__device__ __forceinline__ void group_accumulate(const float (&c)[KB], float& Sum, float thr, int nin, int& send, int& term)
{
    if constexpr (S < KB * G) {
        if (S >= nin) { term = -1; send = S; return; }
        Sum = Sum + group_bcast<G, S % G>(c[S / G]);
        if (!(Sum < thr)) { term = 1; send = S; return; }
        group_accumulate<G, KB, S + 1>(c, Sum, thr, nin, send, term);
    }
}

/* One round of a march (SampleVolume's loop, raymarching.h:55-68, primary; the same loop with
 * the shadow end test, raymarching.h:95-113, otherwise) run by an aligned group of G lanes
 * (G = 4 or 16; 64 / G rays per wave): m is the same in every lane of the group. A round either
 * skips the empty macrocell at MinT (march_step's exact skip) or reads B = GroupBatch<G> x G
 * consecutive samples, lane k of the group taking samples k, k + G, ..., k + (B/G - 1) G, so a
 * ray advances B samples per memory round trip instead of ER_BATCH_PRIMARY / ER_BATCH_SHADOW.
 * As in coop_march every position is the sequential loop's (the same float additions of
 * StepSize), the optical depth is accumulated in sample order (one addition per sample, its
 * addend fetched from the lane that loaded it) with the loop's end test before each sample and
 * threshold test after it; so the event, its position and the sums are the sequential
 * march's. Control flow is uniform within a group (it depends only on the group's ray), so the
 * cross-lane reads stay inside active lanes. Returns 0 to go on, 1 on an event (Pout = the
 * sample, m.MinT one step beyond it), -1 when the ray leaves the volume first; counts go to
 * the group's first lane. iters: the iterations the one-lane-per-ray loop would have taken for
 * this round (1 for a skip, else the samples consumed over ER_BATCH_PRIMARY, rounded up).
 * group_step_v also reports the round's fetched samples to an observer: each lane passes its
 * own samples that the round consumed (vis.sample, in sample order within the lane, so each
 * lane's observer sees a subsequence of the march), and on an event every lane of the group
 * passes the event sample (vis.event). */
template<int FULL, int COUNT, int G, bool PRIMARY, class Visit>
__device__ __forceinline__ int group_step_v(const DevParams& p, const Tables& T, MarchState& m, float StepSize, int lane,
                                            Cnt& cnt, f3& Pout, int& iters, Visit& vis)
{
    constexpr int KB = GroupBatch<G>::value;
    constexpr int NS = KB * G;
    const int k = lane & (G - 1);
    const int gb = lane & ~(G - 1);
    const unsigned long long gmask = (1ull << G) - 1ull;
    if (PRIMARY ? (m.MinT >= m.MaxT) : (m.MinT > m.MaxT)) return -1;
    const int sh = p.mc_shift;
    {
        const Lattice l0 = to_lattice(p, add(m.O, scl(m.MinT, m.D)));
        if (!mc_occupied(T, mc_index(p, l0, sh))) {
            /* the sequential march would skip: the same skip in every lane of the group */
            Cnt c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            NoVisit nv;
            f3 P;
            (void)march_step<FULL, COUNT, 1, NoVisit>(p, T, m, PRIMARY, StepSize, c, nv, P);
            if (COUNT && k == 0) {
                cnt.primary += c.primary; cnt.shadow += c.shadow; cnt.iters += c.iters; cnt.skipped += c.skipped;
                cnt.undone += c.undone;
            }
            iters = 1;
            return 0;
        }
    }
    /* sample j G + k: t after j G + k additions of StepSize to MinT */
    float t = m.MinT;
    for (int s = 0; s < k; ++s) t += StepSize;
    float tk[KB], c[KB];
    bool ins[KB], live[KB];
    CellFetch f[KB];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
        if (j > 0) {
#pragma unroll
            for (int s = 0; s < G; ++s) t += StepSize;
        }
        tk[j] = t;
        ins[j] = PRIMARY ? (t < m.MaxT) : (t <= m.MaxT);
        const f3 Pk = add(m.O, scl(t, m.D));
        const Lattice lk = to_lattice(p, Pk);
        bool cut = false;   /* as in march_step: no branch on the occupancy bit */
        if (FULL && p.clip_count > 0) cut = clip_possible(p, T, lk) && clipped(p, Pk);
        live[j] = ins[j] & mc_occupied(T, mc_index(p, lk, sh)) & !cut;
        f[j] = cell_fetch(p, lk);
    }
    int ivs[KB];
    float sg[KB];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
        ivs[j] = (int)cell_value(f[j]);
        sg[j] = live[j] ? T.lut(min(ivs[j], p.lut_max)) : 0.0f;
        c[j] = sg[j] * StepSize;
    }
    /* samples 0 .. nin-1 lie before the end of the segment (t grows with the sample index) */
    int nin = 0;
#pragma unroll
    for (int j = 0; j < KB; ++j) nin += __popcll((__ballot(ins[j]) >> gb) & gmask);
    float Sum = m.Sum;
    int send = NS, term = 0;
    if constexpr (G == 4) {
        group_accumulate<G, KB>(c, Sum, m.S, nin, send, term);
    } else {
        /* ds_bpermute has LDS latency: every addend is fetched before the chain starts */
        float v[NS];
        group_gather<G, KB>(c, v);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s >= nin) { term = -1; send = s; break; }
            Sum = Sum + v[s];
            if (!(Sum < m.S)) { term = 1; send = s; break; }
        }
    }
    const int ns = term == 1 ? send + 1 : send;
    iters = (ns + ER_BATCH_PRIMARY - 1) / ER_BATCH_PRIMARY;
    if constexpr (!IsNoVisit<Visit>::value) {
#pragma unroll
        for (int j = 0; j < KB; ++j)
            if (live[j] && j * G + k < ns) vis.sample(ivs[j], sg[j]);
        if (term == 1) {
            const int je = send / G;
            int ei = ivs[0];
            float es = sg[0];
#pragma unroll
            for (int j = 1; j < KB; ++j) if (j == je) { ei = ivs[j]; es = sg[j]; }
            vis.event(__shfl(ei, send - je * G, G), __shfl(es, send - je * G, G));
        }
    }
    if (COUNT) {
        uint32_t fetched = 0;
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const int lim = min(max(ns - j * G, 0), G);
            const unsigned long long within = (1ull << lim) - 1ull;
            fetched += (uint32_t)__popcll(((__ballot(live[j]) >> gb) & gmask) & within);
        }
        if (k == 0) {
            if (PRIMARY) { cnt.primary += (uint32_t)ns; cnt.pfetch += fetched; }
            else { cnt.shadow += (uint32_t)ns; cnt.sfetch += fetched; }
            cnt.iters += (uint32_t)ns;
        }
    }
    if (term == 1) {
        const int je = send / G;
        float te = tk[0];
#pragma unroll
        for (int j = 1; j < KB; ++j) if (j == je) te = tk[j];
        te = __shfl(te, send - je * G, G);
        Pout = add(m.O, scl(te, m.D));
        m.MinT = te + StepSize;
        m.Sum = Sum;
        return 1;
    }
    if (term == -1) return -1;
    m.Sum = Sum;
    m.MinT = __shfl(tk[KB - 1], G - 1, G) + StepSize;
    return 0;
}

template<int FULL, int COUNT, int G, bool PRIMARY>
__device__ __forceinline__ int group_step(const DevParams& p, const Tables& T, MarchState& m, float StepSize, int lane,
                                          Cnt& cnt, f3& Pout, int& iters)
{
    NoVisit nv;
    return group_step_v<FULL, COUNT, G, PRIMARY, NoVisit>(p, T, m, StepSize, lane, cnt, Pout, iters, nv);
}

/* A whole primary march by an aligned group of G lanes, round by round (group_step); steps
 * counts the iterations the one-lane-per-ray loop would have taken, so that tile costs compare
 * across lane counts. */
template<int FULL, int COUNT, int G>
__device__ int group_march(const DevParams& p, const Tables& T, MarchState& m, float StepSize, int lane, Cnt& cnt, f3& Pout,
                           int& steps)
{
    for (;;) {
        int iters = 0;
        const int st = group_step<FULL, COUNT, G, true>(p, T, m, StepSize, lane, cnt, Pout, iters);
        steps += iters;
        if (st != 0) return st;
    }
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

/* Buffer2D<ColorRGBAuc>::operator()(Vec2f, true) (buffer2d.h:227-254) and
 * ColorXYZf::FromRGBAuc (color.h:195-211). In the reference the bilinear weights meet the
 * colour type's operators: float * ColorRGBAuc converts the weight to unsigned char (a weight
 * in (0, 1) becomes 0) and scales in place with unsigned-char wrap, and the sums wrap too; the
 * compat estimator keeps that, so only a sample on an exact texel reads the texel. The full
 * estimator interpolates the four texels in float. */
__device__ __forceinline__ f3 bitmap_lookup(const DevParams& p, const DevBitmap& B, float u, float v)
{
    if (!B.pixels || B.w <= 0 || B.h <= 0) return mk(0.0f, 0.0f, 0.0f);
    const float X = u * (float)B.w, Y = v * (float)B.h;
    const int x0i = (int)floorf(X), x1i = (int)ceilf(X), y0i = (int)floorf(Y), y1i = (int)ceilf(Y);
    const float du = X - (float)x0i, dv = Y - (float)y0i;
    const int x0 = min(max(x0i, 0), B.w - 1), x1 = min(max(x1i, 0), B.w - 1);
    const int y0 = min(max(y0i, 0), B.h - 1), y1 = min(max(y1i, 0), B.h - 1);
    const uchar4 V0 = B.pixels[(size_t)y0 * B.w + x0], V1 = B.pixels[(size_t)y0 * B.w + x1];
    const uchar4 V2 = B.pixels[(size_t)y1 * B.w + x0], V3 = B.pixels[(size_t)y1 * B.w + x1];
    float rgb[3];
    const unsigned char c0[3] = {V0.x, V0.y, V0.z}, c1[3] = {V1.x, V1.y, V1.z};
    const unsigned char c2[3] = {V2.x, V2.y, V2.z}, c3[3] = {V3.x, V3.y, V3.z};
    if (p.estimator == ER_ESTIMATOR_FULL) {
        for (int k = 0; k < 3; ++k)
            rgb[k] = (1.0f - dv) * ((1.0f - du) * (float)c0[k] + du * (float)c1[k]) +
                     dv * ((1.0f - du) * (float)c2[k] + du * (float)c3[k]);
    } else {
        const unsigned char a = (unsigned char)(1.0f - dv), b = (unsigned char)(1.0f - du);
        const unsigned char c = (unsigned char)du, d = (unsigned char)dv;
        for (int k = 0; k < 3; ++k) {
            const unsigned char i0 = (unsigned char)((unsigned char)(b * c0[k]) + (unsigned char)(c * c1[k]));
            const unsigned char i1 = (unsigned char)((unsigned char)(b * c2[k]) + (unsigned char)(c * c3[k]));
            rgb[k] = (float)(unsigned char)((unsigned char)(a * i0) + (unsigned char)(d * i1));
        }
    }
    const float ONE_OVER_255 = 1.0f / 255.0f;
    const float r = ONE_OVER_255 * rgb[0], g = ONE_OVER_255 * rgb[1], bl = ONE_OVER_255 * rgb[2];
    return mk(0.412453f * r + 0.357580f * g + 0.180423f * bl, 0.212671f * r + 0.715160f * g + 0.072169f * bl,
              0.019334f * r + 0.119193f * g + 0.950227f * bl);
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
    } else if (T.type == ER_TEXTURE_BITMAP && T.bitmap_id >= 0 && T.bitmap_id < ER_LIST_CAPACITY) {
        L = bitmap_lookup(p, p.bitmaps[T.bitmap_id], u, v);
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

/* IntersectObjects (objects.h:46-66): the nearest object of the tracer's list; returns its
 * list position, or -1 */
__device__ int intersect_objects(const DevParams& p, const Ray& r, float& T, f3& P, f3& N, float& U, float& V)
{
    float best = 3.402823466e+38f;
    int hit = -1;
    for (int i = 0; i < p.object_count; ++i) {
        const DevObject& O = p.objects[i];
        if (!O.valid) continue;
        const bool uv = texture_needs_uv(p, O.diffuse_tex) || texture_needs_uv(p, O.specular_tex) ||
                        texture_needs_uv(p, O.gloss_tex);
        Isect I;
        isect_init(I);
        intersect_shape(O.shape, xform_ray(O.shape.inv_tm, r), uv, I);
        if (!I.Valid) continue;
        const f3 Pw = xform_point(O.shape.tm, I.P);
        const float t = len(sub(Pw, r.O));
        if (t < best) {
            best = t; hit = i; T = t; P = Pw;
            N = xform_vector(O.shape.tm, I.N);
            U = I.U; V = I.V;
        }
    }
    return hit;
}

/* IntersectsObject (objects.h:73-82) */
__device__ bool intersects_any_object(const DevParams& p, const Ray& r)
{
    for (int i = 0; i < p.object_count; ++i) {
        const DevObject& O = p.objects[i];
        if (!O.valid) continue;
        Isect I;
        isect_init(I);
        intersect_shape(O.shape, xform_ray(O.shape.inv_tm, r), false, I);
        if (I.Valid) return true;
    }
    return false;
}

/* the shadow ray of Visible (transport.h:50-53) */
__device__ __forceinline__ Ray shadow_ray(const DevParams& p, f3 P1, f3 P2)
{
    const f3 W = nrm(sub(P2, P1));
    Ray r;
    r.O = add(P1, scl(ER_RAY_EPS, W));
    r.D = W;
    r.MinT = 0.0f;
    r.MaxT = fminf(len(sub(P2, P1)) - (2.0f * ER_RAY_EPS), p.max_shadow);
    return r;
}

/* A shadow ray whose volume march is deferred to k_shadow (48 bytes). */
struct ShadowJob {
    float Ox, Oy, Oz, Dx, Dy, Dz, MaxT;
    uint32_t s0, s1, event, slot, pad;
};

/* Visible (transport.h:46-56) up to the volume march. Returns 0 (occluded by a light shape),
 * 1 (visible without marching) or 2 (march needed: the ray and the RNG state the march starts
 * from are written to job, and rng is advanced past the march's two draws, raymarching.h:89,95;
 * a Woodcock march starts from shadow_stream of that state instead). */
template<int FULL, int COUNT>
__device__ int visible_defer(const DevParams& p, f3 P1, f3 P2, Rng& rng, Cnt& cnt, ShadowJob& job)
{
    if (!p.shadows) return 1;
    const Ray r = shadow_ray(p, P1, P2);
    if (COUNT) cnt.shadow_rays++;
    if (intersects_any_light(p, r)) return 0;
    if (p.object_count > 0 && intersects_any_object(p, r)) return 0;
    float nearT, farT;
    if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return 1;
    job.Ox = r.O.x; job.Oy = r.O.y; job.Oz = r.O.z;
    job.Dx = r.D.x; job.Dy = r.D.y; job.Dz = r.D.z;
    job.MaxT = r.MaxT;
    if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) {
        const Rng rs = shadow_stream(rng);
        job.s0 = rs.s0; job.s1 = rs.s1;
    } else {
        job.s0 = rng.s0; job.s1 = rng.s1;
    }
    (void)rng.next_bits();
    (void)rng.next_bits();
    return 2;
}

/* Visible (transport.h:46-56) -> Intersect (transport.h:30-44) */
template<int FULL, int COUNT>
__device__ bool visible(const DevParams& p, const Tables& T, f3 P1, f3 P2, Rng& rng, Cnt& cnt)
{
    if (!p.shadows) return true;
    const Ray r = shadow_ray(p, P1, P2);
    if (COUNT) cnt.shadow_rays++;
    if (intersects_any_light(p, r)) return false;
    if (p.object_count > 0 && intersects_any_object(p, r)) return false;
    float Tt; f3 P;
    const Tables& Tb = T;
    if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) {
        /* own stream for the march; the path's stream takes the two draws of a ray-march test
         * when the ray meets the volume (Q18) */
        float nearT, farT;
        if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return true;
        Rng rs = shadow_stream(rng);
        (void)rng.next_bits();
        (void)rng.next_bits();
        return !march<FULL, COUNT>(p, Tb, r, false, rs, Tt, P, cnt);
    }
    if (march<FULL, COUNT>(p, Tb, r, false, rng, Tt, P, cnt)) return false;
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
__device__ f3 estimate_direct(const DevParams& p, const Tables& T, const DevLight& L, int light_pos, const Sample& smp, f3 SEP, f3 SEN, f3 Wo,
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
    if (!black(Li) && !black(F) && BsdfPdf > 0.0f && visible<FULL, COUNT>(p, T, SEP, SP, rng, cnt)) {
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
    if (!black(Le2) && visible<FULL, COUNT>(p, T, SEP, P2, rng, cnt)) {
        const f3 d = sub(SEP, P2);
        const float LightPdf = dot(d, d) / (absdot(N2, neg(Wi)) * L.area);
        const float Weight = power_heuristic(BsdfPdf, LightPdf);
        const f3 FL = mul(F, Le2);
        if (sh.brdf) Ld = add(Ld, scl(absdot(Wi, SEN) * Weight / BsdfPdf, FL));
        else Ld = add(Ld, scl(Weight / BsdfPdf, FL));
    }
    return Ld;
}

/* The shader of an event (transport.h:133-150). Volume event (obj < 0): BRDF in compat (Q9);
 * in full mode the shading type picks BRDF or phase function, the stochastic types with one
 * extra draw (design: scatterevent.h:155-212). Object event: BRDF with diffuse, specular and
 * glossiness (its Y) from the object's textures at (U, V). */
template<int FULL, int COUNT>
__device__ __forceinline__ void event_shader(const DevParams& p, float I, int obj, float U, float V, f3 SEP, f3 SEN, f3 Wo,
                                             Rng& rng, Cnt& cnt, Shader& sh, const CD6& cd = CD6{{}, false})
{
    bool brdf = true;
    if (obj >= 0) {
        const DevObject& O = p.objects[obj];
        sh.Kd = eval_texture(p, O.diffuse_tex, U, V);
        sh.Ks = eval_texture(p, O.specular_tex, U, V);
        sh.exponent = 1000000.0f * erm::powf_(eval_texture(p, O.gloss_tex, U, V).y, 7.0f);
    } else {
        if (FULL) {
            const int type = p.shading_type;
            if (type == ER_SHADING_PHASE_ONLY) {
                brdf = false;
            } else if (type != ER_SHADING_BRDF_ONLY) {
                /* cd: the event's central-difference samples when the caller has them (the same
                 * positions as GradientMagnitude's), so no sample is fetched twice */
                const float gm = cd.valid ? gradient_magnitude_cd(p, cd.g) : gradient_magnitude(p, SEP);
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
                    const float PdfBrdf = p.opacity_modulated ? opacity_at(p, I) * base : base;
                    brdf = rng.u() <= PdfBrdf;
                }
            }
        }
        sh.Kd = diffuse_at(p, I);
        sh.Ks = ctf_eval(p.specular, I);
        sh.exponent = 1000000.0f * erm::powf_(plf_eval(p.glossiness, I), 7.0f);   /* utilities.h:28-31 */
    }
    sh.brdf = brdf;
    sh.Nn = nrm(SEN);
    sh.Nu = nrm(crs(SEN, Wo));
    sh.Nv = nrm(crs(SEN, sh.Nu));
}

/* UniformSampleOneLight (transport.h:115-156) with the full-mode shader selection
 * (design at scatterevent.h:155-212) */
template<int FULL, int COUNT>
__device__ f3 uniform_sample_one_light(const DevParams& p, const Tables& T, const Sample& smp, f3 SEP, f3 SEN, f3 Wo, Rng& rng, Cnt& cnt,
                                       int obj = -1, float U = 0.0f, float V = 0.0f)
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
    Shader sh;
    event_shader<FULL, COUNT>(p, I, obj, U, V, SEP, SEN, Wo, rng, cnt, sh);
    Ld = add(Ld, estimate_direct<FULL, COUNT>(p, T, *L, pos, smp, SEP, SEN, Wo, sh, rng, cnt));
    return scl((float)count, Ld);
}

/* Shading of one volume event with the shadow-ray marches deferred (ray-march sampler only,
 * whose shadow march consumes exactly two draws whenever it starts). Same operations and
 * order as uniform_sample_one_light / estimate_direct; the contributions that wait for a
 * visibility result are kept with their state. */
struct EventRec {
    uint32_t pixel, state;        /* state: 2 bits per contribution: 0 none, 1 add, 2 add if visible */
    float count;                  /* light count scaling, or 0 when the result is not scaled */
    float pad;
    float Ld0x, Ld0y, Ld0z;       /* 0 + emission */
    float c1x, c1y, c1z, c2x, c2y, c2z;
    float pad2;
};

template<int FULL, int COUNT>
__device__ void shade_split(const DevParams& p, const Sample& smp, f3 SEP, f3 SEN, f3 Wo, Rng& rng, Cnt& cnt,
                            EventRec& ev, ShadowJob& job0, ShadowJob& job1, bool& has0, bool& has1, int obj,
                            float U, float V, float I, CD6 cd)
{
    /* I: the intensity at SEP (GetIntensity, transport.h:119), fetched by the caller together
     * with the gradient samples */
    f3 Ld = mk(0.0f, 0.0f, 0.0f);
    if (COUNT) cnt.shading += 1 + gradient_samples(p);
    Ld = add(Ld, ctf_eval(p.emission, I));
    ev.Ld0x = Ld.x; ev.Ld0y = Ld.y; ev.Ld0z = Ld.z;
    ev.state = 0;
    ev.count = 0.0f;
    has0 = false;
    has1 = false;
    const int count = p.light_count;
    if (count <= 0) return;
    const int pos = (int)floorf(smp.lightnum * (float)count);
    const int LightID = p.light_ids[pos];
    const DevLight* Lp = light_by_id(p, LightID);
    if (LightID < 0 || !Lp) return;
    const DevLight& L = *Lp;
    ev.count = (float)count;
    Shader sh;
    event_shader<FULL, COUNT>(p, I, obj, U, V, SEP, SEN, Wo, rng, cnt, sh, cd);

    /* light-sampling half */
    f3 SP, SN; float su, sv;
    sample_light_surface(p, L, smp.uvw0, smp.uvw1, smp.uvw2, SP, SN, su, sv);
    f3 Wi = nrm(sub(SP, SEP));
    f3 Li = scl(L.multiplier, eval_texture(p, L.texture_id, su, sv));
    if (L.one_sided && dot(sub(SEP, SP), SN) < 0.0f) Li = mk(0.0f, 0.0f, 0.0f);
    if (L.unit == ER_UNIT_LUX) Li = mk(Li.x / L.area, Li.y / L.area, Li.z / L.area);
    f3 F = shader_f<FULL>(sh, Wo, Wi);
    float BsdfPdf = shader_pdf<FULL>(sh, Wo, Wi);
    if (!black(Li) && !black(F) && BsdfPdf > 0.0f) {
        const int v = visible_defer<FULL, COUNT>(p, SEP, SP, rng, cnt, job0);
        if (v) {
            const f3 d = sub(SEP, SP);
            const float LightPdf = dot(d, d) / (absdot(SN, neg(Wi)) * L.area);
            const float Weight = power_heuristic(LightPdf, BsdfPdf);
            const f3 FL = mul(F, Li);
            f3 c;
            if (sh.brdf) c = scl(absdot(Wi, SEN) * Weight / LightPdf, FL);
            else if (FULL) c = scl(Weight / LightPdf, FL);
            else c = mk(FL.x / LightPdf, FL.y / LightPdf, FL.z / LightPdf);
            ev.c1x = c.x; ev.c1y = c.y; ev.c1z = c.z;
            ev.state |= (uint32_t)v;
            if (v == 2) { job0.slot = 0; has0 = true; }
        }
    }
    if (!FULL) return;

    /* BSDF-sampling half */
    F = shader_sample_f(sh, Wo, Wi, BsdfPdf, smp.comp, smp.dir0, smp.dir1);
    if (black(F) || BsdfPdf <= 0.0f) return;
    Ray r2;
    r2.O = SEP; r2.D = Wi; r2.MinT = 0.0f; r2.MaxT = 1000000.0f;
    float T2; f3 P2, N2, Le2;
    const int idx2 = intersect_lights(p, r2, false, T2, P2, N2, Le2);
    if (idx2 < 0 || idx2 != pos) return;
    if (!black(Le2)) {
        const int v = visible_defer<FULL, COUNT>(p, SEP, P2, rng, cnt, job1);
        if (v) {
            const f3 d = sub(SEP, P2);
            const float LightPdf = dot(d, d) / (absdot(N2, neg(Wi)) * L.area);
            const float Weight = power_heuristic(BsdfPdf, LightPdf);
            const f3 FL = mul(F, Le2);
            f3 c;
            if (sh.brdf) c = scl(absdot(Wi, SEN) * Weight / BsdfPdf, FL);
            else c = scl(Weight / BsdfPdf, FL);
            ev.c2x = c.x; ev.c2y = c.y; ev.c2z = c.z;
            ev.state |= (uint32_t)v << 2;
            if (v == 2) { job1.slot = 1; has1 = true; }
        }
    }
}

/* The 11 sample draws of one path (sample.h:258-262, in the reference's order) and its camera
 * ray (singlescattering.h:29-50, with the thin lens of montecarlo.h:86-137). */
__device__ __forceinline__ Ray camera_ray(const DevParams& p, int px, int py, Rng& rng, Sample& smp)
{
    smp.comp = rng.u();
    smp.dir0 = rng.u(); smp.dir1 = rng.u();
    smp.uvw0 = rng.u(); smp.uvw1 = rng.u(); smp.uvw2 = rng.u();
    smp.lightnum = rng.u();
    smp.film0 = rng.u(); smp.film1 = rng.u();
    smp.lens0 = rng.u(); smp.lens1 = rng.u();
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
    return r;
}

/* One scattering pixel handed from the primary kernel to the shading kernel (96 bytes). */
struct ScatterRec {
    uint32_t pixel, s0, s1;
    int obj;                      /* object list position, or -1 for a volume event */
    float Px, Py, Pz, Wox, Woy, Woz;
    float comp, dir0, dir1, uvw0, uvw1, uvw2;
    float lightnum, Nx, Ny, Nz;   /* N: object normal (object events) */
    float U, V, pad0, pad1;       /* object surface coordinates */
};

/* One tracer of a batch (er_render_batch): its parameter block and per-pass buffers, its
 * share of the batch's 8 x 8 primary tiles (tile_base .. tile_base + ntiles) and 16 x 16
 * post tiles, and its pass count. Jobs are sorted by pass count, largest first, so the jobs
 * still active in pass round r are a prefix of the table. */
struct BatchJob {
    const DevParams* params;
    float4* frame;
    float4* running;
    uchar4* display;
    uint32_t* seeds1;
    uint32_t* seeds2;
    ScatterRec* queue;
    EventRec* events;
    ShadowJob* jobs;
    unsigned char* vis;
    unsigned int* counts;       /* [0] scatter queue, [1] shadow jobs, [2] primary tiles claimed */
    int tiles_x, ntiles, tile_base;
    int post_tiles_x, post_ntiles, post_base;
    int passes, no_iter0;
};

} /* namespace erd */

#endif /* ER_PATH_H */
