/*
 * er_adjoint.hip - path-replay adjoint of the node-table transfer function (er_set_tf_lut).
 *
 * No counterpart in the reference. For each pass, every path is replayed from its seed with
 * the forward pass's exact decisions (same draws, same scatter sample, same light and
 * direction choices; the march code is the forward march). The estimator that is
 * differentiated is the forward estimator with its two discrete tests replaced by their
 * expected values on the same samples:
 *   - shadow test -> transmittance T = exp(-DS * dt * sum sigma_j) over every sample of the
 *     shadow segment (the probability that the reference's ScatterEventInVolume,
 *     raymarching.h:73-113, reports no event);
 *   - primary free flight -> the path's value times P_k(theta) / P_k(theta0), where
 *     P_k = exp(-tau_<k) - exp(-tau_<=k), tau_j = DS * sigma_j * dt, is the probability that
 *     SampleVolume (raymarching.h:30-71) stops at the sampled sample k (for a path that leaves
 *     the volume, P = exp(-tau_all)).
 * Both have the forward's expectation for every theta. With sigma = DS * max(0, lerp(node
 * table)) and diffuse = max(0, lerp(node table)), the derivative of one path's value v with
 * respect to one sample's tau is -v for samples before k (and for shadow samples, the
 * shadowed part of v), v * exp(-tau_k) / (1 - exp(-tau_k)) at k; the diffuse colour enters
 * v linearly. Samples whose interpolated opacity is not positive have zero derivative, which
 * is what lets the replay skip empty macrocells exactly as the forward march does.
 *
 * Per path: march the primary ray (decisions), shade with both MIS halves marching their
 * shadow segments to the end (transmittance), then - knowing the path's value and the
 * adjoint of its pixel - replay the same marches once more, accumulating
 * dL/dtau_j * dtau_j/dnode into a per-workgroup LDS histogram of the node table. Workgroups
 * write their histograms to a partials array that a second kernel sums in FP64.
 *
 * Value mode (er_adjoint_value): the smooth estimator itself, evaluated with another node
 * table while the decisions keep the current one; it reads every sample (no skipping), since
 * a macrocell empty under one table need not be empty under the other.
 */
#include "er_path.h"

namespace erd {

#define ADJ_NODES (ER_TF_LUT_NODES * 4)
/* dense samples per march iteration in the replay (loads in flight together) */
#ifndef ER_BATCH_PRIMARY_ADJ
#define ER_BATCH_PRIMARY_ADJ 4
#endif
#ifndef ER_BATCH_SHADOW_ADJ
#define ER_BATCH_SHADOW_ADJ 4
#endif

struct AdjArgs {
    const float* lut_eval;      /* node table the values use (value mode) or the current one */
    const float* sigma_eval;    /* DS * opacity_eval(i), i in [0, lut_max] (value mode) */
    const float4* adj;          /* grad mode: dL/d(frame XYZ) of this pass */
    float4* value;              /* value mode: the smooth frame estimate */
    float* partials;            /* grad mode: ADJ_NODES floats per workgroup */
};

/* sums sigma over the visited samples; EVAL also sums the evaluation table's sigma */
template<int EVAL>
struct VisSum {
    const float* sig_eval;
    int lut_max;
    double sd, se;
    float last_d, last_e;
    __device__ __forceinline__ void begin(float) {}
    __device__ __forceinline__ void sample(int i, float sigma)
    {
        const float e = EVAL ? gmem(sig_eval)[min(i, lut_max)] : sigma;
        sd += (double)sigma;
        se += (double)e;
        last_d = sigma;
        last_e = e;
    }
};

/* scatters w * dtau/dnode of each visited sample into the LDS node histogram; consecutive
 * samples in the same node interval are summed in registers first */
struct VisGrad {
    float* hist;
    float scale;
    float w;             /* dL/d(opacity) per sample: dL/dtau * DS^2 * dt */
    int cur;
    float a0, a1;
    int last_i;
    float last_sigma;
    __device__ __forceinline__ void begin(float) {}
    __device__ __forceinline__ void flush()
    {
        if (cur >= 0) {
            atomicAdd(&hist[cur * 4], a0);
            atomicAdd(&hist[(cur + 1) * 4], a1);
        }
        cur = -1;
        a0 = a1 = 0.0f;
    }
    __device__ __forceinline__ void sample(int i, float sigma)
    {
        last_i = i;
        last_sigma = sigma;
        if (!(sigma > 0.0f)) return;   /* max(0, .) clamps: zero derivative */
        int b; float f;
        tf_lut_coord((float)i, scale, b, f);
        if (b != cur) { flush(); cur = b; }
        a0 += w * (1.0f - f);
        a1 += w * f;
    }
};

/* Visible (transport.h:46-56) in expected-value form: the transmittance of the shadow segment,
 * with the forward's draw consumption (none when shadows are off, a light shape blocks, or the
 * segment misses the volume; two otherwise). Returns the ray and the RNG state the march starts
 * from, for the gradient replay. */
template<int EVAL>
__device__ __forceinline__ float transmittance(const DevParams& p, const Tables& T, const AdjArgs& a, f3 P1, f3 P2, Rng& rng,
                               Ray& r, Rng& rng_at, bool& marched)
{
    marched = false;
    if (!p.shadows) return 1.0f;
    r = shadow_ray(p, P1, P2);
    if (intersects_any_light(p, r)) return 0.0f;
    rng_at = rng;
    VisSum<EVAL> vs;
    vs.sig_eval = a.sigma_eval; vs.lut_max = p.lut_max;
    vs.sd = vs.se = 0.0; vs.last_d = vs.last_e = 0.0f;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float Tt; f3 Pt;
    float nearT, farT;
    if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return 1.0f;
    marched = true;
    march_v<1, 0, ER_BATCH_SHADOW_ADJ, VisSum<EVAL>>(p, T, r, false, rng, Tt, Pt, cnt, vs, true);
    const double dt = (double)((p.step_shadow) * p.minstep);
    return (float)exp(-(double)p.density_scale * dt * vs.se);
}

__device__ __forceinline__ f3 lut_diffuse(const float* lut, float scale, float I)
{
    return mk(tf_lut_eval(lut, scale, I, 1), tf_lut_eval(lut, scale, I, 2), tf_lut_eval(lut, scale, I, 3));
}

template<int MODE>   /* 0: gradient, 1: value */
__global__ __launch_bounds__(256) void k_adjoint(const DevParams* __restrict__ pp, AdjArgs a,
                                                uint32_t* __restrict__ seeds1, uint32_t* __restrict__ seeds2)
{
    extern __shared__ uint32_t lds[];
    const DevParams& p = *pp;
    __shared__ float hist[ADJ_NODES];
    /* occupancy bits (value mode: all set) and the decision sigma table */
    const int nw = p.mc_words;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = MODE ? 0xffffffffu : gmem(p.mc_bits)[k];
    float* lut = (float*)(lds + nw);
    for (int k = threadIdx.x; k <= p.lut_max; k += blockDim.x) lut[k] = gmem(p.sigma_lut)[k];
    for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) hist[k] = 0.0f;
    __syncthreads();
    Tables T;
    T.bits = lds;
    T.bits_glb = nullptr;
    T.lut_lds = lut;
    T.lut_glb = nullptr;

    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int px = blockIdx.x * 16 + tx, py = blockIdx.y * 16 + ty;
    if (px < p.film_w && py < p.film_h) {
        const size_t i = (size_t)py * p.film_w + px;
        Rng rng;
        rng.s0 = seeds1[i];
        rng.s1 = seeds2[i];
        Sample smp;
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
        const Rng rng_primary = rng;
        VisSum<MODE> vp;
        vp.sig_eval = a.sigma_eval; vp.lut_max = p.lut_max;
        vp.sd = vp.se = 0.0; vp.last_d = vp.last_e = 0.0f;
        Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        float Tv = 0.0f;
        f3 Pv = mk(0.0f, 0.0f, 0.0f);
        const bool vol = march_v<1, 0, ER_BATCH_PRIMARY_ADJ, VisSum<MODE>>(p, T, r, true, rng, Tv, Pv, cnt, vp, false);
        float Tl = 0.0f; f3 Pl, Nl, Lel = mk(0.0f, 0.0f, 0.0f);
        const int il = intersect_lights(p, r, true, Tl, Pl, Nl, Lel);
        int type = -1;
        float Tn = 3.402823466e+38f;
        if (vol && Tv < Tn) { type = 0; Tn = Tv; }
        if (il >= 0 && Tl < Tn) { type = 1; Tn = Tl; }

        /* value of the path: count * (Ld0 + Kd (q1 T1 + q2 T2)) for a volume event, Le or 0 else */
        f3 value = mk(0.0f, 0.0f, 0.0f);
        f3 q1 = value, q2 = value, Kd = value;
        float T1 = 0.0f, T2 = 0.0f, count = 0.0f, Iev = 0.0f;
        bool m1 = false, m2 = false;
        Ray r1, r2s;
        Rng rng1 = rng, rng2 = rng;
        if (type == 0) {
            const f3 SEP = Pv;
            Iev = intensity(p, SEP);
            const f3 Ld0 = ctf_eval(p.emission, Iev);
            value = Ld0;
            const int nl = p.light_count;
            const int pos = nl > 0 ? (int)floorf(smp.lightnum * (float)nl) : 0;
            const int LightID = nl > 0 ? p.light_ids[pos] : -1;
            const DevLight* Lp = light_by_id(p, LightID);
            if (nl > 0 && LightID >= 0 && Lp) {
                const DevLight& L = *Lp;
                count = (float)nl;
                const f3 Kd_dec = diffuse_at(p, Iev);
                Kd = MODE ? lut_diffuse(a.lut_eval, p.tf_lut_scale, Iev) : Kd_dec;
                /* light-sampling half (phase function: F = Kd / pi, pdf 1 / (4 pi)) */
                f3 SP, SN; float su, sv;
                sample_light_surface(p, L, smp.uvw0, smp.uvw1, smp.uvw2, SP, SN, su, sv);
                const f3 Wi = nrm(sub(SP, SEP));
                f3 Li = scl(L.multiplier, eval_texture(p, L.texture_id, su, sv));
                if (L.one_sided && dot(sub(SEP, SP), SN) < 0.0f) Li = mk(0.0f, 0.0f, 0.0f);
                if (L.unit == ER_UNIT_LUX) Li = mk(Li.x / L.area, Li.y / L.area, Li.z / L.area);
                const float PhasePdf = 0.0795774715459476678844f;
                if (!black(Li) && !black(scl(ER_INV_PI_F, Kd_dec))) {
                    rng1 = rng;
                    T1 = transmittance<MODE>(p, T, a, SEP, SP, rng, r1, rng1, m1);
                    const f3 d = sub(SEP, SP);
                    const float LightPdf = dot(d, d) / (absdot(SN, neg(Wi)) * L.area);
                    const float Weight = power_heuristic(LightPdf, PhasePdf);
                    q1 = scl(Weight / LightPdf * ER_INV_PI_F, Li);
                }
                /* phase-sampling half */
                if (!black(scl(ER_INV_PI_F, Kd_dec))) {
                    const float z = 1.0f - 2.0f * smp.dir0;
                    const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                    const float phi = (2.0f * ER_PI_F) * smp.dir1;
                    float sn, cs;
                    erm::sincosf_(phi, &sn, &cs);
                    const f3 W2 = mk(rr * cs, rr * sn, z);
                    Ray rl;
                    rl.O = SEP; rl.D = W2; rl.MinT = 0.0f; rl.MaxT = 1000000.0f;
                    float T2l; f3 P2, N2, Le2;
                    const int idx2 = intersect_lights(p, rl, false, T2l, P2, N2, Le2);
                    if (idx2 >= 0 && idx2 == pos && !black(Le2)) {
                        rng2 = rng;
                        T2 = transmittance<MODE>(p, T, a, SEP, P2, rng, r2s, rng2, m2);
                        const f3 d = sub(SEP, P2);
                        const float LightPdf = dot(d, d) / (absdot(N2, neg(W2)) * L.area);
                        const float Weight = power_heuristic(PhasePdf, LightPdf);
                        q2 = scl(Weight / PhasePdf * ER_INV_PI_F, Le2);
                    }
                }
                value = add(value, add(scl(T1, mul(Kd, q1)), scl(T2, mul(Kd, q2))));
                value = scl(count, value);
            }
        } else if (type == 1) {
            value = Lel;
        }

        const double DS = (double)p.density_scale;
        const double dtp = (double)(p.step_primary * p.minstep);
        if (MODE) {
            /* P_k(theta) / P_k(theta0) of the sampled free flight */
            double w = 1.0;
            if (vol) {
                const double d_before = (vp.se - (double)vp.last_e) - (vp.sd - (double)vp.last_d);
                const double num = -expm1(-DS * dtp * (double)vp.last_e);
                const double den = -expm1(-DS * dtp * (double)vp.last_d);
                w = exp(-DS * dtp * d_before) * (den > 0.0 ? num / den : 1.0);
            } else {
                w = exp(-DS * dtp * (vp.se - vp.sd));
            }
            const float wf = (float)w;
            a.value[i] = make_float4(wf * value.x, wf * value.y, wf * value.z, type >= 0 ? 1.0f : 0.0f);
        } else {
            const float4 A4 = a.adj[i];
            const f3 A = mk(A4.x, A4.y, A4.z);
            const float g_total = dot(A, value);
            const float scale = p.tf_lut_scale;
            VisGrad vg;
            vg.hist = hist; vg.scale = scale; vg.cur = -1; vg.a0 = vg.a1 = 0.0f; vg.last_i = 0; vg.last_sigma = 0.0f;
            /* primary free flight */
            if (g_total != 0.0f) {
                Rng rr = rng_primary;
                vg.w = (float)(-(double)g_total * DS * DS * dtp);
                float Tq; f3 Pq;
                const bool v2 = march_v<1, 0, ER_BATCH_PRIMARY_ADJ, VisGrad>(p, T, r, true, rr, Tq, Pq, cnt, vg, false);
                vg.flush();
                if (v2 && vg.last_sigma > 0.0f) {
                    /* the sampled sample k: replace -g by g exp(-tau) / (1 - exp(-tau)) */
                    const double tau = DS * dtp * (double)vg.last_sigma;
                    const double r_k = 1.0 / expm1(tau);
                    vg.w = (float)(((double)g_total * r_k + (double)g_total) * DS * DS * dtp);
                    vg.sample(vg.last_i, vg.last_sigma);
                    vg.flush();
                }
            }
            if (type == 0 && count > 0.0f) {
                const double dts = (double)(p.step_shadow * p.minstep);
                const float g1 = count * dot(A, scl(T1, mul(Kd, q1)));
                const float g2 = count * dot(A, scl(T2, mul(Kd, q2)));
                if (m1 && g1 != 0.0f) {
                    Rng rr = rng1;
                    vg.w = (float)(-(double)g1 * DS * DS * dts);
                    float Tq; f3 Pq;
                    march_v<1, 0, ER_BATCH_SHADOW_ADJ, VisGrad>(p, T, r1, false, rr, Tq, Pq, cnt, vg, true);
                    vg.flush();
                }
                if (m2 && g2 != 0.0f) {
                    Rng rr = rng2;
                    vg.w = (float)(-(double)g2 * DS * DS * dts);
                    float Tq; f3 Pq;
                    march_v<1, 0, ER_BATCH_SHADOW_ADJ, VisGrad>(p, T, r2s, false, rr, Tq, Pq, cnt, vg, true);
                    vg.flush();
                }
                /* diffuse colour at the event, channel c enters channel c linearly */
                int b; float f;
                tf_lut_coord(Iev, scale, b, f);
                const f3 gk = scl(count, add(scl(T1, q1), scl(T2, q2)));
                const float gc0 = A.x * gk.x, gc1 = A.y * gk.y, gc2 = A.z * gk.z;
                const float* tl = p.tf_lut;
                if (gc0 != 0.0f && tf_lut_raw(tl, b, f, 1) > 0.0f) {
                    atomicAdd(&hist[b * 4 + 1], gc0 * (1.0f - f));
                    atomicAdd(&hist[(b + 1) * 4 + 1], gc0 * f);
                }
                if (gc1 != 0.0f && tf_lut_raw(tl, b, f, 2) > 0.0f) {
                    atomicAdd(&hist[b * 4 + 2], gc1 * (1.0f - f));
                    atomicAdd(&hist[(b + 1) * 4 + 2], gc1 * f);
                }
                if (gc2 != 0.0f && tf_lut_raw(tl, b, f, 3) > 0.0f) {
                    atomicAdd(&hist[b * 4 + 3], gc2 * (1.0f - f));
                    atomicAdd(&hist[(b + 1) * 4 + 3], gc2 * f);
                }
            }
        }
        seeds1[i] = rng.s0;
        seeds2[i] = rng.s1;
    }
    if (!MODE) {
        __syncthreads();
        float* out = a.partials + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * ADJ_NODES;
        for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) out[k] = hist[k];
    }
}

/* dL/dframe of one pass from dL/drunning: the transpose of the 3x3 frame filter (weights
 * exp(-d^2 / 2) normalised by the tap sum at the output pixel, filterframeestimate.cuh:28-64)
 * times the pass's weight in the running mean. */
__global__ __launch_bounds__(256) void k_adj_filter_t(const float* __restrict__ dldr, float4* __restrict__ out, int W,
                                                     int H, float weight)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float g[3] = {1.0f, erm::expf_(-0.5f), erm::expf_(-1.0f)};
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int yy = max(y - 1, 0); yy <= min(y + 1, H - 1); ++yy)
        for (int xx = max(x - 1, 0); xx <= min(x + 1, W - 1); ++xx) {
            /* total tap weight at the output pixel (xx, yy) */
            float tw = 0.0f;
            for (int vy = max(yy - 1, 0); vy <= min(yy + 1, H - 1); ++vy)
                for (int vx = max(xx - 1, 0); vx <= min(xx + 1, W - 1); ++vx)
                    tw += g[(vx - xx) * (vx - xx) + (vy - yy) * (vy - yy)];
            const float w = g[(xx - x) * (xx - x) + (yy - y) * (yy - y)] / tw;
            const float* d = dldr + ((size_t)yy * W + xx) * 3;
            s0 += w * d[0]; s1 += w * d[1]; s2 += w * d[2];
        }
    out[(size_t)y * W + x] = make_float4(weight * s0, weight * s1, weight * s2, 0.0f);
}

/* out[k] += sum over workgroups of partials[.][k], in FP64 */
__global__ __launch_bounds__(256) void k_adj_reduce(const float* __restrict__ partials, int nblocks, double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ADJ_NODES) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += (double)partials[(size_t)b * ADJ_NODES + k];
    out[k] += s;
}

} /* namespace erd */

extern "C" size_t erd_adj_partials_per_block(void) { return ADJ_NODES; }

extern "C" hipError_t erd_launch_adjoint(const erd::DevParams* p, int mode, const float* lut_eval, const float* sigma_eval,
                                         const float4* adj, float4* value, float* partials, uint32_t* s1, uint32_t* s2,
                                         int W, int H, size_t lds_bytes, hipStream_t st)
{
    erd::AdjArgs a;
    a.lut_eval = lut_eval; a.sigma_eval = sigma_eval; a.adj = adj; a.value = value; a.partials = partials;
    const dim3 grid((W + 15) / 16, (H + 15) / 16);
    if (mode) hipLaunchKernelGGL(erd::k_adjoint<1>, grid, dim3(256), lds_bytes, st, p, a, s1, s2);
    else hipLaunchKernelGGL(erd::k_adjoint<0>, grid, dim3(256), lds_bytes, st, p, a, s1, s2);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_adj_filter_t(const float* dldr, float4* out, int W, int H, float weight, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_adj_filter_t, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, dldr, out, W, H, weight);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_adj_reduce(const float* partials, int nblocks, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_adj_reduce, dim3((ADJ_NODES + 255) / 256), dim3(256), 0, st, partials, nblocks, out);
    return hipGetLastError();
}
