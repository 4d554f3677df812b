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
 *     raymarching.h:73-113, reports no event), up to where the optical depth reaches
 *     ADJ_TAU_CAP;
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
 * Gradient mode (er_adjoint_lut) runs per pass the forward primary and shading kernels
 * (er_trace.hip) and then the wavefront kernels below: shadow transmittances and their
 * gradients, the smooth values and diffuse-colour gradients, a list of pixels with a non-zero
 * adjoint, and the replay of their primary marches. Gradients go into per-workgroup LDS
 * histograms of the node table; a last kernel sums the workgroups' histograms in FP64.
 *
 * Value mode (er_adjoint_value, k_adj_value): the smooth estimator itself, evaluated with
 * another node table while the decisions keep the current one, one thread per path; it reads
 * every sample (no skipping), since a macrocell empty under one table need not be empty under
 * the other.
 */
#include "er_path.h"

namespace erd {

#define ADJ_NODES (ER_TF_LUT_NODES * 4)
/* the smooth shadow test stops where the segment's optical depth reaches this value: what
 * lies beyond is attenuated below exp(-12) = 6e-6 (the value and the gradient both stop there,
 * so they stay each other's derivative; the expectation changes by at most that factor) */
#define ADJ_TAU_CAP 12.0f
/* dense samples per march iteration in the replay (loads in flight together) */
#ifndef ER_BATCH_PRIMARY_ADJ
#define ER_BATCH_PRIMARY_ADJ 8
#endif
#ifndef ER_BATCH_SHADOW_ADJ
#define ER_BATCH_SHADOW_ADJ 8
#endif

struct AdjArgs {
    const float* lut_eval;      /* node table the values use (value mode) or the current one */
    const float* sigma_eval;    /* DS * opacity_eval(i), i in [0, lut_max] (value mode) */
    const float4* adj;          /* unused by the value kernel */
    float4* value;              /* the smooth frame estimate */
    float* partials;            /* unused by the value kernel */
};

/* s += x with Kahan's compensation c: the error of a float sum over thousands of samples stays
 * at a few ulp of the total instead of growing with the count (needs -ffp-contract=off, no
 * reassociation) */
__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

/* sums sigma over the visited samples (sd); EVAL also sums the evaluation table's sigma minus
 * the decision table's (dd), so that the difference of the two optical depths is accumulated
 * directly instead of as the difference of two large sums */
template<int EVAL>
struct VisSum {
    const float* sig_eval;
    int lut_max;
    float sd, cd, dd, cdd;
    float last_d, last_e;
    __device__ __forceinline__ void init(const float* sig, int lmax)
    {
        sig_eval = sig; lut_max = lmax;
        sd = cd = dd = cdd = 0.0f;
        last_d = last_e = 0.0f;
    }
    __device__ __forceinline__ void begin(float) {}
    __device__ __forceinline__ void sample(int i, float sigma)
    {
        const float e = EVAL ? gmem(sig_eval)[min(i, lut_max)] : sigma;
        kahan_add(sd, cd, sigma);
        if (EVAL) kahan_add(dd, cdd, e - sigma);
        last_d = sigma;
        last_e = e;
    }
    __device__ __forceinline__ void event(int i, float sigma)
    {
        last_d = sigma;
        last_e = EVAL ? gmem(sig_eval)[min(i, lut_max)] : sigma;
    }
    /* sum of the evaluation table's sigma */
    __device__ __forceinline__ float se() const { return EVAL ? sd + dd : sd; }
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
    __device__ __forceinline__ void event(int i, float sigma)
    {
        last_i = i;
        last_sigma = sigma;
    }
};

/* ---- marches stepped a wave at a time, their last ones finished by groups of lanes ----- */

/* The adjoint kernels' marches, stepped together by the lanes of a wave (march_step with an
 * observer, the same samples in the same order as march_v) until at most 64 / ADJ_DRAIN_G of
 * them are left; those are then handed to aligned groups of ADJ_DRAIN_G lanes (group i the
 * i-th marching lane's march, its state the same in every lane of the group) and stepped with
 * group_step_v, which makes the same decisions. Each lane of a group observes its own samples;
 * the group's observations are merged back into the marching lane's observer. The longest
 * marches of a wave then advance ADJ_DRAIN_G / K times as many samples per round instead of
 * leaving the wave's other lanes idle. Every lane of the wave must call these. */
#ifndef ADJ_DRAIN_G
#define ADJ_DRAIN_G 4
#endif

/* the group's march: the (lane / G)-th marching lane's; returns its lane (0 if none) */
template<int G>
__device__ __forceinline__ int adj_group_source(unsigned long long busy, int lane, bool& has)
{
    unsigned long long x = busy;
    for (int k = 0; k < lane / G; ++k) x &= x - 1ull;
    has = x != 0ull;
    return has ? __ffsll((long long)x) - 1 : 0;
}

__device__ __forceinline__ MarchState adj_march_from(const MarchState& m, int src)
{
    MarchState g;
    g.O = mk(__shfl(m.O.x, src), __shfl(m.O.y, src), __shfl(m.O.z, src));
    g.D = mk(__shfl(m.D.x, src), __shfl(m.D.y, src), __shfl(m.D.z, src));
    g.MinT = __shfl(m.MinT, src);
    g.MaxT = __shfl(m.MaxT, src);
    g.Sum = __shfl(m.Sum, src);
    g.S = __shfl(m.S, src);
    return g;
}

/* the rest of the marching lanes' transmittance marches (sum of sigma, VisSum<0>): on return
 * busy is false everywhere and each marching lane's vs.sd holds its whole march's sum */
template<bool PRIMARY>
__device__ __forceinline__ void adj_drain_sum(const DevParams& p, const Tables& T, const MarchState& m, bool busy, VisSum<0>& vs,
                                           float StepSize, int lane);
template<bool PRIMARY>
__device__ __forceinline__ void adj_march_sum(const DevParams& p, const Tables& T, MarchState& m, bool busy, VisSum<0>& vs,
                                              float StepSize, int lane)
{
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (;;) {
        const int nb = __popcll(__ballot(busy));
        if (nb == 0) return;
        if (nb <= 64 / ADJ_DRAIN_G) break;
        if (busy) {
            f3 P;
            if (march_step<1, 0, ER_BATCH_SHADOW_ADJ, VisSum<0>>(p, T, m, PRIMARY, StepSize, cnt, vs, P) != 0) busy = false;
        }
    }
    adj_drain_sum<PRIMARY>(p, T, m, busy, vs, StepSize, lane);
}
template<bool PRIMARY>
__device__ __forceinline__ void adj_drain_sum(const DevParams& p, const Tables& T, const MarchState& m, bool busy, VisSum<0>& vs,
                                           float StepSize, int lane)
{
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int G = ADJ_DRAIN_G;
    const unsigned long long bm = __ballot(busy);
    bool has;
    const int src = adj_group_source<G>(bm, lane, has);
    MarchState g = adj_march_from(m, src);
    /* the group's first lane continues the marching lane's compensated sum, the others start at 0 */
    const bool first = (lane & (G - 1)) == 0;
    VisSum<0> gv;
    gv.init(nullptr, vs.lut_max);
    const float sd0 = __shfl(vs.sd, src), cd0 = __shfl(vs.cd, src);
    if (first) { gv.sd = sd0; gv.cd = cd0; }
    bool gb = has;
    while (__ballot(gb)) {
        if (gb) {
            f3 P;
            int iters = 0;
            if (group_step_v<1, 0, G, PRIMARY, VisSum<0>>(p, T, g, StepSize, lane, cnt, P, iters, gv) != 0) gb = false;
        }
    }
    /* the group's sum: the lanes' compensated sums added across the group */
    float v = gv.sd - gv.cd;
    float tot = v;
    for (int o = 1; o < G; o <<= 1) tot += __shfl_xor(tot, o, G);
    const int from = busy ? __popcll(bm & ((1ull << lane) - 1ull)) * G : 0;
    const float total = __shfl(tot, from);
    if (busy) { vs.sd = total; vs.cd = 0.0f; }
}

/* the rest of the marching lanes' gradient marches (VisGrad into the LDS histogram): on return
 * each marching lane's vol is whether its march ended on an event, and vg.last_i / last_sigma
 * are the event sample's */
template<bool PRIMARY>
__device__ __forceinline__ void adj_drain_grad(const DevParams& p, const Tables& T, const MarchState& m, bool busy, bool& vol,
                                            VisGrad& vg, float StepSize, int lane);
template<bool PRIMARY, int K>
__device__ __forceinline__ void adj_march_grad(const DevParams& p, const Tables& T, MarchState& m, bool busy, bool& vol,
                                               VisGrad& vg, float StepSize, int lane)
{
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (;;) {
        const int nb = __popcll(__ballot(busy));
        if (nb == 0) return;
        if (nb <= 64 / ADJ_DRAIN_G) break;
        if (busy) {
            f3 P;
            const int st = march_step<1, 0, K, VisGrad>(p, T, m, PRIMARY, StepSize, cnt, vg, P);
            if (st != 0) { busy = false; vol = st > 0; }
        }
    }
    adj_drain_grad<PRIMARY>(p, T, m, busy, vol, vg, StepSize, lane);
}
template<bool PRIMARY>
__device__ __forceinline__ void adj_drain_grad(const DevParams& p, const Tables& T, const MarchState& m, bool busy, bool& vol,
                                            VisGrad& vg, float StepSize, int lane)
{
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int G = ADJ_DRAIN_G;
    const unsigned long long bm = __ballot(busy);
    bool has;
    const int src = adj_group_source<G>(bm, lane, has);
    MarchState g = adj_march_from(m, src);
    if (busy) vg.flush();
    VisGrad gv;
    gv.hist = vg.hist; gv.scale = vg.scale; gv.cur = -1; gv.a0 = gv.a1 = 0.0f;
    gv.last_i = 0; gv.last_sigma = 0.0f;
    gv.w = __shfl(vg.w, src);
    bool gb = has;
    int gst = 0;
    while (__ballot(gb)) {
        if (gb) {
            f3 P;
            int iters = 0;
            const int st = group_step_v<1, 0, G, PRIMARY, VisGrad>(p, T, g, StepSize, lane, cnt, P, iters, gv);
            if (st != 0) { gst = st; gb = false; }
        }
    }
    gv.flush();
    const int from = busy ? __popcll(bm & ((1ull << lane) - 1ull)) * G : 0;
    const int rst = __shfl(gst, from);
    const int li = __shfl(gv.last_i, from);
    const float ls = __shfl(gv.last_sigma, from);
    if (busy) {
        vol = rst > 0;
        vg.last_i = li;
        vg.last_sigma = ls;
    }
}

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
    if (p.object_count > 0 && intersects_any_object(p, r)) return 0.0f;
    rng_at = rng;
    VisSum<EVAL> vs;
    vs.init(a.sigma_eval, p.lut_max);
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float Tt; f3 Pt;
    float nearT, farT;
    if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) return 1.0f;
    marched = true;
    march_v<1, 0, ER_BATCH_SHADOW_ADJ, VisSum<EVAL>>(p, T, r, false, rng, Tt, Pt, cnt, vs, ADJ_TAU_CAP / p.density_scale);
    const float dt = p.step_shadow * p.minstep;
    return erm::expf_(-((p.density_scale * dt) * vs.se()));
}

__device__ __forceinline__ f3 lut_diffuse(const float* lut, float scale, float I)
{
    return mk(tf_lut_eval(lut, scale, I, 1), tf_lut_eval(lut, scale, I, 2), tf_lut_eval(lut, scale, I, 3));
}

/* value mode: the smooth estimator of one pass with the evaluation table, every sample read */
__global__ __launch_bounds__(256) void k_adj_value(const DevParams* __restrict__ pp, AdjArgs a,
                                                uint32_t* __restrict__ seeds1, uint32_t* __restrict__ seeds2)
{
    extern __shared__ uint32_t lds[];
    const DevParams& p = *pp;
    /* occupancy bits (value mode: all set) and the decision sigma table */
    const int nw = p.mc_words;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = 0xffffffffu;
    float* lut = (float*)(lds + nw);
    for (int k = threadIdx.x; k <= p.lut_max; k += blockDim.x) lut[k] = gmem(p.sigma_lut)[k];
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
        const Ray r = camera_ray(p, px, py, rng, smp);
        VisSum<1> vp;
        vp.init(a.sigma_eval, p.lut_max);
        Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        float Tv = 0.0f;
        f3 Pv = mk(0.0f, 0.0f, 0.0f);
        const bool vol = march_v<1, 0, ER_BATCH_PRIMARY_ADJ, VisSum<1>>(p, T, r, true, rng, Tv, Pv, cnt, vp, -1.0f);
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
                Kd = lut_diffuse(a.lut_eval, p.tf_lut_scale, Iev);
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
                    T1 = transmittance<1>(p, T, a, SEP, SP, rng, r1, rng1, m1);
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
                        T2 = transmittance<1>(p, T, a, SEP, P2, rng, r2s, rng2, m2);
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

        /* P_k(theta) / P_k(theta0) of the sampled free flight: exp(-DS dt (optical depth
         * difference before k)) times the ratio of the two scatter probabilities at k */
        const float kdt = p.density_scale * (p.step_primary * p.minstep);
        float wf = 1.0f;
        if (vol) {
            const float d_before = vp.dd - (vp.last_e - vp.last_d);
            const float num = -erm::expm1f_(-(kdt * vp.last_e));
            const float den = -erm::expm1f_(-(kdt * vp.last_d));
            wf = erm::expf_(-(kdt * d_before)) * (den > 0.0f ? num / den : 1.0f);
        } else {
            wf = erm::expf_(-(kdt * vp.dd));
        }
        a.value[i] = make_float4(wf * value.x, wf * value.y, wf * value.z, type >= 0 ? 1.0f : 0.0f);
        seeds1[i] = rng.s0;
        seeds2[i] = rng.s1;
    }
}

/* ---- gradient mode as a wavefront over the forward kernels ------------------------------
 * Per pass: the forward primary and shading kernels (same decisions, queues and deferred
 * shadow jobs), then
 *   k_adj_shadow   each shadow job: transmittance on the same samples (march to the optical
 *                  depth cap), the job's shadowed value g = A . (count c T) from its event,
 *                  and - when g != 0 - the same march again scattering -g dtau/dnode;
 *   k_adj_finish   each event: the smooth value count (Ld0 + c1 T1 + c2 T2) into the frame,
 *                  and the diffuse-colour gradient A_c count (c1_c T1 + c2_c T2) / Kd_c;
 *   k_adj_list     every pixel with g_total = A . value != 0 into a list;
 *   k_adj_primary  each listed pixel: its camera ray regenerated from the pass's start seeds
 *                  and the primary march replayed, scattering -g_total dtau/dnode, and at the
 *                  sampled scatter sample g_total / (exp(tau) - 1) more. */

__device__ __forceinline__ Tables adj_stage(const DevParams& p, uint32_t* lds)
{
    const int nw = p.mc_words;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = gmem(p.mc_bits)[k];
    float* lut = (float*)(lds + nw);
    for (int k = threadIdx.x; k <= p.lut_max; k += blockDim.x) lut[k] = gmem(p.sigma_lut)[k];
    Tables T;
    T.bits = lds;
    T.bits_glb = nullptr;
    T.lut_lds = lut;
    T.lut_glb = nullptr;
    __syncthreads();
    return T;
}

__device__ __forceinline__ void adj_flush_hist(const float* hist, float* __restrict__ partials)
{
    __syncthreads();
    float* out = partials + (size_t)blockIdx.x * ADJ_NODES;
    for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) out[k] = hist[k];
}

#ifndef ADJ_SHADOW_WAVES_PER_EU
#define ADJ_SHADOW_WAVES_PER_EU 3
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ADJ_SHADOW_WAVES_PER_EU))) void k_adj_shadow(const DevParams* __restrict__ pp, const unsigned int* __restrict__ counts,
                                                   const ShadowJob* __restrict__ jobq, const EventRec* __restrict__ events,
                                                   const float4* __restrict__ adj, float* __restrict__ tvis,
                                                   float* __restrict__ partials)
{
    extern __shared__ uint32_t lds[];
    __shared__ float hist[ADJ_NODES];
    const DevParams& p = *pp;
    for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) hist[k] = 0.0f;
    const Tables T = adj_stage(p, lds);   /* ends with a barrier */
    const unsigned int n = counts[1];
    const float DS = p.density_scale;
    const float dts = p.step_shadow * p.minstep;
    const float k2 = (DS * DS) * dts;          /* d tau / d opacity per sample */
    const float cap = ADJ_TAU_CAP / p.density_scale;
    const int lane = threadIdx.x & 63;
    /* wave-aligned grid stride: the marches of a wave step together (adj_march_sum / _grad) */
    for (unsigned int j0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); j0 < n; j0 += gridDim.x * blockDim.x) {
        const unsigned int j = j0 + (unsigned int)lane;
        const bool has = j < n;
        ShadowJob job;
        Ray r;
        MarchState m;
        VisSum<0> vs;
        vs.init(nullptr, p.lut_max);
        bool busy = false;
        float nearT = 0.0f, farT = 0.0f;
        bool inbox = false;
        if (has) {
            job = jobq[j];
            r.O = mk(job.Ox, job.Oy, job.Oz);
            r.D = mk(job.Dx, job.Dy, job.Dz);
            r.MinT = 0.0f;
            r.MaxT = job.MaxT;
            Rng rng;
            rng.s0 = job.s0;
            rng.s1 = job.s1;
            inbox = intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT);
            if (inbox && march_begin(p, r, nearT, farT, dts, rng, cap, m)) busy = m.Sum < m.S;
        }
        adj_march_sum<false>(p, T, m, busy, vs, dts, lane);
        float g = 0.0f;
        if (has) {
            const float Tr = erm::expf_(-((DS * dts) * vs.sd));
            tvis[2 * job.event + job.slot] = Tr;
            const EventRec& ev = events[job.event];
            const f3 c = job.slot ? mk(ev.c2x, ev.c2y, ev.c2z) : mk(ev.c1x, ev.c1y, ev.c1z);
            const float4 A4 = adj[ev.pixel];
            g = ev.count * Tr * dot(mk(A4.x, A4.y, A4.z), c);
        }
        VisGrad vg;
        vg.hist = hist; vg.scale = p.tf_lut_scale; vg.cur = -1; vg.a0 = vg.a1 = 0.0f;
        vg.last_i = 0; vg.last_sigma = 0.0f;
        vg.w = -(g * k2);
        bool busy2 = false, vol = false;
        if (has && g != 0.0f && inbox) {
            /* the same samples again: the march restarts from the job's RNG state */
            Rng rr;
            rr.s0 = job.s0;
            rr.s1 = job.s1;
            if (march_begin(p, r, nearT, farT, dts, rr, cap, m)) busy2 = m.Sum < m.S;
        }
        adj_march_grad<false, ER_BATCH_SHADOW_ADJ>(p, T, m, busy2, vol, vg, dts, lane);
        vg.flush();
    }
    adj_flush_hist(hist, partials);
}

__global__ __launch_bounds__(256) void k_adj_finish(const DevParams* __restrict__ pp, const unsigned int* __restrict__ counts,
                                                   const EventRec* __restrict__ events, const ScatterRec* __restrict__ queue,
                                                   const float* __restrict__ tvis, const float4* __restrict__ adj,
                                                   float4* __restrict__ frame, float* __restrict__ partials)
{
    __shared__ float hist[ADJ_NODES];
    const DevParams& p = *pp;
    for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) hist[k] = 0.0f;
    __syncthreads();
    const unsigned int n = counts[0];
    for (unsigned int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const EventRec ev = events[q];
        const uint32_t s1 = ev.state & 3u, s2 = (ev.state >> 2) & 3u;
        const float T1 = s1 == 1u ? 1.0f : (s1 == 2u ? tvis[2 * q] : 0.0f);
        const float T2 = s2 == 1u ? 1.0f : (s2 == 2u ? tvis[2 * q + 1] : 0.0f);
        const f3 c1 = mk(ev.c1x, ev.c1y, ev.c1z), c2 = mk(ev.c2x, ev.c2y, ev.c2z);
        f3 Ld = mk(ev.Ld0x, ev.Ld0y, ev.Ld0z);
        if (ev.count > 0.0f) Ld = scl(ev.count, add(Ld, add(scl(T1, c1), scl(T2, c2))));
        frame[ev.pixel] = make_float4(Ld.x, Ld.y, Ld.z, 1.0f);
        if (!(ev.count > 0.0f)) continue;
        const float4 A4 = adj[ev.pixel];
        const ScatterRec& rec = queue[q];
        const float I = intensity(p, mk(rec.Px, rec.Py, rec.Pz));
        int b; float f;
        tf_lut_coord(I, p.tf_lut_scale, b, f);
        const f3 gk = scl(ev.count, add(scl(T1, c1), scl(T2, c2)));   /* = Kd (.) dvalue/dKd */
        const float A[3] = {A4.x, A4.y, A4.z}, G[3] = {gk.x, gk.y, gk.z};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float kd = tf_lut_raw(p.tf_lut, b, f, ch + 1);
            if (!(kd > 0.0f) || G[ch] == 0.0f || A[ch] == 0.0f) continue;
            const float gc = A[ch] * (G[ch] / kd);
            atomicAdd(&hist[b * 4 + ch + 1], gc * (1.0f - f));
            atomicAdd(&hist[(b + 1) * 4 + ch + 1], gc * f);
        }
    }
    adj_flush_hist(hist, partials);
}

/* pixels whose smooth value has a non-zero adjoint, with g_total, into a list */
__global__ __launch_bounds__(256) void k_adj_list(const float4* __restrict__ frame, const float4* __restrict__ adj, int n,
                                                 uint2* __restrict__ list, unsigned int* __restrict__ count)
{
    const int lane = threadIdx.x & 63;
    for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += gridDim.x * blockDim.x) {
        const int i = base + lane;
        float g = 0.0f;
        if (i < n) {
            const float4 v = frame[i], A = adj[i];
            g = A.x * v.x + A.y * v.y + A.z * v.z;
        }
        const bool has = g != 0.0f;
        const unsigned long long m = __ballot(has);
        if (!m) continue;
        const int leader = __ffsll((long long)m) - 1;
        unsigned int slot = 0;
        if (lane == leader) slot = atomicAdd(count, (unsigned int)__popcll(m));
        slot = __shfl(slot, leader);
        if (has) list[slot + (unsigned int)__popcll(m & ((1ull << lane) - 1ull))] = make_uint2((unsigned int)i, __float_as_uint(g));
    }
}

__global__ __launch_bounds__(256) void k_adj_primary(const DevParams* __restrict__ pp, const uint2* __restrict__ list,
                                                    const unsigned int* __restrict__ count,
                                                    const uint32_t* __restrict__ seeds1, const uint32_t* __restrict__ seeds2,
                                                    float* __restrict__ partials)
{
    extern __shared__ uint32_t lds[];
    __shared__ float hist[ADJ_NODES];
    const DevParams& p = *pp;
    for (int k = threadIdx.x; k < ADJ_NODES; k += blockDim.x) hist[k] = 0.0f;
    const Tables T = adj_stage(p, lds);
    const unsigned int n = *count;
    const float kdt = p.density_scale * (p.step_primary * p.minstep);   /* d tau / d sigma per sample */
    const float k2 = p.density_scale * kdt;                             /* d tau / d opacity per sample */
    const float StepSize = p.step_primary * p.minstep;
    const int lane = threadIdx.x & 63;
    /* wave-aligned grid stride: the marches of a wave step together (adj_march_grad) */
    for (unsigned int j0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); j0 < n; j0 += gridDim.x * blockDim.x) {
        const unsigned int j = j0 + (unsigned int)lane;
        const bool has = j < n;
        float g = 0.0f;
        VisGrad vg;
        vg.hist = hist; vg.scale = p.tf_lut_scale; vg.cur = -1; vg.a0 = vg.a1 = 0.0f;
        vg.last_i = 0; vg.last_sigma = 0.0f;
        vg.w = 0.0f;
        MarchState m;
        bool busy = false, vol = false;
        if (has) {
            const uint2 e = list[j];
            const int i = (int)e.x;
            g = __uint_as_float(e.y);
            Rng rng;
            rng.s0 = seeds1[i];
            rng.s1 = seeds2[i];
            Sample smp;
            const Ray r = camera_ray(p, i % p.film_w, i / p.film_w, rng, smp);
            vg.w = -(g * k2);
            float nearT, farT;
            if (intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT) &&
                march_begin(p, r, nearT, farT, StepSize, rng, -1.0f, m)) {
                if (!(m.Sum < m.S)) vol = true;   /* an event at the first sample, no sample visited */
                else busy = true;
            }
        }
        adj_march_grad<true, ER_BATCH_PRIMARY_ADJ>(p, T, m, busy, vol, vg, StepSize, lane);
        vg.flush();
        if (has && vol && vg.last_sigma > 0.0f) {
            const float tau = kdt * vg.last_sigma;
            vg.w = (g / erm::expm1f_(tau) + g) * k2;
            vg.sample(vg.last_i, vg.last_sigma);
            vg.flush();
        }
    }
    adj_flush_hist(hist, partials);
}

__global__ void k_adj_reset(unsigned int* __restrict__ counts, unsigned int* __restrict__ lcount)
{
    if (threadIdx.x < 4) counts[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *lcount = 0u;
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

/* out[k] += sum over workgroups of partials[.][k], in FP64: one workgroup per entry */
__global__ __launch_bounds__(256) void k_adj_reduce(const float* __restrict__ partials, int nblocks, double* __restrict__ out)
{
    __shared__ double red[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += (double)partials[(size_t)b * ADJ_NODES + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] += red[0];
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
    (void)mode;
    hipLaunchKernelGGL(erd::k_adj_value, grid, dim3(256), lds_bytes, st, p, a, s1, s2);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_adj_filter_t(const float* dldr, float4* out, int W, int H, float weight, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_adj_filter_t, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, dldr, out, W, H, weight);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_adj_reduce(const float* partials, int nblocks, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_adj_reduce, dim3(ADJ_NODES), dim3(256), 0, st, partials, nblocks, out);
    return hipGetLastError();
}

/* gradient-mode kernels after the forward primary + shading kernels of one pass; partials gets
 * 3 * blocks workgroup histograms */
extern "C" hipError_t erd_launch_adj_wavefront(const erd::DevParams* p, unsigned int* counts, const void* jobs,
                                               const void* events, const void* queue, const float4* adj, float* tvis,
                                               float4* frame, uint2* list, unsigned int* lcount, const uint32_t* s1start,
                                               const uint32_t* s2start, float* partials, int blocks, int n,
                                               size_t lds_bytes, hipStream_t st)
{
    const size_t per = (size_t)blocks * ADJ_NODES;
    hipLaunchKernelGGL(erd::k_adj_shadow, dim3(blocks), dim3(256), lds_bytes, st, p, (const unsigned int*)counts,
                       (const erd::ShadowJob*)jobs, (const erd::EventRec*)events, adj, tvis, partials);
    hipLaunchKernelGGL(erd::k_adj_finish, dim3(blocks), dim3(256), 0, st, p, (const unsigned int*)counts,
                       (const erd::EventRec*)events, (const erd::ScatterRec*)queue, (const float*)tvis, adj, frame,
                       partials + per);
    hipLaunchKernelGGL(erd::k_adj_list, dim3(blocks), dim3(256), 0, st, (const float4*)frame, adj, n, list, lcount);
    hipLaunchKernelGGL(erd::k_adj_primary, dim3(blocks), dim3(256), lds_bytes, st, p, (const uint2*)list,
                       (const unsigned int*)lcount, s1start, s2start, partials + 2 * per);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_adj_reset(unsigned int* counts, unsigned int* lcount, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_adj_reset, dim3(1), dim3(64), 0, st, counts, lcount);
    return hipGetLastError();
}

namespace erd {
/* p[0 .. words) = 0 */
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* __restrict__ p, size_t words)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}
} /* namespace erd */

extern "C" hipError_t erd_launch_zero_words(uint32_t* p, size_t words, hipStream_t st)
{
    if (words == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(erd::k_zero_words, dim3((unsigned)blocks), dim3(256), 0, st, p, words);
    return hipGetLastError();
}
