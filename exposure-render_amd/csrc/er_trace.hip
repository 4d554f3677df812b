/*
 * er_trace.hip - the forward kernels of one progressive pass on gfx950.
 *
 * A pass is a short pipeline of kernels over the whole film (KrnlSingleScattering,
 * Source/singlescattering.cuh:27-38, split where the paths diverge):
 *   k_primary  persistent waves take 8 x 8 pixel tiles from an atomic counter; each lane
 *              generates its camera ray and marches it (SampleVolume, raymarching.h:30-71);
 *              pixels without a volume event are final, the others are appended to a
 *              scatter queue (wave-aggregated atomics);
 *   k_shade    shades each queued event (UniformSampleOneLight, transport.h:115-156) and
 *              defers its shadow-ray marches to a job queue;
 *   k_shadow   marches each shadow ray (ScatterEventInVolume, raymarching.h:73-113);
 *   k_finish   combines the contributions with their visibility into the frame estimate.
 * RNG state lives in registers for the whole path and is written back once (the reference
 * read-modify-writes it in global memory on every draw, Source/rng.h:37-38); a deferred
 * shadow march takes the RNG state it starts from along and the shading kernel advances past
 * its two draws, so every path consumes exactly the reference draw sequence.
 *
 * Template parameters: FULL selects the full estimator (shading-mode selection, both halves
 * of MIS, clipping objects, Woodcock sampler); COUNT adds the sample counters.
 */
#include "er_path.h"

namespace erd {

/* Stage the macrocell bits (always) and the opacity LUT (when LUT_LDS) in LDS. */
template<int LUT_LDS>
__device__ __forceinline__ Tables stage_tables(const DevParams& p, uint32_t* lds)
{
    const int nw = p.mc_words;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = gmem(p.mc_bits)[k];
    Tables T;
    T.bits = lds;
    T.bits_glb = nullptr;
    if (LUT_LDS) {
        float* lut = (float*)(lds + nw);
        for (int k = threadIdx.x; k <= p.lut_max; k += blockDim.x) lut[k] = gmem(p.sigma_lut)[k];
        T.lut_lds = lut;
        T.lut_glb = nullptr;
    } else {
        T.lut_lds = nullptr;
        T.lut_glb = gmem(p.sigma_lut);
    }
    __syncthreads();
    return T;
}

template<int COUNT>
__device__ __forceinline__ void flush_counters(const Cnt& cnt, bool active, int paths, unsigned long long* counters)
{
    if (!COUNT) return;
    __shared__ unsigned long long acc[12];
    if (threadIdx.x < 12) acc[threadIdx.x] = 0ull;
    __syncthreads();
    if (active) {
        if (paths) atomicAdd(&acc[0], (unsigned long long)paths);
        atomicAdd(&acc[1], (unsigned long long)cnt.primary);
        atomicAdd(&acc[2], (unsigned long long)cnt.shadow);
        atomicAdd(&acc[3], (unsigned long long)cnt.shading);
        atomicAdd(&acc[4], (unsigned long long)cnt.pfetch);
        atomicAdd(&acc[5], (unsigned long long)cnt.sfetch);
        atomicAdd(&acc[6], (unsigned long long)cnt.scatter);
        atomicAdd(&acc[7], (unsigned long long)cnt.shadow_rays);
        atomicAdd(&acc[8], (unsigned long long)cnt.iters);
        atomicAdd(&acc[9], (unsigned long long)cnt.skipped);
        atomicAdd(&acc[10], (unsigned long long)cnt.undone);
    }
    __syncthreads();
    if (threadIdx.x < 12) atomicAdd(&counters[threadIdx.x], acc[threadIdx.x]);
}

/* the primary march's last marches of a wave go to groups of this many lanes once at most
 * 64 / ER_PRIMARY_DRAIN_G are left (0: to coop_march, one march at a time with the whole
 * wave, once at most ER_COOP_LANES are left; a wave marching a single ray still does) */
#ifndef ER_PRIMARY_DRAIN_G
#define ER_PRIMARY_DRAIN_G 4
#endif

/* One wave renders one 8 x 8 pixel tile: the 11 sample draws (sample.h:258-262), the camera
 * ray (singlescattering.h:29-50), SampleRay (singlescattering.h:52-74). Pixels whose nearest
 * event is a light or nothing are finished here; pixels with a volume event are appended to
 * the scatter queue (wave-aggregated atomic on *qcount) with their RNG state and the draws
 * the shading still needs. */
template<int FULL, int COUNT, int G = 1>
__device__ __forceinline__ void primary_tile(const DevParams& p, const Tables& T, unsigned int tile, int tiles_x, int lane,
                                             float4* __restrict__ frame, uint32_t* __restrict__ seeds1,
                                             uint32_t* __restrict__ seeds2, ScatterRec* __restrict__ queue,
                                             unsigned int* __restrict__ qcount, Cnt& cnt, int& paths, int& steps, int ray)
{
    const int ts = p.tile_shift;
    /* G == 1, ray >= 0: the whole wave computes that one pixel of the tile (every lane the same)
     * and marches it with coop_march; lane 0 writes the results. G == 1, ray < 0: lane k takes
     * pixel k. G = 4 or 16: ray is a sub-tile of 64 / G pixels, the aligned group of G lanes
     * lane / G computes its pixel lane / G (every lane of the group the same) and marches it
     * with group_march; the group's first lane writes the results. */
    const int pl = G > 1 ? ray * (64 / G) + lane / G : (ray >= 0 ? ray : lane);
    const bool owner = G > 1 ? (lane & (G - 1)) == 0 : (ray < 0 || lane == 0);
    const int px = (int)(tile % (unsigned int)tiles_x) * (1 << ts) + (pl & ((1 << ts) - 1));
    const int py = (int)(tile / (unsigned int)tiles_x) * (64 >> ts) + (pl >> ts);
    const bool active = px < p.film_w && py < p.film_h;
    const size_t i = (size_t)py * p.film_w + px;
    bool scatter = false;
    Sample smp;
    Rng rng;
    f3 Pv = mk(0.0f, 0.0f, 0.0f), D = mk(0.0f, 0.0f, 1.0f), Nobj = mk(0.0f, 0.0f, 0.0f);
    int obj = -1;
    float Uobj = 0.0f, Vobj = 0.0f;
    Ray r;
    bool vol = false, marching = false;
    float Tv = 0.0f;
    MarchState m;
    const float StepSize = p.step_primary * p.minstep;
    if (active) {
        if (owner) paths++;
        rng.s0 = seeds1[i];
        rng.s1 = seeds2[i];
        r = camera_ray(p, px, py, rng, smp);
        D = r.D;
        if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) {
            vol = march<FULL, COUNT, ER_BATCH_PRIMARY>(p, T, r, true, rng, Tv, Pv, cnt);
        } else {
            /* march_v's set-up; its loop follows, wave by wave */
            float nearT, farT;
            if (intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT) &&
                march_begin(p, r, nearT, farT, StepSize, rng, -1.0f, m)) {
                if (!(m.Sum < m.S)) { vol = true; Tv = m.MinT; Pv = mk(0.0f, 0.0f, 0.0f); }
                else marching = true;
            }
        }
    }
    /* The march loop, in three stages: every lane steps its own march until at most 64 / GD
     * lanes (ER_COOP_LANES without group tails) are still marching; groups of GD lanes then
     * take those marches over; a march left after that (a single-ray wave's) is finished by the
     * whole wave (coop_march). A lone ray through kilo-samples of tissue would otherwise keep its
     * wave for hundreds of microseconds. With G > 1 lanes to a ray, group_march steps every
     * march in its group from the start instead.
     * it: the lane's march iterations (the group and wave stages count their equivalents), for
     * the tile cost. */
    NoVisit nv;
    int it = 0;
    if constexpr (G > 1) {
        if (marching) {
            f3 P = mk(0.0f, 0.0f, 0.0f);
            const int st = group_march<FULL, COUNT, G>(p, T, m, StepSize, lane, cnt, P, it);
            marching = false; vol = st > 0; Tv = m.MinT;
            if (st > 0) Pv = P;
        }
    } else {
    constexpr int GD = ER_PRIMARY_DRAIN_G > 0 ? ER_PRIMARY_DRAIN_G : 4;
    for (;;) {
        const int nm = __popcll(__ballot(marching));
        if (nm == 0 || nm <= (ER_PRIMARY_DRAIN_G > 0 ? 64 / GD : ER_COOP_LANES) || ray >= 0) break;
        if (marching) {
            ++it;
            const int st = march_step<FULL, COUNT, ER_BATCH_PRIMARY, NoVisit>(p, T, m, true, StepSize, cnt, nv, Pv);
            if (st != 0) { marching = false; vol = st > 0; Tv = m.MinT; }
        }
    }
    /* At most 64 / GD marches left: aligned groups of GD lanes take them over (group i the
     * i-th marching lane's, its state the same in every lane of the group) and step them with
     * group_step, which makes the same decisions as the one-lane march; then each marching lane
     * reads its march's end from its group's first lane. */
    if (ER_PRIMARY_DRAIN_G > 0 && ray < 0) {
        const unsigned long long bm = __ballot(marching);
        if (bm) {
            unsigned long long x = bm;
            for (int k = 0; k < lane / GD; ++k) x &= x - 1ull;
            const bool has = x != 0ull;
            const int src = has ? __ffsll((long long)x) - 1 : 0;
            MarchState g;
            g.O = mk(__shfl(m.O.x, src), __shfl(m.O.y, src), __shfl(m.O.z, src));
            g.D = mk(__shfl(m.D.x, src), __shfl(m.D.y, src), __shfl(m.D.z, src));
            g.MinT = __shfl(m.MinT, src);
            g.MaxT = __shfl(m.MaxT, src);
            g.Sum = __shfl(m.Sum, src);
            g.S = __shfl(m.S, src);
            bool gb = has;
            int gst = 0, git = 0;
            f3 gP = mk(0.0f, 0.0f, 0.0f);
            while (__ballot(gb)) {
                if (gb) {
                    f3 P = mk(0.0f, 0.0f, 0.0f);
                    int iters = 0;
                    const int st = group_step<FULL, COUNT, GD, true>(p, T, g, StepSize, lane, cnt, P, iters);
                    git += iters;
                    if (st != 0) { gst = st; gP = P; gb = false; }
                }
            }
            const int from = marching ? __popcll(bm & ((1ull << lane) - 1ull)) * GD : 0;
            const int rst = __shfl(gst, from);
            const float rT = __shfl(g.MinT, from);
            const f3 rP = mk(__shfl(gP.x, from), __shfl(gP.y, from), __shfl(gP.z, from));
            const int rit = __shfl(git, from);
            if (marching) {
                marching = false;
                vol = rst > 0;
                Tv = rT;
                if (rst > 0) Pv = rP;
                it += rit;
            }
        }
    }
    unsigned long long rest = __ballot(marching);
    if (ray >= 0) rest &= 1ull;
    for (; rest; rest &= rest - 1ull) {
        const int L = __ffsll((long long)rest) - 1;
        MarchState w;
        w.O = mk(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.O.x), L)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.O.y), L)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.O.z), L)));
        w.D = mk(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.D.x), L)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.D.y), L)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.D.z), L)));
        w.MinT = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.MinT), L));
        w.MaxT = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.MaxT), L));
        w.Sum = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.Sum), L));
        w.S = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.S), L));
        f3 P = mk(0.0f, 0.0f, 0.0f);
        int csteps = 0;
        const int st = coop_march<FULL, COUNT>(p, T, w, StepSize, lane, cnt, P, csteps);
        if (lane == L || ray >= 0) {
            marching = false; vol = st > 0; Tv = w.MinT;
            if (st > 0) Pv = P;
            it += csteps;
        }
    }
    }
    /* the tile's cost: its longest march */
    for (int o = 32; o > 0; o >>= 1) it = max(it, __shfl_xor(it, o));
    steps = it;
    if (active) {
        float Tl = 0.0f; f3 Pl, Nl, Lel = mk(0.0f, 0.0f, 0.0f);
        const int il = intersect_lights(p, r, true, Tl, Pl, Nl, Lel);
        int type = -1;
        float Tn = 3.402823466e+38f;
        if (vol && Tv < Tn) { type = 0; Tn = Tv; }
        if (il >= 0 && Tl < Tn) { type = 1; Tn = Tl; }
        if (p.object_count > 0) {
            float To = 0.0f, Uo = 0.0f, Vo = 0.0f; f3 Po, No;
            const int io = intersect_objects(p, r, To, Po, No, Uo, Vo);
            if (io >= 0 && To < Tn) { type = 0; Tn = To; obj = io; Pv = Po; Nobj = No; Uobj = Uo; Vobj = Vo; }
        }
        if (type == 0) {
            scatter = owner;
            if (COUNT && owner) cnt.scatter++;
        } else if (owner) {
            f3 Lv = mk(0.0f, 0.0f, 0.0f);
            if (type == 1) Lv = add(Lv, Lel);
            frame[i] = make_float4(Lv.x, Lv.y, Lv.z, type >= 0 ? 1.0f : 0.0f);
            seeds1[i] = rng.s0;
            seeds2[i] = rng.s1;
        }
    }
    /* wave-aggregated append to the scatter queue */
    const unsigned long long mask = __ballot(scatter);
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(qcount, (unsigned int)__popcll(mask));
        base = __shfl(base, leader);
        if (scatter) {
            const unsigned int slot = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
            ScatterRec rec;
            rec.pixel = (uint32_t)i; rec.s0 = rng.s0; rec.s1 = rng.s1; rec.obj = obj;
            rec.Px = Pv.x; rec.Py = Pv.y; rec.Pz = Pv.z;
            rec.Wox = -D.x; rec.Woy = -D.y; rec.Woz = -D.z;
            rec.comp = smp.comp; rec.dir0 = smp.dir0; rec.dir1 = smp.dir1;
            rec.uvw0 = smp.uvw0; rec.uvw1 = smp.uvw1; rec.uvw2 = smp.uvw2;
            rec.lightnum = smp.lightnum;
            rec.Nx = Nobj.x; rec.Ny = Nobj.y; rec.Nz = Nobj.z;
            rec.U = Uobj; rec.V = Vobj; rec.pad0 = 0.0f; rec.pad1 = 0.0f;
            queue[slot] = rec;
        }
    }
}

/* One wave shades queue entries q0 .. q0 + 63 (those below n): NormalizedGradient at the
 * event (raymarching.h:70) and UniformSampleOneLight (transport.h:115-156). SPLIT: the shadow
 * marches are appended to the shadow-job queue (*jcount) and the event waits in an EventRec;
 * otherwise (Woodcock, whose draw count per march is not known in advance) the marches run
 * inline. Every lane of the wave must call this (the append is wave-wide). */
template<int FULL, int COUNT, int SPLIT>
__device__ __forceinline__ void shade_chunk(const DevParams& p, const Tables& T, unsigned int q0, unsigned int n, int lane,
                                            float4* __restrict__ frame, uint32_t* __restrict__ seeds1,
                                            uint32_t* __restrict__ seeds2, const ScatterRec* __restrict__ queue,
                                            EventRec* __restrict__ events, ShadowJob* __restrict__ jobq,
                                            unsigned int* __restrict__ jcount, Cnt& cnt)
{
    const unsigned int q = q0 + lane;
    const bool active = q < n;
    bool has0 = false, has1 = false;
    ShadowJob job0, job1;
    if (active) {
        const ScatterRec rec = queue[q];
        Rng rng;
        rng.s0 = rec.s0;
        rng.s1 = rec.s1;
        Sample smp;
        smp.comp = rec.comp; smp.dir0 = rec.dir0; smp.dir1 = rec.dir1;
        smp.uvw0 = rec.uvw0; smp.uvw1 = rec.uvw1; smp.uvw2 = rec.uvw2;
        smp.lightnum = rec.lightnum;
        const f3 Pv = mk(rec.Px, rec.Py, rec.Pz);
        const f3 Wo = mk(rec.Wox, rec.Woy, rec.Woz);
        /* SE.N: the normalised gradient of a volume event (raymarching.h:70), the object's
         * transformed shape normal of an object event (objects.h:36) */
        /* a volume event with central differences keeps its six gradient samples for the
         * gradient magnitude of the shading-type choice (the same positions) */
        CD6 cd;
        cd.valid = rec.obj < 0 && p.gradient_mode == ER_GRADIENT_CENTRAL;
        const float Iv = intensity(p, Pv);   /* the emission's and shader's intensity, loaded with the gradient */
        f3 Nv;
        if (rec.obj >= 0) {
            Nv = mk(rec.Nx, rec.Ny, rec.Nz);
        } else if (cd.valid) {
            cd_samples(p, Pv, cd.g);
            Nv = nrm(mk(cd.g[1] - cd.g[0], cd.g[3] - cd.g[2], cd.g[5] - cd.g[4]));
        } else {
            Nv = normalized_gradient(p, Pv);
        }
        if (SPLIT) {
            EventRec ev;
            ev.pixel = rec.pixel;
            ev.c1x = ev.c1y = ev.c1z = ev.c2x = ev.c2y = ev.c2z = 0.0f;
            ev.pad = ev.pad2 = 0.0f;
            shade_split<FULL, COUNT>(p, smp, Pv, Nv, Wo, rng, cnt, ev, job0, job1, has0, has1, rec.obj, rec.U, rec.V,
                                     Iv, cd);
            events[q] = ev;
        } else {
            f3 Lv = mk(0.0f, 0.0f, 0.0f);
            Lv = add(Lv, uniform_sample_one_light<FULL, COUNT>(p, T, smp, Pv, Nv, Wo, rng, cnt, rec.obj, rec.U, rec.V));
            frame[rec.pixel] = make_float4(Lv.x, Lv.y, Lv.z, 1.0f);
        }
        seeds1[rec.pixel] = rng.s0;
        seeds2[rec.pixel] = rng.s1;
    }
    if (SPLIT) {
        /* wave-aggregated append of 0..2 shadow jobs per lane */
        const unsigned long long m1 = __ballot(has0), m2 = __ballot(has1);
        const unsigned int total = (unsigned int)(__popcll(m1) + __popcll(m2));
        if (total) {
            const int leader = __ffsll((long long)(m1 | m2)) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(jcount, total);
            base = __shfl(base, leader);
            const unsigned long long below = (1ull << lane) - 1ull;
            const unsigned int slot = base + (unsigned int)(__popcll(m1 & below) + __popcll(m2 & below));
            if (has0) { job0.event = q; job0.pad = 0; jobq[slot] = job0; }
            if (has1) { job1.event = q; job1.pad = 0; jobq[slot + (has0 ? 1u : 0u)] = job1; }
        }
    }
}

/* The event's radiance from its contributions and their visibility (the tail of
 * UniformSampleOneLight, transport.h:152-155, and SingleScattering, singlescattering.h:96-99). */
__device__ __forceinline__ void finish_item(const EventRec& ev, unsigned int q, const unsigned char* __restrict__ vis,
                                            float4* __restrict__ frame)
{
    const uint32_t s1 = ev.state & 3u, s2 = (ev.state >> 2) & 3u;
    const bool use1 = s1 == 1u || (s1 == 2u && vis[2 * q] != 0);
    const bool use2 = s2 == 1u || (s2 == 2u && vis[2 * q + 1] != 0);
    f3 Ld = mk(ev.Ld0x, ev.Ld0y, ev.Ld0z);
    if (ev.count > 0.0f) {
        f3 Ed = mk(0.0f, 0.0f, 0.0f);
        if (use1) Ed = add(Ed, mk(ev.c1x, ev.c1y, ev.c1z));
        if (use2) Ed = add(Ed, mk(ev.c2x, ev.c2y, ev.c2z));
        Ld = scl(ev.count, add(Ld, Ed));
    }
    f3 Lv = mk(0.0f, 0.0f, 0.0f);
    Lv = add(Lv, Ld);
    frame[ev.pixel] = make_float4(Lv.x, Lv.y, Lv.z, 1.0f);
}

/* Primary kernel: persistent waves take 64-pixel tiles (8 x 8 by default, DevParams::tile_shift)
 * from an atomic counter (counts[2]), so the cost of a tile - from a few samples for background
 * pixels to thousands behind low-density tissue - is balanced over the whole device. With an
 * order (k_tile_order, from the previous pass) the tiles are claimed most expensive first, and
 * each tile's cost is recorded for the next order. LPR, the lanes that march one ray (64 / LPR
 * rays per wave): 1, one lane per ray (the heavy-tile claims aside); 4 or 16, each claim is a
 * sub-tile of 64 / LPR pixels marched by groups of LPR lanes (group_march); 64, each claim is
 * one ray marched by the whole wave (coop_march). With LPR > 1 the heavy-tile claims are not
 * used and a tile's cost is its first claim's. */
#ifndef ER_PRIMARY_WAVES_PER_EU
#define ER_PRIMARY_WAVES_PER_EU 1
#endif
/* lanes per primary workgroup (the waves of a workgroup share one LDS copy of the tables) */
#ifndef ER_PRIMARY_WG
#define ER_PRIMARY_WG 256
#endif
template<int FULL, int COUNT, int LUT_LDS, int LPR>
__global__ __launch_bounds__(ER_PRIMARY_WG) __attribute__((amdgpu_waves_per_eu(ER_PRIMARY_WAVES_PER_EU))) void k_primary(const DevParams* __restrict__ pp, float4* __restrict__ frame,
                                                uint32_t* __restrict__ seeds1, uint32_t* __restrict__ seeds2,
                                                ScatterRec* __restrict__ queue, unsigned int* __restrict__ counts,
                                                unsigned long long* __restrict__ counters, int tiles_x, int ntiles,
                                                const uint32_t* __restrict__ order, uint32_t* __restrict__ cost,
                                                int heavy_lanes)
{
    extern __shared__ uint32_t lds[];
    const DevParams& p = *pp;
    const Tables T = stage_tables<LUT_LDS>(p, lds);
    const int lane = threadIdx.x & 63;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int paths = 0;
    /* claims [0, HL hc): the hc heaviest tiles (order[0 .. hc)) in HL claims each, HL =
     * heavy_lanes lanes per ray (64: every ray of the tile marched by a whole wave; 4 or 16:
     * sub-tiles of 64 / HL rays marched by groups of HL lanes); then the other tiles in their
     * order */
    const unsigned int HL = heavy_lanes == 4 || heavy_lanes == 16 ? (unsigned int)heavy_lanes : 64u;
    unsigned int hc = order ? min((unsigned int)__builtin_amdgcn_readfirstlane(order[ntiles]), (unsigned int)ntiles) : 0u;
    if ((FULL && p.sampler == ER_SAMPLER_WOODCOCK) || LPR > 1) hc = 0u;
    const unsigned int nclaims = LPR > 1 ? (unsigned int)ntiles * (unsigned int)LPR : hc * HL + ((unsigned int)ntiles - hc);
    for (;;) {
        unsigned int claim = 0;
        if (lane == 0) claim = atomicAdd(&counts[2], 1u);
        claim = __builtin_amdgcn_readfirstlane(claim);
        if (claim >= nclaims) break;
        if constexpr (LPR > 1) {
            const unsigned int idx = claim / (unsigned int)LPR, sub = claim % (unsigned int)LPR;
            const unsigned int tile = order ? __builtin_amdgcn_readfirstlane(order[idx]) : idx;
            int steps = 0;
            primary_tile<FULL, COUNT, (LPR < 64 ? LPR : 1)>(p, T, tile, tiles_x, lane, frame, seeds1, seeds2, queue,
                                                           &counts[0], cnt, paths, steps, (int)sub);
            if (cost && lane == 0 && sub == 0u) cost[tile] = (uint32_t)steps;
            continue;
        }
        const bool heavy = claim < hc * HL;
        const unsigned int idx = heavy ? claim / HL : hc + (claim - hc * HL);
        const unsigned int tile = order ? __builtin_amdgcn_readfirstlane(order[idx]) : idx;
        int steps = 0;
        if (heavy && HL == 4u)
            primary_tile<FULL, COUNT, 4>(p, T, tile, tiles_x, lane, frame, seeds1, seeds2, queue, &counts[0], cnt, paths,
                                         steps, (int)(claim % HL));
        else if (heavy && HL == 16u)
            primary_tile<FULL, COUNT, 16>(p, T, tile, tiles_x, lane, frame, seeds1, seeds2, queue, &counts[0], cnt, paths,
                                          steps, (int)(claim % HL));
        else
            primary_tile<FULL, COUNT>(p, T, tile, tiles_x, lane, frame, seeds1, seeds2, queue, &counts[0], cnt, paths, steps,
                                      heavy ? (int)(claim % HL) : -1);
        if (cost && lane == 0) {
            if (heavy) atomicMax(&cost[tile], (uint32_t)steps);
            else cost[tile] = (uint32_t)steps;
        }
    }
    flush_counters<COUNT>(cnt, true, paths, counters);
}

/* The claim order of the next pass's tiles: most expensive first by this pass's cost per tile
 * (the march iterations of its longest ray; a progressive render's tiles cost about the same
 * from pass to pass), so that the slowest tiles start at once instead of deciding when the
 * kernel ends. Eight classes by cost relative to the maximum (above 1/2, 1/4, ..., 1/128, the
 * rest), stable within a class (scan order keeps neighbouring tiles together). order[n] = hc:
 * the tiles of the first class (at most heavy_cap and n / 16), whose rays the next pass
 * marches one wave per ray; their costs are cleared for its per-ray maxima. One workgroup;
 * order[0 .. n) is a permutation of 0 .. n-1 whatever the costs hold. */
__global__ __launch_bounds__(1024) void k_tile_order(uint32_t* __restrict__ cost, uint32_t* __restrict__ order, int n,
                                                    int heavy_cap)
{
    __shared__ uint32_t smax, base[8], wcount[16][8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) smax = 0u;
    if (tid < 8) base[tid] = 0u;
    __syncthreads();
    uint32_t mx = 0u;
    for (int i = tid; i < n; i += blockDim.x) mx = max(mx, cost[i]);
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
    if (lane == 0) atomicMax(&smax, mx);
    __syncthreads();
    const uint64_t cmax = smax;
    auto cls = [cmax](uint32_t c) {
        int b = 0;
        while (b < 7 && ((uint64_t)c << (b + 1)) <= cmax) ++b;
        return b;
    };
    /* class counts: one ballot per class and wave instead of an LDS atomic per tile */
    for (int chunk = 0; chunk < n; chunk += blockDim.x) {
        const int i = chunk + tid;
        const int b = i < n ? cls(cost[i]) : 8;
        for (int bb = 0; bb < 8; ++bb) {
            const unsigned long long m = __ballot(b == bb);
            if (lane == 0 && m) atomicAdd(&base[bb], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    __shared__ uint32_t heavy;
    if (tid == 0) {
        heavy = min(base[0], min((uint32_t)max(heavy_cap, 0), (uint32_t)n / 16u));
        uint32_t acc = 0u;
        for (int b = 0; b < 8; ++b) { const uint32_t c = base[b]; base[b] = acc; acc += c; }
    }
    __syncthreads();
    for (int chunk = 0; chunk < n; chunk += blockDim.x) {
        const int i = chunk + tid;
        const int b = i < n ? cls(cost[i]) : 8;
        uint32_t rank = 0u;
        for (int bb = 0; bb < 8; ++bb) {
            const unsigned long long m = __ballot(b == bb);
            if (lane == 0) wcount[wave][bb] = (uint32_t)__popcll(m);
            if (b == bb) rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        }
        __syncthreads();
        if (b < 8) {
            uint32_t off = base[b] + rank;
            for (int w = 0; w < wave; ++w) off += wcount[w][b];
            if (off < (uint32_t)n) order[off] = (uint32_t)i;
        }
        __syncthreads();
        if (tid < 8) {
            uint32_t total = 0u;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += wcount[w][tid];
            base[tid] += total;
        }
        __syncthreads();
    }
    for (uint32_t k = tid; k < heavy; k += blockDim.x) cost[order[k]] = 0u;
    if (tid == 0) order[n] = heavy;
}

#ifndef ER_SHADE_WAVES_PER_EU
#define ER_SHADE_WAVES_PER_EU 1
#endif

/* Shading kernel over the scatter queue (counts[0] entries), wave-aligned grid stride. */
template<int FULL, int COUNT, int LUT_LDS, int SPLIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ER_SHADE_WAVES_PER_EU))) void k_shade(const DevParams* __restrict__ pp, float4* __restrict__ frame,
                                              uint32_t* __restrict__ seeds1, uint32_t* __restrict__ seeds2,
                                              const ScatterRec* __restrict__ queue, unsigned int* __restrict__ counts,
                                              EventRec* __restrict__ events, ShadowJob* __restrict__ jobq,
                                              unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const DevParams& p = *pp;
    const Tables T = stage_tables<SPLIT ? 0 : LUT_LDS>(p, lds);
    const unsigned int n = counts[0];
    const int lane = threadIdx.x & 63;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const unsigned int stride = gridDim.x * blockDim.x;
    const unsigned int base_q = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (unsigned int q0 = base_q; q0 < n; q0 += stride)
        shade_chunk<FULL, COUNT, SPLIT>(p, T, q0, n, lane, frame, seeds1, seeds2, queue, events, jobq, &counts[1], cnt);
    flush_counters<COUNT>(cnt, true, 0, counters);
}

/* Once a wave's job queue is drained and at most 64 / G of its lanes still march, those
 * marches are handed to aligned groups of G lanes (group i takes the march of the i-th busy
 * lane: its state is the same in every lane of the group) and stepped with group_step, which
 * makes the same decisions as the one-lane march, until every group is done; the group's first
 * lane writes the visibility byte. The last, longest marches of a wave then advance G / K
 * times as many samples per round instead of leaving the wave's other lanes idle. Every lane of
 * the wave must call this. */
#ifndef ER_SHADOW_DRAIN_G
#define ER_SHADOW_DRAIN_G 4   /* 0: no hand-off (the lanes finish their marches alone) */
#endif
template<int FULL, int COUNT, int G>
__device__ __forceinline__ void shadow_drain(const DevParams& p, const Tables& T, const MarchState& m, unsigned int slot,
                                             bool busy, unsigned char* __restrict__ vis, float StepSize, int lane, Cnt& cnt)
{
    const unsigned long long bm = __ballot(busy);
    /* the group's march: the (lane / G)-th busy lane's, if there is one */
    unsigned long long x = bm;
    for (int k = 0; k < lane / G; ++k) x &= x - 1ull;
    const bool has = x != 0ull;
    const int src = has ? __ffsll((long long)x) - 1 : 0;
    MarchState g;
    g.O = mk(__shfl(m.O.x, src), __shfl(m.O.y, src), __shfl(m.O.z, src));
    g.D = mk(__shfl(m.D.x, src), __shfl(m.D.y, src), __shfl(m.D.z, src));
    g.MinT = __shfl(m.MinT, src);
    g.MaxT = __shfl(m.MaxT, src);
    g.Sum = __shfl(m.Sum, src);
    g.S = __shfl(m.S, src);
    const unsigned int gslot = (unsigned int)__shfl((int)slot, src);
    const bool first = (lane & (G - 1)) == 0;
    bool gbusy = has;
    while (__ballot(gbusy)) {
        if (gbusy) {
            f3 P;
            int iters = 0;
            const int st = group_step<FULL, COUNT, G, false>(p, T, g, StepSize, lane, cnt, P, iters);
            if (st != 0) {
                if (first) vis[gslot] = st > 0 ? 0 : 1;
                gbusy = false;
            }
        }
    }
}

/* shadow_marches for Woodcock tracking: each lane's walk advances one majorant segment per
 * loop iteration (wc_step) instead of running whole when its job starts, so that idle lanes
 * refill while the others walk. */
template<int FULL, int COUNT>
__device__ __forceinline__ void shadow_marches_wc(const DevParams& p, const Tables& T, const ShadowJob* __restrict__ jobq,
                                                  unsigned char* __restrict__ vis, unsigned int* claim, unsigned int n,
                                                  unsigned int j, unsigned int claim_base, bool drained, int lane, Cnt& cnt)
{
    WcWalk w;
    unsigned int slot = 0;
    bool busy = false;
    for (;;) {
        if (j < n) {
            const ShadowJob job = jobq[j];
            slot = 2u * job.event + job.slot;
            Ray r;
            r.O = mk(job.Ox, job.Oy, job.Oz);
            r.D = mk(job.Dx, job.Dy, job.Dz);
            r.MinT = 0.0f;
            r.MaxT = job.MaxT;
            Rng rng;
            rng.s0 = job.s0;
            rng.s1 = job.s1;
            float nearT, farT;
            if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) {
                vis[slot] = 1;
            } else {
                wc_begin(p, r, nearT, farT, rng, w);
                busy = true;
            }
            j = n;
        }
        if (busy) {
            const int st = wc_step<FULL, COUNT>(p, T, w, false, cnt);
            if (st != 0) {
                vis[slot] = st > 0 ? 0 : 1;
                busy = false;
            }
        }
        const unsigned long long idle = __ballot(!busy);
        const int nidle = __popcll(idle);
        if (drained) {
            if (nidle == 64) break;
        } else if (nidle >= ER_SHADOW_REFILL) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(claim, (unsigned int)nidle);
            base = claim_base + __shfl(base, leader);
            if (base + (unsigned int)nidle >= n) drained = true;
            if (!busy) j = base + (unsigned int)__popcll(idle & ((1ull << lane) - 1ull));
        }
    }
}

/* The shadow marches of one queue of n jobs (jobq), one march per lane. The marches differ in
 * length by orders of magnitude (a ray through empty space ends after a few skips, one through
 * tissue takes hundreds of samples), so a wave does not wait for its longest march: the lane
 * starts job j (none if j >= n), and whenever at least ER_SHADOW_REFILL lanes are idle they
 * claim the next jobs from *claim (job claim_base + the counter's value) with one atomic and
 * start them, and the wave steps all its live marches one iteration at a time (Woodcock
 * walks one majorant segment at a time, in shadow_marches_wc). `drained`:
 * nothing is left to claim. The job order does not matter: each march is independent and
 * writes only its own visibility byte. Every lane of the wave must call this. */
template<int FULL, int COUNT>
__device__ __forceinline__ void shadow_marches(const DevParams& p, const Tables& T, const ShadowJob* __restrict__ jobq,
                                               unsigned char* __restrict__ vis, unsigned int* claim, unsigned int n,
                                               unsigned int j, unsigned int claim_base, bool drained, int lane, Cnt& cnt)
{
    if (FULL && p.sampler == ER_SAMPLER_WOODCOCK) {
        shadow_marches_wc<FULL, COUNT>(p, T, jobq, vis, claim, n, j, claim_base, drained, lane, cnt);
        return;
    }
    const float StepSize = p.step_shadow * p.minstep;
    NoVisit nv;
    MarchState m;
    unsigned int slot = 0;
    bool busy = false;
    for (;;) {
        if (j < n) {
            const ShadowJob job = jobq[j];
            slot = 2u * job.event + job.slot;
            Ray r;
            r.O = mk(job.Ox, job.Oy, job.Oz);
            r.D = mk(job.Dx, job.Dy, job.Dz);
            r.MinT = 0.0f;
            r.MaxT = job.MaxT;
            Rng rng;
            rng.s0 = job.s0;
            rng.s1 = job.s1;
            float nearT, farT;
            if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) {
                vis[slot] = 1;
            } else if (!march_begin(p, r, nearT, farT, StepSize, rng, -1.0f, m)) {
                vis[slot] = 1;
            } else if (!(m.Sum < m.S)) {
                vis[slot] = 0;
            } else {
                busy = true;
            }
            j = n;
        }
        if (busy) {
            f3 P;
            const int st = march_step<FULL, COUNT, ER_BATCH_SHADOW, NoVisit>(p, T, m, false, StepSize, cnt, nv, P);
            if (st != 0) {
                vis[slot] = st > 0 ? 0 : 1;
                busy = false;
            }
        }
        const unsigned long long idle = __ballot(!busy);
        const int nidle = __popcll(idle);
        if (drained) {
            if (nidle == 64) break;
            if (ER_SHADOW_DRAIN_G > 0 && 64 - nidle <= 64 / (ER_SHADOW_DRAIN_G > 0 ? ER_SHADOW_DRAIN_G : 1)) {
                shadow_drain<FULL, COUNT, (ER_SHADOW_DRAIN_G > 0 ? ER_SHADOW_DRAIN_G : 4)>(p, T, m, slot, busy, vis, StepSize,
                                                                                           lane, cnt);
                break;
            }
        } else if (nidle >= ER_SHADOW_REFILL) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(claim, (unsigned int)nidle);
            base = claim_base + __shfl(base, leader);
            if (base + (unsigned int)nidle >= n) drained = true;
            if (!busy) j = base + (unsigned int)__popcll(idle & ((1ull << lane) - 1ull));
        }
    }
}

/* shadow_marches with an aligned group of G lanes (G = 4 or 16) per march, stepped one
 * group_step round at a time: the group starts job j (the same in every lane of the group),
 * and whenever at least ER_SHADOW_REFILL lanes belong to idle groups those groups claim the next
 * jobs together. Woodcock tracking is not marched this way (its shadow jobs take one lane each).
 * The group's first lane writes the visibility byte. Every lane of the wave must call this. */
template<int FULL, int COUNT, int G>
__device__ __forceinline__ void shadow_marches_g(const DevParams& p, const Tables& T, const ShadowJob* __restrict__ jobq,
                                                 unsigned char* __restrict__ vis, unsigned int* claim, unsigned int n,
                                                 unsigned int j, unsigned int claim_base, bool drained, int lane, Cnt& cnt)
{
    const float StepSize = p.step_shadow * p.minstep;
    const bool first = (lane & (G - 1)) == 0;
    /* the first lane of every group */
    const unsigned long long leaders = G == 4 ? 0x1111111111111111ull : 0x0001000100010001ull;
    MarchState m;
    unsigned int slot = 0;
    bool busy = false;
    for (;;) {
        if (j < n) {
            const ShadowJob job = jobq[j];
            slot = 2u * job.event + job.slot;
            Ray r;
            r.O = mk(job.Ox, job.Oy, job.Oz);
            r.D = mk(job.Dx, job.Dy, job.Dz);
            r.MinT = 0.0f;
            r.MaxT = job.MaxT;
            Rng rng;
            rng.s0 = job.s0;
            rng.s1 = job.s1;
            float nearT, farT;
            if (!intersect_box(r, ld3(p.minp), ld3(p.maxp), nearT, farT)) {
                if (first) vis[slot] = 1;
            } else if (!march_begin(p, r, nearT, farT, StepSize, rng, -1.0f, m)) {
                if (first) vis[slot] = 1;
            } else if (!(m.Sum < m.S)) {
                if (first) vis[slot] = 0;
            } else {
                busy = true;
            }
            j = n;
        }
        if (busy) {
            f3 P;
            int iters = 0;
            const int st = group_step<FULL, COUNT, G, false>(p, T, m, StepSize, lane, cnt, P, iters);
            if (st != 0) {
                if (first) vis[slot] = st > 0 ? 0 : 1;
                busy = false;
            }
        }
        const unsigned long long idle = __ballot(!busy) & leaders;
        const int nidle = __popcll(idle);
        if (drained) {
            if (nidle == 64 / G) break;
        } else if (nidle * G >= ER_SHADOW_REFILL) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(claim, (unsigned int)nidle);
            base = claim_base + __shfl(base, leader);
            if (base + (unsigned int)nidle >= n) drained = true;
            if (!busy) j = base + (unsigned int)__popcll(idle & ((1ull << (lane & ~(G - 1))) - 1ull));
        }
    }
}

/* Shadow-march kernel over the shadow-job queue (counts[1] entries). The grid is persistent
 * (its size, in waves per CU, is a launch parameter: fewer waves leave the CUs to the primary
 * kernel of the next pass, which runs beside it); with LPR lanes per march (1, 4 or 16) wave w
 * first takes jobs [R w, R w + R), R = 64 / LPR, and idle lanes or groups then claim the jobs
 * beyond the static chunks from counts[3]. */
template<int FULL, int COUNT, int LUT_LDS, int LPR>
__global__ __launch_bounds__(256) void k_shadow(const DevParams* __restrict__ pp, unsigned int* __restrict__ counts,
                                               const ShadowJob* __restrict__ jobq, unsigned char* __restrict__ vis,
                                               unsigned long long* __restrict__ counters)
{
    extern __shared__ uint32_t lds[];
    const DevParams& p = *pp;
    const unsigned int n = counts[1];
    constexpr unsigned int R = 64u / LPR;
    const unsigned int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned int static_end = ((gridDim.x * blockDim.x) >> 6) * R;
    if ((blockIdx.x * blockDim.x >> 6) * R >= n) return;   /* no static chunk in the workgroup (workgroup-uniform) */
    const Tables T = stage_tables<LUT_LDS>(p, lds);
    const int lane = threadIdx.x & 63;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (LPR > 1)
        shadow_marches_g<FULL, COUNT, LPR>(p, T, jobq, vis, &counts[3], n, wave * R + (unsigned int)lane / LPR, static_end,
                                           static_end >= n, lane, cnt);
    else
        shadow_marches<FULL, COUNT>(p, T, jobq, vis, &counts[3], n, wave * 64u + (unsigned int)lane, static_end,
                                    static_end >= n, lane, cnt);
    flush_counters<COUNT>(cnt, true, 0, counters);
}

__global__ __launch_bounds__(256) void k_finish(const unsigned int* __restrict__ counts, const EventRec* __restrict__ events,
                                               const unsigned char* __restrict__ vis, float4* __restrict__ frame)
{
    const unsigned int n = counts[0];
    for (unsigned int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        finish_item(events[q], q, vis, frame);
}

/* ComputeAutoFocusDistance (autofocus.cuh:26-63), one lane: 100 rays through film point
 * (u, v), each jittered by 0.01 x a concentric-disk sample of the screen (Get2 in clang's
 * order), marched with SampleVolume on one RNG stream; the mean distance of the rays' volume
 * events, or the distance to the last event point (origin) when none. */
template<int FULL>
__device__ void autofocus(const DevParams& p, int u, int v, uint32_t s0, uint32_t s1, float* out)
{
    Tables T;
    T.bits = nullptr; T.bits_glb = gmem(p.mc_bits); T.lut_lds = nullptr; T.lut_glb = gmem(p.sigma_lut);
    Rng rng;
    rng.s0 = s0;
    rng.s1 = s1;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const f3 O = ld3(p.cam_pos), N = ld3(p.cam_N), U = ld3(p.cam_U), V = ld3(p.cam_V);
    float sum = 0.0f, w = 0.0f;
    f3 lastP = mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 100; ++i) {
        float sx = p.screen00 + (p.invscreen0 * (float)u);
        float sy = p.screen10 + (p.invscreen1 * (float)v);
        const float d0 = rng.u(), d1 = rng.u();
        float dx, dy;
        concentric_disk(d0, d1, dx, dy);
        sx += 0.01f * dx;
        sy += 0.01f * dy;
        Ray r;
        r.O = O;
        r.D = nrm(sub(add(N, scl(sx, U)), scl(sy, V)));
        r.MinT = p.clip_near;
        r.MaxT = p.clip_far;
        float Tv; f3 Pv;
        if (march<FULL, 0, 1>(p, T, r, true, rng, Tv, Pv, cnt)) {
            sum += len(sub(Pv, O));
            w += 1.0f;
            lastP = Pv;
        }
    }
    *out = sum <= 0.0f ? len(sub(lastP, O)) : sum / w;
}

__global__ __launch_bounds__(64) void k_autofocus(const DevParams* __restrict__ pp, int u, int v, uint32_t s0, uint32_t s1,
                                                 float* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const DevParams& p = *pp;
    if (p.estimator == ER_ESTIMATOR_FULL) autofocus<1>(p, u, v, s0, s1, out);
    else autofocus<0>(p, u, v, s0, s1, out);
}

/* ---- batches: the same pass over many tracers per launch (er_render_batch) ------------ */

/* tables of a batch job: occupancy bits and opacity table read from global memory (the
 * jobs of one workgroup change from tile to tile) */
__device__ __forceinline__ Tables global_tables(const DevParams& p)
{
    Tables T;
    T.bits = nullptr;
    T.bits_glb = gmem(p.mc_bits);
    T.lut_lds = nullptr;
    T.lut_glb = gmem(p.sigma_lut);
    return T;
}

/* The parameter blocks of a batch's jobs, in job-table order. Read from the constant address
 * space, their fields are fetched with scalar loads, as the single-pass kernels fetch their
 * kernel-argument block; through the job table's pointers the compiler would have to use
 * vector loads (it cannot rule out that the kernel's own stores alias the block), which took
 * 40 more VGPRs in the primary kernel. er_render_batch runs larger batches in chunks. */
__constant__ DevParams c_batch_params[ER_BATCH_CONST_JOBS];

/* A job's tables for a batch kernel: LDS = 1 copies the macrocell bits and the opacity table
 * into the workgroup's LDS (the host launches this form only when every job's tables fit),
 * LDS = 0 reads both from global memory. Called by the whole workgroup after every wave has
 * finished with the previous job's tables. */
template<int LDS>
__device__ __forceinline__ Tables stage_job_tables(const DevParams& p, uint32_t* lds)
{
    if (!LDS) return global_tables(p);
    const int nw = p.mc_words;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = gmem(p.mc_bits)[k];
    float* lut = (float*)(lds + nw);
    for (int k = threadIdx.x; k <= p.lut_max; k += blockDim.x) lut[k] = gmem(p.sigma_lut)[k];
    __syncthreads();
    Tables T;
    T.bits = lds;
    T.bits_glb = nullptr;
    T.lut_lds = lut;
    T.lut_glb = nullptr;
    return T;
}

/* The next job after j (cyclically) whose work counter counts[which] is still below its item
 * count (ntiles x claims_per_tile for tiles, counts[1] for shadow jobs); -1 when every item is
 * claimed. Counters only grow, so a stale read can only send a workgroup to a job that has just
 * run dry. */
__device__ __forceinline__ int next_job_with_work(const BatchJob* __restrict__ J, int na, int j, int which,
                                                  unsigned int claims_per_tile = 1u)
{
    for (int k = 1; k <= na; ++k) {
        const int c = (j + k) % na;
        const unsigned int claimed = __hip_atomic_load(&J[c].counts[which], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int items = which == 2 ? (unsigned int)J[c].ntiles * claims_per_tile
                                              : __hip_atomic_load(&J[c].counts[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (claimed < items) return c;
    }
    return -1;
}

/* Primary kernel of a batch round. Workgroups go job by job, so that the job's tables can sit
 * in LDS as in a single pass: a workgroup starts at the job that holds its proportional share
 * of the round's tiles, its waves take that job's 64-pixel tiles from the job's own counter
 * (counts[2]), and when the job has no tiles left the workgroup moves on to the next job that
 * has (a counter only grows, so a workgroup leaves only when every tile is claimed). LPR: the
 * lanes that march one ray, as in k_primary (with LPR > 1 a claim is a sub-tile). */
template<int FULL, int LDS, int LPR>
__global__ __launch_bounds__(256) void k_primary_b(const BatchJob* __restrict__ J, int na)
{
    extern __shared__ uint32_t lds[];
    __shared__ int next_job;
    const int lane = threadIdx.x & 63;
    if (na == 0) return;
    const unsigned int total = (unsigned int)(J[na - 1].tile_base + J[na - 1].ntiles);
    const unsigned int pos = (unsigned int)(((unsigned long long)blockIdx.x * total) / gridDim.x);
    int j = 0;
    while (j < na - 1 && pos >= (unsigned int)(J[j].tile_base + J[j].ntiles)) ++j;
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int paths = 0;
    for (;;) {
        const BatchJob& B = J[j];
        const DevParams& p = c_batch_params[j];
        const Tables T = stage_job_tables<LDS>(p, lds);
        for (;;) {
            unsigned int claim = 0;
            if (lane == 0) claim = atomicAdd(&B.counts[2], 1u);
            claim = __builtin_amdgcn_readfirstlane(claim);
            if (claim >= (unsigned int)B.ntiles * (unsigned int)LPR) break;
            int steps = 0;
            if constexpr (LPR > 1)
                primary_tile<FULL, 0, (LPR < 64 ? LPR : 1)>(p, T, claim / (unsigned int)LPR, B.tiles_x, lane, B.frame,
                                                           B.seeds1, B.seeds2, B.queue, &B.counts[0], cnt, paths, steps,
                                                           (int)(claim % (unsigned int)LPR));
            else
                primary_tile<FULL, 0>(p, T, claim, B.tiles_x, lane, B.frame, B.seeds1, B.seeds2, B.queue, &B.counts[0], cnt,
                                      paths, steps, -1);
        }
        __syncthreads();   /* every wave is done with this job's tables */
        if (threadIdx.x == 0) next_job = next_job_with_work(J, na, j, 2, (unsigned int)LPR);
        __syncthreads();
        j = __builtin_amdgcn_readfirstlane(next_job);   /* uniform: the job's fields load as scalars */
        if (j < 0) break;
    }
}

/* The wave slots of a per-job work list of counts[which] items per job (64 items a slot), for
 * the na active jobs: wb[k] = the first slot of job k, wb[na] = the total. Built in LDS by the
 * first wave of the workgroup (a wave-wide scan per 64 jobs); the whole workgroup calls this. */
__device__ __forceinline__ void batch_wave_table(const BatchJob* __restrict__ J, int na, int which, unsigned int* wb)
{
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        unsigned int carry = 0;
        for (int c0 = 0; c0 < na; c0 += 64) {
            const int k = c0 + lane;
            unsigned int v = k < na ? (J[k].counts[which] + 63u) >> 6 : 0u;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned int y = __shfl_up(v, o);
                if (lane >= o) v += y;
            }
            if (k < na) wb[k + 1] = carry + v;
            carry += __shfl(v, 63);
        }
        if (lane == 0) wb[0] = 0u;
    }
    __syncthreads();
}

/* wave slot w of that work list: job index (the last job whose first slot is at most w, by
 * binary search in wb), the slot's first item and the job's item count; false when w is past
 * the last job */
__device__ __forceinline__ bool batch_slot(const BatchJob* __restrict__ J, const unsigned int* wb, int na, int which,
                                           unsigned int w, int& job, unsigned int& first, unsigned int& n)
{
    if (na == 0 || w >= wb[na]) return false;
    int lo = 0, hi = na - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wb[mid] <= w) lo = mid; else hi = mid - 1;
    }
    job = lo;
    first = (w - wb[lo]) << 6;
    n = J[lo].counts[which];
    return true;
}

template<int FULL>
__global__ __launch_bounds__(256) void k_shade_b(const BatchJob* __restrict__ J, int na)
{
    __shared__ unsigned int wb[ER_BATCH_CONST_JOBS + 1];
    batch_wave_table(J, na, 0, wb);
    const int lane = threadIdx.x & 63;
    const unsigned int waves = gridDim.x * (blockDim.x >> 6);
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (unsigned int w = blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);; w += waves) {
        int job; unsigned int q0, n;
        if (!batch_slot(J, wb, na, 0, w, job, q0, n)) break;
        job = __builtin_amdgcn_readfirstlane(job);   /* uniform (the counts are vector loads) */
        q0 = __builtin_amdgcn_readfirstlane(q0);
        n = __builtin_amdgcn_readfirstlane(n);
        const BatchJob& B = J[job];
        const DevParams& p = c_batch_params[job];
        const Tables T = global_tables(p);
        shade_chunk<FULL, 0, 1>(p, T, q0, n, lane, B.frame, B.seeds1, B.seeds2, B.queue, B.events, B.jobs, &B.counts[1], cnt);
    }
}

/* Shadow marches of a batch round, job by job like k_primary_b: a workgroup starts at the job
 * holding its proportional share of the round's shadow jobs, its waves' idle lanes claim that
 * job's shadow jobs (counts[3], shadow_marches, or shadow_marches_g with LPR = 4 or 16 lanes per
 * march), then it moves on to the next job with shadow jobs left. */
template<int FULL, int LDS, int LPR>
__global__ __launch_bounds__(256) void k_shadow_b(const BatchJob* __restrict__ J, int na)
{
    extern __shared__ uint32_t lds[];
    __shared__ int next_job;
    if (na == 0) return;
    const int lane = threadIdx.x & 63;
    unsigned int total = 0;
    for (int k = 0; k < na; ++k) total += J[k].counts[1];
    if (total == 0) return;
    const unsigned int pos = (unsigned int)(((unsigned long long)blockIdx.x * total) / gridDim.x);
    int j = 0;
    for (unsigned int base = 0; j < na - 1; ++j) {
        base += J[j].counts[1];
        if (pos < base) break;
    }
    j = __builtin_amdgcn_readfirstlane(j);   /* uniform (the counts are vector loads) */
    Cnt cnt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (;;) {
        const BatchJob& B = J[j];
        const DevParams& p = c_batch_params[j];
        const Tables T = stage_job_tables<LDS>(p, lds);
        const unsigned int n = __builtin_amdgcn_readfirstlane(B.counts[1]);
        if constexpr (LPR > 1)
            shadow_marches_g<FULL, 0, LPR>(p, T, B.jobs, B.vis, &B.counts[3], n, n, 0u, false, lane, cnt);
        else
            shadow_marches<FULL, 0>(p, T, B.jobs, B.vis, &B.counts[3], n, n, 0u, false, lane, cnt);
        __syncthreads();
        if (threadIdx.x == 0) next_job = next_job_with_work(J, na, j, 3);
        __syncthreads();
        j = __builtin_amdgcn_readfirstlane(next_job);   /* uniform: the job's fields load as scalars */
        if (j < 0) break;
    }
}

__global__ __launch_bounds__(256) void k_finish_b(const BatchJob* __restrict__ J, int na)
{
    __shared__ unsigned int wb[ER_BATCH_CONST_JOBS + 1];
    batch_wave_table(J, na, 0, wb);
    const int lane = threadIdx.x & 63;
    const unsigned int waves = gridDim.x * (blockDim.x >> 6);
    for (unsigned int w = blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);; w += waves) {
        int job; unsigned int q0, n;
        if (!batch_slot(J, wb, na, 0, w, job, q0, n)) break;
        const BatchJob& B = J[job];
        const unsigned int q = q0 + lane;
        if (q < n) finish_item(B.events[q], q, B.vis, B.frame);
    }
}

/* end of a round: empty the queues of this buffer set and its tile counter (one workgroup) */
__global__ __launch_bounds__(64) void k_round_end(const BatchJob* __restrict__ J, int njobs)
{
    for (int k = threadIdx.x; k < njobs; k += blockDim.x) {
        J[k].counts[0] = 0u;
        J[k].counts[1] = 0u;
        J[k].counts[2] = 0u;
        J[k].counts[3] = 0u;
    }
}

} /* namespace erd */

/* launchers used by the runtime */
extern "C" void erd_pass_record_sizes(size_t* scatter, size_t* event, size_t* job)
{
    *scatter = sizeof(erd::ScatterRec);
    *event = sizeof(erd::EventRec);
    *job = sizeof(erd::ShadowJob);
}

template<int FULL, int COUNT, int LUT_LDS>
static void launch_all(const erd::DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2, erd::ScatterRec* queue,
                       unsigned int* counts, erd::EventRec* events, erd::ShadowJob* jobs, unsigned char* vis,
                       unsigned long long* counters, int tiles_x, int ntiles, int nblocks, int shade_blocks,
                       int shadow_blocks, int split, size_t lds, size_t lds_bits, hipStream_t st, hipStream_t st2,
                       hipEvent_t shaded, hipEvent_t after_primary, const uint32_t* order_in, uint32_t* cost,
                       uint32_t* order_out, int heavy_cap, int lpr, int shadow_lpr, int heavy_lanes)
{
#define ER_PRIMARY(G) hipLaunchKernelGGL((erd::k_primary<FULL, COUNT, LUT_LDS, G>), dim3(nblocks), dim3(ER_PRIMARY_WG), lds, st, p, \
                                         frame, s1, s2, queue, counts, counters, tiles_x, ntiles, order_in, cost, heavy_lanes)
    if (lpr == 4) ER_PRIMARY(4);
    else if (lpr == 16) ER_PRIMARY(16);
    else if (lpr == 64) ER_PRIMARY(64);
    else ER_PRIMARY(1);
#undef ER_PRIMARY
    if (after_primary) (void)hipEventRecord(after_primary, st);
    if (split) {
        hipLaunchKernelGGL((erd::k_shade<FULL, COUNT, LUT_LDS, 1>), dim3(shade_blocks), dim3(256), lds_bits, st, p, frame, s1,
                           s2, (const erd::ScatterRec*)queue, counts, events, jobs, counters);
        (void)hipEventRecord(shaded, st);
        (void)hipStreamWaitEvent(st2, shaded, 0);
#define ER_SHADOW(G) hipLaunchKernelGGL((erd::k_shadow<FULL, COUNT, LUT_LDS, G>), dim3(shadow_blocks), dim3(256), lds, st2, \
                                        p, counts, (const erd::ShadowJob*)jobs, vis, counters)
        if (shadow_lpr == 4) ER_SHADOW(4);
        else if (shadow_lpr == 16) ER_SHADOW(16);
        else ER_SHADOW(1);
#undef ER_SHADOW
        hipLaunchKernelGGL(erd::k_finish, dim3(shade_blocks), dim3(256), 0, st2, (const unsigned int*)counts,
                           (const erd::EventRec*)events, (const unsigned char*)vis, frame);
    } else {
        hipLaunchKernelGGL((erd::k_shade<FULL, COUNT, LUT_LDS, 0>), dim3(shade_blocks), dim3(256), lds, st, p, frame, s1, s2,
                           (const erd::ScatterRec*)queue, counts, events, jobs, counters);
        (void)hipEventRecord(shaded, st);
        (void)hipStreamWaitEvent(st2, shaded, 0);
    }
    /* the next pass's tile order from this pass's tile costs (after the primary kernel has
     * read the current order; stream order keeps the next primary kernel behind it) */
    if (cost && order_out)
        hipLaunchKernelGGL(erd::k_tile_order, dim3(1), dim3(1024), 0, st, cost, order_out, ntiles, heavy_cap);
}

/* One pass's trace kernels: primary and shading on st; shadow marches and combine on st2
 * after the shading (event `shaded`). The caller runs post-processing on st2. */
extern "C" hipError_t erd_launch_trace(const erd::DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2, void* queue,
                                       unsigned int* counts, void* events, void* jobs, unsigned char* vis,
                                       unsigned long long* counters, int W, int H, int full, int count, int lut_lds,
                                       int split, size_t lds_bytes, size_t lds_bits, int shade_blocks,
                                       int primary_blocks, int shadow_blocks, int tile_shift, hipStream_t st,
                                       hipStream_t st2, hipEvent_t shaded, hipEvent_t after_primary,
                                       const uint32_t* order_in, uint32_t* cost, uint32_t* order_out, int heavy_cap,
                                       int lanes_per_ray, int shadow_lanes_per_ray, int heavy_lanes_per_ray)
{
    const int tw = 1 << tile_shift, th = 64 >> tile_shift;
    const int tiles_x = (W + tw - 1) / tw, tiles_y = (H + th - 1) / th;
    const int ntiles = tiles_x * tiles_y;
    const int lpr = (lanes_per_ray == 4 || lanes_per_ray == 16 || lanes_per_ray == 64) ? lanes_per_ray : 1;
    /* persistent workgroups: enough to fill every CU, never more than there are claims */
    const long long nclaims = (long long)ntiles * lpr;
    constexpr long long wpb = ER_PRIMARY_WG / 64;   /* waves per workgroup */
    const int nblocks = (int)std::max(1LL, std::min((nclaims + wpb - 1) / wpb, (long long)primary_blocks));
    erd::ScatterRec* q = (erd::ScatterRec*)queue;
    erd::EventRec* ev = (erd::EventRec*)events;
    erd::ShadowJob* jb = (erd::ShadowJob*)jobs;
#define ER_ALL(F, C, L) launch_all<F, C, L>(p, frame, s1, s2, q, counts, ev, jb, vis, counters, tiles_x, ntiles, nblocks, \
                                            shade_blocks, std::max(1, shadow_blocks), split, lds_bytes, lds_bits, st, st2, \
                                            shaded, after_primary, order_in, cost, order_out, heavy_cap, lpr, \
                                            (shadow_lanes_per_ray == 4 || shadow_lanes_per_ray == 16) ? shadow_lanes_per_ray : 1, \
                                            heavy_lanes_per_ray)
    if (full) {
        if (count) { if (lut_lds) ER_ALL(1, 1, 1); else ER_ALL(1, 1, 0); }
        else { if (lut_lds) ER_ALL(1, 0, 1); else ER_ALL(1, 0, 0); }
    } else {
        if (count) { if (lut_lds) ER_ALL(0, 1, 1); else ER_ALL(0, 1, 0); }
        else { if (lut_lds) ER_ALL(0, 0, 1); else ER_ALL(0, 0, 0); }
    }
#undef ER_ALL
    return hipGetLastError();
}

/* One batch round, split like a single pass: primary and shading on st; shadow marches and
 * combine on st2 after the event `shaded`. lpr / shadow_lpr: lanes per primary / shadow march
 * (1, 4, 16 or 64 / 1, 4 or 16), for every job of the batch. */
template<int F, int L>
static void launch_primary_b(const erd::BatchJob* J, int na, int lpr, int blocks, size_t lds, hipStream_t stream)
{
    if (lpr == 4) hipLaunchKernelGGL((erd::k_primary_b<F, L, 4>), dim3(blocks), dim3(256), lds, stream, J, na);
    else if (lpr == 16) hipLaunchKernelGGL((erd::k_primary_b<F, L, 16>), dim3(blocks), dim3(256), lds, stream, J, na);
    else if (lpr == 64) hipLaunchKernelGGL((erd::k_primary_b<F, L, 64>), dim3(blocks), dim3(256), lds, stream, J, na);
    else hipLaunchKernelGGL((erd::k_primary_b<F, L, 1>), dim3(blocks), dim3(256), lds, stream, J, na);
}

template<int F, int L>
static void launch_shadow_b(const erd::BatchJob* J, int na, int lpr, int blocks, size_t lds, hipStream_t stream)
{
    if (lpr == 4) hipLaunchKernelGGL((erd::k_shadow_b<F, L, 4>), dim3(blocks), dim3(256), lds, stream, J, na);
    else if (lpr == 16) hipLaunchKernelGGL((erd::k_shadow_b<F, L, 16>), dim3(blocks), dim3(256), lds, stream, J, na);
    else hipLaunchKernelGGL((erd::k_shadow_b<F, L, 1>), dim3(blocks), dim3(256), lds, stream, J, na);
}

/* na: the jobs taking part in this round, the first na of the job table (the table is ordered
 * by pass count, largest first) */
extern "C" hipError_t erd_launch_batch_trace(const erd::BatchJob* J, int na, int full, int primary_blocks,
                                             int lds_words, int shade_blocks, hipStream_t stream,
                                             hipStream_t stream2, hipEvent_t shaded, int lpr, int shadow_lpr)
{
    /* lds_words > 0: every job's tables fit in that many words of LDS */
    const size_t lds = (size_t)lds_words * 4;
#define ER_PRIMARY_B(F, L) launch_primary_b<F, L>(J, na, lpr, primary_blocks, lds, stream)
    if (full) {
        if (lds_words > 0) ER_PRIMARY_B(1, 1); else ER_PRIMARY_B(1, 0);
        hipLaunchKernelGGL(erd::k_shade_b<1>, dim3(shade_blocks), dim3(256), 0, stream, J, na);
    } else {
        if (lds_words > 0) ER_PRIMARY_B(0, 1); else ER_PRIMARY_B(0, 0);
        hipLaunchKernelGGL(erd::k_shade_b<0>, dim3(shade_blocks), dim3(256), 0, stream, J, na);
    }
#undef ER_PRIMARY_B
    hipError_t e = hipEventRecord(shaded, stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream2, shaded, 0);
    if (e != hipSuccess) return e;
#define ER_SHADOW_B(F, L) launch_shadow_b<F, L>(J, na, shadow_lpr, primary_blocks, lds, stream2)
    if (full) { if (lds_words > 0) ER_SHADOW_B(1, 1); else ER_SHADOW_B(1, 0); }
    else { if (lds_words > 0) ER_SHADOW_B(0, 1); else ER_SHADOW_B(0, 0); }
#undef ER_SHADOW_B
    hipLaunchKernelGGL(erd::k_finish_b, dim3(shade_blocks), dim3(256), 0, stream2, J, na);
    return hipGetLastError();
}

/* the jobs' parameter blocks for the batch kernels (at most ER_BATCH_CONST_JOBS) */
extern "C" hipError_t erd_batch_params_upload(const erd::DevParams* host, int n)
{
    if (n < 0 || n > ER_BATCH_CONST_JOBS) return hipErrorInvalidValue;
    return hipMemcpyToSymbol(HIP_SYMBOL(erd::c_batch_params), host, (size_t)n * sizeof(erd::DevParams), 0,
                             hipMemcpyHostToDevice);
}

extern "C" hipError_t erd_launch_round_end(const erd::BatchJob* J, int njobs, hipStream_t stream)
{
    hipLaunchKernelGGL(erd::k_round_end, dim3(1), dim3(64), 0, stream, J, njobs);
    return hipGetLastError();
}

/* the primary and shading kernels of a full-estimator pass alone (the path-replay adjoint
 * follows them with its own kernels) */
extern "C" hipError_t erd_launch_primary_shade(const erd::DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2,
                                               void* queue, unsigned int* counts, void* events, void* jobs, int W, int H,
                                               int lut_lds, size_t lds_bytes, size_t lds_bits, int shade_blocks,
                                               int primary_blocks, int tile_shift, hipStream_t st)
{
    const int tw = 1 << tile_shift, th = 64 >> tile_shift;
    const int tiles_x = (W + tw - 1) / tw, tiles_y = (H + th - 1) / th;
    const int ntiles = tiles_x * tiles_y;
    constexpr int wpb = ER_PRIMARY_WG / 64;   /* waves per workgroup */
    const int nblocks = std::max(1, std::min((ntiles + wpb - 1) / wpb, primary_blocks));
    erd::ScatterRec* q = (erd::ScatterRec*)queue;
    if (lut_lds)
        hipLaunchKernelGGL((erd::k_primary<1, 0, 1, 1>), dim3(nblocks), dim3(ER_PRIMARY_WG), lds_bytes, st, p, frame, s1, s2, q, counts,
                           (unsigned long long*)nullptr, tiles_x, ntiles, (const uint32_t*)nullptr, (uint32_t*)nullptr, 64);
    else
        hipLaunchKernelGGL((erd::k_primary<1, 0, 0, 1>), dim3(nblocks), dim3(ER_PRIMARY_WG), lds_bytes, st, p, frame, s1, s2, q, counts,
                           (unsigned long long*)nullptr, tiles_x, ntiles, (const uint32_t*)nullptr, (uint32_t*)nullptr, 64);
    hipLaunchKernelGGL((erd::k_shade<1, 0, 1, 1>), dim3(shade_blocks), dim3(256), lds_bits, st, p, frame, s1, s2,
                       (const erd::ScatterRec*)q, counts, (erd::EventRec*)events, (erd::ShadowJob*)jobs,
                       (unsigned long long*)nullptr);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_autofocus(const erd::DevParams* p, int u, int v, uint32_t s0, uint32_t s1, float* out,
                                           hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_autofocus, dim3(1), dim3(64), 0, st, p, u, v, s0, s1, out);
    return hipGetLastError();
}
