/*
 * er_device.h - device data layout and the per-path estimator functions of the gfx950 tracer.
 *
 * Data layout in HBM (per volume):
 *   cells   : one 16-byte record per lattice cell (vx, vy, vz) in [-1, R-1]^3 holding the eight
 *             clamped uint16 corner voxels a trilinear sample at that cell reads
 *             (Source/buffer3d.h:210-242). A sample is one 128-bit load instead of eight
 *             2-byte gathers; the clamp-to-edge of the reference is folded into the records.
 *             (Rx+1)(Ry+1)(Rz+1) x 16 B (8x the voxel bytes; 1.9 GB for 512x512x460).
 *   mc_*    : macrocells of 8^3 cells (16^3 or 32^3 for volumes whose occupancy bits and
 *             opacity table would not fit 52.5 KB of LDS). For each, the
 *             min/max voxel over the voxels its samples can
 *             read, and (per tracer) the max of DensityScale * Opacity over the intensities a
 *             truncating trilinear sample can produce from that range ([min-3, max]). A sample
 *             whose macrocell maximum is 0 contributes exactly 0 to the optical depth, so the
 *             march skips its fetch and keeps the exact sequence of MinT additions.
 * Per tracer: XYZA float4 frame / running estimates, RGBA8 display (+ bilateral temporaries),
 * two uint32 seed planes (SoA), all row-major y * W + x.
 *
 * Operation order follows the reference expression by expression (compiled with
 * -ffp-contract=off and correctly rounded division / square root), so decisions and sample
 * values match the CPU restatement bit for bit.
 */
#ifndef ER_DEVICE_H
#define ER_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/er_capi.h"
#include "er_fmath.h"

#define ER_NEGLOG_BITS 23          /* -log(u) table over u = m * 2^-23 */
#ifndef ER_GROUP_SHIFT
#define ER_GROUP_SHIFT 2           /* skip groups of 2^ER_GROUP_SHIFT macrocells a side */
#endif

namespace erd {

/* reference constants (Source/defines.h:58-77) */
#define ER_PI_F          3.141592654f
#define ER_INV_PI_F      0.31830988618379067154f
#define ER_INV_TWO_PI_F  0.15915494309189533577f
#define ER_RAY_EPS       0.0001f

struct DevLight {
    float tm[3][4];
    float inv_tm[3][4];
    int   type, one_sided, visible, unit;
    float size[3];
    float area, inner_radius, outer_radius, multiplier;
    int   texture_id;
    int   valid;
};

struct DevTexture {
    int   type, proc_type, bitmap_id, valid;
    float output_level;
    float offset[2], repeat[2];
    int   flip[2];
    float uniform[3], checker1[3], checker2[3];
    int   gradient_index;            /* index into DevParams::tex_gradients, -1 if none */
};

/* a surface object (Source/object.h): its shape in the light layout, and the texture slots of
 * its diffuse, specular and glossiness (objects.h:26-82, transport.h:141-149) */
struct DevObject {
    DevLight shape;
    int   diffuse_tex, specular_tex, gloss_tex;
    int   valid;
};

/* an RGBA8 bitmap (Source/erbitmap.h), row-major y * W + x */
struct DevBitmap {
    const uchar4* pixels;
    int   w, h;
};

struct DevClip {
    float inv_tm[3][4];
    int   type, invert;
    float size[3];
    float outer_radius;
};

/* Everything a pass reads besides the frame buffers; lives in device memory, read through a
 * uniform pointer (scalar loads). */
struct DevParams {
    /* camera (Source/camera.h:61-96) */
    float cam_pos[3], cam_N[3], cam_U[3], cam_V[3];
    float screen00, screen10, invscreen0, invscreen1;
    float focal, aperture, clip_near, clip_far, exposure;
    int   film_w, film_h;
    int   tile_shift;                /* primary tiles are 2^tile_shift x 2^(6 - tile_shift) pixels */
    /* settings */
    float step_primary, step_shadow, max_shadow, density_scale;
    float grad_threshold, grad_factor, gm_inv;
    int   shadows, estimator, shading_type, opacity_modulated, gradient_mode, sampler;
    /* Woodcock delta tracking: majorants on a fixed grid of 2^ER_WC_SHIFT-cell macrocells
     * (independent of the tuned brick size, so the image does not depend on it); wc_sigma[m]
     * is the largest sigma_lut value over the intensities a sample in macrocell m can take,
     * and density_scale * wc_sigma[m] bounds the extinction DS^2 * opacity there */
    const float* wc_sigma;
    int   wc_dim[3];
    /* volume (Source/volume.h:100-138) */
    int   res[3];
    float minp[3], maxp[3], size[3], invsize[3], resf[3], spacing[3], minstep;
    const uint4* cells;
    int   cstride_y, cstride_z;      /* (Rx+1), (Rx+1)(Ry+1) */
    const float* sigma_lut;          /* DensityScale * Opacity(i), i in [0, lut_max] */
    int   lut_max;
    const float* mc_sigma;           /* per macrocell max sigma */
    const uint32_t* mc_bits;         /* bit m set <=> mc_sigma[m] > 0 */
    int   mc_dim[3];
    int   mc_shift;                  /* macrocell edge = 2^mc_shift cells */
    int   mc_words;                  /* 32-bit words of mc_bits: macrocell bits, group bits, group clip-free bits */
    int   mc2_dim[3];                /* groups of 4 x 4 x 4 macrocells */
    int   mc2_base;                  /* first word of the group radii (4 bits a group) in mc_bits */
    int   mc3_base;                  /* first word of the group clip-free bits in mc_bits */
    const float* neglog;             /* -log(m 2^-23), m in [0, 2^23) */
    /* the tracer's lights and textures by slot (light_ids maps list position -> slot), clips */
    int   light_count;
    int   light_ids[ER_MAX_INDICES];
    DevLight   lights[ER_LIST_CAPACITY];
    DevTexture textures[ER_LIST_CAPACITY];
    int   clip_count;
    DevClip    clips[ER_LIST_CAPACITY];
    int   object_count;
    DevObject  objects[ER_LIST_CAPACITY];
    DevBitmap  bitmaps[ER_LIST_CAPACITY];
    /* transfer functions (Source/ertracer.h:105-109) */
    er_plf      opacity;
    er_plf      glossiness;
    er_color_tf diffuse, specular, emission;
    er_color_tf tex_gradients[4];
    /* node-table transfer function (er_set_tf_lut): when set, opacity and diffuse colour come
     * from ER_TF_LUT_NODES nodes x 4 channels (opacity, diffuse X, Y, Z) instead of the PLFs */
    const float* tf_lut;
    float tf_lut_scale;              /* (ER_TF_LUT_NODES - 1) / intensity range */
};

/* ---- small vector type; component-wise, reference operand order ---------------------- */
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 scl(float s, f3 a) { return mk(a.x * s, a.y * s, a.z * s); }   /* float * Vec3f */
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 nrm(f3 a) { const float L = len(a); return mk(a.x / L, a.y / L, a.z / L); }
__device__ __forceinline__ f3 crs(f3 a, f3 b)
{
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__device__ __forceinline__ float absdot(f3 a, f3 b) { return fabsf(dot(a, b)); }
__device__ __forceinline__ bool black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return fmaxf(lo, fminf(v, hi)); }

__device__ __forceinline__ f3 xform_point(const float m[3][4], f3 p)
{
    return mk(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
              m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
              m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]);
}
__device__ __forceinline__ f3 xform_vector(const float m[3][4], f3 v)
{
    return mk(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
              m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}

struct Ray { f3 O, D; float MinT, MaxT; };
__device__ __forceinline__ f3 ray_at(const Ray& r, float t) { return add(r.O, scl(t, nrm(r.D))); }   /* ray.h:47-50 */

/* TransformRay (geometry.h:54-68) */
__device__ __forceinline__ Ray xform_ray(const float m[3][4], const Ray& r)
{
    const f3 P = xform_point(m, r.O);
    const f3 MinP = xform_point(m, ray_at(r, r.MinT));
    const f3 MaxP = xform_point(m, ray_at(r, r.MaxT));
    Ray t;
    t.O = P;
    t.D = nrm(sub(MaxP, P));
    t.MinT = len(sub(MinP, P));
    t.MaxT = len(sub(MaxP, P));
    return t;
}

/* ---- RNG: register-resident pair of 16-bit multiply-with-carry generators (rng.h:35-51) */
struct Rng {
    uint32_t s0, s1;
    __device__ __forceinline__ uint32_t next_bits()
    {
        s0 = 36969u * (s0 & 65535u) + (s0 >> 16);
        s1 = 18000u * (s1 & 65535u) + (s1 >> 16);
        return ((s0 << 16) + s1) & 0x007fffffu;
    }
    /* u = ((r & 0x7fffff) | 0x40000000 as float - 2) / 2 = (r & 0x7fffff) * 2^-23 exactly */
    __device__ __forceinline__ float u() { return (float)next_bits() * 1.1920928955078125e-07f; }
};

/* The stream of a Woodcock shadow march (full estimator), whose draw count is not known before
 * it runs: SplitMix64's finaliser of the path's state (s0 high, s1 low, xor a constant) split
 * in two non-zero halves. The path's own stream advances by the two draws a ray-march shadow
 * test takes, so shading can hand the march to a later kernel. */
__device__ __forceinline__ Rng shadow_stream(const Rng& r)
{
    uint64_t z = (((uint64_t)r.s0 << 32) | (uint64_t)r.s1) ^ 0x5368616457636b21ull;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    Rng s;
    s.s0 = (uint32_t)z ? (uint32_t)z : 1u;
    s.s1 = (uint32_t)(z >> 32) ? (uint32_t)(z >> 32) : 1u;
    return s;
}

/* ---- transfer functions (Source/plf.h:68-91) ----------------------------------------- */
__device__ __forceinline__ float plf_eval(const er_plf& p, float x)
{
    const int n = p.count;
    if (n <= 0) return 0.0f;
    if (x < p.node_range[0]) return p.value[0];
    if (x > p.node_range[1]) return p.value[n - 1];
    for (int i = 1; i < n; ++i) {
        const float P1 = p.position[i - 1];
        const float P2 = p.position[i];
        const float LerpT = (x - P1) / (P2 - P1);
        if (x >= P1 && x < P2) return p.value[i - 1] + LerpT * (p.value[i] - p.value[i - 1]);
    }
    return 0.0f;
}
__device__ __forceinline__ f3 ctf_eval(const er_color_tf& c, float x)
{
    return mk(plf_eval(c.plf[0], x), plf_eval(c.plf[1], x), plf_eval(c.plf[2], x));
}

/* Node-table transfer function: node n sits at intensity n / scale; between nodes the value
 * is linear, below zero it is clamped to zero (the derivative there is zero). */
__host__ __device__ inline void tf_lut_coord(float I, float scale, int& b, float& f)
{
    const float x = fminf(I * scale, (float)(ER_TF_LUT_NODES - 1));
    b = (int)x;
    if (b > ER_TF_LUT_NODES - 2) b = ER_TF_LUT_NODES - 2;
    f = x - (float)b;
}
__host__ __device__ inline float tf_lut_raw(const float* lut, int b, float f, int ch)
{
    return (1.0f - f) * lut[b * 4 + ch] + f * lut[(b + 1) * 4 + ch];
}
__host__ __device__ inline float tf_lut_eval(const float* lut, float scale, float I, int ch)
{
    int b; float f;
    tf_lut_coord(I, scale, b, f);
    return fmaxf(tf_lut_raw(lut, b, f, ch), 0.0f);
}
__device__ __forceinline__ float opacity_at(const DevParams& p, float I)
{
    if (p.tf_lut) return tf_lut_eval(p.tf_lut, p.tf_lut_scale, I, 0);
    return plf_eval(p.opacity, I);
}
__device__ __forceinline__ f3 diffuse_at(const DevParams& p, float I)
{
    if (p.tf_lut)
        return mk(tf_lut_eval(p.tf_lut, p.tf_lut_scale, I, 1), tf_lut_eval(p.tf_lut, p.tf_lut_scale, I, 2),
                  tf_lut_eval(p.tf_lut, p.tf_lut_scale, I, 3));
    return ctf_eval(p.diffuse, I);
}

/* DevParams lives in device memory, so the compiler sees the pointers stored in it as generic
 * and would access them with flat instructions. A flat load also counts against the LDS
 * counter, so every later wait for an LDS read would drain all loads in flight. These pointers
 * always address global memory; gmem() says so. */
template<class T> __device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* ptr)
{
    return (const __attribute__((address_space(1))) T*)ptr;
}

/* ---- volume sampling ----------------------------------------------------------------- */
struct Lattice { float Lx, Ly, Lz; };

/* Volume::operator() (volume.h:131-138): lattice coordinates of a world point */
__device__ __forceinline__ Lattice to_lattice(const DevParams& p, f3 P)
{
    Lattice l;
    l.Lx = ((P.x - p.minp[0]) * p.invsize[0]) * p.resf[0];
    l.Ly = ((P.y - p.minp[1]) * p.invsize[1]) * p.resf[1];
    l.Lz = ((P.z - p.minp[2]) * p.invsize[2]) * p.resf[2];
    return l;
}

/* (float)(unsigned short)v, the reference's truncating store of a lerp into the uint16 voxel
 * type (geometry.h:87-90). The lerped values are never negative and stay below 65536 (at most
 * 65535 (1 + 2^-24)), where that truncation is the floor: one instruction instead of a
 * float -> int -> float round trip. */
__device__ __forceinline__ float trunc16(float v) { return floorf(v); }

/* A trilinear sample split in two halves so that several cell loads can be in flight before
 * their values are used: the cell record and the fractional lattice offsets. */
struct CellFetch { uint4 c; float dx, dy, dz; };

__device__ __forceinline__ CellFetch cell_fetch(const DevParams& p, const Lattice& l)
{
    const float fx = floorf(l.Lx), fy = floorf(l.Ly), fz = floorf(l.Lz);
    const int vx = (int)fx, vy = (int)fy, vz = (int)fz;
    CellFetch f;
    f.dx = l.Lx - (float)vx; f.dy = l.Ly - (float)vy; f.dz = l.Lz - (float)vz;
    const int cx = min(max(vx, -1), p.res[0] - 1) + 1;
    const int cy = min(max(vy, -1), p.res[1] - 1) + 1;
    const int cz = min(max(vz, -1), p.res[2] - 1) + 1;
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    /* a cell index is below 2^32 (the record array stays below 64 GB) and each factor below
     * 2^24: full-rate 24-bit multiplies */
    const unsigned idx = __umul24((unsigned)cz, (unsigned)p.cstride_z) + __umul24((unsigned)cy, (unsigned)p.cstride_y) +
                         (unsigned)cx;
    const u32x4 v = gmem((const u32x4*)p.cells)[idx];
    f.c = make_uint4(v.x, v.y, v.z, v.w);
    return f;
}

/* Buffer3D<ushort>::operator()(Vec3f) (buffer3d.h:222-242) from one cell record: 7 lerps,
 * each truncated to uint16 (geometry.h:87-90). */
__device__ __forceinline__ float cell_value(const CellFetch& f)
{
    const uint4 c = f.c;
    const float dx = f.dx, dy = f.dy, dz = f.dz;
    const float v000 = (float)(c.x & 0xffffu), v100 = (float)(c.x >> 16);
    const float v010 = (float)(c.y & 0xffffu), v110 = (float)(c.y >> 16);
    const float v001 = (float)(c.z & 0xffffu), v101 = (float)(c.z >> 16);
    const float v011 = (float)(c.w & 0xffffu), v111 = (float)(c.w >> 16);
    const float ox = 1.0f - dx, oy = 1.0f - dy, oz = 1.0f - dz;
    const float d00 = trunc16(ox * v000 + dx * v100);
    const float d10 = trunc16(ox * v010 + dx * v110);
    const float d01 = trunc16(ox * v001 + dx * v101);
    const float d11 = trunc16(ox * v011 + dx * v111);
    const float d0 = trunc16(oy * d00 + dy * d10);
    const float d1 = trunc16(oy * d01 + dy * d11);
    return trunc16(oz * d0 + dz * d1);
}

__device__ __forceinline__ float trilinear_cell(const DevParams& p, const Lattice& l) { return cell_value(cell_fetch(p, l)); }

__device__ __forceinline__ float intensity(const DevParams& p, f3 P) { return trilinear_cell(p, to_lattice(p, P)); }

/* GradientCD / FD / Filtered (volumes.h:31-74); Normalize (Q20: zero gradient -> NaN) */
/* the six central-difference samples of GradientCD (volumes.h:31-41): I(P + dX), I(P - dX),
 * I(P + dY), I(P - dY), I(P + dZ), I(P - dZ), d = MinStep on every axis (Q23) */
__device__ __forceinline__ void cd_samples(const DevParams& p, f3 P, float g[6])
{
    const float d = p.minstep;
    const f3 dX = mk(d, 0.0f, 0.0f), dY = mk(0.0f, d, 0.0f), dZ = mk(0.0f, 0.0f, d);
    g[0] = intensity(p, add(P, dX)); g[1] = intensity(p, sub(P, dX));
    g[2] = intensity(p, add(P, dY)); g[3] = intensity(p, sub(P, dY));
    g[4] = intensity(p, add(P, dZ)); g[5] = intensity(p, sub(P, dZ));
}
__device__ __forceinline__ f3 grad_cd(const DevParams& p, f3 P)
{
    float g[6];
    cd_samples(p, P, g);
    return mk(g[1] - g[0], g[3] - g[2], g[5] - g[4]);
}
__device__ __forceinline__ f3 grad_fd(const DevParams& p, f3 P)
{
    const float d = p.minstep;
    const float i0 = intensity(p, P);
    const float i1 = intensity(p, add(P, mk(d, 0.0f, 0.0f)));
    const float i2 = intensity(p, add(P, mk(0.0f, d, 0.0f)));
    const float i3 = intensity(p, add(P, mk(0.0f, 0.0f, d)));
    return mk(i0 - i1, i0 - i2, i0 - i3);
}
__device__ __forceinline__ f3 vlerp(f3 a, f3 b, float c) { return add(a, scl(c, sub(b, a))); }
/* grad_cd as a call of its own: grad_filtered's nine gradients then hold six cell records in
 * registers at a time instead of 54, and the kernels that can call grad_filtered keep the
 * register budget (and so the occupancy) of their other paths */
__device__ __noinline__ f3 grad_cd_call(const DevParams& p, f3 P) { return grad_cd(p, P); }
__device__ __noinline__ f3 grad_filtered(const DevParams& p, f3 P)
{
    const float o = p.minstep;
    const f3 G0 = grad_cd_call(p, P);
    const f3 G1 = grad_cd_call(p, add(P, mk(-o, -o, -o)));
    const f3 G2 = grad_cd_call(p, add(P, mk(o, o, o)));
    const f3 G3 = grad_cd_call(p, add(P, mk(-o, o, -o)));
    const f3 G4 = grad_cd_call(p, add(P, mk(o, -o, o)));
    const f3 G5 = grad_cd_call(p, add(P, mk(-o, -o, o)));
    const f3 G6 = grad_cd_call(p, add(P, mk(o, o, -o)));
    const f3 G7 = grad_cd_call(p, add(P, mk(-o, o, o)));
    const f3 G8 = grad_cd_call(p, add(P, mk(o, -o, -o)));
    const f3 L0 = vlerp(vlerp(G1, G2, 0.5f), vlerp(G3, G4, 0.5f), 0.5f);
    const f3 L1 = vlerp(vlerp(G5, G6, 0.5f), vlerp(G7, G8, 0.5f), 0.5f);
    return vlerp(G0, vlerp(L0, L1, 0.5f), 0.75f);
}
__device__ __forceinline__ int gradient_samples(const DevParams& p)
{
    return p.gradient_mode == ER_GRADIENT_CENTRAL ? 6 : (p.gradient_mode == ER_GRADIENT_FILTERED ? 54 : 4);
}
__device__ __forceinline__ f3 normalized_gradient(const DevParams& p, f3 P)
{
    if (p.gradient_mode == ER_GRADIENT_CENTRAL) return nrm(grad_cd(p, P));
    if (p.gradient_mode == ER_GRADIENT_FILTERED) return nrm(grad_filtered(p, P));
    return nrm(grad_fd(p, P));
}
/* the six central-difference samples of an event, carried by value into the shading (valid:
 * taken at the event's point) */
struct CD6 {
    float g[6];
    bool valid;
};

/* GradientMagnitude (volumes.h:93-114) from the six central-difference samples at its point
 * (cd_samples: the same sample positions as GradientCD's) */
__device__ __forceinline__ float gradient_magnitude_cd(const DevParams& p, const float g[6])
{
    float Sum = 0.0f;
    {
        float D = g[1] - g[0];
        D *= 0.5f / p.spacing[0];
        Sum += D * D;
    }
    {
        float D = g[3] - g[2];
        D *= 0.5f / p.spacing[1];
        Sum += D * D;
    }
    {
        float D = g[5] - g[4];
        D *= 0.5f / p.spacing[2];
        Sum += D * D;
    }
    return sqrtf(Sum);
}
/* GradientMagnitude (volumes.h:93-114) */
__device__ __forceinline__ float gradient_magnitude(const DevParams& p, f3 P)
{
    const float d = p.minstep;
    float Sum = 0.0f;
    {
        float D = intensity(p, sub(P, mk(d, 0.0f, 0.0f))) - intensity(p, add(P, mk(d, 0.0f, 0.0f)));
        D *= 0.5f / p.spacing[0];
        Sum += D * D;
    }
    {
        float D = intensity(p, sub(P, mk(0.0f, d, 0.0f))) - intensity(p, add(P, mk(0.0f, d, 0.0f)));
        D *= 0.5f / p.spacing[1];
        Sum += D * D;
    }
    {
        float D = intensity(p, sub(P, mk(0.0f, 0.0f, d))) - intensity(p, add(P, mk(0.0f, 0.0f, d)));
        D *= 0.5f / p.spacing[2];
        Sum += D * D;
    }
    return sqrtf(Sum);
}

/* IntersectBox (box.h:51-83) */
__device__ __forceinline__ bool intersect_box(const Ray& r, f3 mn, f3 mx, float& nearT, float& farT)
{
    const f3 InvR = mk(1.0f / r.D.x, 1.0f / r.D.y, 1.0f / r.D.z);
    const f3 B = mul(InvR, sub(mn, r.O));
    const f3 T = mul(InvR, sub(mx, r.O));
    const float mn0 = fminf(T.x, B.x), mn1 = fminf(T.y, B.y), mn2 = fminf(T.z, B.z);
    const float mx0 = fmaxf(T.x, B.x), mx1 = fmaxf(T.y, B.y), mx2 = fmaxf(T.z, B.z);
    const float LargestMinT = fmaxf(fmaxf(mn0, mn1), fmaxf(mn0, mn2));
    const float LargestMaxT = fminf(fminf(mx0, mx1), fminf(mx0, mx2));
    if (LargestMaxT < LargestMinT) return false;
    nearT = LargestMinT > 0.0f ? LargestMinT : 0.0f;
    farT = LargestMaxT;
    if (nearT < r.MinT || nearT > r.MaxT) return false;
    return true;
}

} /* namespace erd */

#endif /* ER_DEVICE_H */
