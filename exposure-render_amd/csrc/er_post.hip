/*
 * er_post.hip - per-pass post-processing and the volume preprocessing kernels.
 *
 * k_post fuses, per pixel, the three remaining kernels of a pass:
 *   KrnlFilterFrameEstimate (Source/filterframeestimate.cuh:33-64): 3x3 Gaussian, sigma 1,
 *     weights renormalised at the border; alpha is the centre alpha divided by the total
 *     weight (the code sets Sum[3] and then divides all four channels);
 *   KrnlComputeEstimate (Source/estimate.cuh:27-32): R += (F - R) / max(N, 1) with N the
 *     passes completed before this one (Q1);
 *   KrnlToneMap (Source/tonemap.cuh:30-59): XYZ -> linear sRGB, 1 - exp(-c / Exposure),
 *     clamp, x255 truncated; A = R.a x 255.
 * The filtered frame is not written back (the reference copies it over the frame estimate,
 * filterframeestimate.cuh:71-73); the fused kernel reads the 3x3 neighbourhood through an
 * LDS tile with a one-pixel halo.
 *
 * k_bilateral_{h,v}: the separable bilateral filter on the RGBA8 display image (design:
 * Source/filterrunningestimate.cuh:55-191, Source/filter.h:34-40) with taps
 * [max(0, i-R), min(i+R, n-1)) - exclusive upper bound as written (Q27).
 *
 * k_build_cells / k_macro_minmax / k_macro_sigma / k_macro_bits / k_gm_max: volume preprocessing.
 */
#include "er_path.h"

namespace erd {

#define POST_TILE 16

/* One 16 x 16 tile of one pass's post-processing; every thread of the workgroup calls it. */
__device__ __forceinline__ void post_tile(const DevParams& p, const float4* __restrict__ frame, float4* __restrict__ running,
                                          uchar4* __restrict__ display, int no_iterations, int bx, int by,
                                          float4 (*tile)[POST_TILE + 2])
{
    const int W = p.film_w, H = p.film_h;
    for (int k = threadIdx.x; k < (POST_TILE + 2) * (POST_TILE + 2); k += blockDim.x) {
        const int ty = k / (POST_TILE + 2), tx = k % (POST_TILE + 2);
        const int gx = min(max(bx + tx - 1, 0), W - 1), gy = min(max(by + ty - 1, 0), H - 1);
        tile[ty][tx] = frame[(size_t)gy * W + gx];
    }
    __syncthreads();
    const int lx = threadIdx.x % POST_TILE, ly = threadIdx.x / POST_TILE;
    const int x = bx + lx, y = by + ly;
    if (x >= W || y >= H) return;

    /* Gauss2D(1, X, Y) = exp(-(X^2 + Y^2) / 2) */
    const float g0 = erm::expf_(-0.0f), g1 = erm::expf_(-0.5f), g2 = erm::expf_(-1.0f);
    const int x0 = max(x - 1, 0), x1 = min(x + 1, W - 1);
    const int y0 = max(y - 1, 0), y1 = min(y + 1, H - 1);
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, tw = 0.0f;
    for (int yy = y0; yy <= y1; ++yy) {
        for (int xx = x0; xx <= x1; ++xx) {
            const int d2 = (xx - x) * (xx - x) + (yy - y) * (yy - y);
            const float w = d2 == 0 ? g0 : (d2 == 1 ? g1 : g2);
            const float4 f = tile[yy - by + 1][xx - bx + 1];
            s0 += f.x * w; s1 += f.y * w; s2 += f.z * w; s3 += f.w * w;
            tw += w;
        }
    }
    s3 = tile[ly + 1][lx + 1].w;
    float4 fe;
    if (tw > 0.0f) { fe.x = s0 / tw; fe.y = s1 / tw; fe.z = s2 / tw; fe.w = s3 / tw; }
    else { fe.x = fe.y = fe.z = fe.w = 0.0f; }

    const size_t i = (size_t)y * W + x;
    const float n = fmaxf((float)no_iterations, 1.0f);
    float4 r = running[i];
    r.x = r.x + (fe.x - r.x) / n;
    r.y = r.y + (fe.y - r.y) / n;
    r.z = r.z + (fe.z - r.z) / n;
    r.w = r.w + (fe.w - r.w) / n;
    running[i] = r;

    const float R = 3.240479f * r.x - 1.537150f * r.y - 0.498535f * r.z;
    const float G = -0.969256f * r.x + 1.875991f * r.y + 0.041556f * r.z;
    const float B = 0.055648f * r.x - 0.204043f * r.y + 1.057311f * r.z;
    const float e = p.exposure;
    const float tr = fmaxf(0.0f, fminf(1.0f - erm::expf_(-(R / e)), 1.0f));
    const float tg = fmaxf(0.0f, fminf(1.0f - erm::expf_(-(G / e)), 1.0f));
    const float tb = fmaxf(0.0f, fminf(1.0f - erm::expf_(-(B / e)), 1.0f));
    uchar4 d;
    d.x = (unsigned char)(255.0f * tr);
    d.y = (unsigned char)(255.0f * tg);
    d.z = (unsigned char)(255.0f * tb);
    d.w = (unsigned char)(r.w * 255.0f);
    display[i] = d;
}

__global__ __launch_bounds__(256) void k_post(const DevParams* __restrict__ pp, const float4* __restrict__ frame,
                                             float4* __restrict__ running, uchar4* __restrict__ display, int no_iterations,
                                             unsigned int* __restrict__ qcount)
{
    /* the scatter queue of this pass has been consumed: empty it for the next pass */
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) qcount[threadIdx.x] = 0u;
    __shared__ float4 tile[POST_TILE + 2][POST_TILE + 2];
    post_tile(*pp, frame, running, display, no_iterations, blockIdx.x * POST_TILE, blockIdx.y * POST_TILE, tile);
}

/* post-processing of a batch round: one workgroup per 16 x 16 tile of each of the na jobs
 * taking part (the first na of the table, whose tiles come first); the job by binary search
 * over the tiles' ascending first indices */
__global__ __launch_bounds__(256) void k_post_b(const BatchJob* __restrict__ J, int na, unsigned int round)
{
    const int b = blockIdx.x;
    if (na == 0) return;
    int lo = 0, hi = na - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J[mid].post_base <= b) lo = mid; else hi = mid - 1;
    }
    const int job = lo;
    if (b < J[job].post_base || b >= J[job].post_base + J[job].post_ntiles) return;
    const BatchJob& B = J[job];
    const int t = b - B.post_base;
    __shared__ float4 tile[POST_TILE + 2][POST_TILE + 2];
    post_tile(*B.params, B.frame, B.running, B.display, B.no_iter0 + (int)round, (t % B.post_tiles_x) * POST_TILE,
              (t / B.post_tiles_x) * POST_TILE, tile);
}

struct BilateralParams {
    int radius;
    float kd[65];      /* spatial weights, index R + k - i */
    float gs[256];     /* range weights, index (int)|dY| */
};

__device__ __forceinline__ float lum(uchar4 c)
{
    return 0.212671f * (float)c.x + 0.715160f * (float)c.y + 0.072169f * (float)c.z;
}

template<int HORIZONTAL>
__global__ __launch_bounds__(256) void k_bilateral(const BilateralParams* __restrict__ bp, const uchar4* __restrict__ in,
                                                  uchar4* __restrict__ out, int W, int H)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const int R = bp->radius;
    const int i = HORIZONTAL ? x : y;
    const int n = HORIZONTAL ? W : H;
    const int lo = max(0, i - R), hi = min(i + R, n - 1);
    const uchar4 c = in[(size_t)y * W + x];
    const float yc = lum(c);
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, tw = 0.0f;
    for (int k = lo; k < hi; ++k) {
        const uchar4 q = HORIZONTAL ? in[(size_t)y * W + k] : in[(size_t)k * W + x];
        const float w = bp->kd[R + k - i] * bp->gs[(int)fabsf(lum(q) - yc)];
        tw += w;
        s0 += (float)q.x * w; s1 += (float)q.y * w; s2 += (float)q.z * w;
    }
    uchar4 d;
    if (tw > 0.0f) {
        d.x = (unsigned char)fmaxf(0.0f, fminf(s0 / tw, 255.0f));
        d.y = (unsigned char)fmaxf(0.0f, fminf(s1 / tw, 255.0f));
        d.z = (unsigned char)fmaxf(0.0f, fminf(s2 / tw, 255.0f));
    } else {
        d.x = d.y = d.z = 0;
    }
    d.w = 255;
    out[(size_t)y * W + x] = d;
}

/* cell (cx, cy, cz) in [0, R]^3 holds the corners of lattice cell (cx-1, cy-1, cz-1) */
__global__ __launch_bounds__(256) void k_build_cells(const uint16_t* __restrict__ vox, uint4* __restrict__ cells,
                                                    int rx, int ry, int rz)
{
    const int sx = rx + 1, sy = ry + 1, sz = rz + 1;
    const size_t n = (size_t)sx * sy * sz;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const int cx = (int)(k % sx);
        const int cy = (int)((k / sx) % sy);
        const int cz = (int)(k / ((size_t)sx * sy));
        const int x0 = max(cx - 1, 0), x1 = min(cx, rx - 1);
        const int y0 = max(cy - 1, 0), y1 = min(cy, ry - 1);
        const int z0 = max(cz - 1, 0), z1 = min(cz, rz - 1);
        auto at = [&](int x, int y, int z) -> uint32_t { return vox[((size_t)z * ry + y) * rx + x]; };
        uint4 c;
        c.x = at(x0, y0, z0) | (at(x1, y0, z0) << 16);
        c.y = at(x0, y1, z0) | (at(x1, y1, z0) << 16);
        c.z = at(x0, y0, z1) | (at(x1, y0, z1) << 16);
        c.w = at(x0, y1, z1) | (at(x1, y1, z1) << 16);
        cells[k] = c;
    }
}

/* min/max voxel over [m*E, m*E + E] (E = 2^shift, clamped) on each axis: every voxel a sample
 * whose floor(L) >> shift == m (after clamping) can read */
__global__ __launch_bounds__(256) void k_macro_minmax(const uint16_t* __restrict__ vox, uint32_t* __restrict__ mm,
                                                     int rx, int ry, int rz, int mx, int my, int mz, int shift)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mx * my * mz) return;
    const int E = 1 << shift;
    const int bx = m % mx, by = (m / mx) % my, bz = m / (mx * my);
    const int x0 = bx * E, x1 = min(bx * E + E, rx - 1);
    const int y0 = by * E, y1 = min(by * E + E, ry - 1);
    const int z0 = bz * E, z1 = min(bz * E + E, rz - 1);
    uint32_t lo = 0xffff, hi = 0;
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const uint32_t v = vox[((size_t)z * ry + y) * rx + x];
                lo = min(lo, v);
                hi = max(hi, v);
            }
    mm[m] = lo | (hi << 16);
}

/* max sigma over the intensities a truncating trilinear sample can take in the macrocell:
 * each of the three lerp levels can lose at most one unit below the smallest corner */
__global__ __launch_bounds__(256) void k_macro_sigma(const uint32_t* __restrict__ mm, const float* __restrict__ lut,
                                                    int lut_max, float* __restrict__ out, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int lo = max((int)(mm[m] & 0xffff) - 3, 0);
    const int hi = min((int)(mm[m] >> 16), lut_max);
    float s = 0.0f;
    for (int i = lo; i <= hi; ++i) s = fmaxf(s, lut[i]);
    out[m] = s;
}

/* group bits: bit g set iff any macrocell of group g (2^ER_GROUP_SHIFT macrocells a side, clipped to the
 * grid) has its bit set, from word `occ_base` of the same array; the groups' clip-free bits
 * from word `free_base`: bit g set iff every occupied macrocell of group g is wholly unclipped
 * (bit m of `unclipped`, the host's conservative test; NULL: no clipping objects, every bit
 * clear), so the march need not test its samples against them. */
__global__ __launch_bounds__(256) void k_group_bits(uint32_t* __restrict__ bits, const uint32_t* __restrict__ unclipped,
                                                   int mx, int my, int mz, int gx, int gy, int gz, int occ_base,
                                                   int free_base, int words)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t v = 0, f = 0;
    for (int b = 0; b < 32; ++b) {
        const int g = w * 32 + b;
        if (g >= gx * gy * gz) break;
        const int x = g % gx, y = (g / gx) % gy, z = g / (gx * gy);
        bool any = false, free = unclipped != nullptr;
        const int E = 1 << ER_GROUP_SHIFT;
        for (int k = E * z; k < min(E * z + E, mz); ++k)
            for (int j = E * y; j < min(E * y + E, my); ++j)
                for (int i = E * x; i < min(E * x + E, mx); ++i) {
                    const int m = (k * my + j) * mx + i;
                    if ((bits[m >> 5] >> (m & 31)) & 1u) {
                        any = true;
                        free = free && ((unclipped[m >> 5] >> (m & 31)) & 1u);
                    }
                }
        if (any) v |= 1u << b;
        if (free) f |= 1u << b;
    }
    bits[occ_base + w] = v;
    bits[free_base + w] = f;
}

/* occupancy bit of macrocell m: it can hold a sample with sigma > 0 that is not clipped away;
 * clipped (may be NULL) has bit m set when every sample of macrocell m is clipped away (the
 * host's conservative test; clipping objects apply in the full estimator only) */
__global__ __launch_bounds__(256) void k_macro_bits(const float* __restrict__ mc_sigma, const uint32_t* __restrict__ clipped,
                                                   uint32_t* __restrict__ bits, int n, int words)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    const uint32_t gone = clipped ? clipped[w] : 0u;
    uint32_t v = 0;
    for (int b = 0; b < 32; ++b) {
        const int m = w * 32 + b;
        if (m < n && mc_sigma[m] > 0.0f && !((gone >> b) & 1u)) v |= 1u << b;
    }
    bits[w] = v;
}

/* maximum GradientMagnitude over voxel centres (design: gradientmagnitude.cuh:28-70) */
__global__ __launch_bounds__(256) void k_gm_max(const DevParams* __restrict__ pp, unsigned int* __restrict__ out)
{
    const DevParams& p = *pp;
    const size_t n = (size_t)p.res[0] * p.res[1] * p.res[2];
    float best = 0.0f;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(k % p.res[0]);
        const int y = (int)((k / p.res[0]) % p.res[1]);
        const int z = (int)(k / ((size_t)p.res[0] * p.res[1]));
        const f3 P = mk(p.minp[0] + p.size[0] * (((float)x + 0.5f) / (float)p.res[0]),
                        p.minp[1] + p.size[1] * (((float)y + 0.5f) / (float)p.res[1]),
                        p.minp[2] + p.size[2] * (((float)z + 0.5f) / (float)p.res[2]));
        best = fmaxf(best, gradient_magnitude(p, P));
    }
    /* non-negative floats order like their bit patterns */
    for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(best));
}

} /* namespace erd */

extern "C" hipError_t erd_launch_post(const erd::DevParams* p, const float4* frame, float4* running, uchar4* display,
                                      int no_iterations, unsigned int* qcount, int W, int H, hipStream_t st)
{
    dim3 grid((W + POST_TILE - 1) / POST_TILE, (H + POST_TILE - 1) / POST_TILE), block(256);
    hipLaunchKernelGGL(erd::k_post, grid, block, 0, st, p, frame, running, display, no_iterations, qcount);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_bilateral(const erd::BilateralParams* bp, const uchar4* in, uchar4* tmp, uchar4* out,
                                           int W, int H, hipStream_t st)
{
    dim3 grid((W + 15) / 16, (H + 15) / 16), block(256);
    hipLaunchKernelGGL(erd::k_bilateral<1>, grid, block, 0, st, bp, in, tmp, W, H);
    hipLaunchKernelGGL(erd::k_bilateral<0>, grid, block, 0, st, bp, (const uchar4*)tmp, out, W, H);
    return hipGetLastError();
}

extern "C" size_t erd_bilateral_params_size(void) { return sizeof(erd::BilateralParams); }

extern "C" hipError_t erd_launch_build_cells(const uint16_t* vox, uint4* cells, int rx, int ry, int rz, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_build_cells, dim3(2048), dim3(256), 0, st, vox, cells, rx, ry, rz);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_macro_minmax(const uint16_t* vox, uint32_t* mm, int rx, int ry, int rz, int mx, int my,
                                              int mz, int shift, hipStream_t st)
{
    const int n = mx * my * mz;
    hipLaunchKernelGGL(erd::k_macro_minmax, dim3((n + 255) / 256), dim3(256), 0, st, vox, mm, rx, ry, rz, mx, my, mz, shift);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_macro_bits(const float* mc_sigma, const uint32_t* clipped, uint32_t* bits, int n, int words,
                                            hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_macro_bits, dim3((words + 255) / 256), dim3(256), 0, st, mc_sigma, clipped, bits, n, words);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_group_bits(uint32_t* bits, const uint32_t* unclipped, int mx, int my, int mz, int gx, int gy,
                                            int gz, int occ_base, int free_base, int words, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_group_bits, dim3((words + 255) / 256), dim3(256), 0, st, bits, unclipped, mx, my, mz, gx, gy,
                       gz, occ_base, free_base, words);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_macro_sigma(const uint32_t* mm, const float* lut, int lut_max, float* out, int n, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_macro_sigma, dim3((n + 255) / 256), dim3(256), 0, st, mm, lut, lut_max, out, n);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_gm_max(const erd::DevParams* p, unsigned int* out, hipStream_t st)
{
    hipLaunchKernelGGL(erd::k_gm_max, dim3(2048), dim3(256), 0, st, p, out);
    return hipGetLastError();
}

extern "C" hipError_t erd_launch_batch_post(const erd::BatchJob* J, int na, unsigned int round, int active_tiles,
                                            hipStream_t stream)
{
    if (na == 0 || active_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(erd::k_post_b, dim3(active_tiles), dim3(256), 0, stream, J, na, round);
    return hipGetLastError();
}
