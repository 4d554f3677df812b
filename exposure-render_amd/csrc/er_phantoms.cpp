/*
 * er_phantoms.cpp - seeded synthetic volumes for the benchmark scenes (no GPU involved).
 *
 * Random numbers come from SplitMix64: one stream for the shape parameters, and a counter-based
 * hash of (seed, voxel index) for the per-voxel noise, so the result does not depend on the
 * number of threads.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/er_capi.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

inline uint64_t hash64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double hash_uniform(uint64_t seed, uint64_t i, uint64_t k)
{
    return (double)(hash64(seed * 0x632BE59BD9B4E019ull ^ (i * 4 + k)) >> 11) * (1.0 / 9007199254740992.0);
}

template<class F> void parallel_z(int rz, F f)
{
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() { for (int z = (int)t; z < rz; z += (int)nt) f(z); });
    for (auto& x : th) x.join();
}

inline uint16_t to_u16(double v)
{
    v = std::floor(v + 0.5);
    if (v < 0.0) v = 0.0;
    if (v > 65535.0) v = 65535.0;
    return (uint16_t)v;
}

struct Ellipsoid {
    double c[3], a[3], R[3][3];   /* centre, semi-axes (voxels), rotation (rows = local axes) */
    double value;
};

struct Tube {
    double p[3], d[3], r, value;
};

void random_rotation(SplitMix& rng, double R[3][3])
{
    const double a = rng.range(0, 2 * M_PI), b = rng.range(0, 2 * M_PI), c = rng.range(0, 2 * M_PI);
    const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cc = cos(c), sc = sin(c);
    const double Rz[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    const double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    const double Rx[3][3] = {{1, 0, 0}, {0, cc, -sc}, {0, sc, cc}};
    double T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0;
            for (int k = 0; k < 3; ++k) T[i][j] += Rz[i][k] * Ry[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += T[i][k] * Rx[k][j];
        }
}

}  // namespace

extern "C" {

ER_API int er_phantom_blobs(uint16_t* out, int rx, int ry, int rz, int n_blobs, float sigma_lo, float sigma_hi,
                            float peak_lo, float peak_hi, float noise, uint64_t seed)
{
    if (!out || rx <= 0 || ry <= 0 || rz <= 0 || n_blobs < 0) return ER_ERROR;
    SplitMix rng{seed};
    const double ext = (double)std::max(rx, std::max(ry, rz));
    std::vector<double> cx(n_blobs), cy(n_blobs), cz(n_blobs), s2(n_blobs), pk(n_blobs);
    for (int b = 0; b < n_blobs; ++b) {
        cx[b] = rng.uniform() * rx; cy[b] = rng.uniform() * ry; cz[b] = rng.uniform() * rz;
        const double s = rng.range(sigma_lo, sigma_hi) * ext;
        s2[b] = 2.0 * s * s;
        pk[b] = rng.range(peak_lo, peak_hi);
    }
    parallel_z(rz, [&](int z) {
        for (int y = 0; y < ry; ++y)
            for (int x = 0; x < rx; ++x) {
                double v = 0.0;
                for (int b = 0; b < n_blobs; ++b) {
                    const double dx = x + 0.5 - cx[b], dy = y + 0.5 - cy[b], dz = z + 0.5 - cz[b];
                    v += pk[b] * std::exp(-(dx * dx + dy * dy + dz * dz) / s2[b]);
                }
                const uint64_t i = ((uint64_t)z * ry + y) * rx + x;
                v += noise * (2.0 * hash_uniform(seed, i, 0) - 1.0);
                out[i] = to_u16(v);
            }
    });
    return ER_OK;
}

ER_API int er_phantom_ellipsoids(uint16_t* out, int rx, int ry, int rz, int n_ellipsoids, int n_tubes, float axis_lo,
                                 float axis_hi, int n_bands, float max_intensity, float noise_sigma, uint64_t seed)
{
    if (!out || rx <= 0 || ry <= 0 || rz <= 0 || n_ellipsoids < 0 || n_tubes < 0 || n_bands <= 0) return ER_ERROR;
    SplitMix rng{seed};
    const double ext[3] = {(double)rx, (double)ry, (double)rz};
    std::vector<Ellipsoid> E(n_ellipsoids);
    for (auto& e : E) {
        for (int i = 0; i < 3; ++i) {
            e.c[i] = ext[i] * (0.5 + rng.range(-0.12, 0.12));
            e.a[i] = ext[i] * rng.range(axis_lo, axis_hi);
        }
        random_rotation(rng, e.R);
    }
    /* nested: larger ellipsoids first and in lower (softer) bands; smaller ones overwrite */
    std::sort(E.begin(), E.end(), [](const Ellipsoid& a, const Ellipsoid& b) {
        return a.a[0] * a.a[1] * a.a[2] > b.a[0] * b.a[1] * b.a[2];
    });
    const double band_w = max_intensity / n_bands;
    for (int k = 0; k < n_ellipsoids; ++k) {
        const int band = std::min(n_bands - 1, k * n_bands / std::max(1, n_ellipsoids));
        E[k].value = band_w * (band + rng.range(0.25, 0.75));
    }
    std::vector<Tube> T(n_tubes);
    for (auto& t : T) {
        for (int i = 0; i < 3; ++i) t.p[i] = ext[i] * (0.5 + rng.range(-0.25, 0.25));
        double d[3] = {rng.range(-1, 1), rng.range(-1, 1), rng.range(-1, 1)};
        const double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-12;
        for (int i = 0; i < 3; ++i) t.d[i] = d[i] / l;
        t.r = std::max(ext[0], std::max(ext[1], ext[2])) * rng.range(0.015, 0.05);
        const int band = (int)(rng.uniform() * n_bands) % n_bands;
        t.value = band_w * (band + rng.range(0.25, 0.75));
    }
    parallel_z(rz, [&](int z) {
        for (int y = 0; y < ry; ++y)
            for (int x = 0; x < rx; ++x) {
                const double p[3] = {x + 0.5, y + 0.5, z + 0.5};
                double v = 0.0;
                for (const auto& e : E) {
                    double q = 0.0;
                    const double d[3] = {p[0] - e.c[0], p[1] - e.c[1], p[2] - e.c[2]};
                    for (int i = 0; i < 3; ++i) {
                        const double l = (e.R[0][i] * d[0] + e.R[1][i] * d[1] + e.R[2][i] * d[2]) / e.a[i];
                        q += l * l;
                    }
                    if (q <= 1.0) v = e.value;
                }
                for (const auto& t : T) {
                    const double d[3] = {p[0] - t.p[0], p[1] - t.p[1], p[2] - t.p[2]};
                    const double s = d[0] * t.d[0] + d[1] * t.d[1] + d[2] * t.d[2];
                    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] - s * s;
                    if (r2 <= t.r * t.r) v = t.value;
                }
                const uint64_t i = ((uint64_t)z * ry + y) * rx + x;
                const double u1 = std::max(hash_uniform(seed, i, 1), 1e-300), u2 = hash_uniform(seed, i, 2);
                v += noise_sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                out[i] = to_u16(v);
            }
    });
    return ER_OK;
}

}  /* extern "C" */
