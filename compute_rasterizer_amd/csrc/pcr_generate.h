// pcr_generate.h — synthetic scenes from a counter-based PRNG.
//
// Every point depends only on (seed, index), so a cloud of any size can be
// produced on the device in one pass, and any sub-range reproduced on the
// host. Host (building table) and device (points) share this header;
// compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace pcr {

__host__ __device__ inline uint64_t mix64(uint64_t z) {   // SplitMix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rnd(uint64_t key, uint64_t i) { return mix64(key ^ (i * 0xD6E8FEB86659FD93ull)); }
__host__ __device__ inline float u01(uint64_t r) { return (float)(uint32_t)(r >> 40) * 5.9604644775390625e-08f; }
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t s) { return mix64(mix64(seed) + s); }

__host__ __device__ inline float lattice(uint64_t kn, int32_t ix, int32_t iy, uint32_t oct) {
    const uint64_t h = mix64(kn ^ ((uint64_t)(uint32_t)ix * 0x9E3779B97F4A7C15ull)
                                ^ ((uint64_t)(uint32_t)iy * 0xC2B2AE3D27D4EB4Full)
                                ^ ((uint64_t)oct << 59));
    return u01(h) * 2.0f - 1.0f;
}

__host__ __device__ inline float value_noise(uint64_t kn, float x, float y, uint32_t oct) {
    const float fx = floorf(x), fy = floorf(y);
    const int32_t ix = (int32_t)fx, iy = (int32_t)fy;
    const float tx = x - fx, ty = y - fy;
    const float sx = (tx * tx) * (3.0f - 2.0f * tx);
    const float sy = (ty * ty) * (3.0f - 2.0f * ty);
    const float v00 = lattice(kn, ix, iy, oct), v10 = lattice(kn, ix + 1, iy, oct);
    const float v01 = lattice(kn, ix, iy + 1, oct), v11 = lattice(kn, ix + 1, iy + 1, oct);
    const float a = v00 + (v10 - v00) * sx;
    const float b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sy;
}

// 6-octave fBm, base wavelength 512 m; height in [0, amplitude].
__host__ __device__ inline float terrain_height(uint64_t seed, float amplitude, float x, float y) {
    const uint64_t kn = stream_key(seed, 7);
    float sum = 0.0f, amp = 1.0f, fx = x * (1.0f / 512.0f), fy = y * (1.0f / 512.0f);
    for (uint32_t o = 0; o < 6; o++) {
        sum = sum + value_noise(kn, fx, fy, o) * amp;
        fx = fx * 2.0f; fy = fy * 2.0f; amp = amp * 0.5f;
    }
    const float f = sum / 1.96875f;
    return amplitude * (f * 0.5f + 0.5f);
}

__host__ __device__ inline uint32_t clamp255(int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

struct GenParams {
    uint64_t seed;
    uint64_t n_terrain;     // CITY: points [0, n_terrain) are terrain
    float extent, amplitude, jitter;
    int32_t scene, num_buildings;
};

struct Building {
    float x0, y0, z0, sx, sy, sz;
    uint32_t rgb;
    uint32_t pad;
    uint64_t end;           // exclusive prefix count of building points
};

__host__ __device__ inline void cube_point(const GenParams& g, uint64_t i, float p[3], uint32_t& rgba) {
    const uint64_t k = stream_key(g.seed, 0);
    p[0] = u01(rnd(k, 3 * i)); p[1] = u01(rnd(k, 3 * i + 1)); p[2] = u01(rnd(k, 3 * i + 2));
    rgba = (uint32_t)(uint8_t)(255.0f * p[0]) | ((uint32_t)(uint8_t)(255.0f * p[1]) << 8)
         | ((uint32_t)(uint8_t)(255.0f * p[2]) << 16) | 0xFF000000u;
}

__host__ __device__ inline void terrain_point(const GenParams& g, uint64_t i, float p[3], uint32_t& rgba) {
    const uint64_t k = stream_key(g.seed, 1);
    const uint64_t r0 = rnd(k, 8 * i + 0), r1 = rnd(k, 8 * i + 1), r2 = rnd(k, 8 * i + 2);
    const uint64_t r3 = rnd(k, 8 * i + 3), r4 = rnd(k, 8 * i + 4), r5 = rnd(k, 8 * i + 5);
    const float u = u01(r0) * g.extent, v = u01(r1) * g.extent;
    const float h = terrain_height(g.seed, g.amplitude, u, v);
    const float jx = (u01(r2) * 2.0f - 1.0f) * g.jitter;
    const float jy = (u01(r3) * 2.0f - 1.0f) * g.jitter;
    const float jz = (u01(r4) * 2.0f - 1.0f) * g.jitter;
    p[0] = u + jx; p[1] = v + jy; p[2] = h + jz;
    const float t = h / g.amplitude;
    float cr, cg, cb;
    if (t < 0.5f) {
        const float s = t * 2.0f;
        cr = 60.0f + (140.0f - 60.0f) * s; cg = 110.0f + (110.0f - 110.0f) * s; cb = 40.0f + (70.0f - 40.0f) * s;
    } else {
        const float s = t * 2.0f - 1.0f;
        cr = 140.0f + (235.0f - 140.0f) * s; cg = 110.0f + (235.0f - 110.0f) * s; cb = 70.0f + (235.0f - 70.0f) * s;
    }
    const int n = (int)(r5 >> 59) - 16;
    rgba = clamp255((int)cr + n) | (clamp255((int)cg + n) << 8) | (clamp255((int)cb + n) << 16) | 0xFF000000u;
}

__host__ __device__ inline void building_point(const GenParams& g, const Building* __restrict__ t, uint64_t j,
                                               float p[3], uint32_t& rgba) {
    int lo = 0, hi = g.num_buildings - 1;
    while (lo < hi) { const int m = (lo + hi) / 2; if (j < t[m].end) hi = m; else lo = m + 1; }
    const Building B = t[lo];
    const uint64_t kp = stream_key(g.seed, 3);
    const uint64_t r0 = rnd(kp, 4 * j + 0), r1 = rnd(kp, 4 * j + 1);
    const uint64_t r2 = rnd(kp, 4 * j + 2), r3 = rnd(kp, 4 * j + 3);
    const float ax = B.sx * B.sz, ay = B.sy * B.sz, ar = B.sx * B.sy;
    const float t0 = ax, t1 = t0 + ax, t2 = t1 + ay, t3 = t2 + ay, t4 = t3 + ar;
    const float a = u01(r0) * t4, s = u01(r1), u = u01(r2);
    if (a < t0)      { p[0] = B.x0 + s * B.sx; p[1] = B.y0;            p[2] = B.z0 + u * B.sz; }
    else if (a < t1) { p[0] = B.x0 + s * B.sx; p[1] = B.y0 + B.sy;     p[2] = B.z0 + u * B.sz; }
    else if (a < t2) { p[0] = B.x0;            p[1] = B.y0 + s * B.sy; p[2] = B.z0 + u * B.sz; }
    else if (a < t3) { p[0] = B.x0 + B.sx;     p[1] = B.y0 + s * B.sy; p[2] = B.z0 + u * B.sz; }
    else             { p[0] = B.x0 + s * B.sx; p[1] = B.y0 + u * B.sy; p[2] = B.z0 + B.sz; }
    const int n = (int)(r3 >> 59) - 16;
    rgba = clamp255((int)(B.rgb & 255u) + n) | (clamp255((int)((B.rgb >> 8) & 255u) + n) << 8)
         | (clamp255((int)((B.rgb >> 16) & 255u) + n) << 16) | 0xFF000000u;
}

}  // namespace pcr
