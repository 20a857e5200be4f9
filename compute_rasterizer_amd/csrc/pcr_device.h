// pcr_device.h — device-side numerics of the rasterizer (gfx950).
//
// Fixed fp32 operation order, compiled with -ffp-contract=off and correctly
// rounded division/sqrt, so that every result is a pure function of its
// inputs:
//   mat4*(x,y,z,1): r_i = fma(M[8+i], z, fma(M[4+i], y, fma(M[i], x, M[12+i])))
//   mat4*(x,y,z,w): r_i = fma(M[12+i], w, fma(M[8+i], z, fma(M[4+i], y, M[i]*x)))
//   length(v)     : sqrt(fma(z,z, fma(y,y, x*x)))
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pcr.h"

namespace pcr {

constexpr float kInvSteps30 = 9.3132257461547852e-10f;  // 2^-30
constexpr float kInvSteps10 = 9.765625e-04f;            // 2^-10
constexpr uint8_t kCulled = 255;

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Frustum test of an AABB against the six Gribb-Hartmann planes of `m`
// (compute_loop_las/render.cs:134-197): p-vertex per plane, cull if d < 0.
__device__ inline bool frustum_visible(const float* __restrict__ m, const float mn[3], const float mx[3]) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int axis = i >> 1;
        const float sgn = (i == 0 || i == 3 || i == 4) ? -1.0f : 1.0f;
        float p[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float a = m[4 * k + 3], b = m[4 * k + axis];
            p[k] = sgn < 0.0f ? a - b : a + b;
        }
        const float len = __builtin_sqrtf(fma_(p[2], p[2], fma_(p[1], p[1], p[0] * p[0])));
        const float nx = p[0] / len, ny = p[1] / len, nz = p[2] / len, c = p[3] / len;
        const float vx = nx > 0.0f ? mx[0] : mn[0];
        const float vy = ny > 0.0f ? mx[1] : mn[1];
        const float vz = nz > 0.0f ? mx[2] : mn[2];
        const float d = fma_(nz, vz, fma_(ny, vy, nx * vx)) + c;
        if (d < 0.0f) return false;
    }
    return true;
}

// The same test split in two: plane i of the view (normal and constant, four
// floats), computed once per workgroup, then the p-vertex test per AABB. The
// operations are those of frustum_visible(), so the results are identical.
__device__ inline void frustum_plane(const float* __restrict__ m, int i, float out[4]) {
    const int axis = i >> 1;
    const bool minus = (i == 0 || i == 3 || i == 4);
    float p[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float a = m[4 * k + 3], b = m[4 * k + axis];
        p[k] = minus ? a - b : a + b;
    }
    const float len = __builtin_sqrtf(fma_(p[2], p[2], fma_(p[1], p[1], p[0] * p[0])));
    out[0] = p[0] / len; out[1] = p[1] / len; out[2] = p[2] / len; out[3] = p[3] / len;
}
__device__ inline bool planes_visible(const float (*pl)[4], const float mn[3], const float mx[3]) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float nx = pl[i][0], ny = pl[i][1], nz = pl[i][2], c = pl[i][3];
        const float vx = nx > 0.0f ? mx[0] : mn[0];
        const float vy = ny > 0.0f ? mx[1] : mn[1];
        const float vz = nz > 0.0f ? mx[2] : mn[2];
        const float d = fma_(nz, vz, fma_(ny, vy, nx * vx)) + c;
        if (d < 0.0f) return false;
    }
    return true;
}

// Screen-space length (pixels) of a view-space x offset `len` placed at the
// AABB centre (compute_loop_las/render.cs:199-215).
__device__ inline float projected_px(const float mn[3], const float mx[3], float len,
                                     const float* __restrict__ vw, const float* __restrict__ pr,
                                     int w, int h) {
    const float cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
    float vc[4], pc[4], pe[4];
#pragma unroll
    for (int i = 0; i < 4; i++) vc[i] = fma_(vw[8 + i], cz, fma_(vw[4 + i], cy, fma_(vw[i], cx, vw[12 + i])));
    const float vex = vc[0] + len;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        pc[i] = fma_(pr[12 + i], vc[3], fma_(pr[8 + i], vc[2], fma_(pr[4 + i], vc[1], pr[i] * vc[0])));
        pe[i] = fma_(pr[12 + i], vc[3], fma_(pr[8 + i], vc[2], fma_(pr[4 + i], vc[1], pr[i] * vex)));
    }
    const float pcx = pc[0] / pc[3], pcy = pc[1] / pc[3], pex = pe[0] / pe[3], pey = pe[1] / pe[3];
    const float fw = (float)w, fh = (float)h;
    const float scx = (fw * (pcx + 1.0f)) * 0.5f, scy = (fh * (pcy + 1.0f)) * 0.5f;
    const float sex = (fw * (pex + 1.0f)) * 0.5f, sey = (fh * (pey + 1.0f)) * 0.5f;
    const float dx = sex - scx, dy = sey - scy;
    return __builtin_sqrtf(fma_(dy, dy, dx * dx));
}

// Precision level 0..4 from the projected AABB diagonal
// (compute_loop_las/render.cs:217-230; first threshold 80 for LOD nodes,
// compute_loop_nodes/render.cs:199).
__device__ inline int precision_level(const float mn[3], const float mx[3], const pcr_view& v, float first) {
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    const float radius = __builtin_sqrtf(fma_(dz, dz, fma_(dy, dy, dx * dx)));
    const float px = projected_px(mn, mx, radius, v.view_world, v.proj, v.width, v.height);
    if (px < first) return 4;
    if (px < 200.0f) return 3;
    if (px < 500.0f) return 2;
    if (px < 10000.0f) return 1;
    return 0;
}

// Number of 4-byte position planes a level reads (levels 2-4 read only the
// 10-bit plane).
__host__ __device__ inline int planes_for_level(int level) { return level == 0 ? 3 : (level == 1 ? 2 : 1); }

// Per-batch decode constants: the scale of one integer step and the AABB min.
struct Decode {
    float scale[3];
    float mn[3];
};
__device__ inline Decode make_decode(int level, const float mn[3], const float mx[3]) {
    Decode d;
    const float f = level <= 1 ? kInvSteps30 : kInvSteps10;
#pragma unroll
    for (int a = 0; a < 3; a++) { d.scale[a] = (mx[a] - mn[a]) * f; d.mn[a] = mn[a]; }
    return d;
}
// Decode one point from its plane words (las2/render.cs:340-458).
template <int LEVEL_CLASS>  // 0: 30 bit, 1: 20 bit, 2: 10 bit
__device__ __forceinline__ void decode(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi, float p[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t l = (lo >> (10 * a)) & 1023u;
        uint32_t X;
        if (LEVEL_CLASS == 0) X = (l << 20) | (((med >> (10 * a)) & 1023u) << 10) | ((hi >> (10 * a)) & 1023u);
        else if (LEVEL_CLASS == 1) X = (l << 20) | (((med >> (10 * a)) & 1023u) << 10);
        else X = l;
        p[a] = fma_((float)X, d.scale[a], d.mn[a]);
    }
}

// px/pw and py/pw for pw > 0, each the correctly rounded IEEE quotient.
//
// gfx950 divides in fp32 as: scale the operands (v_div_scale), y = rcp(d),
// y = fma(fma(-d, y, 1), y, y), q = n*y, q = fma(fma(-d, q, n), y, q), then
// q = fma(fma(-d, q, n), y, q) (v_div_fmas), then patch special values
// (v_div_fixup). When d and |n| lie in [2^-40, 2^40] the scaling and the
// patch are identities: operands, reciprocal and quotient are normal, the
// exponent gap is below 96 and the numerator exponent above 23. There the
// code below is the same operations on the same values, with the reciprocal
// step shared by the two numerators. Every other operand (a zero or tiny
// numerator, an extreme depth) takes the compiler's division.
__device__ __forceinline__ void div_fast(float px, float py, float pw, float& nx, float& ny) {
    const float y0 = __builtin_amdgcn_rcpf(pw);
    const float y = fma_(fma_(-pw, y0, 1.0f), y0, y0);
    float q = px * y;
    q = fma_(fma_(-pw, q, px), y, q);
    nx = fma_(fma_(-pw, q, px), y, q);
    q = py * y;
    q = fma_(fma_(-pw, q, py), y, q);
    ny = fma_(fma_(-pw, q, py), y, q);
}
__device__ __forceinline__ void div_by_w(float px, float py, float pw, float& nx, float& ny) {
    const float lo = fminf(fminf(fabsf(px), fabsf(py)), pw);
    const float hi = fmaxf(fmaxf(fabsf(px), fabsf(py)), pw);
    if (lo >= 0x1p-40f && hi <= 0x1p40f) {
        div_fast(px, py, pw, nx, ny);
    } else {
        nx = px / pw;
        ny = py / pw;
    }
}

// Project a point with `t` (column-major), returning the pixel and the depth
// bits (clip w = view-space depth) when it lands inside the image
// (compute_loop_las/render.cs:400-409; truncated pixel == width/height and
// NaN positions are rejected).
__device__ __forceinline__ bool project(const float* __restrict__ t, float x, float y, float z, int w, int h,
                                        uint32_t& pixel, uint32_t& depth_bits) {
    const float px = fma_(t[8], z, fma_(t[4], y, fma_(t[0], x, t[12])));
    const float py = fma_(t[9], z, fma_(t[5], y, fma_(t[1], x, t[13])));
    const float pw = fma_(t[11], z, fma_(t[7], y, fma_(t[3], x, t[15])));
    if (!(pw > 0.0f)) return false;
    float nx, ny;
    div_by_w(px, py, pw, nx, ny);
    if (!(nx >= -1.0f && nx <= 1.0f && ny >= -1.0f && ny <= 1.0f)) return false;
    const float sx = fma_(nx, 0.5f, 0.5f) * (float)w;
    const float sy = fma_(ny, 0.5f, 0.5f) * (float)h;
    const int ix = (int)sx, iy = (int)sy;
    if (ix >= w || iy >= h) return false;
    pixel = (uint32_t)ix + (uint32_t)iy * (uint32_t)w;
    depth_bits = __float_as_uint(pw);
    return true;
}

// The projection constants of one view, with the three addends (t[12],
// t[13], t[15]) copied into VGPRs once per batch: a gfx950 VALU instruction
// reads at most one SGPR, so fma(sgpr, v, sgpr) would cost an extra v_mov per
// point. The arithmetic is the same as project().
struct Proj {
    float t0, t4, t8, t1, t5, t9, t3, t7, t11;   // uniform (SGPR)
    float c12, c13, c15;                        // VGPR copies
    float fw, fh;
    int w, h;
};
__device__ __forceinline__ float in_vgpr(float x) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
__device__ __forceinline__ Proj make_proj(const float* __restrict__ t, int w, int h) {
    Proj p;
    p.t0 = t[0]; p.t4 = t[4]; p.t8 = t[8];
    p.t1 = t[1]; p.t5 = t[5]; p.t9 = t[9];
    p.t3 = t[3]; p.t7 = t[7]; p.t11 = t[11];
    p.c12 = in_vgpr(t[12]); p.c13 = in_vgpr(t[13]); p.c15 = in_vgpr(t[15]);
    p.fw = (float)w; p.fh = (float)h; p.w = w; p.h = h;
    return p;
}
__device__ __forceinline__ bool project_xy(const Proj& p, float x, float y, float z, int& ix, int& iy,
                                           uint32_t& depth_bits) {
    const float px = fma_(p.t8, z, fma_(p.t4, y, fma_(p.t0, x, p.c12)));
    const float py = fma_(p.t9, z, fma_(p.t5, y, fma_(p.t1, x, p.c13)));
    const float pw = fma_(p.t11, z, fma_(p.t7, y, fma_(p.t3, x, p.c15)));
    if (!(pw > 0.0f)) return false;
    float nx, ny;
    div_by_w(px, py, pw, nx, ny);
    if (!(nx >= -1.0f && nx <= 1.0f && ny >= -1.0f && ny <= 1.0f)) return false;
    ix = (int)(fma_(nx, 0.5f, 0.5f) * p.fw);
    iy = (int)(fma_(ny, 0.5f, 0.5f) * p.fh);
    if (ix >= p.w || iy >= p.h) return false;
    depth_bits = __float_as_uint(pw);
    return true;
}

// project_xy() for a batch whose screen_rect() reported `interior`, without
// its tests. There every point's pw lies in [2^-13, 2^40] and |px|, |py| in
// [0, 2^40], so the division is div_by_w()'s fast sequence: correctly rounded
// for a numerator of at least 2^-40, and below that |nx| < 2^-26, where
// fma(nx, 0.5, 0.5) is exactly 0.5 however the quotient rounds. The NDC and
// image tests fall to the caller's rectangle test: the rectangle keeps a
// pixel from every image edge, and a quotient below -1 (above 1) gives a
// pixel coordinate of -1 or 0 (the width or height) after the clamp. Returns
// the pixel relative to (x0, y0) as unsigned, so both tests are `< size`.
// Wherever project_xy() accepts a point, the pixel and depth are its own.
__device__ __forceinline__ void project_interior(const Proj& p, float x, float y, float z, int x0, int y0,
                                                 uint32_t& ux, uint32_t& uy, uint32_t& depth_bits) {
    const float px = fma_(p.t8, z, fma_(p.t4, y, fma_(p.t0, x, p.c12)));
    const float py = fma_(p.t9, z, fma_(p.t5, y, fma_(p.t1, x, p.c13)));
    const float pw = fma_(p.t11, z, fma_(p.t7, y, fma_(p.t3, x, p.c15)));
    float nx, ny;
    div_fast(px, py, pw, nx, ny);
    const float sx = __builtin_amdgcn_fmed3f(fma_(nx, 0.5f, 0.5f) * p.fw, -1.0f, p.fw);
    const float sy = __builtin_amdgcn_fmed3f(fma_(ny, 0.5f, 0.5f) * p.fh, -1.0f, p.fh);
    ux = (uint32_t)((int)sx - x0);
    uy = (uint32_t)((int)sy - y0);
    depth_bits = __float_as_uint(pw);
}

// Screen rectangle (inclusive, clamped to the image, 1-pixel margin) covered
// by the projection of an AABB; false when a corner is not in front of the
// eye or the rectangle is empty. It only routes points to a tile: a point
// whose pixel falls outside it takes the direct path, so the result never
// depends on this bound being conservative.
//
// Called by every lane of a wave with the same arguments: lane & 7 projects
// one corner, and the 8-lane groups reduce with xor shuffles, so a wave
// spends one corner's arithmetic instead of eight (the rectangle is computed
// once per batch by every wave of the workgroup).
//
// `interior` (optional) reports whether project_interior() may replace
// project_xy() for the batch's points: the rectangle keeps a pixel from
// every image edge, each row of the transform has terms of at most 2^38 in
// sum over the box (ex, ey, ew), and every corner's pw is at least
// 2^-12 + 2^-16 * ew. A decoded point lies within a few float steps of the
// box, and its computed pw within about 10 * 2^-24 * ew of the smallest
// corner's, so it keeps pw >= 2^-13, and |px|, |py|, pw <= 2^40.
//
// `wlo` (optional, with `interior`) is a depth no point of the batch can
// undercut, for the occlusion test of rect_open(), or 0 when the test does
// not apply. It applies to an interior batch whose smallest corner pw is at
// least max(W, H) * 2^-21 * (max(ex, ey) + ew). Then:
//   - a decoded point lies in the box (x = fma(X, scale, mn) <= mx, since
//     X * scale < mx - mn) or, for the 20- and 30-bit levels whose float(X)
//     rounds up, within an ulp of it;
//   - its computed pw differs from the affine w at that point, and each
//     corner's computed pw from w at the corner, by at most 3 roundings of
//     terms summing to ew, and w over the box is smallest at a corner, so the
//     point's pw >= min corner pw - 6 * 2^-24 * ew, and wlo = min corner pw -
//     2^-18 * ew lies below that with a margin of 58 * 2^-24 * ew for the
//     rounding of the subtraction itself;
//   - its screen coordinate differs from the exact projection of the point by
//     at most (W / 2) * (3 * 2^-24 * (ex + ew) / pw + 2^-24) plus the
//     roundings of fma(nx, 0.5, 0.5) * W (3 * 2^-25 * W), and so do the
//     corners'; with pw >= W * 2^-21 * (ex + ew) the first term is at most
//     0.1875 px and all of it below 0.2 px for W <= 32768. The exact
//     projection of a point of the box lies in the hull of the corners' (all
//     in front of the eye), so a point's computed coordinate lies within
//     0.4 px of the corners' computed range, and its pixel inside the
//     rectangle with its 1-pixel margin (the same holds for y with H).
__device__ inline bool screen_rect(const float* __restrict__ t, const float mn[3], const float mx[3], int w, int h,
                                   int& x0, int& y0, int& x1, int& y1, bool* interior = nullptr,
                                   float* wlo = nullptr) {
    if (interior) *interior = false;
    if (wlo) *wlo = 0.0f;
    const int c = (int)(__lane_id() & 7u);
    const float x = (c & 1) ? mx[0] : mn[0], y = (c & 2) ? mx[1] : mn[1], z = (c & 4) ? mx[2] : mn[2];
    const float px = fma_(t[8], z, fma_(t[4], y, fma_(t[0], x, t[12])));
    const float py = fma_(t[9], z, fma_(t[5], y, fma_(t[1], x, t[13])));
    const float pw = fma_(t[11], z, fma_(t[7], y, fma_(t[3], x, t[15])));
    // every 8-lane group holds the same 8 corners, so any lane not in front
    // of the eye means a corner that is not
    if (__ballot(!(pw > 0.0f)) != 0ull) return false;
    const float sx = fma_(px / pw, 0.5f, 0.5f) * (float)w, sy = fma_(py / pw, 0.5f, 0.5f) * (float)h;
    float sx0 = sx, sx1 = sx, sy0 = sy, sy1 = sy;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        sx0 = fminf(sx0, __shfl_xor(sx0, m));
        sx1 = fmaxf(sx1, __shfl_xor(sx1, m));
        sy0 = fminf(sy0, __shfl_xor(sy0, m));
        sy1 = fmaxf(sy1, __shfl_xor(sy1, m));
    }
    sx0 = fmaxf(sx0, -2.0f); sy0 = fmaxf(sy0, -2.0f);
    sx1 = fminf(sx1, (float)w + 2.0f); sy1 = fminf(sy1, (float)h + 2.0f);
    // every lane holds the same reduced values: make them wave-uniform (scalar
    // registers) for the callers' per-batch state
    x0 = __builtin_amdgcn_readfirstlane(max((int)floorf(sx0) - 1, 0));
    y0 = __builtin_amdgcn_readfirstlane(max((int)floorf(sy0) - 1, 0));
    x1 = __builtin_amdgcn_readfirstlane(min((int)floorf(sx1) + 1, w - 1));
    y1 = __builtin_amdgcn_readfirstlane(min((int)floorf(sy1) + 1, h - 1));
    if (interior) {
        const float ax = fmaxf(fabsf(mn[0]), fabsf(mx[0])), ay = fmaxf(fabsf(mn[1]), fabsf(mx[1])),
                    az = fmaxf(fabsf(mn[2]), fabsf(mx[2]));
        const float ex = fabsf(t[12]) + fabsf(t[0]) * ax + fabsf(t[4]) * ay + fabsf(t[8]) * az;
        const float ey = fabsf(t[13]) + fabsf(t[1]) * ax + fabsf(t[5]) * ay + fabsf(t[9]) * az;
        const float ew = fabsf(t[15]) + fabsf(t[3]) * ax + fabsf(t[7]) * ay + fabsf(t[11]) * az;
        const bool ok = ex <= 0x1p38f && ey <= 0x1p38f && ew <= 0x1p38f && pw >= 0x1p-12f + 0x1p-16f * ew;
        *interior = __ballot(!ok) == 0ull && x0 >= 1 && y0 >= 1 && x1 <= w - 2 && y1 <= h - 2;
        if (wlo && *interior) {
            float pwmin = pw;
#pragma unroll
            for (int m = 1; m < 8; m <<= 1) pwmin = fminf(pwmin, __shfl_xor(pwmin, m));
            if (pwmin >= fmaxf((float)w, (float)h) * 0x1p-21f * (fmaxf(ex, ey) + ew))
                *wlo = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(pwmin - 0x1p-18f * ew)));
        }
    }
    return x1 >= x0 && y1 >= y0;
}

// A batch's screen rectangle for one view, as screen_rect() reports it.
struct RectInfo {
    int x0, y0, x1, y1;
    float wlo;
    int interior;
    int on;   // screen_rect()'s return value
};

// screen_rect() for up to eight views of one box at once: the lanes of each
// 8-lane group (lane & ~7) project the corners for the group's own view (t,
// w, h may differ between groups), one corner per lane, and reduce within the
// group, so every lane of a group ends with that view's rectangle. The
// operations per view are screen_rect()'s, with the ballots restricted to
// the group and no early return, so the results are the same. Called with
// whole groups active.
__device__ inline void screen_rect_group(const float* __restrict__ t, const float mn[3], const float mx[3], int w,
                                         int h, RectInfo& r) {
    const uint32_t lane = __lane_id();
    const unsigned long long gm = 0xFFull << (lane & ~7u);
    const int c = (int)(lane & 7u);
    const float x = (c & 1) ? mx[0] : mn[0], y = (c & 2) ? mx[1] : mn[1], z = (c & 4) ? mx[2] : mn[2];
    const float px = fma_(t[8], z, fma_(t[4], y, fma_(t[0], x, t[12])));
    const float py = fma_(t[9], z, fma_(t[5], y, fma_(t[1], x, t[13])));
    const float pw = fma_(t[11], z, fma_(t[7], y, fma_(t[3], x, t[15])));
    const bool front = (__ballot(!(pw > 0.0f)) & gm) == 0ull;
    const float sx = fma_(px / pw, 0.5f, 0.5f) * (float)w, sy = fma_(py / pw, 0.5f, 0.5f) * (float)h;
    float sx0 = sx, sx1 = sx, sy0 = sy, sy1 = sy;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        sx0 = fminf(sx0, __shfl_xor(sx0, m));
        sx1 = fmaxf(sx1, __shfl_xor(sx1, m));
        sy0 = fminf(sy0, __shfl_xor(sy0, m));
        sy1 = fmaxf(sy1, __shfl_xor(sy1, m));
    }
    sx0 = fmaxf(sx0, -2.0f); sy0 = fmaxf(sy0, -2.0f);
    sx1 = fminf(sx1, (float)w + 2.0f); sy1 = fminf(sy1, (float)h + 2.0f);
    r.x0 = max((int)floorf(sx0) - 1, 0);
    r.y0 = max((int)floorf(sy0) - 1, 0);
    r.x1 = min((int)floorf(sx1) + 1, w - 1);
    r.y1 = min((int)floorf(sy1) + 1, h - 1);
    const float ax = fmaxf(fabsf(mn[0]), fabsf(mx[0])), ay = fmaxf(fabsf(mn[1]), fabsf(mx[1])),
                az = fmaxf(fabsf(mn[2]), fabsf(mx[2]));
    const float ex = fabsf(t[12]) + fabsf(t[0]) * ax + fabsf(t[4]) * ay + fabsf(t[8]) * az;
    const float ey = fabsf(t[13]) + fabsf(t[1]) * ax + fabsf(t[5]) * ay + fabsf(t[9]) * az;
    const float ew = fabsf(t[15]) + fabsf(t[3]) * ax + fabsf(t[7]) * ay + fabsf(t[11]) * az;
    const bool ok = ex <= 0x1p38f && ey <= 0x1p38f && ew <= 0x1p38f && pw >= 0x1p-12f + 0x1p-16f * ew;
    const bool interior = front && (__ballot(!ok) & gm) == 0ull && r.x0 >= 1 && r.y0 >= 1 && r.x1 <= w - 2 &&
                          r.y1 <= h - 2;
    float pwmin = pw;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) pwmin = fminf(pwmin, __shfl_xor(pwmin, m));
    r.wlo = interior && pwmin >= fmaxf((float)w, (float)h) * 0x1p-21f * (fmaxf(ex, ey) + ew) ? pwmin - 0x1p-18f * ew
                                                                                            : 0.0f;
    r.interior = interior;
    r.on = front && r.x1 >= r.x0 && r.y1 >= r.y0;
    if (!front) {   // screen_rect() returns before the rectangle: the caller's empty one
        r.x0 = 0; r.y0 = 0; r.x1 = -1; r.y1 = -1;
    }
}

// 64-bit fetch-min with the benign early-out compare of render.cs:423-426.
// The plain load can only be stale-high within a frame (values only
// decrease), so skipping when key >= old never loses a minimum. Returns 1
// when it issued the atomic (the reference's numPointsRendered count,
// render.cs:428-430).
__device__ __forceinline__ uint32_t fb_min(unsigned long long* fb, uint32_t pixel, unsigned long long key) {
    const unsigned long long old = fb[pixel];
    if (key < old) {
        atomicMin(fb + pixel, key);
        return 1u;
    }
    return 0u;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// hqs pass 1 (hqs/depth.cs:371-379): 32-bit atomicMin of the depth bits,
// with the same early-out compare as fb_min.
__device__ __forceinline__ void hqs_depth(uint32_t* __restrict__ depth, uint32_t pix, uint32_t d) {
    if (d < depth[pix]) atomicMin(depth + pix, d);
}
// Adds (lo0, hi0) to the uint32 sums p[0], p[1] and (lo1, hi1) to p[2], p[3]
// (p 16-byte aligned) with one 64-bit atomic per pair, keeping every sum
// independent mod 2^32 as four uint32 atomics would: when a pair's low sum
// wraps, the 64-bit add carries 1 into its high sum, and a second atomic
// takes it back. The low sum's running value only grows, so it wraps exactly
// floor(total / 2^32) times whatever the order of the adds, and each wrap is
// seen by the one atomic whose returned old value plus lo overflows. The hqs
// accumulators use it for (R, G) and (B, count): half the atomics of four
// uint32 adds (the reference packs R | G << 32 the same way,
// hqs/color.cs:432-436, but lets the carry stand). Both atomics are issued
// before either return value is checked, so a point waits for one round trip.
__device__ __forceinline__ void add_pairs(uint32_t* p, uint32_t lo0, uint32_t hi0, uint32_t lo1, uint32_t hi1) {
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    const unsigned long long old0 = atomicAdd(q, (unsigned long long)lo0 | ((unsigned long long)hi0 << 32));
    const unsigned long long old1 = atomicAdd(q + 1, (unsigned long long)lo1 | ((unsigned long long)hi1 << 32));
    if ((uint32_t)((uint32_t)old0 + lo0) < lo0) atomicAdd(q, 0xFFFFFFFF00000000ull);
    if ((uint32_t)((uint32_t)old1 + lo1) < lo1) atomicAdd(q + 1, 0xFFFFFFFF00000000ull);
}

// hqs pass 2 (hqs/color.cs:398-437): a point within 1.01x of the pixel's
// minimum depth adds R, G, B and 1 to the pixel's four uint32 accumulators.
__device__ __forceinline__ void hqs_accum(const uint32_t* __restrict__ depth, uint32_t* __restrict__ acc,
                                          uint32_t pix, uint32_t d, uint32_t color) {
    const float bd = __uint_as_float(depth[pix]);
    if (!(__uint_as_float(d) <= bd * 1.01f)) return;
    add_pairs(acc + 4ull * pix, color & 255u, (color >> 8) & 255u, (color >> 16) & 255u, 1u);
}
// hqs/resolve.cs:74-129: integer averages r | g << 8 | b << 16 with alpha 0
// (resolve.cs:119), the background where count == 0.
__device__ __forceinline__ uint32_t hqs_average(const u32x4 a, uint32_t bg) {
    return a.w == 0 ? bg : ((a.x / a.w) | ((a.y / a.w) << 8) | ((a.z / a.w) << 16));
}
// compute_loop_nodes_hqs/resolve.cs:92-158: a float average over the
// (2*window+1)^2 neighbourhood of pixel (x, y), weights 100 (centre),
// 2 (ox <= 1 && oy <= 1) or 1, dropping neighbours deeper than 1.01 x the
// window's smallest depth. Neighbours outside the image are skipped. Sums
// run ox-major in the shader's order, so the float results are reproducible.
__device__ inline uint32_t hqs_window_pixel(const uint32_t* __restrict__ depth, const u32x4* __restrict__ acc, int w,
                                            int h, int x, int y, int window) {
    float dmin = 1000000.0f;
    for (int ox = -window; ox <= window; ox++)
        for (int oy = -window; oy <= window; oy++) {
            const int px = x + ox, py = y + oy;
            if (px < 0 || py < 0 || px >= w || py >= h) continue;
            const float pd = __uint_as_float(depth[(uint64_t)py * w + px]);
            if (pd >= 0.0f) dmin = pd < dmin ? pd : dmin;
        }
    float R = 0.0f, G = 0.0f, B = 0.0f, C = 0.0f;
    for (int ox = -window; ox <= window; ox++)
        for (int oy = -window; oy <= window; oy++) {
            const int px = x + ox, py = y + oy;
            if (px < 0 || py < 0 || px >= w || py >= h) continue;
            const uint64_t q = (uint64_t)py * w + px;
            const float pd = __uint_as_float(depth[q]);
            float wt = (ox == 0 && oy == 0) ? 100.0f : (ox <= 1 && oy <= 1) ? 2.0f : 1.0f;
            if (pd > dmin * 1.01f) wt = 0.0f;
            const u32x4 a = acc[q];
            R += (float)a.x * wt;
            G += (float)a.y * wt;
            B += (float)a.z * wt;
            C += (float)a.w * wt;
        }
    uint32_t color = 0u;
    if (C != 0.0f) color = (uint32_t)(R / C) | ((uint32_t)(G / C) << 8) | ((uint32_t)(B / C) << 16);
    if (dmin < 0.0f) color = 0x0000FF00u;
    return color;
}

// e^x for x <= 0, for the eye-dome lighting shade. The oracle runs the same
// sequence (orc_edl_exp), so the shade has the same bits on both sides.
// Method: Cody-Waite range reduction x = k ln2 + r, with ln2 split into a
// high part of 17 significant bits (k * ln2_hi is exact for |k| < 128) and a
// low part; e^r from its Taylor series to r^7 in Horner form (|r| <= ln2/2,
// truncation error below 2^-27 relative); then scaled by 2^k with ldexp.
// x <= -86 returns 0: e^-86 < 2^-124, and an 8-bit channel times it
// truncates to 0 like the true value.
__device__ inline float edl_exp(float x) {
    if (!(x > -86.0f)) return 0.0f;
    const float k = __builtin_rintf(x * 1.44269504f);
    float r = fma_(-k, 0.693145751953125f, x);
    r = fma_(-k, 1.42860677e-06f, r);
    float p = 1.98412698e-04f;           // 1/7!
    p = fma_(p, r, 1.38888889e-03f);     // 1/6!
    p = fma_(p, r, 8.33333333e-03f);     // 1/5!
    p = fma_(p, r, 4.16666667e-02f);     // 1/4!
    p = fma_(p, r, 1.66666667e-01f);     // 1/3!
    p = fma_(p, r, 0.5f);
    p = fma_(p, r, 1.0f);
    p = fma_(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)k);
}

// getDepthAt(pixel, window 0) of compute_loop_las/resolve.cs:43-64: the
// pixel's depth if it is in (0, 1e8), else 1e8. A neighbour outside the
// image counts as background (the shader reads past the row or buffer).
__device__ __forceinline__ float edl_depth(const unsigned long long* __restrict__ fb, int w, int h, int x, int y) {
    if (x < 0 || y < 0 || x >= w || y >= h) return 100000000.0f;
    const float d = __uint_as_float((uint32_t)(fb[(uint64_t)y * w + x] >> 32));
    return (d > 0.0f && d < 100000000.0f) ? d : 100000000.0f;
}

// Resolve kernels walk the image in 16x16-pixel tiles, one 8x8 quadrant per
// wave of a 256-thread workgroup: a wave's colour gathers then come from a
// few neighbouring batches, whose colour-plane lines it shares, rather than
// from a whole row of them. Sets (x, y) of this thread's pixel in `tile`.
__device__ __forceinline__ void tile_pixel(uint32_t tile, uint32_t tiles_x, uint32_t& x, uint32_t& y) {
    const uint32_t lane = threadIdx.x & 63u, wv = (threadIdx.x >> 6) & 3u;
    x = (tile % tiles_x) * 16u + (wv & 1u) * 8u + (lane & 7u);
    y = (tile / tiles_x) * 16u + (wv >> 1) * 8u + (lane >> 3);
}

// The colour of the point in pixel (x, y) for an index-payload framebuffer
// (compute_loop_las/resolve.cs:100-146): bg for a clear key or an index
// outside the cloud. With strength > 0, the EDL block the shader carries
// switched off (resolve.cs:148-191): shade = exp(-(sum / 9) * 300 *
// strength), sum adding max(0, depth - neighbour depth) over the 3x3 window
// in ox-major order; channels truncate to integers and alpha is 0.
__device__ inline uint32_t resolve_index_pixel(const unsigned long long* __restrict__ fb, int w, int h, int x, int y,
                                               const uint32_t* __restrict__ rgba, uint64_t npoints, uint32_t bg,
                                               float strength) {
    const unsigned long long key = fb[(uint64_t)y * w + x];
    const uint32_t id = (uint32_t)key;
    if (key == PCR_FB_CLEAR || id >= npoints) return bg;
    const uint32_t color = rgba[id];
    if (!(strength > 0.0f)) return color;
    const float depth = __uint_as_float((uint32_t)(key >> 32));
    float sum = 0.0f;
    for (int ox = -1; ox <= 1; ox++)
        for (int oy = -1; oy <= 1; oy++) {
            const float dd = depth - edl_depth(fb, w, h, x + ox, y + oy);
            sum += dd > 0.0f ? dd : 0.0f;
        }
    const float response = sum / 9.0f;
    const float shade = edl_exp(-response * 300.0f * strength);
    const uint32_t r = (uint32_t)((float)(color & 255u) * shade);
    const uint32_t g = (uint32_t)((float)((color >> 8) & 255u) * shade);
    const uint32_t b = (uint32_t)((float)((color >> 16) & 255u) * shade);
    return r | (g << 8) | (b << 16);
}

}  // namespace pcr
