/*
 * pcr_oracle.c — CPU oracle (TEST INFRASTRUCTURE: only tests/, smoke() and
 * bench.py's cpu_baseline leg load it, to check the GPU's results against).
 *
 * Restates, in plain C and a fixed fp32 operation order (see pcr_oracle.h),
 * the reference's per-frame logic:
 *   encode        modules/compute/LasLoaderSparse.cpp:37-212
 *   cull          modules/compute_loop_las/render.cs:122-197 (three.js planes)
 *   precision     modules/compute_loop_las/render.cs:199-231
 *   decode        modules/compute_loop_las2/render.cs:319-469 (Y/Z of level 0 as las2)
 *   project/pack  modules/compute_loop_las/render.cs:399-426,
 *                 modules/compute_2021_earlyz/render.cs:35-63
 *   resolve       modules/compute_loop_las/resolve.cs:100-146,
 *                 modules/compute_2021_earlyz/resolve.cs:41-78
 *   hqs           modules/compute_loop_las_hqs/{depth.cs:242-389, color.cs:242-448, resolve.cs:62-129}
 *   nodes         modules/compute_loop_nodes/render.cs:144-330
 *   morton        include/utils.h:14-27
 *   camera        include/Camera.h:34-38, libs/glm/glm/gtc/matrix_transform.inl:258-274
 * Decisions where the reference is undefined or inconsistent (each pinned by
 * a test): level-0 Y/Z use their own low bits; the batch loop bound is
 * `i < numPoints`; a point whose truncated pixel equals width/height is
 * rejected; NaN positions are outside; background is `key == clear value`;
 * a zero-extent AABB axis quantises to 0; the early-out compare is on the
 * full 64-bit key so the result is the order-independent minimum; hqs
 * accumulates four separate uint32 sums.
 */
#include "pcr_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>

/* ------------------------------------------------------------------ PRNG */
uint64_t orc_mix64(uint64_t z) {               /* SplitMix64 finaliser */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t orc_rand(uint64_t key, uint64_t i) { return orc_mix64(key ^ (i * 0xD6E8FEB86659FD93ull)); }
float orc_u01(uint64_t r) { return (float)(uint32_t)(r >> 40) * 5.9604644775390625e-08f; }
static uint64_t skey(uint64_t seed, uint64_t stream) { return orc_mix64(orc_mix64(seed) + stream); }

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------ generators */
static float lattice(uint64_t kn, int32_t ix, int32_t iy, uint32_t oct) {
    uint64_t h = orc_mix64(kn ^ ((uint64_t)(uint32_t)ix * 0x9E3779B97F4A7C15ull)
                              ^ ((uint64_t)(uint32_t)iy * 0xC2B2AE3D27D4EB4Full)
                              ^ ((uint64_t)oct << 59));
    return orc_u01(h) * 2.0f - 1.0f;
}
static float value_noise(uint64_t kn, float x, float y, uint32_t oct) {
    float fx = floorf(x), fy = floorf(y);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy;
    float tx = x - fx, ty = y - fy;
    float sx = (tx * tx) * (3.0f - 2.0f * tx);
    float sy = (ty * ty) * (3.0f - 2.0f * ty);
    float v00 = lattice(kn, ix, iy, oct), v10 = lattice(kn, ix + 1, iy, oct);
    float v01 = lattice(kn, ix, iy + 1, oct), v11 = lattice(kn, ix + 1, iy + 1, oct);
    float a = v00 + (v10 - v00) * sx;
    float b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sy;
}
/* 6-octave fBm, base wavelength 512 m, amplitude halves per octave; height in [0, amplitude] */
float orc_terrain_height(uint64_t seed, float extent, float amplitude, float x, float y) {
    (void)extent;
    uint64_t kn = skey(seed, 7);
    float sum = 0.0f, amp = 1.0f, fx = x * (1.0f / 512.0f), fy = y * (1.0f / 512.0f);
    for (uint32_t o = 0; o < 6; o++) {
        sum = sum + value_noise(kn, fx, fy, o) * amp;
        fx = fx * 2.0f; fy = fy * 2.0f; amp = amp * 0.5f;
    }
    float f = sum / 1.96875f;
    return amplitude * (f * 0.5f + 0.5f);
}
static uint32_t clamp255(int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

static void terrain_point(const pcr_gen_cfg* c, uint64_t i, float* p, uint32_t* rgba) {
    uint64_t k = skey(c->seed, 1);
    uint64_t r0 = orc_rand(k, 8 * i + 0), r1 = orc_rand(k, 8 * i + 1), r2 = orc_rand(k, 8 * i + 2);
    uint64_t r3 = orc_rand(k, 8 * i + 3), r4 = orc_rand(k, 8 * i + 4), r5 = orc_rand(k, 8 * i + 5);
    float u = orc_u01(r0) * c->extent, v = orc_u01(r1) * c->extent;
    float h = orc_terrain_height(c->seed, c->extent, c->amplitude, u, v);
    float jx = (orc_u01(r2) * 2.0f - 1.0f) * c->jitter;
    float jy = (orc_u01(r3) * 2.0f - 1.0f) * c->jitter;
    float jz = (orc_u01(r4) * 2.0f - 1.0f) * c->jitter;
    p[0] = u + jx; p[1] = v + jy; p[2] = h + jz;
    float t = h / c->amplitude;
    float cr, cg, cb;
    if (t < 0.5f) {
        float s = t * 2.0f;
        cr = 60.0f + (140.0f - 60.0f) * s; cg = 110.0f + (110.0f - 110.0f) * s; cb = 40.0f + (70.0f - 40.0f) * s;
    } else {
        float s = t * 2.0f - 1.0f;
        cr = 140.0f + (235.0f - 140.0f) * s; cg = 110.0f + (235.0f - 110.0f) * s; cb = 70.0f + (235.0f - 70.0f) * s;
    }
    int n = (int)(r5 >> 59) - 16;
    *rgba = clamp255((int)cr + n) | (clamp255((int)cg + n) << 8) | (clamp255((int)cb + n) << 16) | 0xFF000000u;
}

typedef struct { float x0, y0, z0, sx, sy, sz; uint32_t rgb; uint64_t end; } orc_building;

static orc_building* city_table(const pcr_gen_cfg* c, uint64_t n_build) {
    int nb = c->num_buildings;
    orc_building* t = (orc_building*)calloc((size_t)nb, sizeof(orc_building));
    double* w = (double*)calloc((size_t)nb, sizeof(double));
    uint64_t kb = skey(c->seed, 2);
    double wsum = 0.0;
    for (int b = 0; b < nb; b++) {
        uint64_t r[7];
        for (int k = 0; k < 7; k++) r[k] = orc_rand(kb, 8 * (uint64_t)b + (uint64_t)k);
        float cx = (0.05f + 0.9f * orc_u01(r[0])) * c->extent;
        float cy = (0.05f + 0.9f * orc_u01(r[1])) * c->extent;
        float sx = 10.0f + 70.0f * orc_u01(r[2]);
        float sy = 10.0f + 70.0f * orc_u01(r[3]);
        float sz = 10.0f + 70.0f * orc_u01(r[4]);
        float sp = 0.02f + 0.03f * orc_u01(r[5]);
        t[b].x0 = cx - sx * 0.5f; t[b].y0 = cy - sy * 0.5f;
        t[b].z0 = orc_terrain_height(c->seed, c->extent, c->amplitude, cx, cy);
        t[b].sx = sx; t[b].sy = sy; t[b].sz = sz;
        t[b].rgb = (80u + (uint32_t)(r[6] & 127u)) | ((80u + (uint32_t)((r[6] >> 8) & 127u)) << 8)
                 | ((80u + (uint32_t)((r[6] >> 16) & 127u)) << 16);
        double area = 2.0 * ((double)sx * (double)sz) + 2.0 * ((double)sy * (double)sz) + (double)sx * (double)sy;
        w[b] = area / ((double)sp * (double)sp);
        wsum += w[b];
    }
    double cum = 0.0;
    for (int b = 0; b < nb; b++) {
        cum += w[b];
        t[b].end = (b == nb - 1) ? n_build : (uint64_t)floor((double)n_build * (cum / wsum));
    }
    free(w);
    return t;
}

static void building_point(const pcr_gen_cfg* c, const orc_building* t, uint64_t j, float* p, uint32_t* rgba) {
    int lo = 0, hi = c->num_buildings - 1;
    while (lo < hi) { int m = (lo + hi) / 2; if (j < t[m].end) hi = m; else lo = m + 1; }
    const orc_building* B = &t[lo];
    uint64_t kp = skey(c->seed, 3);
    uint64_t r0 = orc_rand(kp, 4 * j + 0), r1 = orc_rand(kp, 4 * j + 1);
    uint64_t r2 = orc_rand(kp, 4 * j + 2), r3 = orc_rand(kp, 4 * j + 3);
    float ax = B->sx * B->sz, ay = B->sy * B->sz, ar = B->sx * B->sy;
    float t0 = ax, t1 = t0 + ax, t2 = t1 + ay, t3 = t2 + ay, t4 = t3 + ar;
    float a = orc_u01(r0) * t4, s = orc_u01(r1), u = orc_u01(r2);
    if (a < t0)      { p[0] = B->x0 + s * B->sx; p[1] = B->y0;          p[2] = B->z0 + u * B->sz; }
    else if (a < t1) { p[0] = B->x0 + s * B->sx; p[1] = B->y0 + B->sy;  p[2] = B->z0 + u * B->sz; }
    else if (a < t2) { p[0] = B->x0;          p[1] = B->y0 + s * B->sy; p[2] = B->z0 + u * B->sz; }
    else if (a < t3) { p[0] = B->x0 + B->sx;  p[1] = B->y0 + s * B->sy; p[2] = B->z0 + u * B->sz; }
    else             { p[0] = B->x0 + s * B->sx; p[1] = B->y0 + u * B->sy; p[2] = B->z0 + B->sz; }
    int n = (int)(r3 >> 59) - 16;
    *rgba = clamp255((int)(B->rgb & 255u) + n) | (clamp255((int)((B->rgb >> 8) & 255u) + n) << 8)
          | (clamp255((int)((B->rgb >> 16) & 255u) + n) << 16) | 0xFF000000u;
}

void orc_generate(const pcr_gen_cfg* c, uint64_t first, uint64_t count, float* xyz, uint32_t* rgba) {
    if (c->scene == PCR_SCENE_CUBE) {
        uint64_t k = skey(c->seed, 0);
        for (uint64_t q = 0; q < count; q++) {
            uint64_t i = first + q;
            float x = orc_u01(orc_rand(k, 3 * i)), y = orc_u01(orc_rand(k, 3 * i + 1)), z = orc_u01(orc_rand(k, 3 * i + 2));
            xyz[3 * q] = x; xyz[3 * q + 1] = y; xyz[3 * q + 2] = z;
            rgba[q] = (uint32_t)(uint8_t)(255.0f * x) | ((uint32_t)(uint8_t)(255.0f * y) << 8)
                    | ((uint32_t)(uint8_t)(255.0f * z) << 16) | 0xFF000000u;
        }
    } else if (c->scene == PCR_SCENE_TERRAIN) {
        for (uint64_t q = 0; q < count; q++) terrain_point(c, first + q, xyz + 3 * q, rgba + q);
    } else {
        uint64_t n_terrain = (uint64_t)((double)c->num_points * (double)c->terrain_frac);
        orc_building* t = city_table(c, c->num_points - n_terrain);
        for (uint64_t q = 0; q < count; q++) {
            uint64_t i = first + q;
            if (i < n_terrain) terrain_point(c, i, xyz + 3 * q, rgba + q);
            else building_point(c, t, i - n_terrain, xyz + 3 * q, rgba + q);
        }
        free(t);
    }
}

/* ---------------------------------------------------------------- morton */
static uint64_t spread3(uint32_t v) {            /* 21 bits -> every third bit */
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8))  & 0x100F00F00F00F00Full;
    x = (x | (x << 4))  & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2))  & 0x1249249249249249ull;
    return x;
}
/* include/utils.h:24-27 convention: x << 2 | y << 1 | z */
uint64_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

typedef struct { uint64_t key; uint32_t idx; } kv;
static int kv_cmp(const void* a, const void* b) {
    const kv* x = (const kv*)a; const kv* y = (const kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static uint32_t quant21(float x, float mn, double scale) {
    double q = ((double)x - (double)mn) * scale;
    return q >= 2097151.0 ? 2097151u : (uint32_t)q;
}
void orc_morton_order(const float* xyz, uint64_t n, uint32_t* perm, uint64_t* keys_out) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) { float v = xyz[3 * i + a]; mn[a] = v < mn[a] ? v : mn[a]; mx[a] = v > mx[a] ? v : mx[a]; }
    float side = mx[0] - mn[0];
    if (mx[1] - mn[1] > side) side = mx[1] - mn[1];
    if (mx[2] - mn[2] > side) side = mx[2] - mn[2];
    double scale = side > 0.0f ? 2097152.0 / (double)side : 0.0;
    kv* a = (kv*)malloc(sizeof(kv) * (n ? n : 1));
    for (uint64_t i = 0; i < n; i++) {
        a[i].idx = (uint32_t)i;
        a[i].key = orc_morton3(quant21(xyz[3 * i], mn[0], scale), quant21(xyz[3 * i + 1], mn[1], scale),
                               quant21(xyz[3 * i + 2], mn[2], scale));
    }
    qsort(a, n, sizeof(kv), kv_cmp);
    for (uint64_t i = 0; i < n; i++) { perm[i] = a[i].idx; if (keys_out) keys_out[i] = a[i].key; }
    free(a);
}

/* ---------------------------------------------------------------- encode */
/* LasLoaderSparse.cpp:146-152 with the zero-extent decision; a position
 * below the box (possible against a Potree node's float box) clamps to 0
 * instead of converting a negative double to uint32 */
void orc_quantize30(double x, double mn, double size, uint32_t* q) {
    if (size == 0.0) { *q = 0; return; }
    const double f = ((x - mn) / size) * 1073741824.0;
    if (!(f > 0.0)) { *q = 0; return; }
    uint32_t v = f < 1073741823.0 ? (uint32_t)f : 1073741823u;
    *q = v;
}
void orc_encode_batches(const float* xyz, uint64_t n, uint32_t bs, pcr_batch* batches,
                        uint32_t* lo, uint32_t* med, uint32_t* hi) {
    uint64_t nb = (n + bs - 1) / bs;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t first = b * bs, cnt = (n - first) < bs ? (n - first) : bs;
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint64_t i = first; i < first + cnt; i++)
            for (int a = 0; a < 3; a++) { float v = xyz[3 * i + a]; mn[a] = v < mn[a] ? v : mn[a]; mx[a] = v > mx[a] ? v : mx[a]; }
        pcr_batch* h = &batches[b];
        memset(h, 0, sizeof(*h));
        for (int a = 0; a < 3; a++) { h->min[a] = mn[a]; h->max[a] = mx[a]; }
        h->num_points = (uint32_t)cnt; h->first_point = (uint32_t)first; h->parent = -1;
        double size[3];
        for (int a = 0; a < 3; a++) size[a] = (double)mx[a] - (double)mn[a];
        for (uint64_t i = first; i < first + cnt; i++) {
            uint32_t q[3];
            for (int a = 0; a < 3; a++) orc_quantize30((double)xyz[3 * i + a], (double)mn[a], size[a], &q[a]);
            lo[i]  = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
            med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
            hi[i]  = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
        }
    }
}

/* ------------------------------------------------------------ LAS ingest */
static uint16_t rd_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }
static uint32_t rd_u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rd_u64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
static double   rd_f64(const uint8_t* p) { double v; memcpy(&v, p, 8); return v; }

/* LasLoaderSparse.cpp:495-526 (header fields), :123-132 (colour offset) */
int orc_las_parse_header(const uint8_t* b, uint64_t size, pcr_las_info* out) {
    memset(out, 0, sizeof(*out));
    if (size < 227 || memcmp(b, "LASF", 4) != 0) return -1;
    out->version_major = b[24];
    out->version_minor = b[25];
    const int v14 = !(out->version_major == 1 && out->version_minor < 4);
    if (v14 && size < 255) return -1;
    uint64_t n = v14 ? rd_u64(b + 247) : rd_u32(b + 107);
    out->num_points = n < 1000000000ull ? n : 1000000000ull;
    out->offset_to_points = rd_u32(b + 96);
    out->point_format = (uint8_t)(b[104] % 128);
    out->record_length = rd_u16(b + 105);
    const int f = out->point_format;
    out->rgb_offset = (uint8_t)(f == 2 ? 20 : f == 3 ? 28 : (f == 7 || f == 8) ? 30 : 0);
    for (int a = 0; a < 3; a++) {
        out->scale[a] = rd_f64(b + 131 + 8 * a);
        out->offset[a] = rd_f64(b + 155 + 8 * a);
        out->max[a] = rd_f64(b + 179 + 16 * a);
        out->min[a] = rd_f64(b + 187 + 16 * a);
    }
    if (out->record_length < 12 || out->record_length < out->rgb_offset + 6) return -1;
    if ((uint64_t)out->offset_to_points + out->num_points * out->record_length > size) return -1;
    return 0;
}

/* LasLoaderSparse.cpp:37-212: positions in double relative to the header
 * min, per-batch double AABB, float header, double 30-bit quantisation,
 * 16-to-8-bit colour. */
void orc_las_encode(const uint8_t* rec, uint64_t n, const pcr_las_info* info, uint32_t bs, uint64_t sparse,
                    uint32_t file_index, pcr_batch* batches, uint32_t* lo, uint32_t* med, uint32_t* hi,
                    uint32_t* rgba) {
    const uint64_t nb = (n + bs - 1) / bs, L = info->record_length;
    for (uint64_t bi = 0; bi < nb; bi++) {
        const uint64_t first = bi * bs, cnt = (n - first) < bs ? (n - first) : bs;
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint64_t i = first; i < first + cnt; i++)
            for (int a = 0; a < 3; a++) {
                const double x = (double)(int32_t)rd_u32(rec + i * L + 4 * a) * info->scale[a] + info->offset[a] -
                                 info->min[a];
                mn[a] = x < mn[a] ? x : mn[a];
                mx[a] = x > mx[a] ? x : mx[a];
            }
        pcr_batch* h = &batches[bi];
        memset(h, 0, sizeof(*h));
        for (int a = 0; a < 3; a++) { h->min[a] = (float)mn[a]; h->max[a] = (float)mx[a]; }
        h->num_points = (uint32_t)cnt;
        h->first_point = (uint32_t)(sparse + first);
        h->level = file_index;
        h->parent = -1;
        const double size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
        for (uint64_t i = first; i < first + cnt; i++) {
            const uint8_t* r = rec + i * L;
            uint32_t q[3];
            for (int a = 0; a < 3; a++) {
                const double x = (double)(int32_t)rd_u32(r + 4 * a) * info->scale[a] + info->offset[a] - info->min[a];
                orc_quantize30(x, mn[a], size[a], &q[a]);
            }
            lo[i]  = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
            med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
            hi[i]  = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
            uint32_t c[3];
            for (int k = 0; k < 3; k++) {
                const uint32_t v = rd_u16(r + info->rgb_offset + 2 * k);
                c[k] = v < 256 ? v : v / 256;
            }
            rgba[i] = c[0] | (c[1] << 8) | (c[2] << 16);
        }
    }
}

/* ---------------------------------------------------------- Potree 2.0 */
typedef struct {
    double box[6];
    float spacing;
    int level, type, mask, parent, child[8];
    uint32_t num_points;
    int64_t byte_offset, byte_size, hier_offset, hier_size;
    uint64_t dfs;
} potree_node;

typedef struct {
    potree_node* n;
    uint64_t count, cap;
} potree_pool;

static int pool_add(potree_pool* p) {
    if (p->count == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 256;
        p->n = (potree_node*)realloc(p->n, p->cap * sizeof(potree_node));
    }
    memset(&p->n[p->count], 0, sizeof(potree_node));
    for (int k = 0; k < 8; k++) p->n[p->count].child[k] = -1;
    p->n[p->count].parent = -1;
    return (int)p->count++;
}

/* PotreeData.h:115-139: child box halved in float, bit 0 = z, 1 = y, 2 = x */
static void child_box(const double* b, int index, double* out) {
    float mn[3] = {(float)b[0], (float)b[1], (float)b[2]}, mx[3] = {(float)b[3], (float)b[4], (float)b[5]};
    float size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    for (int bit = 0; bit < 3; bit++) {
        const int a = 2 - bit;
        if (index & (1 << bit)) mn[a] += size[a] / 2.0f; else mx[a] -= size[a] / 2.0f;
    }
    for (int a = 0; a < 3; a++) { out[a] = mn[a]; out[3 + a] = mx[a]; }
}

/* PotreeData.h:185-248: one hierarchy chunk, whose first record is `root` */
static int parse_chunk(potree_pool* p, int root, const uint8_t* h, uint64_t hsize) {
    const int64_t off = p->n[root].hier_offset, size = p->n[root].hier_size;
    if (off < 0 || size < 0 || size % 22 || (uint64_t)(off + size) > hsize || size == 0) return -1;
    const int64_t nrec = size / 22;
    int* list = (int*)malloc(sizeof(int) * nrec);
    int64_t fill = 1;
    list[0] = root;
    int rc = 0;
    for (int64_t i = 0; i < nrec && rc == 0; i++) {
        if (i >= fill) { rc = -1; break; }
        potree_node* c = &p->n[list[i]];
        const uint8_t* r = h + off + 22 * i;
        const int type = r[0], mask = r[1];
        const uint32_t np = rd_u32(r + 2);
        const int64_t bo = (int64_t)rd_u64(r + 6), bsz = (int64_t)rd_u64(r + 14);
        if (c->type == 2) { c->byte_offset = bo; c->byte_size = bsz; c->num_points = np; }
        else if (type == 2) { c->hier_offset = bo; c->hier_size = bsz; c->num_points = np; }
        else { c->byte_offset = bo; c->byte_size = bsz; c->num_points = np; }
        c->type = type;
        c->mask = mask;
        if (type == 2) continue;
        for (int k = 0; k < 8; k++) {
            if (!(mask & (1 << k))) continue;
            if (fill >= nrec) { rc = -1; break; }
            const int ci = pool_add(p);
            c = &p->n[list[i]];   /* the pool may have moved */
            potree_node* ch = &p->n[ci];
            child_box(c->box, k, ch->box);
            ch->spacing = c->spacing / 2.0f;
            ch->level = c->level + 1;
            ch->parent = list[i];
            c->child[k] = ci;
            list[fill++] = ci;
        }
    }
    for (int64_t i = 0; i < fill && rc == 0; i++)
        if (p->n[list[i]].type == 2) rc = parse_chunk(p, list[i], h, hsize);
    free(list);
    return rc;
}

static void dfs_order(potree_pool* p, int i, uint64_t* counter) {
    p->n[i].dfs = (*counter)++;
    for (int k = 0; k < 8; k++) if (p->n[i].child[k] >= 0) dfs_order(p, p->n[i].child[k], counter);
}

static const potree_node* g_sort_nodes;
static int cmp_offset(const void* a, const void* b) {
    const potree_node* x = &g_sort_nodes[*(const int*)a];
    const potree_node* y = &g_sort_nodes[*(const int*)b];
    if (x->byte_offset != y->byte_offset) return x->byte_offset < y->byte_offset ? -1 : 1;
    return x->dfs < y->dfs ? -1 : (x->dfs > y->dfs);
}

/* PotreeData.h:250-317 (traverse, then sort by byteOffset; ties keep the
 * traversal order) and :368-380 (the batch header) */
int64_t orc_potree_nodes(const pcr_potree_info* info, const uint8_t* h, uint64_t hsize, pcr_batch* nodes,
                         double* boxes, uint64_t max_nodes) {
    potree_pool p = {0, 0, 0};
    const int root = pool_add(&p);
    p.n[root].type = 2;
    p.n[root].hier_offset = 0;
    p.n[root].hier_size = (int64_t)info->first_chunk_size;
    p.n[root].spacing = (float)info->spacing;
    for (int a = 0; a < 3; a++) { p.n[root].box[a] = 0.0; p.n[root].box[3 + a] = info->max[a] - info->min[a]; }
    int rc = parse_chunk(&p, root, h, hsize);
    int64_t out = -1;
    if (rc == 0 && p.count <= max_nodes && info->bytes_per_point > 0) {
        uint64_t counter = 0;
        dfs_order(&p, root, &counter);
        int* idx = (int*)malloc(sizeof(int) * p.count);
        for (uint64_t i = 0; i < p.count; i++) idx[i] = (int)i;
        g_sort_nodes = p.n;
        qsort(idx, p.count, sizeof(int), cmp_offset);
        int* pos = (int*)malloc(sizeof(int) * p.count);
        for (uint64_t i = 0; i < p.count; i++) pos[idx[i]] = (int)i;
        out = (int64_t)p.count;
        for (uint64_t i = 0; i < p.count && out >= 0; i++) {
            const potree_node* n = &p.n[idx[i]];
            if (n->byte_offset < 0 || n->byte_offset % info->bytes_per_point ||
                (uint64_t)n->byte_offset / info->bytes_per_point + n->num_points > info->num_points) {
                out = -1;
                break;
            }
            pcr_batch* b = &nodes[i];
            memset(b, 0, sizeof(*b));
            for (int a = 0; a < 3; a++) { b->min[a] = (float)n->box[a]; b->max[a] = (float)n->box[3 + a]; }
            b->num_points = n->num_points;
            b->first_point = (uint32_t)(n->byte_offset / info->bytes_per_point);
            b->level = (uint32_t)n->level;
            b->parent = n->parent >= 0 ? pos[n->parent] : -1;
            b->spacing = n->spacing;
            for (int a = 0; a < 6; a++) boxes[6 * i + a] = n->box[a];
        }
        free(pos);
        free(idx);
    }
    free(p.n);
    return out;
}

/* PotreeData.h:439-511 with the clamped quantisation and bits 9..0 in the
 * fine plane */
void orc_potree_encode(const pcr_potree_info* info, const uint8_t* oct, const pcr_batch* nodes,
                       const double* boxes, uint64_t nn, uint32_t* lo, uint32_t* med, uint32_t* hi, uint32_t* rgba) {
    const uint64_t L = info->bytes_per_point;
    for (uint64_t k = 0; k < nn; k++) {
        const double* bx = boxes + 6 * k;
        for (uint32_t j = 0; j < nodes[k].num_points; j++) {
            const uint64_t i = (uint64_t)nodes[k].first_point + j;
            const uint8_t* r = oct + i * L;
            uint32_t q[3];
            for (int a = 0; a < 3; a++) {
                const double x = (double)(int32_t)rd_u32(r + 4 * a) * info->scale[a] + info->offset[a] - info->min[a];
                orc_quantize30(x, bx[a], bx[3 + a] - bx[a], &q[a]);
            }
            lo[i]  = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
            med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
            hi[i]  = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
            uint32_t c[3];
            for (int m = 0; m < 3; m++) {
                const uint32_t v = rd_u16(r + info->rgb_offset + 2 * m);
                c[m] = v < 256 ? v : v / 256;
            }
            rgba[i] = c[0] | (c[1] << 8) | (c[2] << 16);
        }
    }
}

/* ------------------------------------------------------- per-batch stages */
static void mul_point(const float m[16], float x, float y, float z, float r[4]) {
    for (int i = 0; i < 4; i++) r[i] = fmaf(m[8 + i], z, fmaf(m[4 + i], y, fmaf(m[i], x, m[12 + i])));
}
static void mul_vec4(const float m[16], const float v[4], float r[4]) {
    for (int i = 0; i < 4; i++) r[i] = fmaf(m[12 + i], v[3], fmaf(m[8 + i], v[2], fmaf(m[4 + i], v[1], m[i] * v[0])));
}

/* render.cs:134-197 */
int orc_frustum_visible(const float m[16], const float mn[3], const float mx[3]) {
    float pl[6][4];
    for (int k = 0; k < 4; k++) {
        pl[0][k] = m[4 * k + 3] - m[4 * k + 0];
        pl[1][k] = m[4 * k + 3] + m[4 * k + 0];
        pl[2][k] = m[4 * k + 3] + m[4 * k + 1];
        pl[3][k] = m[4 * k + 3] - m[4 * k + 1];
        pl[4][k] = m[4 * k + 3] - m[4 * k + 2];
        pl[5][k] = m[4 * k + 3] + m[4 * k + 2];
    }
    for (int i = 0; i < 6; i++) {
        float len = sqrtf(fmaf(pl[i][2], pl[i][2], fmaf(pl[i][1], pl[i][1], pl[i][0] * pl[i][0])));
        float nx = pl[i][0] / len, ny = pl[i][1] / len, nz = pl[i][2] / len, c = pl[i][3] / len;
        float vx = nx > 0.0f ? mx[0] : mn[0];
        float vy = ny > 0.0f ? mx[1] : mn[1];
        float vz = nz > 0.0f ? mx[2] : mn[2];
        float d = fmaf(nz, vz, fmaf(ny, vy, nx * vx)) + c;
        if (d < 0.0f) return 0;
    }
    return 1;
}

/* render.cs:199-215: screen length of a view-space x offset `len` at the box centre */
float orc_projected_px(const float mn[3], const float mx[3], float len,
                       const float vw[16], const float proj[16], int w, int h) {
    float cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
    float vc[4], ve[4], pc[4], pe[4];
    mul_point(vw, cx, cy, cz, vc);
    ve[0] = vc[0] + len; ve[1] = vc[1]; ve[2] = vc[2]; ve[3] = vc[3];
    mul_vec4(proj, vc, pc);
    mul_vec4(proj, ve, pe);
    float pcx = pc[0] / pc[3], pcy = pc[1] / pc[3], pex = pe[0] / pe[3], pey = pe[1] / pe[3];
    float fw = (float)w, fh = (float)h;
    float scx = (fw * (pcx + 1.0f)) * 0.5f, scy = (fh * (pcy + 1.0f)) * 0.5f;
    float sex = (fw * (pex + 1.0f)) * 0.5f, sey = (fh * (pey + 1.0f)) * 0.5f;
    float dx = sex - scx, dy = sey - scy;
    return sqrtf(fmaf(dy, dy, dx * dx));
}

/* render.cs:199-231 (first_threshold 100); nodes/render.cs:180-212 (80) */
int orc_precision_level(const float mn[3], const float mx[3], const pcr_view* v, float first_threshold) {
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    float radius = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    float px = orc_projected_px(mn, mx, radius, v->view_world, v->proj, v->width, v->height);
    if (px < first_threshold) return 4;
    if (px < 200.0f) return 3;
    if (px < 500.0f) return 2;
    if (px < 10000.0f) return 1;
    return 0;
}

/* las2/render.cs:340-366 (level 0), :402-424 (level 1), :450-458 (levels 2-4) */
void orc_decode(int level, uint32_t lo, uint32_t med, uint32_t hi,
                const float mn[3], const float mx[3], float out[3]) {
    for (int a = 0; a < 3; a++) {
        float size = mx[a] - mn[a];
        uint32_t sh = 10u * (uint32_t)a;
        uint32_t l = (lo >> sh) & 1023u, m = (med >> sh) & 1023u, hh = (hi >> sh) & 1023u;
        if (level == 0) {
            uint32_t X = (l << 20) | (m << 10) | hh;
            out[a] = fmaf((float)X, size * 9.3132257461547852e-10f, mn[a]);
        } else if (level == 1) {
            uint32_t X = (l << 20) | (m << 10);
            out[a] = fmaf((float)X, size * 9.3132257461547852e-10f, mn[a]);
        } else {
            out[a] = fmaf((float)l, size * 9.765625e-04f, mn[a]);
        }
    }
}

/* orc_decode over n points sharing one box (out: 3 floats per point) */
void orc_decode_n(int level, uint64_t n, const uint32_t* lo, const uint32_t* med, const uint32_t* hi,
                  const float mn[3], const float mx[3], float* out) {
    for (uint64_t i = 0; i < n; i++) orc_decode(level, lo[i], med ? med[i] : 0u, hi ? hi[i] : 0u, mn, mx, out + 3 * i);
}

/* render.cs:400-409 with the edge decision (pixel == width/height rejected) */
int orc_project(const float t[16], const float p[3], int w, int h, uint32_t* pixel, uint32_t* depth_bits) {
    float px = fmaf(t[8], p[2], fmaf(t[4], p[1], fmaf(t[0], p[0], t[12])));
    float py = fmaf(t[9], p[2], fmaf(t[5], p[1], fmaf(t[1], p[0], t[13])));
    float pw = fmaf(t[11], p[2], fmaf(t[7], p[1], fmaf(t[3], p[0], t[15])));
    float nx = px / pw, ny = py / pw;
    if (!(pw > 0.0f && nx >= -1.0f && nx <= 1.0f && ny >= -1.0f && ny <= 1.0f)) return 0;
    float sx = fmaf(nx, 0.5f, 0.5f) * (float)w;
    float sy = fmaf(ny, 0.5f, 0.5f) * (float)h;
    int ix = (int)sx, iy = (int)sy;
    if (ix >= w || iy >= h) return 0;
    *pixel = (uint32_t)ix + (uint32_t)iy * (uint32_t)w;
    *depth_bits = fbits(pw);
    return 1;
}

/* ------------------------------------------------------------ full frames */
void orc_clear64(uint64_t* fb, uint64_t npix) { for (uint64_t i = 0; i < npix; i++) fb[i] = PCR_FB_CLEAR; }

/* compute_2021_earlyz/render.cs:35-63 (payload = RGBA) */
void orc_render_points16(const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_view* v, uint64_t* fb) {
    for (uint64_t i = 0; i < n; i++) {
        uint32_t pix, d;
        if (!orc_project(v->transform, xyz + 3 * i, v->width, v->height, &pix, &d)) continue;
        uint64_t key = ((uint64_t)d << 32) | rgba[i];
        if (key < fb[pix]) fb[pix] = key;
    }
}

static void render_batch(const pcr_batch* h, const uint32_t* lo, const uint32_t* med, const uint32_t* hi,
                         const pcr_view* v, int level, uint64_t* fb, int atomic) {
    for (uint32_t i = 0; i < h->num_points; i++) {
        uint32_t idx = h->first_point + i;
        float p[3];
        orc_decode(level, lo[idx], level <= 1 ? med[idx] : 0u, level == 0 ? hi[idx] : 0u, h->min, h->max, p);
        uint32_t pix, d;
        if (!orc_project(v->transform, p, v->width, v->height, &pix, &d)) continue;
        uint64_t key = ((uint64_t)d << 32) | idx;
        if (!atomic) {
            if (key < fb[pix]) fb[pix] = key;
        } else {
            uint64_t old = __atomic_load_n(&fb[pix], __ATOMIC_RELAXED);
            while (key < old && !__atomic_compare_exchange_n(&fb[pix], &old, key, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
        }
    }
}

/* compute_loop_las/render.cs:233-428 */
void orc_render_batched(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                        const uint32_t* hi, const pcr_view* v, int frustum_culling,
                        uint64_t* fb, uint8_t* levels_out) {
    for (uint32_t k = 0; k < nb; k++) {
        if (frustum_culling && !orc_frustum_visible(v->transform_frustum, b[k].min, b[k].max)) {
            if (levels_out) levels_out[k] = 255;
            continue;
        }
        int level = orc_precision_level(b[k].min, b[k].max, v, 100.0f);
        if (levels_out) levels_out[k] = (uint8_t)level;
        render_batch(&b[k], lo, med, hi, v, level, fb, 0);
    }
}

/* compute_loop_las/resolve.cs:100-146: the colour of the pixel's point,
 * gathered by the index in the key's low 32 bits (:137), all 32 bits kept.
 * Background iff the key is the clear value (point index 0 is a point and no
 * depth is capped), or the index is outside the cloud. */
void orc_resolve_index(const uint64_t* fb, uint64_t npix, const uint32_t* rgba, uint64_t npoints, uint32_t bg,
                       uint32_t* out) {
    for (uint64_t i = 0; i < npix; i++) {
        const uint32_t id = (uint32_t)fb[i];
        out[i] = (fb[i] == PCR_FB_CLEAR || id >= npoints) ? bg : rgba[id];
    }
}
/* e^x for x <= 0, for the eye-dome lighting shade. It is not libm's expf:
 * the device must produce the same bits, so both sides run this sequence.
 * Method: Cody-Waite range reduction x = k ln2 + r, with ln2 split into a
 * high part of 17 significant bits (k * ln2_hi is exact for |k| < 128) and a
 * low part; e^r from its Taylor series to r^7 in Horner form (|r| <= ln2/2,
 * truncation error below 2^-27 relative); then scaled by 2^k with ldexp.
 * Every step is one IEEE operation or an explicit fma. x <= -86 returns 0:
 * e^-86 < 2^-124, and an 8-bit channel times it truncates to 0 like the
 * true value. */
float orc_edl_exp(float x) {
    if (!(x > -86.0f)) return 0.0f;
    const float k = rintf(x * 1.44269504f);
    float r = fmaf(-k, 0.693145751953125f, x);
    r = fmaf(-k, 1.42860677e-06f, r);
    float p = 1.98412698e-04f;            /* 1/7! */
    p = fmaf(p, r, 1.38888889e-03f);      /* 1/6! */
    p = fmaf(p, r, 8.33333333e-03f);      /* 1/5! */
    p = fmaf(p, r, 4.16666667e-02f);      /* 1/4! */
    p = fmaf(p, r, 1.66666667e-01f);      /* 1/3! */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

/* getDepthAt(pixel, window 0) of compute_loop_las/resolve.cs:43-64: the
 * pixel's depth if it is in (0, 1e8), else 1e8. A neighbour outside the
 * image counts as background (the shader reads past the row or buffer). */
static float edl_depth(const uint64_t* fb, int w, int h, int x, int y) {
    if (x < 0 || y < 0 || x >= w || y >= h) return 100000000.0f;
    const float d = bitsf((uint32_t)(fb[(uint64_t)y * w + x] >> 32));
    return (d > 0.0f && d < 100000000.0f) ? d : 100000000.0f;
}

/* compute_loop_las/resolve.cs:100-191 with its EDL block switched on: the
 * colour of the pixel's point, shaded by exp(-(sum/9) * 300 * strength),
 * where sum adds max(0, depth - neighbour depth) over the 3x3 window in the
 * shader's ox-major order; channels truncate to integers and alpha is 0.
 * Background pixels (clear key, or an index outside the cloud) get bg. */
void orc_resolve_edl(const uint64_t* fb, int w, int h, const uint32_t* rgba, uint64_t npoints, uint32_t bg,
                     float strength, uint32_t* out) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint64_t i = (uint64_t)y * w + x;
            const uint64_t key = fb[i];
            const uint32_t id = (uint32_t)key;
            if (key == PCR_FB_CLEAR || id >= npoints) { out[i] = bg; continue; }
            const uint32_t color = rgba[id];
            const float depth = bitsf((uint32_t)(key >> 32));
            float sum = 0.0f;
            for (int ox = -1; ox <= 1; ox++)
                for (int oy = -1; oy <= 1; oy++) {
                    const float dd = depth - edl_depth(fb, w, h, x + ox, y + oy);
                    sum += dd > 0.0f ? dd : 0.0f;
                }
            const float response = sum / 9.0f;
            const float shade = orc_edl_exp(-response * 300.0f * strength);
            const uint32_t r = (uint32_t)((float)(color & 255u) * shade);
            const uint32_t g = (uint32_t)((float)((color >> 8) & 255u) * shade);
            const uint32_t b = (uint32_t)((float)((color >> 16) & 255u) * shade);
            out[i] = r | (g << 8) | (b << 16);
        }
}

/* compute_2021_earlyz/resolve.cs:41-78: colour carried in the key */
void orc_resolve_color(const uint64_t* fb, uint64_t npix, uint32_t bg, uint32_t* out) {
    for (uint64_t i = 0; i < npix; i++) out[i] = fb[i] == PCR_FB_CLEAR ? bg : (uint32_t)fb[i];
}

/* ------------------------------------------------------------------- hqs */
static void hqs_depth_point(const pcr_view* v, const float p[3], uint32_t* depth) {
    uint32_t pix, d;
    if (!orc_project(v->transform, p, v->width, v->height, &pix, &d)) return;
    if (d < depth[pix]) depth[pix] = d;                         /* depth.cs:371-379 */
}
static void hqs_color_point(const pcr_view* v, const float p[3], uint32_t color, const uint32_t* depth, uint32_t* acc) {
    uint32_t pix, d;
    if (!orc_project(v->transform, p, v->width, v->height, &pix, &d)) return;
    float bd = bitsf(depth[pix]);
    if (!(bitsf(d) <= bd * 1.01f)) return;                     /* color.cs:400-403 */
    acc[4 * (uint64_t)pix + 0] += color & 255u;
    acc[4 * (uint64_t)pix + 1] += (color >> 8) & 255u;
    acc[4 * (uint64_t)pix + 2] += (color >> 16) & 255u;
    acc[4 * (uint64_t)pix + 3] += 1u;
}
void orc_hqs_batched(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                     const uint32_t* hi, const uint32_t* rgba, const pcr_view* v,
                     int frustum_culling, uint32_t* depth, uint32_t* acc) {
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t k = 0; k < nb; k++) {
            if (frustum_culling && !orc_frustum_visible(v->transform_frustum, b[k].min, b[k].max)) continue;
            int level = orc_precision_level(b[k].min, b[k].max, v, 100.0f);
            for (uint32_t i = 0; i < b[k].num_points; i++) {
                uint32_t idx = b[k].first_point + i;
                float p[3];
                orc_decode(level, lo[idx], level <= 1 ? med[idx] : 0u, level == 0 ? hi[idx] : 0u, b[k].min, b[k].max, p);
                if (pass == 0) hqs_depth_point(v, p, depth);
                else hqs_color_point(v, p, rgba[idx], depth, acc);
            }
        }
    }
}
void orc_hqs_points16(const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_view* v,
                      uint32_t* depth, uint32_t* acc) {
    for (uint64_t i = 0; i < n; i++) hqs_depth_point(v, xyz + 3 * i, depth);
    for (uint64_t i = 0; i < n; i++) hqs_color_point(v, xyz + 3 * i, rgba[i], depth, acc);
}
/* hqs/resolve.cs:74-129: truncating integer averages, r | g << 8 | b << 16
 * with alpha 0 (resolve.cs:119); bg where count == 0 (:121-122) */
void orc_hqs_resolve(const uint32_t* acc, uint64_t npix, uint32_t bg, uint32_t* out) {
    for (uint64_t i = 0; i < npix; i++) {
        uint32_t c = acc[4 * i + 3];
        if (c == 0) { out[i] = bg; continue; }
        uint32_t r = acc[4 * i] / c, g = acc[4 * i + 1] / c, bb = acc[4 * i + 2] / c;
        out[i] = r | (g << 8) | (bb << 16);
    }
}

/* ------------------------------------------------------------------- LOD */
static int node_visible(const pcr_batch* n, const pcr_view* v, int culling) {
    return !culling || orc_frustum_visible(v->transform_frustum, n->min, n->max);
}
static int node_refines(const pcr_batch* n, const pcr_view* v, float lod_px) {
    return orc_projected_px(n->min, n->max, n->spacing, v->view_world, v->proj, v->width, v->height) > lod_px;
}
/* nodes/render.cs:267-275 (SIZE80) or the spacing cut: a node is drawn when it
 * and all its ancestors are visible and every ancestor's projected spacing
 * exceeds lod_px, which makes the drawn set a top-down cut of the tree. */
void orc_lod_select(const pcr_batch* nodes, uint32_t nn, const pcr_view* v, int culling,
                    int rule, float lod_px, uint8_t* flags) {
    for (uint32_t i = 0; i < nn; i++) {
        const pcr_batch* n = &nodes[i];
        if (rule == PCR_LOD_SIZE80) {
            flags[i] = node_visible(n, v, culling) && orc_precision_level(n->min, n->max, v, 80.0f) < 4;
            continue;
        }
        int ok = node_visible(n, v, culling);
        int32_t p = n->parent;
        while (ok && p >= 0) {
            ok = node_visible(&nodes[p], v, culling) && node_refines(&nodes[p], v, lod_px);
            p = nodes[p].parent;
        }
        flags[i] = (uint8_t)ok;
    }
}
/* nodes/render.cs:244-479: drawn nodes use the nodes precision thresholds */
void orc_render_nodes(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                      const uint32_t* hi, const pcr_view* v, int culling, int rule,
                      float lod_px, uint64_t* fb, uint8_t* flags_out) {
    uint8_t* flags = flags_out ? flags_out : (uint8_t*)malloc(nn ? nn : 1);
    orc_lod_select(nodes, nn, v, culling, rule, lod_px, flags);
    for (uint32_t i = 0; i < nn; i++) {
        if (!flags[i]) continue;
        int level = orc_precision_level(nodes[i].min, nodes[i].max, v, 80.0f);
        render_batch(&nodes[i], lo, med, hi, v, level, fb, 0);
    }
    if (!flags_out) free(flags);
}

/* compute_loop_nodes_hqs/depth.cs:225-360 and color.cs:600-760: the hqs
 * passes over the nodes of the LOD cut, each at its node precision level
 * (first threshold 80 px); the interactive selection/highlight colouring is
 * not part of it. */
void orc_hqs_nodes(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                   const uint32_t* hi, const uint32_t* rgba, const pcr_view* v, int culling, int rule,
                   float lod_px, uint32_t* depth, uint32_t* acc) {
    uint8_t* flags = (uint8_t*)malloc(nn ? nn : 1);
    orc_lod_select(nodes, nn, v, culling, rule, lod_px, flags);
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t k = 0; k < nn; k++) {
            if (!flags[k]) continue;
            const int level = orc_precision_level(nodes[k].min, nodes[k].max, v, 80.0f);
            for (uint32_t i = 0; i < nodes[k].num_points; i++) {
                const uint32_t idx = nodes[k].first_point + i;
                float p[3];
                orc_decode(level, lo[idx], level <= 1 ? med[idx] : 0u, level == 0 ? hi[idx] : 0u, nodes[k].min,
                           nodes[k].max, p);
                if (pass == 0) hqs_depth_point(v, p, depth);
                else hqs_color_point(v, p, rgba[idx], depth, acc);
            }
        }
    }
    free(flags);
}

/* compute_loop_nodes_hqs/resolve.cs:92-158: each pixel averages the window
 * (2*window+1)^2 around it in float, neighbours weighted 100 (centre), 2
 * (ox <= 1 && oy <= 1) or 1, and dropped when deeper than 1.01 x the
 * window's smallest depth; neighbours outside the image are skipped (the
 * reference reads past the buffer). No colour -> 0; alpha 0. */
void orc_hqs_resolve_window(const uint32_t* depth, const uint32_t* acc, int w, int h, int window, uint32_t* out) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float dmin = 1000000.0f;
            for (int ox = -window; ox <= window; ox++)
                for (int oy = -window; oy <= window; oy++) {
                    const int px = x + ox, py = y + oy;
                    if (px < 0 || py < 0 || px >= w || py >= h) continue;
                    const float pd = bitsf(depth[(uint64_t)py * w + px]);
                    if (pd >= 0.0f) dmin = pd < dmin ? pd : dmin;
                }
            float R = 0.0f, G = 0.0f, B = 0.0f, C = 0.0f;
            for (int ox = -window; ox <= window; ox++)
                for (int oy = -window; oy <= window; oy++) {
                    const int px = x + ox, py = y + oy;
                    if (px < 0 || py < 0 || px >= w || py >= h) continue;
                    const uint64_t q = (uint64_t)py * w + px;
                    const float pd = bitsf(depth[q]);
                    float wt = (ox == 0 && oy == 0) ? 100.0f : (ox <= 1 && oy <= 1) ? 2.0f : 1.0f;
                    if (pd > dmin * 1.01f) wt = 0.0f;
                    R += (float)acc[4 * q + 0] * wt;
                    G += (float)acc[4 * q + 1] * wt;
                    B += (float)acc[4 * q + 2] * wt;
                    C += (float)acc[4 * q + 3] * wt;
                }
            uint32_t color = 0u;
            if (C != 0.0f)
                color = (uint32_t)(R / C) | ((uint32_t)(G / C) << 8) | ((uint32_t)(B / C) << 16);
            if (dmin < 0.0f) color = 0x0000FF00u;
            out[(uint64_t)y * w + x] = color;
        }
}

/* ------------------------------------------------------ multithreaded CPU */
typedef struct {
    const pcr_batch* b; uint32_t nb; const uint32_t *lo, *med, *hi; const pcr_view* v;
    int culling; uint64_t* fb; uint32_t* next;
} mt_job;
static void* mt_worker(void* arg) {
    mt_job* j = (mt_job*)arg;
    for (;;) {
        uint32_t k0 = __atomic_fetch_add(j->next, 16u, __ATOMIC_RELAXED);
        if (k0 >= j->nb) break;
        uint32_t k1 = k0 + 16u < j->nb ? k0 + 16u : j->nb;
        for (uint32_t k = k0; k < k1; k++) {
            if (j->culling && !orc_frustum_visible(j->v->transform_frustum, j->b[k].min, j->b[k].max)) continue;
            int level = orc_precision_level(j->b[k].min, j->b[k].max, j->v, 100.0f);
            render_batch(&j->b[k], j->lo, j->med, j->hi, j->v, level, j->fb, 1);
        }
    }
    return NULL;
}
void orc_render_batched_mt(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                           const uint32_t* hi, const pcr_view* v, int culling, uint64_t* fb, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    uint32_t next = 0;
    mt_job job = {b, nb, lo, med, hi, v, culling, fb, &next};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, mt_worker, &job);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
}

/* orc_render_nodes with nthreads threads over the drawn nodes (CAS fetch-min;
 * min is order independent, so the result equals the single-thread one) */
typedef struct {
    const pcr_batch* nodes; uint32_t nn; const uint32_t *lo, *med, *hi; const pcr_view* v;
    const uint8_t* flags; uint64_t* fb; uint32_t* next;
} nodes_job;
static void* nodes_worker(void* arg) {
    nodes_job* j = (nodes_job*)arg;
    for (;;) {
        uint32_t k0 = __atomic_fetch_add(j->next, 8u, __ATOMIC_RELAXED);
        if (k0 >= j->nn) break;
        uint32_t k1 = k0 + 8u < j->nn ? k0 + 8u : j->nn;
        for (uint32_t k = k0; k < k1; k++) {
            if (!j->flags[k]) continue;
            int level = orc_precision_level(j->nodes[k].min, j->nodes[k].max, j->v, 80.0f);
            render_batch(&j->nodes[k], j->lo, j->med, j->hi, j->v, level, j->fb, 1);
        }
    }
    return NULL;
}
void orc_render_nodes_mt(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                         const uint32_t* hi, const pcr_view* v, int culling, int rule, float lod_px, uint64_t* fb,
                         int nthreads) {
    if (nthreads < 1) nthreads = 1;
    uint8_t* flags = (uint8_t*)malloc(nn ? nn : 1);
    orc_lod_select(nodes, nn, v, culling, rule, lod_px, flags);
    uint32_t next = 0;
    nodes_job job = {nodes, nn, lo, med, hi, v, flags, fb, &next};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, nodes_worker, &job);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    free(flags);
}

/* ---------------------------------------------------------------- camera */
static void mat_mul_d(const double a[16], const double b[16], double r[16]) {
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a[k * 4 + rr] * b[c * 4 + k];
            r[c * 4 + rr] = s;
        }
}
void orc_look_at_view(const double eye[3], const double target[3], const double up[3],
                      double fovy_deg, double zn, double zf, const double origin[3],
                      int width, int height, pcr_view* out) {
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int i = 0; i < 3; i++) f[i] /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int i = 0; i < 3; i++) s[i] /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double V[16] = {0}, P[16] = {0}, W[16] = {0}, VW[16], PVW[16];
    V[0] = s[0]; V[4] = s[1]; V[8] = s[2];
    V[1] = u[0]; V[5] = u[1]; V[9] = u[2];
    V[2] = -f[0]; V[6] = -f[1]; V[10] = -f[2];
    V[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    V[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    V[14] = (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
    V[15] = 1.0;
    double th = tan(fovy_deg * 3.14159265358979323846 / 180.0 / 2.0);
    double aspect = (double)width / (double)height;
    P[0] = 1.0 / (aspect * th); P[5] = 1.0 / th;
    P[10] = -(zf + zn) / (zf - zn); P[11] = -1.0; P[14] = -(2.0 * zf * zn) / (zf - zn);
    W[0] = W[5] = W[10] = W[15] = 1.0;
    W[12] = origin[0]; W[13] = origin[1]; W[14] = origin[2];
    mat_mul_d(V, W, VW);
    mat_mul_d(P, VW, PVW);
    memset(out, 0, sizeof(*out));
    for (int i = 0; i < 16; i++) {
        out->transform[i] = (float)PVW[i];
        out->transform_frustum[i] = (float)PVW[i];
        out->view_world[i] = (float)VW[i];
        out->proj[i] = (float)P[i];
    }
    out->width = width; out->height = height;
}
