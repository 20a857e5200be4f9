// pcr_layout.hip — builds the adaptive-precision point layout on the device.
//
// Replaces the CPU loader/encoder of modules/compute/LasLoaderSparse.cpp:37-212
// (batches of 10,240 points, per-batch AABB, 30-bit quantisation relative to
// the AABB in double, three 10:10:10 planes) and the offline Morton sort of
// tools/sort_points/Sort_Frugal/src/main_frugal.cpp:234-295, with the
// include/utils.h:24-27 axis convention (x<<2 | y<<1 | z).
//
// HBM layout of a cloud (structure of arrays, 4-point aligned):
//   batches  pcr_batch[nb]        64 B per batch
//   lo/med/hi uint32[plane_len]   bits 29..20 / 19..10 / 9..0 of x|y<<10|z<<20
//   rgba     uint32[plane_len]    colour, same index
//   points16 uint4[n]             optional x, y, z, rgba for the 16 B/pt path
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include "pcr_device.h"
#include "pcr_generate.h"
#include "pcr_internal.h"

namespace pcr {

// ------------------------------------------------------------- generation
__global__ void k_generate(GenParams g, const Building* __restrict__ table, uint64_t n,
                           float* __restrict__ xyz, uint32_t* __restrict__ rgba) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float p[3];
        uint32_t c;
        if (g.scene == PCR_SCENE_CUBE) cube_point(g, i, p, c);
        else if (g.scene == PCR_SCENE_TERRAIN || i < g.n_terrain) terrain_point(g, i, p, c);
        else building_point(g, table, i - g.n_terrain, p, c);
        xyz[3 * i] = p[0]; xyz[3 * i + 1] = p[1]; xyz[3 * i + 2] = p[2];
        rgba[i] = c;
    }
}

// Building table of the CITY scene, computed on the host (double weights).
static std::vector<Building> city_table(const pcr_gen_cfg* c, uint64_t n_build) {
    const int nb = c->num_buildings;
    std::vector<Building> t((size_t)nb);
    std::vector<double> w((size_t)nb);
    const uint64_t kb = stream_key(c->seed, 2);
    double wsum = 0.0;
    for (int b = 0; b < nb; b++) {
        uint64_t r[7];
        for (int k = 0; k < 7; k++) r[k] = rnd(kb, 8 * (uint64_t)b + (uint64_t)k);
        const float cx = (0.05f + 0.9f * u01(r[0])) * c->extent;
        const float cy = (0.05f + 0.9f * u01(r[1])) * c->extent;
        const float sx = 10.0f + 70.0f * u01(r[2]);
        const float sy = 10.0f + 70.0f * u01(r[3]);
        const float sz = 10.0f + 70.0f * u01(r[4]);
        const float sp = 0.02f + 0.03f * u01(r[5]);
        t[b].x0 = cx - sx * 0.5f; t[b].y0 = cy - sy * 0.5f;
        t[b].z0 = terrain_height(c->seed, c->amplitude, cx, cy);
        t[b].sx = sx; t[b].sy = sy; t[b].sz = sz;
        t[b].rgb = (80u + (uint32_t)(r[6] & 127u)) | ((80u + (uint32_t)((r[6] >> 8) & 127u)) << 8) |
                   ((80u + (uint32_t)((r[6] >> 16) & 127u)) << 16);
        t[b].pad = 0;
        const double area = 2.0 * ((double)sx * (double)sz) + 2.0 * ((double)sy * (double)sz) + (double)sx * (double)sy;
        w[b] = area / ((double)sp * (double)sp);
        wsum += w[b];
    }
    double cum = 0.0;
    for (int b = 0; b < nb; b++) {
        cum += w[b];
        t[b].end = (b == nb - 1) ? n_build : (uint64_t)std::floor((double)n_build * (cum / wsum));
    }
    return t;
}

int generate_points(pcr_ctx* ctx, const pcr_gen_cfg* cfg, float* xyz, uint32_t* rgba) {
    PCR_CHECK(cfg && xyz && rgba, "generate: null argument");
    PCR_CHECK(cfg->scene >= 0 && cfg->scene <= 2, "generate: unknown scene %d", cfg->scene);
    GenParams g{};
    g.seed = cfg->seed; g.extent = cfg->extent; g.amplitude = cfg->amplitude; g.jitter = cfg->jitter;
    g.scene = cfg->scene; g.num_buildings = cfg->num_buildings;
    Building* d_table = nullptr;
    if (cfg->scene == PCR_SCENE_CITY) {
        PCR_CHECK(cfg->num_buildings > 0, "generate: CITY needs num_buildings > 0");
        g.n_terrain = (uint64_t)((double)cfg->num_points * (double)cfg->terrain_frac);
        PCR_CHECK(g.n_terrain < cfg->num_points, "generate: CITY needs building points");
        std::vector<Building> t = city_table(cfg, cfg->num_points - g.n_terrain);
        PCR_HIP(hipMalloc(&d_table, sizeof(Building) * t.size()));
        PCR_HIP(hipMemcpyAsync(d_table, t.data(), sizeof(Building) * t.size(), hipMemcpyHostToDevice, ctx->stream));
    } else if (cfg->scene == PCR_SCENE_TERRAIN) {
        PCR_CHECK(cfg->amplitude > 0.0f, "generate: TERRAIN needs amplitude > 0");
    }
    uint64_t grid = (cfg->num_points + 255) / 256;
    if (grid > (uint64_t)ctx->cu_count * 128) grid = (uint64_t)ctx->cu_count * 128;
    if (grid == 0) grid = 1;
    k_generate<<<(int)grid, 256, 0, ctx->stream>>>(g, d_table, cfg->num_points, xyz, rgba);
    PCR_HIP(hipGetLastError());
    if (d_table) {
        PCR_HIP(hipStreamSynchronize(ctx->stream));
        PCR_HIP(hipFree(d_table));
    }
    return PCR_OK;
}

// ---------------------------------------------------------------- builder
__global__ __launch_bounds__(256) void k_aabb_partial(const float* __restrict__ xyz, uint64_t n, float* __restrict__ part) {
    __shared__ float s[6][256];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * i + a];
            mn[a] = v < mn[a] ? v : mn[a];
            mx[a] = v > mx[a] ? v : mx[a];
        }
    }
    for (int a = 0; a < 3; a++) { s[a][threadIdx.x] = mn[a]; s[3 + a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int a = 0; a < 3; a++) {
                const float o = s[a][threadIdx.x + off];
                s[a][threadIdx.x] = o < s[a][threadIdx.x] ? o : s[a][threadIdx.x];
                const float O = s[3 + a][threadIdx.x + off];
                s[3 + a][threadIdx.x] = O > s[3 + a][threadIdx.x] ? O : s[3 + a][threadIdx.x];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint32_t quant21(float x, float mn, double scale) {
    const double q = ((double)x - (double)mn) * scale;
    return q >= 2097151.0 ? 2097151u : (uint32_t)q;
}
__global__ void k_morton(const float* __restrict__ xyz, uint64_t n, float mn0, float mn1, float mn2, double scale,
                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t qx = quant21(xyz[3 * i], mn0, scale);
        const uint32_t qy = quant21(xyz[3 * i + 1], mn1, scale);
        const uint32_t qz = quant21(xyz[3 * i + 2], mn2, scale);
        keys[i] = (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
        vals[i] = (uint32_t)i;
    }
}
__global__ void k_gather(const float* __restrict__ xyz, const uint32_t* __restrict__ rgba,
                         const uint32_t* __restrict__ perm, uint64_t n, float* __restrict__ xyz_s,
                         uint32_t* __restrict__ rgba_s) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t j = perm ? perm[i] : i;
        xyz_s[3 * i] = xyz[3 * j]; xyz_s[3 * i + 1] = xyz[3 * j + 1]; xyz_s[3 * i + 2] = xyz[3 * j + 2];
        rgba_s[i] = rgba[j];
    }
}
__global__ void k_pack16(const float* __restrict__ xyz, const uint32_t* __restrict__ rgba, uint64_t n,
                         uint4* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = make_uint4(__float_as_uint(xyz[3 * i]), __float_as_uint(xyz[3 * i + 1]),
                            __float_as_uint(xyz[3 * i + 2]), rgba[i]);
}

// Per-batch float AABB (LasLoaderSparse.cpp:88-121), one workgroup per batch.
__global__ __launch_bounds__(256) void k_batch_aabb(const float* __restrict__ xyz, uint64_t n, uint32_t bs,
                                                    pcr_batch* __restrict__ hdr) {
    __shared__ float s[6][256];
    const uint64_t first = (uint64_t)blockIdx.x * bs;
    const uint64_t last = first + bs < n ? first + bs : n;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = first + threadIdx.x; i < last; i += 256) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * i + a];
            mn[a] = v < mn[a] ? v : mn[a];
            mx[a] = v > mx[a] ? v : mx[a];
        }
    }
    for (int a = 0; a < 3; a++) { s[a][threadIdx.x] = mn[a]; s[3 + a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int a = 0; a < 3; a++) {
                const float o = s[a][threadIdx.x + off];
                s[a][threadIdx.x] = o < s[a][threadIdx.x] ? o : s[a][threadIdx.x];
                const float O = s[3 + a][threadIdx.x + off];
                s[3 + a][threadIdx.x] = O > s[3 + a][threadIdx.x] ? O : s[3 + a][threadIdx.x];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pcr_batch h;
        memset(&h, 0, sizeof(h));
        for (int a = 0; a < 3; a++) { h.min[a] = s[a][0]; h.max[a] = s[3 + a][0]; }
        h.num_points = (uint32_t)(last - first);
        h.first_point = (uint32_t)first;
        h.parent = -1;
        hdr[blockIdx.x] = h;
    }
}

// 30-bit quantisation in double relative to the batch AABB, split into three
// 10:10:10 planes (LasLoaderSparse.cpp:146-182); zero extent quantises to 0.
__device__ __forceinline__ uint32_t quant30(float x, float mn, double size) {
    if (size == 0.0) return 0u;
    const uint32_t v = (uint32_t)((((double)x - (double)mn) / size) * 1073741824.0);
    return v < 1073741823u ? v : 1073741823u;
}
__global__ void k_encode(const float* __restrict__ xyz, uint64_t n, uint32_t bs, const pcr_batch* __restrict__ hdr,
                         uint32_t* __restrict__ lo, uint32_t* __restrict__ med, uint32_t* __restrict__ hi) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const pcr_batch& h = hdr[i / bs];
        uint32_t q[3];
#pragma unroll
        for (int a = 0; a < 3; a++) q[a] = quant30(xyz[3 * i + a], h.min[a], (double)h.max[a] - (double)h.min[a]);
        lo[i] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
        med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
        hi[i] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
    }
}

static int grid_n(const pcr_ctx* ctx, uint64_t n) {
    uint64_t g = (n + 255) / 256;
    const uint64_t cap = (uint64_t)ctx->cu_count * 128;
    if (g > cap) g = cap;
    return (int)(g ? g : 1);
}

void cloud_free(pcr_cloud* c) {
    if (!c) return;
    (void)hipFree(c->d_batches); (void)hipFree(c->d_lo); (void)hipFree(c->d_med); (void)hipFree(c->d_hi);
    (void)hipFree(c->d_rgba); (void)hipFree(c->d_points16); (void)hipFree(c->d_levels); (void)hipFree(c->d_list);
    c->d_batches = nullptr; c->d_lo = c->d_med = c->d_hi = c->d_rgba = nullptr;
    c->d_points16 = nullptr; c->d_levels = nullptr; c->d_list = nullptr;
}

int cloud_alloc_batched(pcr_ctx* ctx, uint32_t nb, uint64_t npoints, pcr_cloud* c) {
    c->ctx = ctx;
    c->num_batches = nb;
    c->num_points = npoints;
    c->plane_len = (npoints + 3) & ~3ull;
    if (c->plane_len == 0) c->plane_len = 4;
    const size_t pb = c->plane_len * sizeof(uint32_t);
    const size_t nbb = (size_t)(nb ? nb : 1);
    if (hipMalloc(&c->d_batches, nbb * sizeof(pcr_batch)) != hipSuccess ||
        hipMalloc(&c->d_lo, pb) != hipSuccess || hipMalloc(&c->d_med, pb) != hipSuccess ||
        hipMalloc(&c->d_hi, pb) != hipSuccess || hipMalloc(&c->d_rgba, pb) != hipSuccess ||
        hipMalloc(&c->d_levels, nbb) != hipSuccess || hipMalloc(&c->d_list, nbb * 4) != hipSuccess) {
        (void)hipGetLastError();
        cloud_free(c);
        set_error("device allocation of %llu points / %u batches failed", (unsigned long long)npoints, nb);
        return PCR_E_OOM;
    }
    // zero the padding words so vector loads past the last point read zeros
    const uint64_t pad = c->plane_len - npoints;
    if (pad) {
        PCR_HIP(hipMemsetAsync(c->d_lo + npoints, 0, pad * 4, ctx->stream));
        PCR_HIP(hipMemsetAsync(c->d_med + npoints, 0, pad * 4, ctx->stream));
        PCR_HIP(hipMemsetAsync(c->d_hi + npoints, 0, pad * 4, ctx->stream));
        PCR_HIP(hipMemsetAsync(c->d_rgba + npoints, 0, pad * 4, ctx->stream));
    }
    c->device_bytes = nbb * (sizeof(pcr_batch) + 5) + 4 * pb;
    return PCR_OK;
}

// Float AABB of device points (partial reductions on the device, final on the host).
int device_aabb(pcr_ctx* ctx, const float* xyz, uint64_t n, float mn[3], float mx[3]) {
    const int gp = 1024;
    float* d_part = nullptr;
    PCR_HIP(hipMalloc(&d_part, sizeof(float) * 6 * gp));
    k_aabb_partial<<<gp, 256, 0, ctx->stream>>>(xyz, n, d_part);
    PCR_HIP(hipGetLastError());
    std::vector<float> part(6 * gp);
    PCR_HIP(hipMemcpyAsync(part.data(), d_part, sizeof(float) * 6 * gp, hipMemcpyDeviceToHost, ctx->stream));
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    PCR_HIP(hipFree(d_part));
    for (int a = 0; a < 3; a++) { mn[a] = INFINITY; mx[a] = -INFINITY; }
    for (int b = 0; b < gp; b++)
        for (int a = 0; a < 3; a++) {
            const float o = part[6 * b + a], O = part[6 * b + 3 + a];
            mn[a] = o < mn[a] ? o : mn[a];
            mx[a] = O > mx[a] ? O : mx[a];
        }
    return PCR_OK;
}

int build_cloud_device(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n,
                       const pcr_build_opts* opts, pcr_cloud** out) {
    PCR_CHECK(ctx && xyz && rgba && out, "build: null argument");
    PCR_CHECK(n > 0 && n < 0x7FFFFFFFull, "build: point count %llu out of range (1 .. 2^31-1)", (unsigned long long)n);
    const uint32_t bs = opts && opts->batch_size ? opts->batch_size : PCR_POINTS_PER_BATCH;
    PCR_CHECK(bs % 4 == 0, "build: batch_size %u must be a multiple of 4", bs);
    const bool sort = !opts || opts->morton_sort;
    const bool keep16 = opts && opts->keep_points16;
    hipStream_t s = ctx->stream;

    float mn[3], mx[3];
    const int rc0 = device_aabb(ctx, xyz, n, mn, mx);
    if (rc0) return rc0;

    uint32_t* d_perm = nullptr;
    if (sort) {
        float side = mx[0] - mn[0];
        if (mx[1] - mn[1] > side) side = mx[1] - mn[1];
        if (mx[2] - mn[2] > side) side = mx[2] - mn[2];
        const double scale = side > 0.0f ? 2097152.0 / (double)side : 0.0;
        uint64_t *k_in = nullptr, *k_out = nullptr;
        uint32_t* v_in = nullptr;
        PCR_HIP(hipMalloc(&k_in, n * 8));
        PCR_HIP(hipMalloc(&k_out, n * 8));
        PCR_HIP(hipMalloc(&v_in, n * 4));
        PCR_HIP(hipMalloc(&d_perm, n * 4));
        k_morton<<<grid_n(ctx, n), 256, 0, s>>>(xyz, n, mn[0], mn[1], mn[2], scale, k_in, v_in);
        PCR_HIP(hipGetLastError());
        size_t temp_bytes = 0;
        PCR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, k_in, k_out, v_in, d_perm, (int)n, 0, 63, s));
        void* temp = nullptr;
        PCR_HIP(hipMalloc(&temp, temp_bytes));
        PCR_HIP(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k_in, k_out, v_in, d_perm, (int)n, 0, 63, s));
        PCR_HIP(hipStreamSynchronize(s));
        PCR_HIP(hipFree(temp)); PCR_HIP(hipFree(k_in)); PCR_HIP(hipFree(k_out)); PCR_HIP(hipFree(v_in));
    }

    const uint32_t nb = (uint32_t)((n + bs - 1) / bs);
    pcr_cloud* c = new pcr_cloud();
    int rc = cloud_alloc_batched(ctx, nb, n, c);
    if (rc != PCR_OK) { delete c; (void)hipFree(d_perm); return rc; }
    c->batch_size = bs;
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = mn[a]; c->aabb_max[a] = mx[a]; }

    float* xyz_s = nullptr;
    PCR_HIP(hipMalloc(&xyz_s, n * 12));
    k_gather<<<grid_n(ctx, n), 256, 0, s>>>(xyz, rgba, d_perm, n, xyz_s, c->d_rgba);
    PCR_HIP(hipGetLastError());
    k_batch_aabb<<<nb, 256, 0, s>>>(xyz_s, n, bs, c->d_batches);
    PCR_HIP(hipGetLastError());
    k_encode<<<grid_n(ctx, n), 256, 0, s>>>(xyz_s, n, bs, c->d_batches, c->d_lo, c->d_med, c->d_hi);
    PCR_HIP(hipGetLastError());
    if (keep16) {
        PCR_HIP(hipMalloc(&c->d_points16, n * 16));
        c->points16_len = n;
        c->device_bytes += n * 16;
        k_pack16<<<grid_n(ctx, n), 256, 0, s>>>(xyz_s, c->d_rgba, n, c->d_points16);
        PCR_HIP(hipGetLastError());
    }
    PCR_HIP(hipStreamSynchronize(s));
    PCR_HIP(hipFree(xyz_s));
    if (d_perm) PCR_HIP(hipFree(d_perm));
    c->h_batches.resize(nb);
    PCR_HIP(hipMemcpy(c->h_batches.data(), c->d_batches, sizeof(pcr_batch) * nb, hipMemcpyDeviceToHost));
    *out = c;
    return PCR_OK;
}

// ----------------------------------------------------------- host encoder
static uint32_t quant30_host(float x, float mn, double size) {
    if (size == 0.0) return 0u;
    const uint32_t v = (uint32_t)((((double)x - (double)mn) / size) * 1073741824.0);
    return v < 1073741823u ? v : 1073741823u;
}
void encode_batches_host(const float* xyz, uint64_t n, uint32_t bs, pcr_batch* hdr, uint32_t* lo, uint32_t* med,
                         uint32_t* hi) {
    const uint64_t nb = (n + bs - 1) / bs;
    auto work = [&](uint64_t b0, uint64_t b1) {
        for (uint64_t b = b0; b < b1; b++) {
            const uint64_t first = b * bs, last = std::min<uint64_t>(first + bs, n);
            float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint64_t i = first; i < last; i++)
                for (int a = 0; a < 3; a++) {
                    const float v = xyz[3 * i + a];
                    mn[a] = v < mn[a] ? v : mn[a];
                    mx[a] = v > mx[a] ? v : mx[a];
                }
            pcr_batch h;
            memset(&h, 0, sizeof(h));
            for (int a = 0; a < 3; a++) { h.min[a] = mn[a]; h.max[a] = mx[a]; }
            h.num_points = (uint32_t)(last - first);
            h.first_point = (uint32_t)first;
            h.parent = -1;
            hdr[b] = h;
            double size[3];
            for (int a = 0; a < 3; a++) size[a] = (double)mx[a] - (double)mn[a];
            for (uint64_t i = first; i < last; i++) {
                uint32_t q[3];
                for (int a = 0; a < 3; a++) q[a] = quant30_host(xyz[3 * i + a], mn[a], size[a]);
                lo[i] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
                med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
                hi[i] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nb < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) th.emplace_back(work, nb * t / nt, nb * (t + 1) / nt);
    for (auto& x : th) x.join();
}

}  // namespace pcr
