// pcr_kernels.hip — per-frame kernels for gfx950.
//
// Stage map (reference file:line it replaces):
//   k_select          cull + precision select + LOD cut, one thread per batch
//                     (compute_loop_las/render.cs:245-254, compute_loop_nodes/render.cs:267-275)
//   k_raster_batched  decode + project + 64-bit atomicMin, one workgroup per batch
//                     (compute_loop_las/render.cs:296-428, compute_loop_las2/render.cs:319-469)
//   k_raster_points16 16 B/pt points, colour payload (compute_2021_earlyz/render.cs:35-63)
//   k_hqs_*           two-pass depth + accumulate (compute_loop_las_hqs/depth.cs, color.cs)
//   k_resolve_*       to RGBA8 (compute_loop_las/resolve.cs:100-146, hqs/resolve.cs:62-129,
//                     compute_2021_earlyz/resolve.cs:41-78)
#include "pcr_device.h"
#include "pcr_internal.h"

namespace pcr {


__device__ __forceinline__ u32x4 ld_stream(const u32x4* p) { return __builtin_nontemporal_load(p); }

// 64-bit atomicMin on LDS (ds_min_u64): the pointer is cast to the LDS
// address space so that the compiler does not emit a flat atomic.
//
// The tile word is read first, and only a smaller value issues the atomic.
// A far batch puts hundreds of points on each pixel of its tile, so the
// lanes of one ds_min often hit the same few words, and the LDS performs
// same-word atomics one after the other; a read of one word is broadcast to
// every lane that asks for it. A tile word only decreases, so a value not
// below the word read is not below the word later either, and the result is
// the same. This halves the time of a frame that draws all of a 1e9-point
// city (2.73 -> 1.38 ms) and takes 8-9% off the flythrough frames.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
__device__ __forceinline__ void lds_min(unsigned long long* p, unsigned long long v) {
    if (v < __hip_atomic_load((lds_u64*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
        __hip_atomic_fetch_min((lds_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// ds_min_u32 (read first, as lds_min) and ds_add_u64, for the hqs tiles.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ void lds_min32(uint32_t* p, uint32_t v) {
    if (v < __hip_atomic_load((lds_u32*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
        __hip_atomic_fetch_min((lds_u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add((lds_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// ds_add_u32: the tiled rasters count their framebuffer atomics in an LDS
// word at the point of issue (a per-thread count would hold a register for
// the whole point loop).
__device__ __forceinline__ void lds_add32(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_add((lds_u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// ------------------------------------------------------------------ fills
// A frame's counters are zeroed by the select launch of the frame before it
// (k_select resets the other of the context's two frame states).
__device__ __forceinline__ void frame_reset(FrameState* st) {
    st->list_count = 0; st->queue_head = 0; st->queue_head2 = 0; st->culled = 0;
    for (int i = 0; i < 5; i++) st->level_count[i] = 0;
    st->points_drawn = 0; st->plane_bytes = 0;
    st->occluded_batches = 0; st->occluded_points = 0; st->occluded_bytes = 0; st->atomics = 0;
    st->pixels_visible = 0;
}
// 16-byte stores (p is hipMalloc-aligned); the last element of an odd-length
// 64-bit fill, or the last 1-3 of a 32-bit fill, are stored singly.
__global__ void k_fill64(unsigned long long* __restrict__ p, uint64_t n, unsigned long long v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    u64x2* p2 = reinterpret_cast<u64x2*>(p);
    for (uint64_t i = t; i < n / 2; i += stride) p2[i] = u64x2{v, v};
    if (t == 0 && (n & 1)) p[n - 1] = v;
}
__global__ void k_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    u32x4* p4 = reinterpret_cast<u32x4*>(p);
    for (uint64_t i = t; i < n / 4; i += stride) p4[i] = u32x4{v, v, v, v};
    if (t < (n & 3)) p[(n & ~3ull) + t] = v;
}
__global__ void k_frame_reset(FrameState* st) {
    if (threadIdx.x == 0) frame_reset(st);
}

static int grid_for(uint64_t n, int block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

int launch_fill64(hipStream_t s, unsigned long long* p, uint64_t n, unsigned long long v) {
    k_fill64<<<grid_for(n / 2 + 1, 256), 256, 0, s>>>(p, n, v);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}
int launch_fill32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v) {
    k_fill32<<<grid_for(n / 4 + 1, 256), 256, 0, s>>>(p, n, v);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}
int launch_frame_reset(hipStream_t s, FrameState* st) {
    k_frame_reset<<<1, 64, 0, s>>>(st);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// ----------------------------------------------------------------- select
// One thread per batch/node: frustum cull, LOD cut, precision level. Writes
// levels[i] (255 = not drawn), appends drawn batches to `list` and
// accumulates per-frame statistics. Counts are combined across the
// workgroup's 16 waves in LDS first, so the frame's shared counters see one
// list atomic and eight statistics atomics per 1024 batches.
constexpr int kSelectBlock = 1024;
// The same launch clears the frame's framebuffer (fill0, fill1: 16-byte
// stores over the whole grid; the raster runs after it on the stream) and
// resets the other frame state for the next frame, so a frame needs no
// launch before its select.
__device__ __forceinline__ void fill_words(const FillRegion& f, uint64_t t, uint64_t stride) {
    u32x4* p4 = reinterpret_cast<u32x4*>(f.p);
    for (uint64_t i = t; i < f.n / 4; i += stride) p4[i] = u32x4{f.v, f.v, f.v, f.v};
    if (t < (f.n & 3)) f.p[(f.n & ~3ull) + t] = f.v;
}

template <bool LOD>
__global__ __launch_bounds__(kSelectBlock) void k_select(const pcr_batch* __restrict__ b, uint32_t nb, pcr_view v,
                                                         int culling, int rule, float lod_px, float first,
                                                         uint32_t b0, uint32_t b1,
                                                         uint8_t* __restrict__ levels, uint32_t* __restrict__ list,
                                                         FrameState* __restrict__ st, FrameState* __restrict__ st_next,
                                                         FillRegion fill0, FillRegion fill1) {
    constexpr int kWaves = kSelectBlock / 64;
    {
        const uint64_t t = (uint64_t)blockIdx.x * kSelectBlock + threadIdx.x;
        const uint64_t stride = (uint64_t)gridDim.x * kSelectBlock;
        if (t == 0) frame_reset(st_next);
        fill_words(fill0, t, stride);
        fill_words(fill1, t, stride);
    }
    __shared__ float planes[6][4];   // the view's frustum planes, computed once per workgroup
    __shared__ uint32_t s_keep[kWaves];
    __shared__ unsigned long long s_stat[8][kWaves];   // culled, level 0..4, points, bytes
    __shared__ uint32_t s_base;
    if (threadIdx.x < 6) frustum_plane(v.transform_frustum, (int)threadIdx.x, planes[threadIdx.x]);
    __syncthreads();
    const uint32_t i = blockIdx.x * kSelectBlock + threadIdx.x;
    const bool in_range = i < nb && i >= b0 && i < b1;   // [b0, b1): the batches of this draw
    int level = kCulled;
    uint32_t np = 0;
    if (in_range) {
        const pcr_batch h = b[i];
        bool vis = !culling || planes_visible(planes, h.min, h.max);
        if (LOD && vis) {
            if (rule == PCR_LOD_SIZE80) {
                vis = precision_level(h.min, h.max, v, 80.0f) < 4;
            } else {
                int32_t p = h.parent;
                while (vis && p >= 0) {
                    const pcr_batch a = b[p];
                    vis = (!culling || planes_visible(planes, a.min, a.max)) &&
                          projected_px(a.min, a.max, a.spacing, v.view_world, v.proj, v.width, v.height) > lod_px;
                    p = a.parent;
                }
            }
        }
        if (vis) { level = precision_level(h.min, h.max, v, first); np = h.num_points; }
    }
    if (i < nb) levels[i] = (uint8_t)level;
    const bool keep = level != kCulled;
    const unsigned long long m = __ballot(keep);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint32_t culled = (uint32_t)__popcll(__ballot(in_range && !keep));
    uint32_t cnt[5];
#pragma unroll
    for (int L = 0; L < 5; L++) cnt[L] = (uint32_t)__popcll(__ballot(level == L));
    const unsigned long long pts = wave_sum_u64(np);
    const unsigned long long bytes =
        wave_sum_u64(keep ? (unsigned long long)np * 4ull * planes_for_level(level) : 0ull);
    if (lane == 0) {
        s_keep[wave] = (uint32_t)__popcll(m);
        s_stat[0][wave] = culled;
#pragma unroll
        for (int L = 0; L < 5; L++) s_stat[1 + L][wave] = cnt[L];
        s_stat[6][wave] = pts;
        s_stat[7][wave] = bytes;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWaves; w++) total += s_keep[w];
        s_base = total ? atomicAdd(&st->list_count, total) : 0u;
    } else if (threadIdx.x >= 64 && threadIdx.x < 72) {
        const int k = (int)threadIdx.x - 64;
        unsigned long long sum = 0;
        for (int w = 0; w < kWaves; w++) sum += s_stat[k][w];
        if (sum) {
            if (k == 0) atomicAdd(&st->culled, (uint32_t)sum);
            else if (k <= 5) atomicAdd(&st->level_count[k - 1], (uint32_t)sum);
            else if (k == 6) atomicAdd(&st->points_drawn, sum);
            else atomicAdd(&st->plane_bytes, sum);
        }
    }
    __syncthreads();
    if (keep) {
        uint32_t base = s_base;
        for (uint32_t w = 0; w < wave; w++) base += s_keep[w];
        list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    }
}

int launch_select(hipStream_t s, const pcr_cloud* c, const pcr_view& v, const pcr_render_params& p,
                  FrameState* st, FrameState* st_next, const FillRegion& fill0, const FillRegion& fill1,
                  float first, bool lod) {
    // enough workgroups for the batches, and for the fill (at most 512)
    const uint64_t sel = (c->num_batches + kSelectBlock - 1) / kSelectBlock;
    const uint64_t words = fill0.n > fill1.n ? fill0.n : fill1.n;
    uint64_t fg = (words / 4 + kSelectBlock - 1) / kSelectBlock;
    if (fg > 512) fg = 512;
    const int grid = (int)(sel > fg ? sel : (fg ? fg : 1));
    const uint32_t b0 = p.batch_begin, b1 = p.batch_end ? p.batch_end : c->num_batches;
    if (lod)
        k_select<true><<<grid, kSelectBlock, 0, s>>>(c->d_batches, c->num_batches, v, p.frustum_culling, p.lod_rule,
                                                     p.lod_px, first, b0, b1, c->d_levels, c->d_list, st, st_next,
                                                     fill0, fill1);
    else
        k_select<false><<<grid, kSelectBlock, 0, s>>>(c->d_batches, c->num_batches, v, p.frustum_culling,
                                                      p.lod_rule, p.lod_px, first, b0, b1, c->d_levels, c->d_list,
                                                      st, st_next, fill0, fill1);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// ----------------------------------------------------------------- raster
template <int LC>
__device__ __forceinline__ void raster_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi,
                                             uint32_t local, uint32_t np, uint32_t first, const pcr_view& v,
                                             unsigned long long* __restrict__ fb, uint32_t& natom) {
    if (local >= np) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    uint32_t pix, depth;
    if (!project(v.transform, p[0], p[1], p[2], v.width, v.height, pix, depth)) return;
    natom += fb_min(fb, pix, ((unsigned long long)depth << 32) | (first + local));
}

// Four consecutive points per thread per iteration, 128-bit streaming loads of
// only the planes the level needs (compute_loop_las2/render.cs:319-469).
template <int BLOCK, int LC>
__device__ __forceinline__ void raster_vec4(const Decode& d, const uint32_t* __restrict__ lo,
                                            const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                            uint32_t first, uint32_t np, const pcr_view& v,
                                            unsigned long long* __restrict__ fb, uint32_t& natom) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 a = ld_stream(lo4 + q);
        const u32x4 m = LC <= 1 ? ld_stream(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 c = LC == 0 ? ld_stream(hi4 + q) : u32x4{0, 0, 0, 0};
        const uint32_t base = q * 4u - lead;
        raster_point<LC>(d, a.x, m.x, c.x, base + 0, np, first, v, fb, natom);
        raster_point<LC>(d, a.y, m.y, c.y, base + 1, np, first, v, fb, natom);
        raster_point<LC>(d, a.z, m.z, c.z, base + 2, np, first, v, fb, natom);
        raster_point<LC>(d, a.w, m.w, c.w, base + 3, np, first, v, fb, natom);
    }
}

// One point per thread per iteration with stride BLOCK
// (compute_loop_las/render.cs:296-318, RENDER_STRIDE_128).
template <int BLOCK, int LC>
__device__ __forceinline__ void raster_strided(const Decode& d, const uint32_t* __restrict__ lo,
                                               const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                               uint32_t first, uint32_t np, const pcr_view& v,
                                               unsigned long long* __restrict__ fb, uint32_t& natom) {
    for (uint32_t i = threadIdx.x; i < np; i += BLOCK) {
        const uint32_t idx = first + i;
        raster_point<LC>(d, lo[idx], LC <= 1 ? med[idx] : 0u, LC == 0 ? hi[idx] : 0u, i, np, first, v, fb, natom);
    }
}

template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void k_raster_batched(const pcr_batch* __restrict__ b,
                                                          const uint8_t* __restrict__ levels,
                                                          const uint32_t* __restrict__ lo,
                                                          const uint32_t* __restrict__ med,
                                                          const uint32_t* __restrict__ hi, pcr_view v,
                                                          unsigned long long* __restrict__ fb,
                                                          FrameState* __restrict__ st) {
    const uint32_t bi = blockIdx.x;
    const int level = levels[bi];
    if (level == kCulled) return;
    const pcr_batch h = b[bi];
    const Decode d = make_decode(level, h.min, h.max);
    uint32_t natom = 0;
    if (VEC == 4) {
        if (level == 0) raster_vec4<BLOCK, 0>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
        else if (level == 1) raster_vec4<BLOCK, 1>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
        else raster_vec4<BLOCK, 2>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
    } else {
        if (level == 0) raster_strided<BLOCK, 0>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
        else if (level == 1) raster_strided<BLOCK, 1>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
        else raster_strided<BLOCK, 2>(d, lo, med, hi, h.first_point, h.num_points, v, fb, natom);
    }
    const unsigned long long wa = wave_sum_u64(natom);
    if (__lane_id() == 0 && wa) atomicAdd(&st->atomics, wa);
}

template <int BLOCK, int VEC>
static void raster_batched_t(hipStream_t s, const pcr_cloud* c, const pcr_view& v, unsigned long long* fb,
                             FrameState* st) {
    k_raster_batched<BLOCK, VEC><<<c->num_batches, BLOCK, 0, s>>>(c->d_batches, c->d_levels, c->d_lo, c->d_med,
                                                                  c->d_hi, v, fb, st);
}

// --------------------------------------------------------- division check
__global__ __launch_bounds__(256) void k_selftest_divide(const float* __restrict__ nx, const float* __restrict__ ny,
                                                         const float* __restrict__ den, uint64_t n,
                                                         float* __restrict__ out2) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const float d = den[i];
        float qx = 0.0f, qy = 0.0f;
        if (d > 0.0f) div_by_w(nx[i], ny[i], d, qx, qy);
        out2[2 * i] = qx;
        out2[2 * i + 1] = qy;
    }
}

int launch_selftest_divide(hipStream_t s, const float* nx, const float* ny, const float* den, uint64_t n,
                           float* out2) {
    if (n == 0) return PCR_OK;
    const uint64_t g = (n + 255) / 256;
    k_selftest_divide<<<(int)(g < 4096 ? g : 4096), 256, 0, s>>>(nx, ny, den, n, out2);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// ------------------------------------------------------------ tiled raster
// A batch whose projected AABB covers at most TILE_PIX pixels first merges
// its points into an LDS copy of that screen rectangle with 64-bit LDS
// atomicMin, then flushes the touched pixels to the framebuffer with the
// usual compare + global atomicMin. min is associative and commutative, so
// the framebuffer is bit-identical to the direct path; the per-point global
// traffic drops to one access per touched pixel per batch.
//
// INTERIOR (a batch for which screen_rect() reported `interior`): points
// project without project_xy()'s tests, and one that lands in the tile is
// done; any other point runs the full test below, so the result is the same.
template <int LC, bool TILE, bool INTERIOR = false>
__device__ __forceinline__ void tile_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi, uint32_t local,
                                           uint32_t np, uint32_t first, const Proj& pj,
                                           unsigned long long* __restrict__ fb, unsigned long long* tile, int x0,
                                           int y0, int tw, int th, uint32_t* natom) {
    if (local >= np) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    if (INTERIOR) {
        uint32_t ux, uy, depth;
        project_interior(pj, p[0], p[1], p[2], x0, y0, ux, uy, depth);
        if (ux < (uint32_t)tw && uy < (uint32_t)th) {
            // uy < th and tw are at most the tile size, so 24-bit products are exact
            lds_min(tile + __umul24(uy, (uint32_t)tw) + ux, ((unsigned long long)depth << 32) | (first + local));
            return;
        }
    }
    int ix, iy;
    uint32_t depth;
    if (!project_xy(pj, p[0], p[1], p[2], ix, iy, depth)) return;
    const unsigned long long key = ((unsigned long long)depth << 32) | (first + local);
    const uint32_t ux = (uint32_t)(ix - x0), uy = (uint32_t)(iy - y0);
    if (TILE && ux < (uint32_t)tw && uy < (uint32_t)th) lds_min(tile + uy * (uint32_t)tw + ux, key);
    else if (fb_min(fb, (uint32_t)ix + (uint32_t)iy * (uint32_t)pj.w, key)) lds_add32(natom, 1u);
}

// NT: non-temporal plane loads (a single view streams each plane once);
// multi-view passes re-read a batch once per view and keep it cached.
template <bool NT>
__device__ __forceinline__ u32x4 ld_plane(const u32x4* p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <int BLOCK, int LC, bool TILE, bool NT, bool INTERIOR = false>
__device__ __forceinline__ void tile_loop(const Decode& d, const uint32_t* __restrict__ lo,
                                          const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                          uint32_t first, uint32_t np, const Proj& pj,
                                          unsigned long long* __restrict__ fb, unsigned long long* tile, int x0,
                                          int y0, int tw, int th, uint32_t* natom) {
    // A batch may start at any word: quads are read from the aligned word
    // at or below `first`, and the `lead` words in front of the batch get a
    // local index that wraps past np, so the local < np test drops them as
    // it drops the padding behind the last point. first + local is the
    // point's index either way.
    const uint32_t lead = first & 3u;
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 a = ld_plane<NT>(lo4 + q);
        const u32x4 m = LC <= 1 ? ld_plane<NT>(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 c = LC == 0 ? ld_plane<NT>(hi4 + q) : u32x4{0, 0, 0, 0};
        const uint32_t base = q * 4u - lead;
        tile_point<LC, TILE, INTERIOR>(d, a.x, m.x, c.x, base + 0, np, first, pj, fb, tile, x0, y0, tw, th,
                                       natom);
        tile_point<LC, TILE, INTERIOR>(d, a.y, m.y, c.y, base + 1, np, first, pj, fb, tile, x0, y0, tw, th,
                                       natom);
        tile_point<LC, TILE, INTERIOR>(d, a.z, m.z, c.z, base + 2, np, first, pj, fb, tile, x0, y0, tw, th,
                                       natom);
        tile_point<LC, TILE, INTERIOR>(d, a.w, m.w, c.w, base + 3, np, first, pj, fb, tile, x0, y0, tw, th,
                                       natom);
    }
}

// The interior loop exists for the 10-bit levels only, which draw nearly all
// points; the 20- and 30-bit levels keep one loop each (code size).
template <int BLOCK, bool TILE, bool NT>
__device__ __forceinline__ void tile_dispatch(int level, const Decode& d, const uint32_t* __restrict__ lo,
                                              const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                              const pcr_batch& h, const Proj& pj, unsigned long long* __restrict__ fb,
                                              unsigned long long* tile, int x0, int y0, int tw, int th,
                                              uint32_t* natom, bool interior = false) {
    if (level == 0)
        tile_loop<BLOCK, 0, TILE, NT>(d, lo, med, hi, h.first_point, h.num_points, pj, fb, tile, x0, y0, tw, th,
                                      natom);
    else if (level == 1)
        tile_loop<BLOCK, 1, TILE, NT>(d, lo, med, hi, h.first_point, h.num_points, pj, fb, tile, x0, y0, tw, th,
                                      natom);
    else if (TILE && interior)
        tile_loop<BLOCK, 2, true, NT, true>(d, lo, med, hi, h.first_point, h.num_points, pj, fb, tile, x0, y0, tw,
                                            th, natom);
    else
        tile_loop<BLOCK, 2, TILE, NT>(d, lo, med, hi, h.first_point, h.num_points, pj, fb, tile, x0, y0, tw, th,
                                      natom);
}

// Occlusion test of one batch draw against the depths already in the
// frame's buffer: true when some pixel of the rectangle [x0, x0 + tw) x
// [y0, y0 + n / tw) may still take a point whose depth is at least wlo.
// Every point of the batch lands in that rectangle with a depth of at least
// wlo (screen_rect's `wlo`). Depths only decrease during a frame, so a value
// read here bounds its pixel's depth from then on (a stale read is only
// larger), and a batch for which the test returns false cannot change the
// buffer:
//   KIND 0 (uint64 keys, depth bits in the high word): each point's key
//          exceeds the pixel's key;
//   KIND 1 (hqs pass 1, uint32 depths): each point's depth bits exceed the
//          pixel's;
//   KIND 2 (hqs pass 2, final depths): each point fails w <= depth * 1.01f.
// A cleared pixel is open for kinds 0 and 1 and closed for kind 2 (its depth
// is NaN, which no point passes). n <= 2^13 and 2 <= tw (the multiply-high row
// split). Every thread of the workgroup calls it (it ends in a barrier).
//
// Most batches are not hidden, and one open pixel settles that: the first
// wave reads 64 pixels spread over the rectangle, and only when all of them
// are closed does the workgroup read every pixel. The decision is taken from
// one block-wide OR, so every wave takes the same branch.
constexpr int kOcclusionMaxPix = 2048;
template <int KIND>
__device__ __forceinline__ bool depth_open(const uint32_t* __restrict__ w32, uint32_t pix, uint32_t wb, float wlo) {
    // pixel < 2^30 (PCR_MAX_DIM^2), so 2 * pixel + 1 fits 32 bits
    const uint32_t d = KIND == 0 ? w32[2u * pix + 1u] : w32[pix];
    return KIND == 2 ? (__uint_as_float(d) * 1.01f >= wlo) : (d >= wb);
}
template <int BLOCK, int KIND>
__device__ __forceinline__ bool rect_open(const void* __restrict__ buf, int width, int x0, int y0, int tw, int n,
                                          float wlo) {
    const uint32_t magic = (uint32_t)((0x100000000ull + (uint64_t)tw - 1) / (uint64_t)tw);
    const uint32_t* w32 = reinterpret_cast<const uint32_t*>(buf);
    const uint32_t wb = __float_as_uint(wlo);
    const auto pixel = [&](uint32_t i) {
        const uint32_t uy = __umulhi(i, magic), ux = i - uy * (uint32_t)tw;
        return (uint32_t)x0 + ux + ((uint32_t)y0 + uy) * (uint32_t)width;
    };
    int open = 0;
    if (threadIdx.x < 64) open = depth_open<KIND>(w32, pixel(((uint32_t)n * threadIdx.x) >> 6), wb, wlo);
    if (__syncthreads_or(open)) return true;
    constexpr int F = 2;   // two loads in flight per thread (register use stays that of the point loop)
    for (int i0 = threadIdx.x; i0 < n; i0 += F * BLOCK) {
        bool o[F];
#pragma unroll
        for (int j = 0; j < F; j++) o[j] = depth_open<KIND>(w32, pixel((uint32_t)min(i0 + j * BLOCK, n - 1)), wb, wlo);
#pragma unroll
        for (int j = 0; j < F; j++) open |= o[j];
    }
    return __syncthreads_or(open) != 0;
}

// Flushes the touched pixels i < n of a 64-bit LDS tile over the rectangle
// at (x0, y0), tw pixels wide, into the framebuffer: four pixels per thread
// at a time (fb_min's compare and atomic, reordered across pixels). The
// compiler keeps each compare load in the branch of its pixel, with a wait
// after it; issuing the four unconditionally, so that they are in flight
// together, measured the same: the other waves of the SIMD already cover the
// wait. row = i / tw as a multiply-high by magic = ceil(2^32 / tw): exact
// because i < 2^13 and 2 <= tw <= 2^13 keep the error below 1/tw.
template <int BLOCK>
__device__ __forceinline__ void tile_flush(const unsigned long long* tile, unsigned long long* __restrict__ fb, int x0,
                                           int y0, int tw, int n, uint32_t magic, uint32_t width, uint32_t* natom) {
    constexpr int F = 4;
    for (int i0 = threadIdx.x; i0 < n; i0 += F * BLOCK) {
        unsigned long long k[F], old[F];
        uint32_t pix[F];
#pragma unroll
        for (int j = 0; j < F; j++) {
            const int i = i0 + j * BLOCK;
            k[j] = i < n ? tile[i] : PCR_FB_CLEAR;
            const uint32_t uy = __umulhi((uint32_t)i, magic), ux = (uint32_t)i - uy * (uint32_t)tw;
            pix[j] = (uint32_t)x0 + ux + ((uint32_t)y0 + uy) * width;
        }
#pragma unroll
        for (int j = 0; j < F; j++) old[j] = k[j] != PCR_FB_CLEAR ? fb[pix[j]] : 0ull;
#pragma unroll
        for (int j = 0; j < F; j++)
            if (k[j] < old[j]) {
                atomicMin(fb + pix[j], k[j]);
                lds_add32(natom, 1u);
            }
    }
}

// ceil(2^32 / tw) for the rows of a rectangle tw pixels wide, 2 <= tw.
__device__ __forceinline__ uint32_t row_magic(int tw) {
    return (uint32_t)((0x100000000ull + (uint64_t)tw - 1) / (uint64_t)tw);
}

// A rectangle read from LDS by every lane, in scalar registers.
__device__ __forceinline__ RectInfo uniform_rect(const RectInfo& s) {
    RectInfo r;
    r.x0 = __builtin_amdgcn_readfirstlane(s.x0);
    r.y0 = __builtin_amdgcn_readfirstlane(s.y0);
    r.x1 = __builtin_amdgcn_readfirstlane(s.x1);
    r.y1 = __builtin_amdgcn_readfirstlane(s.y1);
    r.wlo = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(s.wlo)));
    r.interior = __builtin_amdgcn_readfirstlane(s.interior);
    r.on = __builtin_amdgcn_readfirstlane(s.on);
    return r;
}

// One batch of the tiled raster; `bi` and `level` are workgroup-uniform.
// Returns the points skipped by the occlusion test (0 if the batch drew).
// PRE: the batch's screen rectangle for the view comes in `pre` (the
// multi-view kernel computes it for all its views at once) instead of from
// screen_rect() here.
template <int BLOCK, int TILE_PIX, bool NT, bool PRE = false>
__device__ __forceinline__ uint32_t tiled_batch(uint32_t bi, int level, const pcr_batch* __restrict__ b,
                                                const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                const uint32_t* __restrict__ hi, const pcr_view& v,
                                                unsigned long long* __restrict__ fb, unsigned long long* tile,
                                                uint32_t* natom, bool occl, uint32_t part = 0, uint32_t nparts = 1,
                                                const RectInfo* pre = nullptr) {
    pcr_batch h = b[bi];
    if (nparts > 1) {   // points [part * k, (part + 1) * k) of the batch, k = ceil(n / nparts)
        const uint32_t k = (h.num_points + nparts - 1) / nparts, off = min(part * k, h.num_points);
        h.first_point += off;
        h.num_points = min(k, h.num_points - off);
    }
    int x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    bool interior, use;
    float wlo;
    if constexpr (PRE) {
        x0 = pre->x0; y0 = pre->y0; x1 = pre->x1; y1 = pre->y1;
        interior = pre->interior != 0;
        wlo = pre->wlo;
        use = pre->on != 0;
    } else {
        use = screen_rect(v.transform, h.min, h.max, v.width, v.height, x0, y0, x1, y1, &interior, &wlo);
    }
    const int tw = x1 - x0 + 1, th = y1 - y0 + 1;
    // a batch behind everything already drawn over its rectangle draws nothing
    // (wlo > 0 only for interior batches, whose rectangle is not empty); the
    // test runs before the decode and projection constants exist, so that
    // its loads do not add to the registers of the point loop
    if (occl && wlo > 0.0f && tw >= 2 && tw * th <= kOcclusionMaxPix &&
        !rect_open<BLOCK, 0>(fb, v.width, x0, y0, tw, tw * th, wlo))
        return h.num_points;
    const Decode d = make_decode(level, h.min, h.max);
    const Proj pj = make_proj(v.transform, v.width, v.height);
    // A rectangle one column wide (clamped at the image's left or right edge,
    // where only rounding can put a point inside) draws directly: the tile's
    // row split needs tw >= 2 (ceil(2^32 / tw) must fit 32 bits).
    use = use && tw >= 2 && tw <= TILE_PIX && th <= TILE_PIX && tw * th <= TILE_PIX;
    // every thread evaluates `use` from the same header and view, so the
    // branch (and the barriers inside it) is uniform across the workgroup
    if (use) {
        const int n = tw * th;
        for (int i = threadIdx.x; i < n; i += BLOCK) tile[i] = PCR_FB_CLEAR;
        __syncthreads();
        tile_dispatch<BLOCK, true, NT>(level, d, lo, med, hi, h, pj, fb, tile, x0, y0, tw, th, natom, interior);
        __syncthreads();
        tile_flush<BLOCK>(tile, fb, x0, y0, tw, n, row_magic(tw), (uint32_t)v.width, natom);
    } else {
        tile_dispatch<BLOCK, false, NT>(level, d, lo, med, hi, h, pj, fb, nullptr, 0, 0, 0, 0, natom);
    }
    return 0;
}

// A persistent workgroup's next queue item. Every workgroup increments the
// one queue counter, and the memory performs atomics on one address one after
// the other, so a workgroup takes two items per atomic and draws them one
// after the other (with the split halves ordered as in k_raster_tiled, the
// 1e9-point 100 m flythrough's raster is 4.7% faster than with one item per
// atomic, and three or four items are no better). take() is called by every
// thread of the workgroup; the returned item is workgroup-uniform. `s_item`
// must not be rewritten before the workgroup's next barrier.
struct QueueTake {
    uint32_t next = 0xFFFFFFFFu;   // the second item of the last pair, not yet drawn
    __device__ __forceinline__ uint32_t take(uint32_t* head, uint32_t& s_item) {
        if (next != 0xFFFFFFFFu) {
            const uint32_t it = next;
            next = 0xFFFFFFFFu;
            return it;
        }
        if (threadIdx.x == 0) s_item = atomicAdd(head, 2u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
        next = it + 1u;
        return it;
    }
};

// Persistent workgroups take drawn batches from the select stage's compacted
// list through an atomic queue, so no workgroup is launched for a culled
// batch. The loop ends when the queue passes list_count, which every
// workgroup reaches.
template <int BLOCK, int TILE_PIX>
__global__ __launch_bounds__(BLOCK) void k_raster_tiled(const pcr_batch* __restrict__ b,
                                                        const uint8_t* __restrict__ levels,
                                                        const uint32_t* __restrict__ list,
                                                        FrameState* __restrict__ st,
                                                        const uint32_t* __restrict__ lo,
                                                        const uint32_t* __restrict__ med,
                                                        const uint32_t* __restrict__ hi, pcr_view v,
                                                        unsigned long long* __restrict__ fb, int occl) {
    __shared__ unsigned long long tile[TILE_PIX];
    __shared__ uint32_t s_item, s_occ_batches, s_natom;   // s_natom: framebuffer atomics issued
    __shared__ unsigned long long s_occ_points, s_occ_bytes;   // what the occlusion test skipped
    const uint32_t count = st->list_count;
    const uint32_t nsplit = min(count, gridDim.x), whole = count - nsplit;
    if (threadIdx.x == 0) { s_occ_batches = 0; s_occ_points = 0; s_occ_bytes = 0; s_natom = 0; }
    QueueTake q;
    for (;;) {
        uint32_t item = q.take(&st->queue_head, s_item);
        if (item >= count + nsplit) break;
        // The last nsplit batches (one per workgroup) are drawn as two halves
        // each: workgroups otherwise idle 43 us on average between their
        // last batch and the end of the launch, and the halves shorten the
        // batches taken last (splitting twice as many items costs more in
        // extra flushes than it saves). The first halves of all of them come
        // first, then the second halves, so that a pair of queue items never
        // holds both halves of one batch.
        uint32_t part = 0, nparts = 1;
        if (item >= whole) {
            const uint32_t h = item - whole;
            part = h >= nsplit ? 1u : 0u;
            nparts = 2;
            item = whole + (part ? h - nsplit : h);
        }
        // Consecutive queue items map to list slots spread over the whole list
        // (multiplying by a prime above any count, modulo the count, permutes
        // the slots), so the batches in flight at one time come from distant
        // parts of the Morton order and cover different screen regions: fewer
        // of their atomics meet on the same framebuffer lines.
        const uint32_t slot = (uint32_t)(((uint64_t)item * 2654435761ull) % count);
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[slot]);
        const int level = __builtin_amdgcn_readfirstlane(levels[bi]);
        const uint32_t skipped = tiled_batch<BLOCK, TILE_PIX, true>(bi, level, b, lo, med, hi, v, fb, tile, &s_natom,
                                                                    occl != 0, part, nparts);
        if (threadIdx.x == 0 && skipped) {
            s_occ_points += skipped;
            s_occ_bytes += 4ull * planes_for_level(level) * skipped;
            s_occ_batches += part == 0;   // a batch drawn as two halves counts once
        }
        __syncthreads();   // s_item and the tile are reused by the next item
    }
    // the loop ends after a barrier that follows every thread's last atomic
    if (threadIdx.x == 0 && s_natom) atomicAdd(&st->atomics, (unsigned long long)s_natom);
    if (threadIdx.x == 0 && s_occ_points) {
        atomicAdd(&st->occluded_batches, s_occ_batches);
        atomicAdd(&st->occluded_points, s_occ_points);
        atomicAdd(&st->occluded_bytes, s_occ_bytes);
    }
}

template <int BLOCK, int TILE_PIX>
static void tiled_t(hipStream_t s, const pcr_cloud* c, const pcr_view& v, FrameState* st, unsigned long long* fb,
                    bool occl) {
    static const int occ = [] {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_raster_tiled<BLOCK, TILE_PIX>, BLOCK, 0) != hipSuccess) {
            (void)hipGetLastError();
            o = 0;
        }
        return o > 0 ? o : 4;
    }();
    uint64_t grid = (uint64_t)c->ctx->cu_count * (uint64_t)occ;
    if (grid > c->num_batches) grid = c->num_batches;
    k_raster_tiled<BLOCK, TILE_PIX><<<(int)grid, BLOCK, 0, s>>>(c->d_batches, c->d_levels, c->d_list, st, c->d_lo,
                                                               c->d_med, c->d_hi, v, fb, occl ? 1 : 0);
}

int launch_raster_tiled(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int tile_px, int block,
                        FrameState* st, unsigned long long* fb, bool occl) {
    if (c->num_batches == 0) return PCR_OK;
    if (block == 128) {
        if (tile_px == 1024) tiled_t<128, 1024>(s, c, v, st, fb, occl);
        else if (tile_px == 4096) tiled_t<128, 4096>(s, c, v, st, fb, occl);
        else tiled_t<128, 2048>(s, c, v, st, fb, occl);
    } else {
        if (tile_px == 1024) tiled_t<256, 1024>(s, c, v, st, fb, occl);
        else if (tile_px == 4096) tiled_t<256, 4096>(s, c, v, st, fb, occl);
        else tiled_t<256, 2048>(s, c, v, st, fb, occl);
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// Multi-view tiled raster: persistent workgroups take batches kept by any
// view; for each view that draws the batch the workgroup runs the tiled
// batch pass with that view's level. After the first view the batch's planes
// come from cache, so HBM sees each plane once per frame.
// One batch for a pair of views at once, the next two views a multi-view
// launch draws the batch for, at one precision level: each point is decoded
// once and projected for both views, into the two halves of the LDS tile.
// Used when both views' rectangles are interior (screen_rect) and fit half
// the tile; the caller draws the views one at a time otherwise. Per view the
// operations are those of tiled_batch (tile_point's interior path, the
// four-pixel flush), so each framebuffer is the same; each view keeps its own
// occlusion test.
struct RasterPairView {
    Proj pj;
    unsigned long long* fb;
    unsigned long long* tile;
    int x0, y0, tw, th;
    bool draw;
};

// The post-decode half of tile_point (interior batch, tiled) for one view.
__device__ __forceinline__ void raster_pair_point(const float p[3], unsigned long long key_low,
                                                  const RasterPairView& rv, uint32_t* natom) {
    uint32_t ux, uy, depth;
    project_interior(rv.pj, p[0], p[1], p[2], rv.x0, rv.y0, ux, uy, depth);
    if (ux < (uint32_t)rv.tw && uy < (uint32_t)rv.th) {
        lds_min(rv.tile + __umul24(uy, (uint32_t)rv.tw) + ux, ((unsigned long long)depth << 32) | key_low);
        return;
    }
    int ix, iy;
    if (!project_xy(rv.pj, p[0], p[1], p[2], ix, iy, depth)) return;
    const unsigned long long key = ((unsigned long long)depth << 32) | key_low;
    const uint32_t vx = (uint32_t)(ix - rv.x0), vy = (uint32_t)(iy - rv.y0);
    if (vx < (uint32_t)rv.tw && vy < (uint32_t)rv.th) lds_min(rv.tile + vy * (uint32_t)rv.tw + vx, key);
    else if (fb_min(rv.fb, (uint32_t)ix + (uint32_t)iy * (uint32_t)rv.pj.w, key)) lds_add32(natom, 1u);
}

template <int BLOCK, int LC, bool NT>
__device__ __forceinline__ void raster_pair_loop(const Decode& d, const uint32_t* __restrict__ lo,
                                                 const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                                 uint32_t first, uint32_t np, const RasterPairView& a,
                                                 const RasterPairView& c, uint32_t* natom) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 w0 = ld_plane<NT>(lo4 + q);
        const u32x4 w1 = LC <= 1 ? ld_plane<NT>(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 w2 = LC == 0 ? ld_plane<NT>(hi4 + q) : u32x4{0, 0, 0, 0};
        const uint32_t base = q * 4u - lead;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t local = base + (uint32_t)j;
            if (local >= np) continue;
            float p[3];
            decode<LC>(d, w0[j], w1[j], w2[j], p);
            if (a.draw) raster_pair_point(p, first + local, a, natom);
            if (c.draw) raster_pair_point(p, first + local, c, natom);
        }
    }
}

// Returns false, having done nothing, when the pair cannot be drawn together.
template <int BLOCK, int TILE_PIX, bool NT>
__device__ __forceinline__ bool raster_pair_batch(uint32_t bi, int level, const pcr_batch* __restrict__ b,
                                                  const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                  const uint32_t* __restrict__ hi, const pcr_view& va,
                                                  unsigned long long* __restrict__ fb_a, const pcr_view& vb,
                                                  unsigned long long* __restrict__ fb_b, unsigned long long* tile,
                                                  uint32_t* natom, bool occl, uint32_t part, uint32_t nparts,
                                                  const RectInfo& ra, const RectInfo& rb) {
    pcr_batch h = b[bi];
    if (nparts > 1) {   // as in tiled_batch
        const uint32_t k = (h.num_points + nparts - 1) / nparts, off = min(part * k, h.num_points);
        h.first_point += off;
        h.num_points = min(k, h.num_points - off);
    }
    const int ax0 = ra.x0, ay0 = ra.y0, bx0 = rb.x0, by0 = rb.y0;
    const bool aon = ra.on != 0, bon = rb.on != 0, ai = ra.interior != 0, bi_ = rb.interior != 0;
    const float awlo = ra.wlo, bwlo = rb.wlo;
    const int atw = ra.x1 - ax0 + 1, ath = ra.y1 - ay0 + 1, btw = rb.x1 - bx0 + 1, bth = rb.y1 - by0 + 1;
    constexpr int half = TILE_PIX / 2;
    if (!(aon && bon && ai && bi_ && atw >= 2 && btw >= 2 && atw <= half && ath <= half && btw <= half &&
          bth <= half && atw * ath <= half && btw * bth <= half))
        return false;
    // each view's occlusion test, as in tiled_batch
    const bool da = !(occl && awlo > 0.0f && atw * ath <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 0>(fb_a, va.width, ax0, ay0, atw, atw * ath, awlo);
    const bool db = !(occl && bwlo > 0.0f && btw * bth <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 0>(fb_b, vb.width, bx0, by0, btw, btw * bth, bwlo);
    if (!da && !db) return true;
    const Decode d = make_decode(level, h.min, h.max);
    const RasterPairView a{make_proj(va.transform, va.width, va.height), fb_a, tile, ax0, ay0, atw, ath, da};
    const RasterPairView c{make_proj(vb.transform, vb.width, vb.height), fb_b, tile + half, bx0, by0, btw, bth, db};
    for (int i = threadIdx.x; i < atw * ath; i += BLOCK) a.tile[i] = PCR_FB_CLEAR;
    for (int i = threadIdx.x; i < btw * bth; i += BLOCK) c.tile[i] = PCR_FB_CLEAR;
    __syncthreads();
    if (level == 0) raster_pair_loop<BLOCK, 0, NT>(d, lo, med, hi, h.first_point, h.num_points, a, c, natom);
    else if (level == 1) raster_pair_loop<BLOCK, 1, NT>(d, lo, med, hi, h.first_point, h.num_points, a, c, natom);
    else raster_pair_loop<BLOCK, 2, NT>(d, lo, med, hi, h.first_point, h.num_points, a, c, natom);
    __syncthreads();
    // each view's flush, as in tiled_batch
    if (a.draw)
        tile_flush<BLOCK>(a.tile, a.fb, a.x0, a.y0, a.tw, a.tw * a.th, row_magic(a.tw), (uint32_t)a.pj.w, natom);
    if (c.draw)
        tile_flush<BLOCK>(c.tile, c.fb, c.x0, c.y0, c.tw, c.tw * c.th, row_magic(c.tw), (uint32_t)c.pj.w, natom);
    return true;
}

template <int BLOCK, int TILE_PIX>
__global__ __launch_bounds__(BLOCK) void k_mv_raster_tiled(const pcr_batch* __restrict__ b, uint32_t nb,
                                                           const uint8_t* __restrict__ levels,
                                                           const uint32_t* __restrict__ list,
                                                           MVState* __restrict__ st, const uint32_t* __restrict__ lo,
                                                           const uint32_t* __restrict__ med,
                                                           const uint32_t* __restrict__ hi,
                                                           const pcr_view* __restrict__ views, int nviews,
                                                           unsigned long long* const* __restrict__ fbs, int occl) {
    __shared__ unsigned long long tile[TILE_PIX];
    __shared__ uint32_t s_item, s_nv;
    __shared__ uint8_t s_vid[64], s_lv[64];
    __shared__ RectInfo s_rect[64];   // the batch's rectangle in each view that draws it
    const uint32_t count = st->list_count;
    const uint32_t nsplit = min(count, gridDim.x), whole = count - nsplit;
    __shared__ uint32_t s_natom;   // counted, not reported per view
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&st->queue_head, 1u);
        __syncthreads();
        uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= count + nsplit) break;
        // as in k_raster_tiled, the last item per workgroup is drawn as two halves
        uint32_t part = 0, nparts = 1;
        if (item >= whole) { part = (item - whole) & 1u; nparts = 2; item = whole + ((item - whole) >> 1); }
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[item]);
        if (threadIdx.x < 64) {   // compact the views drawing this batch (wave 0)
            const int v = (int)threadIdx.x;
            const int lv = v < nviews ? levels[(uint64_t)v * nb + bi] : kCulled;
            const unsigned long long m = __ballot(lv != kCulled);
            if (lv != kCulled) {
                const uint32_t pos = (uint32_t)__popcll(m & ((1ull << v) - 1ull));
                s_vid[pos] = (uint8_t)v;
                s_lv[pos] = (uint8_t)lv;
            }
            if (v == 0) s_nv = (uint32_t)__popcll(m);
        }
        __syncthreads();
        const uint32_t nv = __builtin_amdgcn_readfirstlane(s_nv);
        {   // every view's rectangle at once: one view per 8-lane group
            const pcr_batch& hb = b[bi];
            for (uint32_t j = threadIdx.x >> 3; j < nv; j += BLOCK / 8) {
                const pcr_view& vw = views[s_vid[j]];
                RectInfo r;
                screen_rect_group(vw.transform, hb.min, hb.max, vw.width, vw.height, r);
                if ((threadIdx.x & 7u) == 0) s_rect[j] = r;
            }
        }
        __syncthreads();
        for (uint32_t k = 0; k < nv;) {
            const int vi = __builtin_amdgcn_readfirstlane(s_vid[k]);
            const int lv = __builtin_amdgcn_readfirstlane(s_lv[k]);
            const RectInfo ra = uniform_rect(s_rect[k]);
            // the next two views at one level in one pass over the points
            if (k + 1 < nv) {
                const int vj = __builtin_amdgcn_readfirstlane(s_vid[k + 1]);
                const int lj = __builtin_amdgcn_readfirstlane(s_lv[k + 1]);
                if (lv == lj && raster_pair_batch<BLOCK, TILE_PIX, false>(bi, lv, b, lo, med, hi, views[vi], fbs[vi],
                                                                           views[vj], fbs[vj], tile, &s_natom,
                                                                           occl != 0, part, nparts, ra,
                                                                           uniform_rect(s_rect[k + 1]))) {
                    __syncthreads();   // the tile is reused by the next views
                    k += 2;
                    continue;
                }
            }
            tiled_batch<BLOCK, TILE_PIX, false, true>(bi, lv, b, lo, med, hi, views[vi], fbs[vi], tile, &s_natom,
                                                      occl != 0, part, nparts, &ra);
            __syncthreads();   // the tile is reused by the next view
            k++;
        }
    }
}

template <int BLOCK, int TILE_PIX>
static void mv_tiled_t(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int cu_count, bool occl) {
    static const int occ = [] {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_mv_raster_tiled<BLOCK, TILE_PIX>, BLOCK, 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            o = 0;
        }
        return o > 0 ? o : 4;
    }();
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    if (grid > c->num_batches) grid = c->num_batches;
    if (grid == 0) grid = 1;
    k_mv_raster_tiled<BLOCK, TILE_PIX><<<(int)grid, BLOCK, 0, s>>>(c->d_batches, c->num_batches, B.d_levels, B.d_list,
                                                                  B.d_state, c->d_lo, c->d_med, c->d_hi, B.d_views,
                                                                  nviews, B.d_fbs, occl ? 1 : 0);
}

int launch_mv_raster_tiled(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int tile_px, int block,
                           int cu_count, bool occl) {
    if (block == 128) {
        if (tile_px == 1024) mv_tiled_t<128, 1024>(s, c, B, nviews, cu_count, occl);
        else if (tile_px == 4096) mv_tiled_t<128, 4096>(s, c, B, nviews, cu_count, occl);
        else mv_tiled_t<128, 2048>(s, c, B, nviews, cu_count, occl);
    } else {
        if (tile_px == 1024) mv_tiled_t<256, 1024>(s, c, B, nviews, cu_count, occl);
        else if (tile_px == 4096) mv_tiled_t<256, 4096>(s, c, B, nviews, cu_count, occl);
        else mv_tiled_t<256, 2048>(s, c, B, nviews, cu_count, occl);
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// ------------------------------------------------------------- tiled hqs
// The LDS tile of the raster, for the two hqs passes. A batch whose
// projected rectangle covers at most TILE_PIX pixels
//   pass 1: min-reduces its depths in an LDS copy of the rectangle
//           (ds_min_u32), then flushes each touched pixel with the compare +
//           global atomicMin of hqs_depth;
//   pass 2: loads the rectangle's final depths into LDS, accumulates its
//           visible points there as two packed 64-bit sums per pixel,
//           R | G << 32 and B | count << 32 (ds_add_u64), then adds each
//           touched pixel's sums to the global uint32 accumulators (add_pairs).
// A batch of at most kHqsTileMaxPoints points cannot carry out of a 32-bit
// half (255 x 16,843,009 < 2^32), and uint32 addition is associative mod
// 2^32, so depth and accumulators are bit-identical to the direct passes.
// Pass 2 is where the direct path spends its time: at a distance most points
// of a pixel lie within 1.01x of its depth, and every one of them made four
// contended global atomics.
constexpr uint32_t kHqsTileMaxPoints = 16843009u;

// A batch's screen rectangle [x0, x0 + tw) x [y0, y0 + th) and the band of
// its rows, [y0 + boff, y0 + boff + bh), that the LDS tile holds in this
// pass over the batch's points. Points in the rectangle but outside the band
// belong to another band; points outside the rectangle (possible only where
// the rounding of the rectangle and of a point's pixel disagree) take the
// direct path once, in the first band.
struct TileBand {
    int x0, y0, tw, th, boff, bh;
    bool first;
};

// Pass 2's sums for one tile pixel, carried in registers over the eight
// points of a thread's two quads. The points are consecutive in the batch's
// Morton order, and in a far batch most of them land on one pixel: a run of
// them is added to the tile as one pair of ds_add_u64 instead of one pair
// per point. The LDS performs adds on one word one after the other, so a far
// batch's adds (hundreds of points per pixel) queued on few words. At most
// 8 x 255 per channel is held in a register, so a half never carries into
// the other, and the tile word gets the same sum.
struct PendingSums {
    uint32_t t = 0xFFFFFFFFu;          // tile pixel, or none
    unsigned long long rg = 0, bc = 0; // R | G << 32, B | count << 32
};
__device__ __forceinline__ void pending_flush(PendingSums& p, unsigned long long* tacc) {
    if (p.t != 0xFFFFFFFFu) {
        lds_add64(tacc + 2 * p.t, p.rg);
        lds_add64(tacc + 2 * p.t + 1, p.bc);
        p.t = 0xFFFFFFFFu;
    }
}

// A point in the band's rows of the rectangle: its tile pixel (pass 1) or,
// if visible, its sums (pass 2, joined to the pending run of its thread).
template <int PASS>
__device__ __forceinline__ void hqs_band_point(uint32_t ux, uint32_t ry, uint32_t dep, uint32_t color,
                                               uint32_t* tdepth, unsigned long long* tacc, const TileBand& tb,
                                               PendingSums& pend) {
    const uint32_t by = ry - (uint32_t)tb.boff;
    if (by >= (uint32_t)tb.bh) return;
    const uint32_t t = by * (uint32_t)tb.tw + ux;
    if (PASS == 0) {
        lds_min32(tdepth + t, dep);
    } else {
        if (!(__uint_as_float(dep) <= __uint_as_float(tdepth[t]) * 1.01f)) return;
        const unsigned long long rg =
            (unsigned long long)(color & 255u) | ((unsigned long long)((color >> 8) & 255u) << 32);
        const unsigned long long bc = (unsigned long long)((color >> 16) & 255u) | (1ull << 32);
        if (t != pend.t) {
            pending_flush(pend, tacc);
            pend.t = t;
            pend.rg = rg;
            pend.bc = bc;
        } else {
            pend.rg += rg;
            pend.bc += bc;
        }
    }
}

// INTERIOR: as in tile_point, a batch whose rectangle screen_rect() reported
// `interior` projects without project_xy()'s tests; a point outside the
// rectangle runs them below.
template <int PASS, int LC, bool TILE, bool INTERIOR = false>
__device__ __forceinline__ void hqs_tile_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi,
                                               uint32_t color, uint32_t local, uint32_t np, const Proj& pj,
                                               uint32_t* __restrict__ depth, uint32_t* __restrict__ acc,
                                               uint32_t* tdepth, unsigned long long* tacc, const TileBand& tb,
                                               PendingSums& pend) {
    if (local >= np) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    if (INTERIOR) {
        uint32_t ux, ry, dep;
        project_interior(pj, p[0], p[1], p[2], tb.x0, tb.y0, ux, ry, dep);
        if (ux < (uint32_t)tb.tw && ry < (uint32_t)tb.th) {
            hqs_band_point<PASS>(ux, ry, dep, color, tdepth, tacc, tb, pend);
            return;
        }
    }
    int ix, iy;
    uint32_t dep;
    if (!project_xy(pj, p[0], p[1], p[2], ix, iy, dep)) return;
    if (TILE) {
        const uint32_t ux = (uint32_t)(ix - tb.x0), ry = (uint32_t)(iy - tb.y0);
        if (ux < (uint32_t)tb.tw && ry < (uint32_t)tb.th) {
            hqs_band_point<PASS>(ux, ry, dep, color, tdepth, tacc, tb, pend);
            return;
        }
        if (!tb.first) return;
    }
    const uint32_t pix = (uint32_t)ix + (uint32_t)iy * (uint32_t)pj.w;
    if (PASS == 0) hqs_depth(depth, pix, dep);
    else hqs_accum(depth, acc, pix, dep, color);
}

template <int BLOCK, int PASS, int LC, bool TILE, bool NT, bool INTERIOR = false>
__device__ __forceinline__ void hqs_tile_loop(const Decode& d, const uint32_t* __restrict__ lo,
                                              const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                              const uint32_t* __restrict__ rgba, uint32_t first, uint32_t np,
                                              const Proj& pj, uint32_t* __restrict__ depth,
                                              uint32_t* __restrict__ acc, uint32_t* tdepth, unsigned long long* tacc,
                                              const TileBand& tb) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const u32x4* c4 = reinterpret_cast<const u32x4*>(rgba + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    // Pass 2 takes two adjacent quads per thread, so that a run of one
    // pixel's points (PendingSums) spans up to eight consecutive points. At
    // the default 2048-px tile its LDS (40 KiB per workgroup) allows four
    // workgroups per CU, which leaves the registers for the second quad.
    // The second quad of a pair past the batch's last is not read; its
    // points have local >= np.
    constexpr uint32_t QPT = PASS == 1 ? 2u : 1u;
    const u32x4 z = u32x4{0, 0, 0, 0};
    for (uint32_t q0 = threadIdx.x * QPT; q0 < nq; q0 += BLOCK * QPT) {
        u32x4 a[QPT], m[QPT], c[QPT], col[QPT];
#pragma unroll
        for (uint32_t k = 0; k < QPT; k++) {
            const uint32_t q = q0 + k;
            const bool in = k == 0 || q < nq;
            a[k] = in ? ld_plane<NT>(lo4 + q) : z;
            m[k] = LC <= 1 && in ? ld_plane<NT>(med4 + q) : z;
            c[k] = LC == 0 && in ? ld_plane<NT>(hi4 + q) : z;
            col[k] = PASS == 1 && in ? ld_plane<NT>(c4 + q) : z;
        }
        PendingSums pend;
#pragma unroll
        for (uint32_t k = 0; k < QPT; k++) {
            const uint32_t base = (q0 + k) * 4u - lead;
            hqs_tile_point<PASS, LC, TILE, INTERIOR>(d, a[k].x, m[k].x, c[k].x, col[k].x, base + 0, np, pj, depth,
                                                     acc, tdepth, tacc, tb, pend);
            hqs_tile_point<PASS, LC, TILE, INTERIOR>(d, a[k].y, m[k].y, c[k].y, col[k].y, base + 1, np, pj, depth,
                                                     acc, tdepth, tacc, tb, pend);
            hqs_tile_point<PASS, LC, TILE, INTERIOR>(d, a[k].z, m[k].z, c[k].z, col[k].z, base + 2, np, pj, depth,
                                                     acc, tdepth, tacc, tb, pend);
            hqs_tile_point<PASS, LC, TILE, INTERIOR>(d, a[k].w, m[k].w, c[k].w, col[k].w, base + 3, np, pj, depth,
                                                     acc, tdepth, tacc, tb, pend);
        }
        if (PASS == 1) pending_flush(pend, tacc);
    }
}

template <int BLOCK, int PASS, bool TILE, bool NT>
__device__ __forceinline__ void hqs_tile_dispatch(int level, const Decode& d, const uint32_t* __restrict__ lo,
                                                  const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                                  const uint32_t* __restrict__ rgba, const pcr_batch& h,
                                                  const Proj& pj, uint32_t* __restrict__ depth,
                                                  uint32_t* __restrict__ acc, uint32_t* tdepth,
                                                  unsigned long long* tacc, const TileBand& tb,
                                                  bool interior = false) {
    if (level == 0)
        hqs_tile_loop<BLOCK, PASS, 0, TILE, NT>(d, lo, med, hi, rgba, h.first_point, h.num_points, pj, depth, acc,
                                                tdepth, tacc, tb);
    else if (level == 1)
        hqs_tile_loop<BLOCK, PASS, 1, TILE, NT>(d, lo, med, hi, rgba, h.first_point, h.num_points, pj, depth, acc,
                                                tdepth, tacc, tb);
    else if (TILE && interior)
        hqs_tile_loop<BLOCK, PASS, 2, true, NT, true>(d, lo, med, hi, rgba, h.first_point, h.num_points, pj, depth,
                                                      acc, tdepth, tacc, tb);
    else
        hqs_tile_loop<BLOCK, PASS, 2, TILE, NT>(d, lo, med, hi, rgba, h.first_point, h.num_points, pj, depth, acc,
                                                tdepth, tacc, tb);
}

// Pass 2 keeps 20 B per tile pixel; its tile holds 24 pixels fewer than
// TILE_PIX so that with the kernels' other shared words (and the block-wide
// OR of the occlusion test) four 256-thread workgroups still fit a CU's
// 160 KB of LDS at TILE_PIX = 2048.
template <int PASS, int TILE_PIX>
struct HqsTile {
    static constexpr int kCap = PASS == 1 ? TILE_PIX - 24 : TILE_PIX;
    uint32_t depth[kCap];
    unsigned long long acc[PASS == 1 ? 2 * kCap : 1];
};

// One batch of one hqs pass; `bi` and `level` are workgroup-uniform, and so
// are the rectangle and the band count (every thread derives them from the
// same header and view). A rectangle of at most the tile's width is cut
// into bands of whole rows that each fit the tile; up to kHqsMaxBands bands
// re-read and re-project the batch's points once per band, which for a dense
// batch costs far less than the scattered global atomics of the direct path,
// and each touched pixel is flushed once, in coalesced rows.
constexpr int kHqsMaxBands = 8;

// The tile loops of the hqs passes over the n pixels of a band, rows from
// by0 of a rectangle at x0, tw pixels wide (row = i / tw as a multiply-high
// by magic = row_magic(tw)). Pass 1 flushes each touched depth with
// hqs_depth; pass 2 first loads the final depths and clears the sums, then
// adds each touched pixel's sums to the accumulators (add_pairs).
template <int BLOCK>
__device__ __forceinline__ void hqs_depth_flush(const uint32_t* tdepth, uint32_t* __restrict__ depth, int x0,
                                                uint32_t by0, int tw, int n, uint32_t magic, uint32_t width) {
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        const uint32_t t = tdepth[i];
        if (t == PCR_DEPTH_CLEAR) continue;
        const uint32_t uy = __umulhi((uint32_t)i, magic), ux = (uint32_t)i - uy * (uint32_t)tw;
        hqs_depth(depth, (uint32_t)x0 + ux + (by0 + uy) * width, t);
    }
}
template <int BLOCK>
__device__ __forceinline__ void hqs_sums_preload(uint32_t* tdepth, unsigned long long* tacc,
                                                 const uint32_t* __restrict__ depth, int x0, uint32_t by0, int tw,
                                                 int n, uint32_t magic, uint32_t width) {
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        const uint32_t uy = __umulhi((uint32_t)i, magic), ux = (uint32_t)i - uy * (uint32_t)tw;
        tdepth[i] = depth[(uint32_t)x0 + ux + (by0 + uy) * width];
        tacc[2 * i] = 0ull;
        tacc[2 * i + 1] = 0ull;
    }
}
template <int BLOCK>
__device__ __forceinline__ void hqs_sums_flush(const unsigned long long* tacc, uint32_t* __restrict__ acc, int x0,
                                               uint32_t by0, int tw, int n, uint32_t magic, uint32_t width) {
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        const unsigned long long rg = tacc[2 * i], bc = tacc[2 * i + 1];
        if ((bc >> 32) == 0ull) continue;
        const uint32_t uy = __umulhi((uint32_t)i, magic), ux = (uint32_t)i - uy * (uint32_t)tw;
        const uint32_t pix = (uint32_t)x0 + ux + (by0 + uy) * width;
        add_pairs(acc + 4ull * pix, (uint32_t)rg, (uint32_t)(rg >> 32), (uint32_t)bc, (uint32_t)(bc >> 32));
    }
}

// Returns the points skipped by the occlusion test (0 if the batch drew).
template <int BLOCK, int TILE_PIX, int PASS, bool NT>
__device__ __forceinline__ uint32_t hqs_tiled_batch(uint32_t bi, int level, const pcr_batch* __restrict__ b,
                                                const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                const uint32_t* __restrict__ hi, const uint32_t* __restrict__ rgba,
                                                const pcr_view& v, uint32_t* __restrict__ depth,
                                                uint32_t* __restrict__ acc, uint32_t* tdepth,
                                                unsigned long long* tacc, bool occl, uint32_t part = 0,
                                                uint32_t nparts = 1) {
    const pcr_batch h = b[bi];
    pcr_batch hp = h;   // the points this item draws; the tile and band choices follow the whole batch
    if (nparts > 1) {   // points [part * k, (part + 1) * k), k = ceil(n / nparts)
        const uint32_t k = (h.num_points + nparts - 1) / nparts, off = min(part * k, h.num_points);
        hp.first_point += off;
        hp.num_points = min(k, h.num_points - off);
    }
    const Decode d = make_decode(level, h.min, h.max);
    const Proj pj = make_proj(v.transform, v.width, v.height);
    int x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    bool interior;
    float wlo;
    const bool on_screen = screen_rect(v.transform, h.min, h.max, v.width, v.height, x0, y0, x1, y1, &interior,
                                       &wlo);
    const int tw = x1 - x0 + 1, th = y1 - y0 + 1;
    // pass 1: behind every depth already drawn over the rectangle; pass 2:
    // behind 1.01x every final depth there (rect_open)
    if (occl && wlo > 0.0f && tw >= 2 && tw * th <= kOcclusionMaxPix &&
        !rect_open<BLOCK, PASS == 0 ? 1 : 2>(depth, v.width, x0, y0, tw, tw * th, wlo))
        return hp.num_points;
    constexpr int cap = HqsTile<PASS, TILE_PIX>::kCap;
    // one-column rectangles draw directly, as in tiled_batch (tw >= 2 keeps
    // the row split's ceil(2^32 / tw) within 32 bits)
    const bool fits = on_screen && tw >= 2 && tw <= cap && h.num_points <= kHqsTileMaxPoints;
    const int bh = fits ? min(th, cap / tw) : 1;    // rows per band
    const int nbands = fits ? (th + bh - 1) / bh : 0;
    // bands pay when the batch is dense (a point or more per rectangle pixel,
    // as near batches of a dense cloud are); a sparse one, such as a near LOD
    // node, costs more to re-project per band than its few atomics do
    if (!fits || nbands > kHqsMaxBands || (nbands > 1 && (uint64_t)h.num_points < (uint64_t)tw * th)) {
        const TileBand none{0, 0, 0, 0, 0, 0, true};
        hqs_tile_dispatch<BLOCK, PASS, false, NT>(level, d, lo, med, hi, rgba, hp, pj, depth, acc, nullptr, nullptr,
                                                  none);
        return 0;
    }
    const uint32_t magic = row_magic(tw), width = (uint32_t)v.width;
    for (int k = 0; k < nbands; k++) {
        const int boff = k * bh, rows = min(bh, th - boff);
        const int n = tw * rows;
        const uint32_t by0 = (uint32_t)(y0 + boff);
        if (k) __syncthreads();   // the previous band's flush has read the tile
        if (PASS == 0)
            for (int i = threadIdx.x; i < n; i += BLOCK) tdepth[i] = PCR_DEPTH_CLEAR;
        else
            hqs_sums_preload<BLOCK>(tdepth, tacc, depth, x0, by0, tw, n, magic, width);
        __syncthreads();
        const TileBand tb{x0, y0, tw, th, boff, rows, k == 0};
        hqs_tile_dispatch<BLOCK, PASS, true, NT>(level, d, lo, med, hi, rgba, hp, pj, depth, acc, tdepth, tacc, tb,
                                                 interior);
        __syncthreads();
        if (PASS == 0) hqs_depth_flush<BLOCK>(tdepth, depth, x0, by0, tw, n, magic, width);
        else hqs_sums_flush<BLOCK>(tacc, acc, x0, by0, tw, n, magic, width);
    }
    return 0;
}

// hqs pass 1 of one batch for a pair of views at once, the next two views a
// multi-view launch draws the batch for, at one precision level: each point
// is decoded once and projected for both views, into two depth tiles. Used
// when both views' rectangles are interior (screen_rect) and fit the tile in
// one band; the caller draws the views one at a time otherwise. Per view the
// operations are those of hqs_tiled_batch's pass 1 (hqs_tile_point with one
// band), so each depth buffer is the same; each view keeps its own
// occlusion test.

// The post-decode half of hqs_tile_point for pass 1 of an interior batch
// with one band: the tile pixel's depth, or the direct path for a point that
// rounding puts outside the rectangle.
__device__ __forceinline__ void hqs1_view_point(const float p[3], const Proj& pj, uint32_t* __restrict__ depth,
                                                uint32_t* tdepth, int x0, int y0, int tw, int th) {
    uint32_t ux, uy, dep;
    project_interior(pj, p[0], p[1], p[2], x0, y0, ux, uy, dep);
    if (ux < (uint32_t)tw && uy < (uint32_t)th) {
        lds_min32(tdepth + uy * (uint32_t)tw + ux, dep);
        return;
    }
    int ix, iy;
    if (!project_xy(pj, p[0], p[1], p[2], ix, iy, dep)) return;
    const uint32_t vx = (uint32_t)(ix - x0), vy = (uint32_t)(iy - y0);
    if (vx < (uint32_t)tw && vy < (uint32_t)th) {
        lds_min32(tdepth + vy * (uint32_t)tw + vx, dep);
        return;
    }
    hqs_depth(depth, (uint32_t)ix + (uint32_t)iy * (uint32_t)pj.w, dep);
}

// One view of a pair: its projection, buffers, rectangle and whether it draws.
struct PairView {
    Proj pj;
    uint32_t* depth;
    uint32_t* tdepth;
    int x0, y0, tw, th;
    bool draw;
};

template <int BLOCK, int LC, bool NT>
__device__ __forceinline__ void hqs1_pair_loop(const Decode& d, const uint32_t* __restrict__ lo,
                                               const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                               uint32_t first, uint32_t np, const PairView& a, const PairView& b) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 w0 = ld_plane<NT>(lo4 + q);
        const u32x4 w1 = LC <= 1 ? ld_plane<NT>(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 w2 = LC == 0 ? ld_plane<NT>(hi4 + q) : u32x4{0, 0, 0, 0};
        const uint32_t base = q * 4u - lead;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (base + (uint32_t)j >= np) continue;
            float p[3];
            decode<LC>(d, w0[j], w1[j], w2[j], p);
            if (a.draw) hqs1_view_point(p, a.pj, a.depth, a.tdepth, a.x0, a.y0, a.tw, a.th);
            if (b.draw) hqs1_view_point(p, b.pj, b.depth, b.tdepth, b.x0, b.y0, b.tw, b.th);
        }
    }
}

// Returns false, having done nothing, when the pair cannot be drawn together.
template <int BLOCK, int TILE_PIX, bool NT>
__device__ __forceinline__ bool hqs1_pair_batch(uint32_t bi, int level, const pcr_batch* __restrict__ b,
                                                const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                const uint32_t* __restrict__ hi, const pcr_view& va,
                                                uint32_t* __restrict__ depth_a, const pcr_view& vb,
                                                uint32_t* __restrict__ depth_b, uint32_t* tdepth_a,
                                                uint32_t* tdepth_b, bool occl, uint32_t part, uint32_t nparts) {
    const pcr_batch h = b[bi];
    pcr_batch hp = h;   // as in hqs_tiled_batch: the points this item draws
    if (nparts > 1) {
        const uint32_t k = (h.num_points + nparts - 1) / nparts, off = min(part * k, h.num_points);
        hp.first_point += off;
        hp.num_points = min(k, h.num_points - off);
    }
    int ax0 = 0, ay0 = 0, ax1 = -1, ay1 = -1, bx0 = 0, by0 = 0, bx1 = -1, by1 = -1;
    bool ai, bi_;
    float awlo, bwlo;
    const bool aon = screen_rect(va.transform, h.min, h.max, va.width, va.height, ax0, ay0, ax1, ay1, &ai, &awlo);
    const bool bon = screen_rect(vb.transform, h.min, h.max, vb.width, vb.height, bx0, by0, bx1, by1, &bi_, &bwlo);
    const int atw = ax1 - ax0 + 1, ath = ay1 - ay0 + 1, btw = bx1 - bx0 + 1, bth = by1 - by0 + 1;
    constexpr int cap = HqsTile<0, TILE_PIX>::kCap;
    if (!(aon && bon && ai && bi_ && atw >= 2 && btw >= 2 && atw * ath <= cap && btw * bth <= cap &&
          h.num_points <= kHqsTileMaxPoints))
        return false;
    // each view's occlusion test, as in hqs_tiled_batch
    const bool da = !(occl && awlo > 0.0f && atw * ath <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 1>(depth_a, va.width, ax0, ay0, atw, atw * ath, awlo);
    const bool db = !(occl && bwlo > 0.0f && btw * bth <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 1>(depth_b, vb.width, bx0, by0, btw, btw * bth, bwlo);
    if (!da && !db) return true;
    const Decode d = make_decode(level, h.min, h.max);
    const PairView a{make_proj(va.transform, va.width, va.height), depth_a, tdepth_a, ax0, ay0, atw, ath, da};
    const PairView c{make_proj(vb.transform, vb.width, vb.height), depth_b, tdepth_b, bx0, by0, btw, bth, db};
    for (int i = threadIdx.x; i < atw * ath; i += BLOCK) tdepth_a[i] = PCR_DEPTH_CLEAR;
    for (int i = threadIdx.x; i < btw * bth; i += BLOCK) tdepth_b[i] = PCR_DEPTH_CLEAR;
    __syncthreads();
    if (level == 0) hqs1_pair_loop<BLOCK, 0, NT>(d, lo, med, hi, hp.first_point, hp.num_points, a, c);
    else if (level == 1) hqs1_pair_loop<BLOCK, 1, NT>(d, lo, med, hi, hp.first_point, hp.num_points, a, c);
    else hqs1_pair_loop<BLOCK, 2, NT>(d, lo, med, hi, hp.first_point, hp.num_points, a, c);
    __syncthreads();
    // each view's flush, as in hqs_tiled_batch
    if (a.draw)
        hqs_depth_flush<BLOCK>(a.tdepth, a.depth, a.x0, (uint32_t)a.y0, a.tw, a.tw * a.th, row_magic(a.tw),
                               (uint32_t)a.pj.w);
    if (c.draw)
        hqs_depth_flush<BLOCK>(c.tdepth, c.depth, c.x0, (uint32_t)c.y0, c.tw, c.tw * c.th, row_magic(c.tw),
                               (uint32_t)c.pj.w);
    return true;
}

// hqs pass 2 of one batch for a pair of views at once, under the same
// conditions as hqs1_pair_batch but with each view's rectangle fitting half
// the pass-2 tile: the two halves hold each view's final depths and sums.
// Per view the operations are those of hqs_tiled_batch's pass 2 with one band
// (hqs_tile_point, the run merge of PendingSums, the add_pairs flush), so
// each view's accumulators are the same; each view keeps its own occlusion
// test.
struct Pair2View {
    Proj pj;
    uint32_t* depth;
    uint32_t* acc;
    uint32_t* tdepth;
    unsigned long long* tacc;
    TileBand tb;
    bool draw;
};

// The post-decode half of hqs_tile_point for pass 2 of an interior batch with
// one band.
__device__ __forceinline__ void hqs2_view_point(const float p[3], uint32_t color, const Pair2View& pv,
                                                PendingSums& pend) {
    uint32_t ux, ry, dep;
    project_interior(pv.pj, p[0], p[1], p[2], pv.tb.x0, pv.tb.y0, ux, ry, dep);
    if (ux < (uint32_t)pv.tb.tw && ry < (uint32_t)pv.tb.th) {
        hqs_band_point<1>(ux, ry, dep, color, pv.tdepth, pv.tacc, pv.tb, pend);
        return;
    }
    int ix, iy;
    if (!project_xy(pv.pj, p[0], p[1], p[2], ix, iy, dep)) return;
    const uint32_t vx = (uint32_t)(ix - pv.tb.x0), vy = (uint32_t)(iy - pv.tb.y0);
    if (vx < (uint32_t)pv.tb.tw && vy < (uint32_t)pv.tb.th) {
        hqs_band_point<1>(vx, vy, dep, color, pv.tdepth, pv.tacc, pv.tb, pend);
        return;
    }
    hqs_accum(pv.depth, pv.acc, (uint32_t)ix + (uint32_t)iy * (uint32_t)pv.pj.w, dep, color);
}

template <int BLOCK, int LC, bool NT>
__device__ __forceinline__ void hqs2_pair_loop(const Decode& d, const uint32_t* __restrict__ lo,
                                               const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                               const uint32_t* __restrict__ rgba, uint32_t first, uint32_t np,
                                               const Pair2View& a, const Pair2View& c) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const u32x4* c4 = reinterpret_cast<const u32x4*>(rgba + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 w0 = ld_plane<NT>(lo4 + q);
        const u32x4 w1 = LC <= 1 ? ld_plane<NT>(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 w2 = LC == 0 ? ld_plane<NT>(hi4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 col = ld_plane<NT>(c4 + q);
        const uint32_t base = q * 4u - lead;
        PendingSums pa, pc;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (base + (uint32_t)j >= np) continue;
            float p[3];
            decode<LC>(d, w0[j], w1[j], w2[j], p);
            if (a.draw) hqs2_view_point(p, col[j], a, pa);
            if (c.draw) hqs2_view_point(p, col[j], c, pc);
        }
        pending_flush(pa, a.tacc);
        pending_flush(pc, c.tacc);
    }
}

// Returns false, having done nothing, when the pair cannot be drawn together.
template <int BLOCK, int TILE_PIX, bool NT>
__device__ __forceinline__ bool hqs2_pair_batch(uint32_t bi, int level, const pcr_batch* __restrict__ b,
                                                const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                const uint32_t* __restrict__ hi, const uint32_t* __restrict__ rgba,
                                                const pcr_view& va, uint32_t* __restrict__ depth_a,
                                                uint32_t* __restrict__ acc_a, const pcr_view& vb,
                                                uint32_t* __restrict__ depth_b, uint32_t* __restrict__ acc_b,
                                                uint32_t* tdepth, unsigned long long* tacc, bool occl, uint32_t part,
                                                uint32_t nparts) {
    const pcr_batch h = b[bi];
    pcr_batch hp = h;   // as in hqs_tiled_batch: the points this item draws
    if (nparts > 1) {
        const uint32_t k = (h.num_points + nparts - 1) / nparts, off = min(part * k, h.num_points);
        hp.first_point += off;
        hp.num_points = min(k, h.num_points - off);
    }
    int ax0 = 0, ay0 = 0, ax1 = -1, ay1 = -1, bx0 = 0, by0 = 0, bx1 = -1, by1 = -1;
    bool ai, bi_;
    float awlo, bwlo;
    const bool aon = screen_rect(va.transform, h.min, h.max, va.width, va.height, ax0, ay0, ax1, ay1, &ai, &awlo);
    const bool bon = screen_rect(vb.transform, h.min, h.max, vb.width, vb.height, bx0, by0, bx1, by1, &bi_, &bwlo);
    const int atw = ax1 - ax0 + 1, ath = ay1 - ay0 + 1, btw = bx1 - bx0 + 1, bth = by1 - by0 + 1;
    constexpr int half = HqsTile<1, TILE_PIX>::kCap / 2;
    if (!(aon && bon && ai && bi_ && atw >= 2 && btw >= 2 && atw * ath <= half && btw * bth <= half &&
          h.num_points <= kHqsTileMaxPoints))
        return false;
    // each view's occlusion test, as in hqs_tiled_batch's pass 2
    const bool da = !(occl && awlo > 0.0f && atw * ath <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 2>(depth_a, va.width, ax0, ay0, atw, atw * ath, awlo);
    const bool db = !(occl && bwlo > 0.0f && btw * bth <= kOcclusionMaxPix) ||
                    rect_open<BLOCK, 2>(depth_b, vb.width, bx0, by0, btw, btw * bth, bwlo);
    if (!da && !db) return true;
    const Decode d = make_decode(level, h.min, h.max);
    const Pair2View a{make_proj(va.transform, va.width, va.height), depth_a, acc_a, tdepth, tacc,
                      TileBand{ax0, ay0, atw, ath, 0, ath, true}, da};
    const Pair2View c{make_proj(vb.transform, vb.width, vb.height), depth_b, acc_b, tdepth + half, tacc + 2 * half,
                      TileBand{bx0, by0, btw, bth, 0, bth, true}, db};
    // each drawing view's final depths into its half, its sums cleared
    const uint32_t am = row_magic(atw), cm = row_magic(btw);
    if (a.draw)
        hqs_sums_preload<BLOCK>(a.tdepth, a.tacc, depth_a, ax0, (uint32_t)ay0, atw, atw * ath, am, (uint32_t)va.width);
    if (c.draw)
        hqs_sums_preload<BLOCK>(c.tdepth, c.tacc, depth_b, bx0, (uint32_t)by0, btw, btw * bth, cm, (uint32_t)vb.width);
    __syncthreads();
    if (level == 0) hqs2_pair_loop<BLOCK, 0, NT>(d, lo, med, hi, rgba, hp.first_point, hp.num_points, a, c);
    else if (level == 1) hqs2_pair_loop<BLOCK, 1, NT>(d, lo, med, hi, rgba, hp.first_point, hp.num_points, a, c);
    else hqs2_pair_loop<BLOCK, 2, NT>(d, lo, med, hi, rgba, hp.first_point, hp.num_points, a, c);
    __syncthreads();
    // each view's flush, as in hqs_tiled_batch
    if (a.draw) hqs_sums_flush<BLOCK>(a.tacc, acc_a, ax0, (uint32_t)ay0, atw, atw * ath, am, (uint32_t)va.width);
    if (c.draw) hqs_sums_flush<BLOCK>(c.tacc, acc_b, bx0, (uint32_t)by0, btw, btw * bth, cm, (uint32_t)vb.width);
    return true;
}

// Persistent workgroups over the select stage's list, in the spread order of
// k_raster_tiled; pass 1 takes items from queue_head, pass 2 (launched after
// pass 1 has finished) from queue_head2. Pass 1 with tiles up to 2048 px is
// held to 7 waves per SIMD (72 registers), which it reaches without spilling.
template <int BLOCK, int TILE_PIX, int PASS>
__global__ __launch_bounds__(BLOCK, PASS == 0 && TILE_PIX <= 2048 ? 7 : 1)
void k_hqs_tiled(const pcr_batch* __restrict__ b,
                 const uint8_t* __restrict__ levels,
                 const uint32_t* __restrict__ list,
                 FrameState* __restrict__ st,
                 const uint32_t* __restrict__ lo,
                 const uint32_t* __restrict__ med,
                 const uint32_t* __restrict__ hi,
                 const uint32_t* __restrict__ rgba, pcr_view v,
                 uint32_t* __restrict__ depth,
                 uint32_t* __restrict__ acc, int occl) {
    __shared__ HqsTile<PASS, TILE_PIX> tile;
    __shared__ uint32_t s_item;
    uint32_t* head = PASS == 0 ? &st->queue_head : &st->queue_head2;
    const uint32_t count = st->list_count;
    const uint32_t nsplit = min(count, gridDim.x), whole = count - nsplit;
    __shared__ uint32_t s_occ_batches;
    __shared__ unsigned long long s_occ_points, s_occ_bytes;   // what the occlusion test skipped
    if (threadIdx.x == 0) { s_occ_batches = 0; s_occ_points = 0; s_occ_bytes = 0; }
    QueueTake q;
    for (;;) {
        uint32_t item = q.take(head, s_item);
        if (item >= count + nsplit) break;
        // as in k_raster_tiled, the last nsplit batches are drawn as two halves,
        // all first halves before the second ones
        uint32_t part = 0, nparts = 1;
        if (item >= whole) {
            const uint32_t h = item - whole;
            part = h >= nsplit ? 1u : 0u;
            nparts = 2;
            item = whole + (part ? h - nsplit : h);
        }
        const uint32_t slot = (uint32_t)(((uint64_t)item * 2654435761ull) % count);
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[slot]);
        const int level = __builtin_amdgcn_readfirstlane(levels[bi]);
        const uint32_t skipped = hqs_tiled_batch<BLOCK, TILE_PIX, PASS, true>(bi, level, b, lo, med, hi, rgba, v,
                                                                              depth, acc, tile.depth, tile.acc,
                                                                              occl != 0, part, nparts);
        if (threadIdx.x == 0 && skipped) {
            s_occ_points += skipped;
            s_occ_bytes += (4ull * planes_for_level(level) + (PASS == 1 ? 4ull : 0ull)) * skipped;
            s_occ_batches += part == 0;
        }
        __syncthreads();   // s_item and the tile are reused by the next item
    }
    if (threadIdx.x == 0 && s_occ_points) {
        atomicAdd(&st->occluded_batches, s_occ_batches);
        atomicAdd(&st->occluded_points, s_occ_points);
        atomicAdd(&st->occluded_bytes, s_occ_bytes);
    }
}

template <int BLOCK, int TILE_PIX, int PASS>
static void hqs_tiled_t(hipStream_t s, const pcr_cloud* c, const pcr_view& v, FrameState* st, pcr_fb* fb,
                        bool occl) {
    static const int occ = [] {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_hqs_tiled<BLOCK, TILE_PIX, PASS>, BLOCK, 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            o = 0;
        }
        return o > 0 ? o : 4;
    }();
    uint64_t grid = (uint64_t)c->ctx->cu_count * (uint64_t)occ;
    if (grid > c->num_batches) grid = c->num_batches;
    k_hqs_tiled<BLOCK, TILE_PIX, PASS><<<(int)grid, BLOCK, 0, s>>>(c->d_batches, c->d_levels, c->d_list, st, c->d_lo,
                                                                   c->d_med, c->d_hi, c->d_rgba, v, fb->d_depth,
                                                                   fb->d_acc, occl ? 1 : 0);
}

template <int BLOCK, int TILE_PIX>
static void hqs_tiled_passes(hipStream_t s, const pcr_cloud* c, const pcr_view& v, FrameState* st, pcr_fb* fb,
                             bool occl) {
    hqs_tiled_t<BLOCK, TILE_PIX, 0>(s, c, v, st, fb, occl);
    hqs_tiled_t<BLOCK, TILE_PIX, 1>(s, c, v, st, fb, occl);
}

int launch_hqs_tiled(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int tile_px, int block, FrameState* st,
                     pcr_fb* fb, bool occl) {
    if (c->num_batches == 0) return PCR_OK;
    if (block == 128) {
        if (tile_px == 1024) hqs_tiled_passes<128, 1024>(s, c, v, st, fb, occl);
        else if (tile_px == 4096) hqs_tiled_passes<128, 4096>(s, c, v, st, fb, occl);
        else hqs_tiled_passes<128, 2048>(s, c, v, st, fb, occl);
    } else {
        if (tile_px == 1024) hqs_tiled_passes<256, 1024>(s, c, v, st, fb, occl);
        else if (tile_px == 4096) hqs_tiled_passes<256, 4096>(s, c, v, st, fb, occl);
        else hqs_tiled_passes<256, 2048>(s, c, v, st, fb, occl);
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// Multi-view tiled hqs: one pass of k_hqs_tiled over every view that draws
// the batch, as k_mv_raster_tiled does for keys. Pass 1 takes items from
// queue_head, pass 2 from queue_head2.
template <int BLOCK, int TILE_PIX, int PASS>
__global__ __launch_bounds__(BLOCK) void k_mv_hqs_tiled(const pcr_batch* __restrict__ b, uint32_t nb,
                                                        const uint8_t* __restrict__ levels,
                                                        const uint32_t* __restrict__ list, MVState* __restrict__ st,
                                                        const uint32_t* __restrict__ lo,
                                                        const uint32_t* __restrict__ med,
                                                        const uint32_t* __restrict__ hi,
                                                        const uint32_t* __restrict__ rgba,
                                                        const pcr_view* __restrict__ views, int nviews,
                                                        uint32_t* const* __restrict__ depths,
                                                        uint32_t* const* __restrict__ accs, int occl) {
    __shared__ HqsTile<PASS, TILE_PIX> tile;
    __shared__ uint32_t tile2[PASS == 0 ? TILE_PIX : 1];   // pass 1: the second view of a pair
    __shared__ uint32_t s_item, s_nv;
    __shared__ uint8_t s_vid[64], s_lv[64];
    uint32_t* head = PASS == 0 ? &st->queue_head : &st->queue_head2;
    const uint32_t count = st->list_count;
    const uint32_t nsplit = min(count, gridDim.x), whole = count - nsplit;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
        __syncthreads();
        uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= count + nsplit) break;
        // as in k_hqs_tiled, the last item per workgroup is drawn as two halves
        uint32_t part = 0, nparts = 1;
        if (item >= whole) { part = (item - whole) & 1u; nparts = 2; item = whole + ((item - whole) >> 1); }
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[item]);
        if (threadIdx.x < 64) {   // compact the views drawing this batch (wave 0)
            const int v = (int)threadIdx.x;
            const int lv = v < nviews ? levels[(uint64_t)v * nb + bi] : kCulled;
            const unsigned long long m = __ballot(lv != kCulled);
            if (lv != kCulled) {
                const uint32_t pos = (uint32_t)__popcll(m & ((1ull << v) - 1ull));
                s_vid[pos] = (uint8_t)v;
                s_lv[pos] = (uint8_t)lv;
            }
            if (v == 0) s_nv = (uint32_t)__popcll(m);
        }
        __syncthreads();
        const uint32_t nv = __builtin_amdgcn_readfirstlane(s_nv);
        for (uint32_t k = 0; k < nv;) {
            const int vi = __builtin_amdgcn_readfirstlane(s_vid[k]);
            const int lv = __builtin_amdgcn_readfirstlane(s_lv[k]);
            // pass 1: the next two views at one level in one pass over the points
            if (PASS == 0 && k + 1 < nv) {
                const int vj = __builtin_amdgcn_readfirstlane(s_vid[k + 1]);
                const int lj = __builtin_amdgcn_readfirstlane(s_lv[k + 1]);
                if (lv == lj && hqs1_pair_batch<BLOCK, TILE_PIX, false>(bi, lv, b, lo, med, hi, views[vi], depths[vi],
                                                                         views[vj], depths[vj], tile.depth, tile2,
                                                                         occl != 0, part, nparts)) {
                    __syncthreads();   // the tiles are reused by the next views
                    k += 2;
                    continue;
                }
            }
            if constexpr (PASS == 1) {
                if (k + 1 < nv) {   // pass 2: the next two views at one level, each in half the tile
                    const int vj = __builtin_amdgcn_readfirstlane(s_vid[k + 1]);
                    const int lj = __builtin_amdgcn_readfirstlane(s_lv[k + 1]);
                    if (lv == lj && hqs2_pair_batch<BLOCK, TILE_PIX, false>(bi, lv, b, lo, med, hi, rgba, views[vi],
                                                                             depths[vi], accs[vi], views[vj],
                                                                             depths[vj], accs[vj], tile.depth,
                                                                             tile.acc, occl != 0, part, nparts)) {
                        __syncthreads();   // the tile is reused by the next views
                        k += 2;
                        continue;
                    }
                }
            }
            hqs_tiled_batch<BLOCK, TILE_PIX, PASS, false>(bi, lv, b, lo, med, hi, rgba, views[vi], depths[vi],
                                                           accs[vi], tile.depth, tile.acc, occl != 0, part, nparts);
            __syncthreads();   // the tile is reused by the next view
            k++;
        }
    }
}

template <int BLOCK, int TILE_PIX, int PASS>
static void mv_hqs_tiled_t(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int cu_count,
                           bool occl) {
    static const int occ = [] {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_mv_hqs_tiled<BLOCK, TILE_PIX, PASS>, BLOCK, 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            o = 0;
        }
        return o > 0 ? o : 4;
    }();
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    if (grid > c->num_batches) grid = c->num_batches;
    if (grid == 0) grid = 1;
    k_mv_hqs_tiled<BLOCK, TILE_PIX, PASS><<<(int)grid, BLOCK, 0, s>>>(c->d_batches, c->num_batches, B.d_levels,
                                                                      B.d_list, B.d_state, c->d_lo, c->d_med, c->d_hi,
                                                                      c->d_rgba, B.d_views, nviews, B.d_depths,
                                                                      B.d_accs, occl ? 1 : 0);
}

template <int BLOCK, int TILE_PIX>
static void mv_hqs_tiled_passes(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int cu_count,
                                bool occl) {
    mv_hqs_tiled_t<BLOCK, TILE_PIX, 0>(s, c, B, nviews, cu_count, occl);
    mv_hqs_tiled_t<BLOCK, TILE_PIX, 1>(s, c, B, nviews, cu_count, occl);
}

int launch_mv_hqs_tiled(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int tile_px, int block,
                        int cu_count, bool occl) {
    if (block == 128) {
        if (tile_px == 1024) mv_hqs_tiled_passes<128, 1024>(s, c, B, nviews, cu_count, occl);
        else if (tile_px == 4096) mv_hqs_tiled_passes<128, 4096>(s, c, B, nviews, cu_count, occl);
        else mv_hqs_tiled_passes<128, 2048>(s, c, B, nviews, cu_count, occl);
    } else {
        if (tile_px == 1024) mv_hqs_tiled_passes<256, 1024>(s, c, B, nviews, cu_count, occl);
        else if (tile_px == 4096) mv_hqs_tiled_passes<256, 4096>(s, c, B, nviews, cu_count, occl);
        else mv_hqs_tiled_passes<256, 2048>(s, c, B, nviews, cu_count, occl);
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

int launch_raster_batched(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int vec, int block,
                          FrameState* st, unsigned long long* fb) {
    if (c->num_batches == 0) return PCR_OK;
    if (vec == 4) {
        switch (block) {
            case 64: raster_batched_t<64, 4>(s, c, v, fb, st); break;
            case 128: raster_batched_t<128, 4>(s, c, v, fb, st); break;
            case 512: raster_batched_t<512, 4>(s, c, v, fb, st); break;
            default: raster_batched_t<256, 4>(s, c, v, fb, st); break;
        }
    } else {
        switch (block) {
            case 64: raster_batched_t<64, 1>(s, c, v, fb, st); break;
            case 256: raster_batched_t<256, 1>(s, c, v, fb, st); break;
            case 512: raster_batched_t<512, 1>(s, c, v, fb, st); break;
            default: raster_batched_t<128, 1>(s, c, v, fb, st); break;
        }
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// 16 B/pt points (x, y, z, rgba), colour payload.
__global__ __launch_bounds__(256) void k_raster_points16(const u32x4* __restrict__ pts, uint64_t n, pcr_view v,
                                                         unsigned long long* __restrict__ fb) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32x4 p = ld_stream(pts + i);
        uint32_t pix, depth;
        if (!project(v.transform, __uint_as_float(p.x), __uint_as_float(p.y), __uint_as_float(p.z), v.width,
                     v.height, pix, depth))
            continue;
        fb_min(fb, pix, ((unsigned long long)depth << 32) | p.w);
    }
}

int launch_raster_points16(hipStream_t s, const pcr_cloud* c, const pcr_view& v, unsigned long long* fb) {
    if (c->points16_len == 0) return PCR_OK;
    uint64_t g = (c->points16_len + 255) / 256;
    if (g > (uint64_t)c->ctx->cu_count * 64) g = (uint64_t)c->ctx->cu_count * 64;
    k_raster_points16<<<(int)g, 256, 0, s>>>(reinterpret_cast<const u32x4*>(c->d_points16), c->points16_len, v, fb);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// -------------------------------------------------------------------- hqs
// The per-pixel operations (hqs_depth, hqs_accum) are in pcr_device.h.
template <int PASS, int LC>
__device__ __forceinline__ void hqs_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi, uint32_t color,
                                          uint32_t local, uint32_t np, const pcr_view& v,
                                          uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    if (local >= np) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    uint32_t pix, dep;
    if (!project(v.transform, p[0], p[1], p[2], v.width, v.height, pix, dep)) return;
    if (PASS == 0) hqs_depth(depth, pix, dep);
    else hqs_accum(depth, acc, pix, dep, color);
}

template <int BLOCK, int PASS, int LC>
__device__ __forceinline__ void hqs_vec4(const Decode& d, const uint32_t* __restrict__ lo,
                                         const uint32_t* __restrict__ med, const uint32_t* __restrict__ hi,
                                         const uint32_t* __restrict__ rgba, uint32_t first, uint32_t np,
                                         const pcr_view& v, uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    const uint32_t lead = first & 3u;   // see tile_loop()
    const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
    const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
    const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
    const u32x4* c4 = reinterpret_cast<const u32x4*>(rgba + (first - lead));
    const uint32_t nq = (np + lead + 3u) >> 2;
    for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
        const u32x4 a = ld_stream(lo4 + q);
        const u32x4 m = LC <= 1 ? ld_stream(med4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 c = LC == 0 ? ld_stream(hi4 + q) : u32x4{0, 0, 0, 0};
        const u32x4 col = PASS == 1 ? ld_stream(c4 + q) : u32x4{0, 0, 0, 0};
        const uint32_t base = q * 4u - lead;
        hqs_point<PASS, LC>(d, a.x, m.x, c.x, col.x, base + 0, np, v, depth, acc);
        hqs_point<PASS, LC>(d, a.y, m.y, c.y, col.y, base + 1, np, v, depth, acc);
        hqs_point<PASS, LC>(d, a.z, m.z, c.z, col.z, base + 2, np, v, depth, acc);
        hqs_point<PASS, LC>(d, a.w, m.w, c.w, col.w, base + 3, np, v, depth, acc);
    }
}

template <int BLOCK, int PASS>
__global__ __launch_bounds__(BLOCK) void k_hqs_batched(const pcr_batch* __restrict__ b,
                                                       const uint8_t* __restrict__ levels,
                                                       const uint32_t* __restrict__ lo,
                                                       const uint32_t* __restrict__ med,
                                                       const uint32_t* __restrict__ hi,
                                                       const uint32_t* __restrict__ rgba, pcr_view v,
                                                       uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    const uint32_t bi = blockIdx.x;
    const int level = levels[bi];
    if (level == kCulled) return;
    const pcr_batch h = b[bi];
    const Decode d = make_decode(level, h.min, h.max);
    if (level == 0) hqs_vec4<BLOCK, PASS, 0>(d, lo, med, hi, rgba, h.first_point, h.num_points, v, depth, acc);
    else if (level == 1) hqs_vec4<BLOCK, PASS, 1>(d, lo, med, hi, rgba, h.first_point, h.num_points, v, depth, acc);
    else hqs_vec4<BLOCK, PASS, 2>(d, lo, med, hi, rgba, h.first_point, h.num_points, v, depth, acc);
}

int launch_hqs_batched(hipStream_t s, const pcr_cloud* c, const pcr_view& v, FrameState* st, pcr_fb* fb) {
    (void)st;
    if (c->num_batches == 0) return PCR_OK;
    k_hqs_batched<256, 0><<<c->num_batches, 256, 0, s>>>(c->d_batches, c->d_levels, c->d_lo, c->d_med, c->d_hi,
                                                         c->d_rgba, v, fb->d_depth, fb->d_acc);
    PCR_HIP(hipGetLastError());
    k_hqs_batched<256, 1><<<c->num_batches, 256, 0, s>>>(c->d_batches, c->d_levels, c->d_lo, c->d_med, c->d_hi,
                                                         c->d_rgba, v, fb->d_depth, fb->d_acc);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

template <int PASS>
__global__ __launch_bounds__(256) void k_hqs_points16(const u32x4* __restrict__ pts, uint64_t n, pcr_view v,
                                                      uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32x4 p = ld_stream(pts + i);
        uint32_t pix, dep;
        if (!project(v.transform, __uint_as_float(p.x), __uint_as_float(p.y), __uint_as_float(p.z), v.width,
                     v.height, pix, dep))
            continue;
        if (PASS == 0) hqs_depth(depth, pix, dep);
        else hqs_accum(depth, acc, pix, dep, p.w);
    }
}

int launch_hqs_points16(hipStream_t s, const pcr_cloud* c, const pcr_view& v, pcr_fb* fb) {
    if (c->points16_len == 0) return PCR_OK;
    uint64_t g = (c->points16_len + 255) / 256;
    if (g > (uint64_t)c->ctx->cu_count * 64) g = (uint64_t)c->ctx->cu_count * 64;
    const u32x4* pts = reinterpret_cast<const u32x4*>(c->d_points16);
    k_hqs_points16<0><<<(int)g, 256, 0, s>>>(pts, c->points16_len, v, fb->d_depth, fb->d_acc);
    PCR_HIP(hipGetLastError());
    k_hqs_points16<1><<<(int)g, 256, 0, s>>>(pts, c->points16_len, v, fb->d_depth, fb->d_acc);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

// ---------------------------------------------------------------- resolve
// Index payload: colour gathered from the colour plane (resolve.cs:117-146).
// An index outside the cloud (a framebuffer filled by another cloud or mode)
// resolves to the background instead of reading out of bounds. strength > 0
// adds the eye-dome lighting shade (resolve_index_pixel). 256 threads per
// workgroup, one 16x16 tile per iteration (tile_pixel).
// `visible` (optional) counts the pixels holding a point, as the reference's
// resolve counts numPointsVisible (resolve.cs:143-145): one atomic per wave.
__global__ __launch_bounds__(256) void k_resolve_index(const unsigned long long* __restrict__ fb, int w, int h,
                                                       const uint32_t* __restrict__ rgba, uint64_t npoints,
                                                       uint32_t bg, float strength, uint32_t* __restrict__ out,
                                                       unsigned long long* __restrict__ visible) {
    const uint32_t tiles_x = ((uint32_t)w + 15u) >> 4, tiles = tiles_x * (((uint32_t)h + 15u) >> 4);
    uint32_t covered = 0;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t x, y;
        tile_pixel(tile, tiles_x, x, y);
        if (x >= (uint32_t)w || y >= (uint32_t)h) continue;
        const uint64_t i = (uint64_t)y * (uint32_t)w + x;
        const unsigned long long k = fb[i];
        const uint32_t id = (uint32_t)k;
        const bool point = k != PCR_FB_CLEAR && id < npoints;
        covered += point;
        if (strength > 0.0f) out[i] = resolve_index_pixel(fb, w, h, (int)x, (int)y, rgba, npoints, bg, strength);
        else out[i] = point ? rgba[id] : bg;
    }
    if (visible) {
        const unsigned long long n = wave_sum_u64(covered);
        if (__lane_id() == 0 && n) atomicAdd(visible, n);
    }
}
// Colour payload: the key's low 32 bits are the RGBA (2021 resolve.cs:45-57).
__global__ void k_resolve_color(const unsigned long long* __restrict__ fb, uint64_t npix, uint32_t bg,
                                uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
        const unsigned long long k = fb[i];
        out[i] = k == PCR_FB_CLEAR ? bg : (uint32_t)k;
    }
}
// hqs/resolve.cs:74-129 (hqs_average).
__global__ void k_resolve_hqs(const u32x4* __restrict__ acc, uint64_t npix, uint32_t bg, uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride)
        out[i] = hqs_average(acc[i], bg);
}

// compute_loop_nodes_hqs/resolve.cs:92-158 (hqs_window_pixel), in 16x16
// tiles as k_resolve_index, so the neighbourhoods a wave reads overlap.
__global__ __launch_bounds__(256) void k_resolve_hqs_window(const uint32_t* __restrict__ depth,
                                                            const u32x4* __restrict__ acc, int w, int h, int window,
                                                            uint32_t* __restrict__ out) {
    const uint32_t tiles_x = ((uint32_t)w + 15u) >> 4, tiles = tiles_x * (((uint32_t)h + 15u) >> 4);
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t x, y;
        tile_pixel(tile, tiles_x, x, y);
        if (x >= (uint32_t)w || y >= (uint32_t)h) continue;
        out[(uint64_t)y * (uint32_t)w + x] = hqs_window_pixel(depth, acc, w, h, (int)x, (int)y, window);
    }
}

int launch_resolve(hipStream_t s, const pcr_cloud* c, int mode, uint32_t bg, int window, float edl, pcr_fb* fb,
                   FrameState* st) {
    const uint64_t npix = (uint64_t)fb->width * fb->height;
    const int g = grid_for(npix, 256);
    if (mode == PCR_MODE_NODES_HQS) {
        k_resolve_hqs_window<<<g, 256, 0, s>>>(fb->d_depth, reinterpret_cast<const u32x4*>(fb->d_acc), fb->width,
                                               fb->height, window, fb->d_rgba);
    } else if (mode == PCR_MODE_HQS || mode == PCR_MODE_HQS16) {
        k_resolve_hqs<<<g, 256, 0, s>>>(reinterpret_cast<const u32x4*>(fb->d_acc), npix, bg, fb->d_rgba);
    } else if (mode == PCR_MODE_PLAIN16) {
        k_resolve_color<<<g, 256, 0, s>>>(fb->d_keys, npix, bg, fb->d_rgba);
    } else {
        k_resolve_index<<<g, 256, 0, s>>>(fb->d_keys, fb->width, fb->height, c->d_rgba, c->num_points, bg, edl,
                                          fb->d_rgba, st ? &st->pixels_visible : nullptr);
    }
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

}  // namespace pcr
