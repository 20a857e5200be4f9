// pcr_multiview.hip — many views of one cloud in one persistent launch.
//
// The reference renders one view per dispatch; its VR methods decode a batch
// once and project it into two views in the same dispatch
// (modules/compute_loop_las_hqs_vr/color.cs:403-510). This generalises that
// to up to 64 views of uneven resolution:
//   k_mv_select  one thread per batch: per-view cull + precision level
//                (levels[v][b], 255 = not drawn); batches kept by any view
//                are appended to a work list with their most precise level.
//   k_mv_raster  persistent workgroups take (batch, chunk) items from an
//                atomic queue; each item's planes are read once, at the most
//                precise level any view needs, and every view decodes the
//                words at its own level, so each framebuffer is bit-identical
//                to a single-view render of that view.
//   k_mv_clear / k_mv_resolve cover all views in one launch each.
//   hqs modes (HQS, NODES_HQS): k_mv_hqs<PASS> (or, with a tile size,
//                k_mv_hqs_tiled in pcr_kernels.hip) runs the depth pass and
//                then the accumulate pass over the same list, reading each
//                batch's planes (and, in pass 2, its colours) once for all
//                views, as the VR method does for its two eyes;
//                k_mv_hqs_clear and k_mv_resolve_hqs cover all views.
// The frame (clear, reset, select, raster, resolve) is captured once in a
// hipGraph; views change through a device buffer copied before each launch.
#include <algorithm>
#include <vector>
#include "pcr_device.h"
#include "pcr_internal.h"

namespace pcr {

constexpr int kMaxViews = 64;

__global__ void k_mv_reset(MVState* st) {
    if (threadIdx.x == 0) {
        st->list_count = 0; st->queue_head = 0; st->queue_head2 = 0;
        st->bytes_batched = 0; st->bytes_per_view = 0; st->points_per_view = 0; st->points_any_view = 0;
    }
}

__device__ __forceinline__ unsigned long long mv_wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off), hi = __shfl_xor((uint32_t)(v >> 32), off);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

__global__ void k_mv_clear(unsigned long long* const* __restrict__ fbs, const pcr_view* __restrict__ views) {
    const uint32_t v = blockIdx.y;
    const uint64_t npix = (uint64_t)views[v].width * views[v].height;
    unsigned long long* fb = fbs[v];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (uint64_t)gridDim.x * blockDim.x)
        fb[i] = PCR_FB_CLEAR;
}

// lod != 0: LOD node clouds; a node is drawn for view v when it and its
// ancestors pass the view's frustum and every ancestor's projected spacing
// exceeds lod_px[v] (or, rule SIZE80, its projected size is >= 80 px).
__global__ __launch_bounds__(256) void k_mv_select(const pcr_batch* __restrict__ b, uint32_t nb,
                                                   const pcr_view* __restrict__ views, int nviews, int culling,
                                                   int lod, int rule, const float* __restrict__ lod_px,
                                                   uint8_t* __restrict__ levels, uint8_t* __restrict__ best,
                                                   uint32_t* __restrict__ list, MVState* __restrict__ st) {
    __shared__ float planes[kMaxViews][6][4];   // every view's frustum planes, once per workgroup
    __shared__ unsigned long long s_bytes[4];   // batched bytes, per-view bytes, per-view points, kept points
    for (int t = (int)threadIdx.x; t < 6 * nviews; t += (int)blockDim.x)
        frustum_plane(views[t / 6].transform_frustum, t % 6, planes[t / 6][t % 6]);
    if (threadIdx.x < 4) s_bytes[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int top = kCulled;
    unsigned long long view_bytes = 0, view_points = 0;
    if (i < nb) {
        const pcr_batch h = b[i];
        const float first = lod ? 80.0f : 100.0f;
        for (int v = 0; v < nviews; v++) {
            const pcr_view& vw = views[v];
            bool vis = !culling || planes_visible(planes[v], h.min, h.max);
            if (lod && vis) {
                if (rule == PCR_LOD_SIZE80) {
                    vis = precision_level(h.min, h.max, vw, 80.0f) < 4;
                } else {
                    int32_t p = h.parent;
                    while (vis && p >= 0) {
                        const pcr_batch a = b[p];
                        vis = (!culling || planes_visible(planes[v], a.min, a.max)) &&
                              projected_px(a.min, a.max, a.spacing, vw.view_world, vw.proj, vw.width, vw.height) >
                                  lod_px[v];
                        p = a.parent;
                    }
                }
            }
            const int level = vis ? precision_level(h.min, h.max, vw, first) : kCulled;
            levels[(uint64_t)v * nb + i] = (uint8_t)level;
            top = level < top ? level : top;
            if (level != kCulled) {
                view_bytes += 64ull + 4ull * planes_for_level(level) * h.num_points;
                view_points += h.num_points;
            }
        }
        best[i] = (uint8_t)top;
    }
    const bool keep = top != kCulled;
    {   // algorithmic bytes of the frame: one atomic per workgroup and counter
        const unsigned long long np = keep ? b[i].num_points : 0ull;
        const unsigned long long bb = mv_wave_sum(keep ? 64ull + 4ull * planes_for_level(top) * np : 0ull);
        const unsigned long long vb = mv_wave_sum(view_bytes), vp = mv_wave_sum(view_points);
        const unsigned long long kp = mv_wave_sum(np);
        if (__lane_id() == 0) {
            if (bb) atomicAdd(&s_bytes[0], bb);
            if (vb) atomicAdd(&s_bytes[1], vb);
            if (vp) atomicAdd(&s_bytes[2], vp);
            if (kp) atomicAdd(&s_bytes[3], kp);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_bytes[0]) atomicAdd(&st->bytes_batched, s_bytes[0]);
            if (s_bytes[1]) atomicAdd(&st->bytes_per_view, s_bytes[1]);
            if (s_bytes[2]) atomicAdd(&st->points_per_view, s_bytes[2]);
            if (s_bytes[3]) atomicAdd(&st->points_any_view, s_bytes[3]);
        }
    }
    const unsigned long long m = __ballot(keep);
    if (m) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&st->list_count, (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader);
        if (keep) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    }
}

template <int LC>
__device__ __forceinline__ void mv_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi, uint32_t local,
                                         uint32_t cnt, uint32_t index, const pcr_view& v,
                                         unsigned long long* __restrict__ fb) {
    if (local >= cnt) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    uint32_t pix, depth;
    if (!project(v.transform, p[0], p[1], p[2], v.width, v.height, pix, depth)) return;
    fb_min(fb, pix, ((unsigned long long)depth << 32) | index);
}

template <int LC>
__device__ __forceinline__ void mv_points4(const Decode& d, const u32x4& a, const u32x4& m, const u32x4& c,
                                           uint32_t base, uint32_t cnt, uint32_t first, const pcr_view& v,
                                           unsigned long long* __restrict__ fb) {
    mv_point<LC>(d, a.x, m.x, c.x, base + 0, cnt, first + base + 0, v, fb);
    mv_point<LC>(d, a.y, m.y, c.y, base + 1, cnt, first + base + 1, v, fb);
    mv_point<LC>(d, a.z, m.z, c.z, base + 2, cnt, first + base + 2, v, fb);
    mv_point<LC>(d, a.w, m.w, c.w, base + 3, cnt, first + base + 3, v, fb);
}

// Persistent raster over (batch, chunk) items; chunk_pts is a multiple of 4.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_mv_raster(const pcr_batch* __restrict__ b, uint32_t nb,
                                                     const uint8_t* __restrict__ levels,
                                                     const uint8_t* __restrict__ best,
                                                     const uint32_t* __restrict__ list, MVState* __restrict__ st,
                                                     const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                     const uint32_t* __restrict__ hi,
                                                     const pcr_view* __restrict__ views, int nviews,
                                                     unsigned long long* const* __restrict__ fbs, uint32_t chunk_pts,
                                                     uint32_t chunks_per_batch) {
    __shared__ uint32_t s_item;
    __shared__ uint8_t s_vid[kMaxViews];
    __shared__ uint8_t s_lv[kMaxViews];
    __shared__ uint32_t s_nv;
    const uint32_t total = st->list_count * chunks_per_batch;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&st->queue_head, 1u);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= total) break;
        // spread order, as in the single-view tiled raster: consecutive items
        // come from distant parts of the list (2654435761 is a prime above
        // any item count, so the map is a permutation)
        const uint32_t sp = (uint32_t)(((uint64_t)item * 2654435761ull) % total);
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[sp / chunks_per_batch]);
        const uint32_t chunk = sp % chunks_per_batch;
        // compact the views that draw this batch (wave 0, one lane per view)
        if (threadIdx.x < 64) {
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
        const pcr_batch h = b[bi];
        const uint32_t start = chunk * chunk_pts;
        const uint32_t nv = __builtin_amdgcn_readfirstlane(s_nv);
        if (start < h.num_points) {
            const uint32_t cnt = min(chunk_pts, h.num_points - start);
            const uint32_t first = h.first_point + start;
            const int top = best[bi];
            // unaligned batch starts as in the single-view tiled raster
            const uint32_t lead = first & 3u;
            const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
            const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
            const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
            const uint32_t nq = (cnt + lead + 3u) >> 2;
            for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
                const u32x4 a = __builtin_nontemporal_load(lo4 + q);
                const u32x4 m = top <= 1 ? __builtin_nontemporal_load(med4 + q) : u32x4{0, 0, 0, 0};
                const u32x4 c = top == 0 ? __builtin_nontemporal_load(hi4 + q) : u32x4{0, 0, 0, 0};
                for (uint32_t k = 0; k < nv; k++) {
                    const int vi = __builtin_amdgcn_readfirstlane(s_vid[k]);
                    const int lv = __builtin_amdgcn_readfirstlane(s_lv[k]);
                    const pcr_view& v = views[vi];
                    const Decode d = make_decode(lv, h.min, h.max);
                    unsigned long long* fb = fbs[vi];
                    if (lv == 0) mv_points4<0>(d, a, m, c, q * 4u - lead, cnt, first, v, fb);
                    else if (lv == 1) mv_points4<1>(d, a, m, c, q * 4u - lead, cnt, first, v, fb);
                    else mv_points4<2>(d, a, m, c, q * 4u - lead, cnt, first, v, fb);
                }
            }
        }
        __syncthreads();   // s_item / s_vid are rewritten by the next item
    }
}

// One view per blockIdx.y, its pixels in 16x16 tiles as k_resolve_index.
__global__ __launch_bounds__(256) void k_mv_resolve(const unsigned long long* const* __restrict__ fbs,
                                                    uint32_t* const* __restrict__ out,
                                                    const pcr_view* __restrict__ views,
                                                    const uint32_t* __restrict__ rgba, uint64_t npoints, uint32_t bg,
                                                    float edl) {
    const uint32_t v = blockIdx.y;
    const int w = views[v].width, h = views[v].height;
    const unsigned long long* fb = fbs[v];
    uint32_t* o = out[v];
    const uint32_t tiles_x = ((uint32_t)w + 15u) >> 4, tiles = tiles_x * (((uint32_t)h + 15u) >> 4);
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t x, y;
        tile_pixel(tile, tiles_x, x, y);
        if (x >= (uint32_t)w || y >= (uint32_t)h) continue;
        const uint64_t i = (uint64_t)y * (uint32_t)w + x;
        if (edl > 0.0f) {
            o[i] = resolve_index_pixel(fb, w, h, (int)x, (int)y, rgba, npoints, bg, edl);
        } else {
            const unsigned long long k = fb[i];
            const uint32_t id = (uint32_t)k;
            o[i] = (k == PCR_FB_CLEAR || id >= npoints) ? bg : rgba[id];
        }
    }
}

// ------------------------------------------------------------ hqs views
__global__ void k_mv_hqs_clear(uint32_t* const* __restrict__ depths, uint32_t* const* __restrict__ accs,
                               const pcr_view* __restrict__ views) {
    const uint32_t v = blockIdx.y;
    const uint64_t npix = (uint64_t)views[v].width * views[v].height;
    uint32_t* dp = depths[v];
    u32x4* ac = reinterpret_cast<u32x4*>(accs[v]);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (uint64_t)gridDim.x * blockDim.x) {
        dp[i] = PCR_DEPTH_CLEAR;
        ac[i] = u32x4{0, 0, 0, 0};
    }
}

template <int PASS, int LC>
__device__ __forceinline__ void mv_hqs_point(const Decode& d, uint32_t lo, uint32_t med, uint32_t hi, uint32_t color,
                                             uint32_t local, uint32_t cnt, const pcr_view& v,
                                             uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    if (local >= cnt) return;
    float p[3];
    decode<LC>(d, lo, med, hi, p);
    uint32_t pix, dep;
    if (!project(v.transform, p[0], p[1], p[2], v.width, v.height, pix, dep)) return;
    if (PASS == 0) hqs_depth(depth, pix, dep);
    else hqs_accum(depth, acc, pix, dep, color);
}

template <int PASS, int LC>
__device__ __forceinline__ void mv_hqs_points4(const Decode& d, const u32x4& a, const u32x4& m, const u32x4& c,
                                               const u32x4& col, uint32_t base, uint32_t cnt, const pcr_view& v,
                                               uint32_t* __restrict__ depth, uint32_t* __restrict__ acc) {
    mv_hqs_point<PASS, LC>(d, a.x, m.x, c.x, col.x, base + 0, cnt, v, depth, acc);
    mv_hqs_point<PASS, LC>(d, a.y, m.y, c.y, col.y, base + 1, cnt, v, depth, acc);
    mv_hqs_point<PASS, LC>(d, a.z, m.z, c.z, col.z, base + 2, cnt, v, depth, acc);
    mv_hqs_point<PASS, LC>(d, a.w, m.w, c.w, col.w, base + 3, cnt, v, depth, acc);
}

// One hqs pass over every view: PASS 0 takes items from queue_head and
// min-reduces depth, PASS 1 (launched after PASS 0 has finished, so every
// view's depth is final) takes items from queue_head2 and accumulates the
// points within 1.01x of that depth. Same (batch, chunk) items, spread order
// and per-view levels as k_mv_raster.
template <int BLOCK, int PASS>
__global__ __launch_bounds__(BLOCK) void k_mv_hqs(const pcr_batch* __restrict__ b, uint32_t nb,
                                                  const uint8_t* __restrict__ levels, const uint8_t* __restrict__ best,
                                                  const uint32_t* __restrict__ list, MVState* __restrict__ st,
                                                  const uint32_t* __restrict__ lo, const uint32_t* __restrict__ med,
                                                  const uint32_t* __restrict__ hi, const uint32_t* __restrict__ rgba,
                                                  const pcr_view* __restrict__ views, int nviews,
                                                  uint32_t* const* __restrict__ depths,
                                                  uint32_t* const* __restrict__ accs, uint32_t chunk_pts,
                                                  uint32_t chunks_per_batch) {
    __shared__ uint32_t s_item;
    __shared__ uint8_t s_vid[kMaxViews];
    __shared__ uint8_t s_lv[kMaxViews];
    __shared__ uint32_t s_nv;
    uint32_t* head = PASS == 0 ? &st->queue_head : &st->queue_head2;
    const uint32_t total = st->list_count * chunks_per_batch;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= total) break;
        const uint32_t sp = (uint32_t)(((uint64_t)item * 2654435761ull) % total);
        const uint32_t bi = __builtin_amdgcn_readfirstlane(list[sp / chunks_per_batch]);
        const uint32_t chunk = sp % chunks_per_batch;
        if (threadIdx.x < 64) {
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
        const pcr_batch h = b[bi];
        const uint32_t start = chunk * chunk_pts;
        const uint32_t nv = __builtin_amdgcn_readfirstlane(s_nv);
        if (start < h.num_points) {
            const uint32_t cnt = min(chunk_pts, h.num_points - start);
            const uint32_t first = h.first_point + start;
            const int top = best[bi];
            const uint32_t lead = first & 3u;
            const u32x4* lo4 = reinterpret_cast<const u32x4*>(lo + (first - lead));
            const u32x4* med4 = reinterpret_cast<const u32x4*>(med + (first - lead));
            const u32x4* hi4 = reinterpret_cast<const u32x4*>(hi + (first - lead));
            const u32x4* c4 = reinterpret_cast<const u32x4*>(rgba + (first - lead));
            const uint32_t nq = (cnt + lead + 3u) >> 2;
            for (uint32_t q = threadIdx.x; q < nq; q += BLOCK) {
                const u32x4 a = __builtin_nontemporal_load(lo4 + q);
                const u32x4 m = top <= 1 ? __builtin_nontemporal_load(med4 + q) : u32x4{0, 0, 0, 0};
                const u32x4 c = top == 0 ? __builtin_nontemporal_load(hi4 + q) : u32x4{0, 0, 0, 0};
                const u32x4 col = PASS == 1 ? __builtin_nontemporal_load(c4 + q) : u32x4{0, 0, 0, 0};
                for (uint32_t k = 0; k < nv; k++) {
                    const int vi = __builtin_amdgcn_readfirstlane(s_vid[k]);
                    const int lv = __builtin_amdgcn_readfirstlane(s_lv[k]);
                    const pcr_view& v = views[vi];
                    const Decode d = make_decode(lv, h.min, h.max);
                    uint32_t* dp = depths[vi];
                    uint32_t* ac = accs[vi];
                    if (lv == 0) mv_hqs_points4<PASS, 0>(d, a, m, c, col, q * 4u - lead, cnt, v, dp, ac);
                    else if (lv == 1) mv_hqs_points4<PASS, 1>(d, a, m, c, col, q * 4u - lead, cnt, v, dp, ac);
                    else mv_hqs_points4<PASS, 2>(d, a, m, c, col, q * 4u - lead, cnt, v, dp, ac);
                }
            }
        }
        __syncthreads();   // s_item / s_vid are rewritten by the next item
    }
}

// window < 0: integer averages with the background (HQS); otherwise the
// weighted neighbourhood average of NODES_HQS.
__global__ __launch_bounds__(256) void k_mv_resolve_hqs(uint32_t* const* __restrict__ depths,
                                                        uint32_t* const* __restrict__ accs,
                                                        uint32_t* const* __restrict__ out,
                                                        const pcr_view* __restrict__ views, int window, uint32_t bg) {
    const uint32_t v = blockIdx.y;
    const int w = views[v].width, h = views[v].height;
    const uint32_t* dp = depths[v];
    const u32x4* ac = reinterpret_cast<const u32x4*>(accs[v]);
    uint32_t* o = out[v];
    if (window < 0) {
        const uint64_t npix = (uint64_t)w * h;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix;
             i += (uint64_t)gridDim.x * blockDim.x)
            o[i] = hqs_average(ac[i], bg);
        return;
    }
    const uint32_t tiles_x = ((uint32_t)w + 15u) >> 4, tiles = tiles_x * (((uint32_t)h + 15u) >> 4);
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {   // as k_resolve_hqs_window
        uint32_t x, y;
        tile_pixel(tile, tiles_x, x, y);
        if (x >= (uint32_t)w || y >= (uint32_t)h) continue;
        o[(uint64_t)y * (uint32_t)w + x] = hqs_window_pixel(dp, ac, w, h, (int)x, (int)y, window);
    }
}

int mv_alloc(MVBuffers& B, int nviews, uint64_t nb) {
    hipError_t e = hipMalloc(&B.d_views, sizeof(pcr_view) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_fbs, sizeof(void*) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_depths, sizeof(void*) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_accs, sizeof(void*) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_out, sizeof(void*) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_lod, sizeof(float) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_levels, (nb ? nb : 1) * nviews);
    if (e == hipSuccess) e = hipMalloc(&B.d_best, nb ? nb : 1);
    if (e == hipSuccess) e = hipMalloc(&B.d_list, (nb ? nb : 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&B.d_state, sizeof(MVState));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        mv_release(B);
        set_error("multiview: device allocation for %d views x %llu batches failed", nviews, (unsigned long long)nb);
        return PCR_E_OOM;
    }
    B.cap_views = nviews;
    B.cap_batches = nb;
    return PCR_OK;
}

void mv_release(MVBuffers& B) {
    (void)hipFree(B.d_views); (void)hipFree(B.d_fbs); (void)hipFree(B.d_depths); (void)hipFree(B.d_accs);
    (void)hipFree(B.d_out); (void)hipFree(B.d_lod);
    (void)hipFree(B.d_levels); (void)hipFree(B.d_best); (void)hipFree(B.d_list); (void)hipFree(B.d_state);
    B = MVBuffers{};
}

// Queues one multi-view frame: clear, reset, select, raster, resolve.
int mv_enqueue_raw(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, uint64_t max_pix,
                   const pcr_render_params& prm, bool lod, const pcr_tune& t, int cu_count) {
    const int groups = t.persistent_groups > 0 ? t.persistent_groups : cu_count * (2048 / t.block_size);
    uint32_t maxnp = 4;
    for (const pcr_batch& h : c->h_batches) maxnp = maxnp > h.num_points ? maxnp : h.num_points;
    const uint32_t chunk_pts = (uint32_t)(t.block_size * t.points_per_thread);
    const uint32_t cpb = (maxnp + chunk_pts - 1) / chunk_pts;
    const uint32_t pix_blocks = (uint32_t)std::min<uint64_t>((max_pix + 255) / 256, 4096);
    dim3 g2(pix_blocks, (uint32_t)nviews);
    const bool hqs = prm.mode == PCR_MODE_HQS || prm.mode == PCR_MODE_NODES_HQS;
    if (hqs) k_mv_hqs_clear<<<g2, 256, 0, s>>>(B.d_depths, B.d_accs, B.d_views);
    else k_mv_clear<<<g2, 256, 0, s>>>(B.d_fbs, B.d_views);
    k_mv_reset<<<1, 64, 0, s>>>(B.d_state);
    k_mv_select<<<(c->num_batches + 255) / 256, 256, 0, s>>>(c->d_batches, c->num_batches, B.d_views, nviews,
                                                             prm.frustum_culling, lod ? 1 : 0, prm.lod_rule, B.d_lod,
                                                             B.d_levels, B.d_best, B.d_list, B.d_state);
    if (hqs) {
#define MV_HQS(BS, PASS)                                                                                        \
    k_mv_hqs<BS, PASS><<<groups, BS, 0, s>>>(c->d_batches, c->num_batches, B.d_levels, B.d_best, B.d_list,       \
                                             B.d_state, c->d_lo, c->d_med, c->d_hi, c->d_rgba, B.d_views, nviews, \
                                             B.d_depths, B.d_accs, chunk_pts, cpb)
        if (t.tile_px) {
            int rc = launch_mv_hqs_tiled(s, c, B, nviews, t.tile_px, t.block_size, cu_count, prm.occlusion_skip == 0);
            if (rc) return rc;
        } else {
            for (int pass = 0; pass < 2; pass++) {
                if (t.block_size == 64) { if (pass) MV_HQS(64, 1); else MV_HQS(64, 0); }
                else if (t.block_size == 128) { if (pass) MV_HQS(128, 1); else MV_HQS(128, 0); }
                else if (t.block_size == 512) { if (pass) MV_HQS(512, 1); else MV_HQS(512, 0); }
                else { if (pass) MV_HQS(256, 1); else MV_HQS(256, 0); }
            }
        }
#undef MV_HQS
        const int window = prm.mode == PCR_MODE_NODES_HQS ? prm.resolve_window : -1;
        k_mv_resolve_hqs<<<g2, 256, 0, s>>>(B.d_depths, B.d_accs, B.d_out, B.d_views, window, prm.background);
        PCR_HIP(hipGetLastError());
        return PCR_OK;
    }
#define MV_RASTER(BS)                                                                                           \
    k_mv_raster<BS><<<groups, BS, 0, s>>>(c->d_batches, c->num_batches, B.d_levels, B.d_best, B.d_list,          \
                                          B.d_state, c->d_lo, c->d_med, c->d_hi, B.d_views, nviews, B.d_fbs,    \
                                          chunk_pts, cpb)
    if (t.tile_px) {
        int rc = launch_mv_raster_tiled(s, c, B, nviews, t.tile_px, t.block_size, cu_count, prm.occlusion_skip == 0);
        if (rc) return rc;
    } else if (t.block_size == 64) MV_RASTER(64);
    else if (t.block_size == 128) MV_RASTER(128);
    else if (t.block_size == 512) MV_RASTER(512);
    else MV_RASTER(256);
#undef MV_RASTER
    k_mv_resolve<<<g2, 256, 0, s>>>(B.d_fbs, B.d_out, B.d_views, c->d_rgba, c->num_points, prm.background,
                                    prm.edl_strength);
    PCR_HIP(hipGetLastError());
    return PCR_OK;
}

}  // namespace pcr

using namespace pcr;

struct pcr_multiview {
    pcr_ctx* ctx = nullptr;
    pcr_cloud* cloud = nullptr;
    int nviews = 0;
    pcr_render_params prm{};
    pcr_tune tune{};
    bool lod = false;
    std::vector<pcr_view> h_views;
    MVBuffers B;
    uint64_t max_pix = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace pcr {

static int mv_enqueue(pcr_multiview* mv, hipStream_t s) {
    return mv_enqueue_raw(s, mv->cloud, mv->B, mv->nviews, mv->max_pix, mv->prm, mv->lod, mv->tune,
                          mv->ctx->cu_count);
}

static void mv_free(pcr_multiview* mv) {
    if (mv->exec) (void)hipGraphExecDestroy(mv->exec);
    if (mv->graph) (void)hipGraphDestroy(mv->graph);
    mv_release(mv->B);
}

}  // namespace pcr

extern "C" {

int pcr_multiview_create(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* views, int nviews, pcr_fb* const* fbs,
                         const pcr_render_params* params, const pcr_tune* tune, int use_graph, pcr_multiview** out) {
    PCR_CHECK(ctx && cloud && views && fbs && params && out, "multiview_create: null argument");
    PCR_CHECK(nviews >= 1 && nviews <= kMaxViews, "multiview_create: %d views (1..%d)", nviews, kMaxViews);
    PCR_CHECK(cloud->d_batches, "multiview_create: cloud has no batched layout");
    const int mode = params->mode;
    const bool hqs = mode == PCR_MODE_HQS || mode == PCR_MODE_NODES_HQS;
    PCR_CHECK(mode == PCR_MODE_LAS || mode == PCR_MODE_LAS2 || mode == PCR_MODE_NODES || hqs,
              "multiview_create: mode %d is not a batched mode (LAS, LAS2, NODES, HQS, NODES_HQS)", mode);
    PCR_CHECK(params->batch_begin == 0 && params->batch_end == 0,
              "multiview_create: a batch range is not supported (every view draws the whole cloud)");
    PCR_CHECK(params->resolve_window >= 0 && params->resolve_window <= 4,
              "multiview_create: resolve_window %d not in 0..4", params->resolve_window);
    PCR_CHECK(params->edl_strength == 0.0f ||
                  (params->edl_strength > 0.0f && params->edl_strength < 1.0e30f && !hqs),
              "multiview_create: edl_strength %g needs a finite positive value and an index mode",
              (double)params->edl_strength);
    PCR_CHECK(mode != PCR_MODE_NODES && mode != PCR_MODE_NODES_HQS ||
                  params->lod_rule == PCR_LOD_SPACING || params->lod_rule == PCR_LOD_SIZE80,
              "multiview_create: bad lod_rule %d", params->lod_rule);
    for (int v = 0; v < nviews; v++) {
        PCR_CHECK(fbs[v] && fbs[v]->kind == (hqs ? PCR_FB_HQS : PCR_FB_KEY64),
                  "multiview_create: view %d needs a %s framebuffer for mode %d", v, hqs ? "HQS" : "KEY64", mode);
        PCR_CHECK(views[v].width == fbs[v]->width && views[v].height == fbs[v]->height,
                  "multiview_create: view %d is %dx%d but its framebuffer is %dx%d", v, views[v].width,
                  views[v].height, fbs[v]->width, fbs[v]->height);
    }
    pcr_tune t = tune ? *tune : pcr_tune{128, 40, 0, 2048};
    if (t.block_size == 0) t.block_size = 256;
    if (t.points_per_thread == 0) t.points_per_thread = 40;
    PCR_CHECK(t.block_size == 64 || t.block_size == 128 || t.block_size == 256 || t.block_size == 512,
              "multiview_create: block_size %d not in {64,128,256,512}", t.block_size);
    PCR_CHECK(t.points_per_thread > 0 && t.points_per_thread % 4 == 0 && t.points_per_thread <= 1024,
              "multiview_create: points_per_thread %d must be a positive multiple of 4", t.points_per_thread);
    PCR_CHECK(t.tile_px == 0 || ((t.tile_px == 1024 || t.tile_px == 2048 || t.tile_px == 4096) &&
                                 (t.block_size == 128 || t.block_size == 256)),
              "multiview_create: tile_px %d needs 1024/2048/4096 with block 128 or 256", t.tile_px);
    pcr_multiview* mv = new pcr_multiview();
    mv->ctx = ctx; mv->cloud = cloud; mv->nviews = nviews; mv->prm = *params; mv->tune = t;
    mv->lod = mode == PCR_MODE_NODES || mode == PCR_MODE_NODES_HQS;
    mv->h_views.assign(views, views + nviews);
    std::vector<unsigned long long*> hf(nviews);
    std::vector<uint32_t*> ho(nviews), hd(nviews), ha(nviews);
    std::vector<float> hl(nviews, params->lod_px);
    for (int v = 0; v < nviews; v++) {
        hf[v] = fbs[v]->d_keys;
        hd[v] = fbs[v]->d_depth;
        ha[v] = fbs[v]->d_acc;
        ho[v] = fbs[v]->d_rgba;
        mv->max_pix = std::max<uint64_t>(mv->max_pix, (uint64_t)views[v].width * views[v].height);
    }
    int rc0 = mv_alloc(mv->B, nviews, cloud->num_batches);
    if (rc0) { mv_release(mv->B); delete mv; return rc0; }
    hipStream_t s = ctx->stream;
    const auto drop = [mv, s] { (void)hipStreamSynchronize(s); mv_free(mv); delete mv; };
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_views, views, sizeof(pcr_view) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_fbs, hf.data(), sizeof(void*) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_depths, hd.data(), sizeof(void*) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_accs, ha.data(), sizeof(void*) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_out, ho.data(), sizeof(void*) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipMemcpyAsync(mv->B.d_lod, hl.data(), sizeof(float) * nviews, hipMemcpyHostToDevice, s), drop());
    PCR_HIP_OR(hipStreamSynchronize(s), drop());
    if (use_graph) {
        PCR_HIP_OR(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), drop());
        int rc = mv_enqueue(mv, s);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(s, &g);
        if (rc != PCR_OK || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            mv_free(mv);
            delete mv;
            if (rc == PCR_OK) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
            return rc;
        }
        mv->graph = g;
        PCR_HIP_OR(hipGraphInstantiate(&mv->exec, g, nullptr, nullptr, 0), (mv->exec = nullptr, drop()));
    }
    *out = mv;
    return PCR_OK;
}

int pcr_multiview_set_views(pcr_multiview* mv, const pcr_view* views, int nviews) {
    PCR_CHECK(mv && views, "multiview_set_views: null argument");
    PCR_CHECK(nviews == mv->nviews, "multiview_set_views: %d views, plan has %d", nviews, mv->nviews);
    for (int v = 0; v < nviews; v++)
        PCR_CHECK(views[v].width == mv->h_views[v].width && views[v].height == mv->h_views[v].height,
                  "multiview_set_views: view %d changed size", v);
    mv->h_views.assign(views, views + nviews);
    PCR_HIP(hipMemcpyAsync(mv->B.d_views, mv->h_views.data(), sizeof(pcr_view) * nviews, hipMemcpyHostToDevice,
                           mv->ctx->stream));
    return PCR_OK;
}

int pcr_multiview_launch(pcr_multiview* mv) {
    PCR_CHECK(mv, "multiview_launch: null plan");
    if (mv->exec) {
        PCR_HIP(hipGraphLaunch(mv->exec, mv->ctx->stream));
        return PCR_OK;
    }
    return mv_enqueue(mv, mv->ctx->stream);
}

int pcr_multiview_stats_get(pcr_multiview* mv, pcr_multiview_stats* out) {
    PCR_CHECK(mv && out, "multiview_stats: null argument");
    PCR_HIP(hipStreamSynchronize(mv->ctx->stream));
    MVState st;
    PCR_HIP(hipMemcpy(&st, mv->B.d_state, sizeof(st), hipMemcpyDeviceToHost));
    out->batches_any_view = st.list_count;
    out->bytes_batched = st.bytes_batched;
    out->bytes_per_view_sum = st.bytes_per_view;
    out->points_per_view_sum = st.points_per_view;
    out->points_any_view = st.points_any_view;
    return PCR_OK;
}

int pcr_multiview_destroy(pcr_multiview* mv) {
    if (!mv) return PCR_OK;
    (void)hipStreamSynchronize(mv->ctx->stream);
    mv_free(mv);
    delete mv;
    return PCR_OK;
}

// Sweeps the untiled kernel over workgroup size {64,128,256,512} x points per
// thread {40,80,160} (hqs modes: one untiled candidate) and the LDS-tiled
// kernel over {128,256} x tile {1024,2048,4096} px for the given views; keeps
// the fastest configuration (median of `iters` graph launches, HIP events).
int pcr_autotune(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* views, int nviews, pcr_fb* const* fbs,
                 const pcr_render_params* params, int iters, pcr_tune* best, float* best_ms) {
    PCR_CHECK(ctx && cloud && views && fbs && params && best && best_ms, "autotune: null argument");
    PCR_CHECK(iters >= 1 && iters <= 1000, "autotune: iters %d out of range", iters);
    std::vector<pcr_tune> space;
    const bool hqs = params->mode == PCR_MODE_HQS || params->mode == PCR_MODE_NODES_HQS;
    if (hqs) {
        // untiled hqs adds every visible point's colour with global atomics and
        // runs about ten times slower than the tiled kernel: one candidate
        space.push_back(pcr_tune{256, 40, 0, 0});
    } else {
        for (int bsz : {64, 128, 256, 512})
            for (int ppt : {40, 80, 160}) space.push_back(pcr_tune{bsz, ppt, 0, 0});
    }
    for (int bsz : {128, 256})
        for (int tile : {1024, 2048, 4096}) space.push_back(pcr_tune{bsz, 40, 0, tile});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const auto drop_events = [&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    PCR_HIP_OR(hipEventCreate(&e0), drop_events());
    PCR_HIP_OR(hipEventCreate(&e1), drop_events());
    float best_t = 3.0e38f;
    pcr_tune bt = space[0];
    for (const pcr_tune& cand : space) {
        pcr_tune t = cand;
        pcr_multiview* mv = nullptr;
        int rc = pcr_multiview_create(ctx, cloud, views, nviews, fbs, params, &t, 1, &mv);
        if (rc) { drop_events(); return rc; }
        const auto drop = [&] { pcr_multiview_destroy(mv); drop_events(); };
        std::vector<float> ts;
        for (int it = 0; it < iters + 1; it++) {
            PCR_HIP_OR(hipEventRecord(e0, ctx->stream), drop());
            rc = pcr_multiview_launch(mv);
            if (rc) { drop(); return rc; }
            PCR_HIP_OR(hipEventRecord(e1, ctx->stream), drop());
            PCR_HIP_OR(hipEventSynchronize(e1), drop());
            float ms = 0;
            PCR_HIP_OR(hipEventElapsedTime(&ms, e0, e1), drop());
            if (it > 0) ts.push_back(ms);   // the first launch warms up
        }
        pcr_multiview_destroy(mv);
        std::sort(ts.begin(), ts.end());
        const float med = ts[ts.size() / 2];
        if (med < best_t) { best_t = med; bt = t; }
    }
    drop_events();
    *best = bt;
    *best_ms = best_t;
    return PCR_OK;
}

}  // extern "C"
