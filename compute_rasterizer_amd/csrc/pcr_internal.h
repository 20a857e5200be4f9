// pcr_internal.h — library-internal host structures and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/pcr.h"

namespace pcr {

// Per-frame counters in device memory, zeroed by a kernel at frame start
// (a kernel rather than a memset so that the frame can be graph-captured).
struct FrameState {
    uint32_t list_count;        // batches appended by the select stage
    uint32_t queue_head;        // work items taken by the raster stage (hqs: pass 1)
    uint32_t queue_head2;       // hqs: work items taken by pass 2
    uint32_t level_count[5];
    uint32_t culled;
    unsigned long long points_drawn;
    unsigned long long plane_bytes;
    uint32_t occluded_batches;          // batch draws skipped by the occlusion test (rect_open)
    unsigned long long occluded_points; // their points
    unsigned long long occluded_bytes;  // the plane (hqs pass 2: and colour) bytes they did not read
    unsigned long long atomics;         // 64-bit framebuffer atomics issued (LAS, LAS2, NODES)
    unsigned long long pixels_visible;  // pixels the index resolve found covered
};

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define PCR_HIP(expr)                                                         \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return ::pcr::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)
// PCR_HIP that runs `cleanup` (statements freeing what the caller built so
// far) before it returns the error.
#define PCR_HIP_OR(expr, cleanup)                                             \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) {                                               \
            cleanup;                                                          \
            return ::pcr::hip_fail(_e, #expr, __FILE__, __LINE__);            \
        }                                                                     \
    } while (0)
#define PCR_CHECK(cond, ...)                                                  \
    do {                                                                      \
        if (!(cond)) { ::pcr::set_error(__VA_ARGS__); return PCR_E_INVALID; } \
    } while (0)

}  // namespace pcr

struct pcr_ctx {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    // Two frame states: a frame uses d_state[cur], and its select launch
    // resets the other one for the next frame, so no launch of its own has
    // to reset the state before the select's atomics
    pcr::FrameState* d_state = nullptr;
    int cur = 0;                          // the state the next frame uses
    int last_state = 0;                   // the state the last frame used (pcr_stats_get)
    bool state_clean[2] = {true, true};   // the state is known to be reset
    hipEvent_t (*ring)[6] = nullptr;      // stage timers, PCR_TIMING_RING frames x 6 events
    uint64_t frames_timed = 0;            // frames that recorded events (ring head)
    int timing = 0;                       // the current frame's time_stages (0, 1, 2)
    uint8_t ring_level[PCR_TIMING_RING] = {};   // time_stages each ring slot was recorded with
    pcr_stats last{};                     // host copy of the last frame's stats
    uint32_t last_nb = 0;
};

struct pcr_cloud {
    pcr_ctx* ctx = nullptr;
    uint64_t num_points = 0;
    uint32_t num_batches = 0;
    uint32_t batch_size = 0;
    uint64_t plane_len = 0;               // words per plane (padded to a multiple of 4)
    pcr_batch* d_batches = nullptr;
    uint32_t* d_lo = nullptr;
    uint32_t* d_med = nullptr;
    uint32_t* d_hi = nullptr;
    uint32_t* d_rgba = nullptr;           // colour plane, same index as the position planes
    uint4* d_points16 = nullptr;          // x, y, z (float bits), rgba
    uint64_t points16_len = 0;
    uint8_t* d_levels = nullptr;          // per-batch level of the last frame (255 = not drawn)
    uint32_t* d_list = nullptr;           // compacted list of drawn batches
    int num_levels = 0;                   // octree depth + 1 (LOD clouds)
    float aabb_min[3] = {0, 0, 0};
    float aabb_max[3] = {0, 0, 0};
    uint64_t device_bytes = 0;
    std::vector<pcr_batch> h_batches;     // host copy of the headers
};

struct pcr_fb {
    pcr_ctx* ctx = nullptr;
    int width = 0, height = 0, kind = 0;
    unsigned long long* d_keys = nullptr; // KEY64
    uint32_t* d_depth = nullptr;          // HQS
    uint32_t* d_acc = nullptr;            // HQS: R,G,B,count per pixel
    uint32_t* d_rgba = nullptr;           // resolved image
    uint64_t bytes = 0;
};

namespace pcr {

// kernels (pcr_kernels.hip)
int launch_fill64(hipStream_t s, unsigned long long* p, uint64_t n, unsigned long long v);
int launch_fill32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v);
int launch_frame_reset(hipStream_t s, FrameState* st);
// A framebuffer region to fill with one 32-bit word value (n words).
struct FillRegion {
    uint32_t* p = nullptr;
    uint64_t n = 0;
    uint32_t v = 0;
};
// st: the frame's state (reset); st_next: reset here for the next frame;
// fill0/fill1: framebuffer regions cleared in the same launch.
int launch_select(hipStream_t s, const pcr_cloud* c, const pcr_view& v, const pcr_render_params& p,
                  FrameState* st, FrameState* st_next, const FillRegion& fill0, const FillRegion& fill1,
                  float first_threshold, bool lod);
int launch_raster_batched(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int vec, int block,
                          FrameState* st, unsigned long long* fb);
// occl: the occlusion skip (rect_open) is on
int launch_raster_tiled(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int tile_px, int block,
                        FrameState* st, unsigned long long* fb, bool occl = true);
int launch_raster_points16(hipStream_t s, const pcr_cloud* c, const pcr_view& v, unsigned long long* fb);
int launch_hqs_batched(hipStream_t s, const pcr_cloud* c, const pcr_view& v, FrameState* st, pcr_fb* fb);
int launch_hqs_tiled(hipStream_t s, const pcr_cloud* c, const pcr_view& v, int tile_px, int block, FrameState* st,
                     pcr_fb* fb, bool occl = true);
int launch_hqs_points16(hipStream_t s, const pcr_cloud* c, const pcr_view& v, pcr_fb* fb);
// st (optional): the frame's state, whose pixels_visible the index resolve counts
int launch_resolve(hipStream_t s, const pcr_cloud* c, int mode, uint32_t bg, int window, float edl, pcr_fb* fb,
                   FrameState* st = nullptr);
int launch_selftest_divide(hipStream_t s, const float* nx, const float* ny, const float* den, uint64_t n,
                           float* out2);

// multi-view frame (pcr_multiview.hip), shared with the server
struct MVState {
    uint32_t list_count;   // batches drawn by at least one view
    uint32_t queue_head;   // work items taken by the persistent raster (hqs: pass 1)
    uint32_t queue_head2;  // hqs: work items taken by pass 2
    uint32_t pad;
    unsigned long long bytes_batched;    // headers + planes at each kept batch's most precise level
    unsigned long long bytes_per_view;   // headers + planes summed over the views that draw each batch
    unsigned long long points_per_view;  // points in drawn batches summed over views
    unsigned long long points_any_view;  // points in batches drawn by at least one view
};
struct MVBuffers {
    pcr_view* d_views = nullptr;
    unsigned long long** d_fbs = nullptr;
    uint32_t** d_depths = nullptr;   // hqs framebuffers: depth and accumulators
    uint32_t** d_accs = nullptr;
    uint32_t** d_out = nullptr;
    float* d_lod = nullptr;
    uint8_t* d_levels = nullptr;
    uint8_t* d_best = nullptr;
    uint32_t* d_list = nullptr;
    MVState* d_state = nullptr;
    int cap_views = 0;
    uint64_t cap_batches = 0;
};
int mv_alloc(MVBuffers& B, int nviews, uint64_t nb);
void mv_release(MVBuffers& B);
int mv_enqueue_raw(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, uint64_t max_pix,
                   const pcr_render_params& prm, bool lod, const pcr_tune& t, int cu_count);
int launch_mv_raster_tiled(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int tile_px, int block,
                           int cu_count, bool occl = true);
int launch_mv_hqs_tiled(hipStream_t s, const pcr_cloud* c, const MVBuffers& B, int nviews, int tile_px, int block,
                        int cu_count, bool occl = true);

// host parsers (pcr_parse.cpp)
int potree_layout(const pcr_potree_info* info, const void* hierarchy, uint64_t hierarchy_size, uint64_t octree_size,
                  std::vector<pcr_batch>& hb, std::vector<double>& boxes, int* max_level);

// layout (pcr_layout.hip)
int generate_points(pcr_ctx* ctx, const pcr_gen_cfg* cfg, float* xyz, uint32_t* rgba);
int build_cloud_device(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n,
                       const pcr_build_opts* opts, pcr_cloud** out);
int cloud_alloc_batched(pcr_ctx* ctx, uint32_t nb, uint64_t npoints, pcr_cloud* c);
int device_aabb(pcr_ctx* ctx, const float* xyz, uint64_t n, float mn[3], float mx[3]);
int build_octree_device(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_octree_opts* o,
                        pcr_cloud** out);
void cloud_free(pcr_cloud* c);
void encode_batches_host(const float* xyz, uint64_t n, uint32_t bs, pcr_batch* hdr,
                         uint32_t* lo, uint32_t* med, uint32_t* hi);

}  // namespace pcr
