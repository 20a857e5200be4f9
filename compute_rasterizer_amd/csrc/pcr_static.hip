// pcr_static.hip — pcr_render_static: the per-frame render of pcr_render for
// the index-payload modes (LAS, LAS2, NODES) without any device allocation.
//
// pcr_render keeps its frame scratch in memory the library allocated earlier:
// the context's two frame states (counters and queue heads) and the cloud's
// drawn-batch list and per-batch levels. Here the same scratch lives in one
// file-scope device array, `scratch`, of a fixed size. Each GPU holds its own
// copy at its own address; the host looks that address up once per device
// with hipGetSymbolAddress and passes it to the kernels, which are the same
// kernels pcr_render launches. The array's name in host code is only a
// host-side stand-in and is never passed to a kernel or a copy.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind;
// the call resets the frame counters with a kernel before the select (a
// kernel rather than a memset, so that the call can be graph-captured), and
// the select rewrites the list and levels of every batch. Calls that overlap
// on one GPU race on the array.
#include <mutex>

#include "pcr_internal.h"

namespace pcr {
namespace {

constexpr uint64_t kStateBytes = 256;                                      // two FrameStates
constexpr uint64_t kListBytes = 4ull * PCR_STATIC_MAX_BATCHES;             // drawn-batch list
constexpr uint64_t kLevelBytes = 1ull * PCR_STATIC_MAX_BATCHES;            // per-batch level
constexpr uint64_t kScratchBytes = kStateBytes + kListBytes + kLevelBytes;
static_assert(2 * sizeof(FrameState) <= kStateBytes, "frame states exceed their slot");
static_assert(kStateBytes % 16 == 0 && kListBytes % 16 == 0, "list and levels keep 16-byte alignment");
constexpr int kMaxDevices = 64;

}  // namespace
}  // namespace pcr

static __device__ __align__(256) unsigned char scratch[pcr::kScratchBytes];

namespace pcr {
namespace {

struct DeviceScratch {
    unsigned char* ptr = nullptr;   // this device's copy of `scratch`
    int cu_count = 0;
};
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The device address of `scratch` on device `dev` (the current device),
// looked up on the first call for that device.
int device_scratch(int dev, DeviceScratch* out) {
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        PCR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        int cus = 0;
        PCR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        PCR_CHECK(p != nullptr && cus > 0, "render_static: no scratch address or CU count on device %d", dev);
        d.cu_count = cus;
        d.ptr = static_cast<unsigned char*>(p);
    }
    *out = d;
    return PCR_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace
}  // namespace pcr

extern "C" int pcr_render_static(const pcr_batch* batches_dev, uint32_t nbatches, const uint32_t* lo_dev,
                                 const uint32_t* med_dev, const uint32_t* hi_dev, uint64_t plane_words,
                                 const pcr_view* view, const pcr_render_params* params, uint64_t* keys_dev,
                                 void* stream) {
    using namespace pcr;
    PCR_CHECK(batches_dev && lo_dev && med_dev && hi_dev && view && params && keys_dev,
              "render_static: null argument");
    PCR_CHECK(nbatches <= PCR_STATIC_MAX_BATCHES,
              "render_static: %u batches exceed the static scratch's limit of %u", nbatches,
              PCR_STATIC_MAX_BATCHES);
    PCR_CHECK(view->width > 0 && view->height > 0 && view->width <= PCR_MAX_DIM && view->height <= PCR_MAX_DIM,
              "render_static: bad view size %dx%d", view->width, view->height);
    const int mode = params->mode;
    PCR_CHECK(mode == PCR_MODE_LAS || mode == PCR_MODE_LAS2 || mode == PCR_MODE_NODES,
              "render_static: mode %d is not LAS, LAS2 or NODES", mode);
    PCR_CHECK(params->time_stages == 0, "render_static: time_stages must be 0 (stage events need a context)");
    PCR_CHECK(mode != PCR_MODE_NODES || params->lod_rule == PCR_LOD_SPACING || params->lod_rule == PCR_LOD_SIZE80,
              "render_static: bad lod_rule %d", params->lod_rule);
    PCR_CHECK(params->tile_px == 0 || params->tile_px == 1024 || params->tile_px == 2048 || params->tile_px == 4096,
              "render_static: tile_px %d not in {0,1024,2048,4096}", params->tile_px);
    PCR_CHECK(params->occlusion_skip == 0 || params->occlusion_skip == -1,
              "render_static: occlusion_skip %d not 0 or -1", params->occlusion_skip);
    const int block = params->block_size;
    PCR_CHECK(block == 0 || block == 64 || block == 128 || block == 256 || block == 512,
              "render_static: block_size %d not in {64,128,256,512}", block);
    PCR_CHECK(params->batch_end == 0 ? params->batch_begin == 0
                                     : params->batch_begin <= params->batch_end && params->batch_end <= nbatches,
              "render_static: batch range [%u, %u) outside the %u batches", params->batch_begin, params->batch_end,
              nbatches);
    PCR_CHECK(plane_words % 4 == 0 && aligned16(lo_dev) && aligned16(med_dev) && aligned16(hi_dev),
              "render_static: planes must be 16-byte aligned and a multiple of 4 words long");
    PCR_CHECK(aligned16(keys_dev), "render_static: keys_dev must be 16-byte aligned");
    int dev = 0;
    PCR_HIP(hipGetDevice(&dev));
    PCR_CHECK(dev >= 0 && dev < kMaxDevices, "render_static: device %d beyond %d", dev, kMaxDevices);
    DeviceScratch ds;
    int rc = device_scratch(dev, &ds);
    if (rc) return rc;
    FrameState* st = reinterpret_cast<FrameState*>(ds.ptr);
    uint32_t* list = reinterpret_cast<uint32_t*>(ds.ptr + kStateBytes);
    uint8_t* levels = ds.ptr + kStateBytes + kListBytes;

    // stand-ins for the context and cloud the launchers read (no allocation:
    // the cloud's host vector stays empty)
    pcr_ctx cx;
    cx.device = dev;
    cx.cu_count = ds.cu_count;
    pcr_cloud c;
    c.ctx = &cx;
    c.num_batches = nbatches;
    c.plane_len = plane_words;
    c.d_batches = const_cast<pcr_batch*>(batches_dev);
    c.d_lo = const_cast<uint32_t*>(lo_dev);
    c.d_med = const_cast<uint32_t*>(med_dev);
    c.d_hi = const_cast<uint32_t*>(hi_dev);
    c.d_levels = levels;
    c.d_list = list;

    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(keys_dev);
    const bool lod = mode == PCR_MODE_NODES;
    if ((rc = launch_frame_reset(s, st))) return rc;
    FillRegion f0, f1;
    if (params->clear_before)
        f0 = FillRegion{reinterpret_cast<uint32_t*>(keys_dev), 2ull * view->width * view->height, 0xFFFFFFFFu};
    if ((rc = launch_select(s, &c, *view, *params, st, st + 1, f0, f1, lod ? 80.0f : 100.0f, lod))) return rc;
    if (mode == PCR_MODE_LAS) return launch_raster_batched(s, &c, *view, 1, block ? block : 128, st, keys);
    return params->tile_px ? launch_raster_tiled(s, &c, *view, params->tile_px, block, st, keys,
                                                 params->occlusion_skip == 0)
                           : launch_raster_batched(s, &c, *view, 4, block ? block : 256, st, keys);
}
