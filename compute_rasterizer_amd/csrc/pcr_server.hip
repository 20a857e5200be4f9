// pcr_server.hip — continuous-batching render server over device-resident clouds.
//
// The reference serves one interactive view per frame (src/main.cpp:342,455).
// Here render requests against several device-resident clouds arrive at any
// time from any host thread:
//   - admission: every request needs a framebuffer (8 B key + 4 B RGBA per
//     pixel) carved from one device arena whose size is the framebuffer
//     budget; a request waits (FIFO) until its framebuffer fits, and is
//     refused only when it alone exceeds the budget;
//   - batching: a worker thread groups admitted requests for the same cloud
//     and mode (up to max_batch) into one multi-view frame (the kernels of
//     pcr_multiview.hip) on one of num_streams HIP streams, so requests that
//     arrive while others run are launched on another stream and overlap;
//   - completion: the resolved image is copied to a pinned host arena on the
//     same stream and an event marks the request done; poll/wait copy it out
//     and release its memory. The pinned arena holds 4 B per pixel and is
//     sized min(budget / 3, kMaxPinnedBytes); a request is admitted when its
//     framebuffer fits the device arena and its image the pinned one.
// Each request's image depends only on its cloud and view (min is order
// independent), so it is the same alone and in any mix.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include "pcr_internal.h"

namespace pcr {

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) in multiples of `gran`, coalescing frees.
struct Arena {
    std::map<uint64_t, uint64_t> free_;   // offset -> size
    uint64_t size = 0, used = 0, peak = 0, gran = 768;
    void init(uint64_t s) { size = s - s % gran; free_.clear(); if (size) free_[0] = size; used = peak = 0; }
    uint64_t round(uint64_t b) const { return (b + gran - 1) / gran * gran; }
    bool alloc(uint64_t bytes, uint64_t& off) {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used += bytes;
            peak = used > peak ? used : peak;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = round(bytes);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

enum ReqState { kPending = 0, kInFlight = 1, kDone = 2, kCancelled = 3, kTaking = 4, kHeld = 5 };

struct Req {
    pcr_request r{};
    int state = kPending;
    uint64_t off = 0, bytes = 0, npix = 0;
    uint64_t hoff = 0, hbytes = 0;           // the image's range of the pinned arena
    double t_submit = 0, t_done = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    MVBuffers B;
    pcr_view* h_views = nullptr;             // pinned staging of the launch's views
    float* h_lod = nullptr;
    unsigned long long** h_fbs = nullptr;
    uint32_t** h_out = nullptr;
    MVState* h_state = nullptr;              // pinned copy of the launch's counters
    std::vector<uint64_t> tickets;
    bool busy = false;
};

static uint64_t fb_bytes(const pcr_view& v) { return 12ull * (uint64_t)v.width * (uint64_t)v.height; }
// Pinned host memory for finished images is capped: page-locked memory is
// taken from the host's RAM for the server's lifetime.
constexpr uint64_t kMaxPinnedBytes = 1ull << 30;

}  // namespace pcr

using namespace pcr;

struct pcr_server {
    pcr_ctx* ctx = nullptr;
    int max_batch = 16;
    unsigned char* d_arena = nullptr;
    unsigned char* h_arena = nullptr;       // pinned result images (harena)
    Arena arena;                            // device framebuffers, 12 B per pixel
    Arena harena;                           // pinned images, 4 B per pixel
    std::vector<pcr_cloud*> clouds;
    uint64_t max_nb = 0;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Req> reqs;
    std::vector<Slot> slots;
    std::thread worker;
    bool stop = false;
    uint64_t next_ticket = 1;
    pcr_server_stats st{};
    int worker_error = 0;
};

namespace pcr {

static int slot_alloc(Slot& s, int max_batch, uint64_t nb) {
    mv_release(s.B);
    int rc = mv_alloc(s.B, max_batch, nb);
    if (rc) return rc;
    if (!s.h_views) {
        PCR_HIP(hipHostMalloc(&s.h_views, sizeof(pcr_view) * max_batch));
        PCR_HIP(hipHostMalloc(&s.h_lod, sizeof(float) * max_batch));
        PCR_HIP(hipHostMalloc(&s.h_fbs, sizeof(void*) * max_batch));
        PCR_HIP(hipHostMalloc(&s.h_out, sizeof(void*) * max_batch));
        PCR_HIP(hipHostMalloc(&s.h_state, sizeof(MVState)));
    }
    return PCR_OK;
}

// Launches a group of requests (same cloud and mode) on a free slot. Caller holds the lock.
static int launch_group(pcr_server* srv, Slot& sl, const std::vector<uint64_t>& group) {
    const Req& head = srv->reqs[group[0]];
    const pcr_cloud* c = srv->clouds[head.r.cloud_id];
    const int n = (int)group.size();
    uint64_t max_pix = 0;
    for (int k = 0; k < n; k++) {
        Req& q = srv->reqs[group[k]];
        sl.h_views[k] = q.r.view;
        sl.h_lod[k] = q.r.lod_px;
        sl.h_fbs[k] = reinterpret_cast<unsigned long long*>(srv->d_arena + q.off);
        sl.h_out[k] = reinterpret_cast<uint32_t*>(srv->d_arena + q.off + 8 * q.npix);
        max_pix = q.npix > max_pix ? q.npix : max_pix;
        q.state = kInFlight;
    }
    hipStream_t s = sl.stream;
    PCR_HIP(hipMemcpyAsync(sl.B.d_views, sl.h_views, sizeof(pcr_view) * n, hipMemcpyHostToDevice, s));
    PCR_HIP(hipMemcpyAsync(sl.B.d_lod, sl.h_lod, sizeof(float) * n, hipMemcpyHostToDevice, s));
    PCR_HIP(hipMemcpyAsync(sl.B.d_fbs, sl.h_fbs, sizeof(void*) * n, hipMemcpyHostToDevice, s));
    PCR_HIP(hipMemcpyAsync(sl.B.d_out, sl.h_out, sizeof(void*) * n, hipMemcpyHostToDevice, s));
    pcr_render_params prm{};
    prm.mode = head.r.mode;
    prm.frustum_culling = 1;
    prm.lod_rule = PCR_LOD_SPACING;
    prm.background = PCR_BG_LAS;
    // measured on the 1e9-point city: LOD node groups run fastest untiled
    // (their nodes are small on screen and spread across views), full-cloud
    // groups with the 1024-px LDS tile
    const pcr_tune t = head.r.mode == PCR_MODE_NODES ? pcr_tune{128, 40, 0, 0} : pcr_tune{128, 40, 0, 1024};
    int rc = mv_enqueue_raw(s, c, sl.B, n, max_pix, prm, head.r.mode == PCR_MODE_NODES, t, srv->ctx->cu_count);
    if (rc) return rc;
    for (int k = 0; k < n; k++) {
        const Req& q = srv->reqs[group[k]];
        PCR_HIP(hipMemcpyAsync(srv->h_arena + q.hoff, sl.h_out[k], 4 * q.npix, hipMemcpyDeviceToHost, s));
    }
    PCR_HIP(hipMemcpyAsync(sl.h_state, sl.B.d_state, sizeof(MVState), hipMemcpyDeviceToHost, s));
    PCR_HIP(hipEventRecord(sl.done, s));
    sl.tickets = group;
    sl.busy = true;
    srv->st.launches++;
    return PCR_OK;
}

static void worker_main(pcr_server* srv) {
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    while (!srv->stop) {
        bool progress = false, busy = false;
        // retire finished launches
        for (Slot& sl : srv->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) { busy = true; continue; }
            const double t = now_ms();
            for (uint64_t tk : sl.tickets) {
                Req& r = srv->reqs[tk];
                r.t_done = t;
                if (r.state == kCancelled) {
                    srv->arena.release(r.off, r.bytes);
                    srv->harena.release(r.hoff, r.hbytes);
                    srv->reqs.erase(tk);
                } else {
                    r.state = kDone;
                    srv->st.completed++;
                    srv->st.points_rendered += srv->clouds[r.r.cloud_id]->num_points;
                }
            }
            if (q != hipSuccess) {
                srv->worker_error = PCR_E_HIP;
            } else {   // the counters' copy precedes the event on the slot's stream
                srv->st.points_drawn += sl.h_state->points_per_view;
                srv->st.bytes_algorithmic += sl.h_state->bytes_batched;
            }
            sl.tickets.clear();
            sl.busy = false;
            progress = true;
        }
        if (progress) srv->done_cv.notify_all();
        // admit and launch: FIFO head decides the cloud and mode of the group
        for (Slot& sl : srv->slots) {
            if (sl.busy || srv->pending.empty()) continue;
            const Req& head = srv->reqs[srv->pending.front()];
            std::vector<uint64_t> group;
            std::vector<size_t> taken;
            for (size_t i = 0; i < srv->pending.size() && (int)group.size() < srv->max_batch; i++) {
                Req& q = srv->reqs[srv->pending[i]];
                if (q.r.cloud_id != head.r.cloud_id || q.r.mode != head.r.mode) continue;
                if (!srv->arena.alloc(q.bytes, q.off)) break;   // wait for memory, keep FIFO order
                if (!srv->harena.alloc(q.hbytes, q.hoff)) {
                    srv->arena.release(q.off, q.bytes);
                    break;
                }
                group.push_back(srv->pending[i]);
                taken.push_back(i);
            }
            if (group.empty()) break;
            for (size_t k = taken.size(); k-- > 0;) srv->pending.erase(srv->pending.begin() + (long)taken[k]);
            if (launch_group(srv, sl, group) != PCR_OK) srv->worker_error = PCR_E_HIP;
            progress = true;
            busy = true;
        }
        if (srv->st.fb_bytes_peak < srv->arena.peak) srv->st.fb_bytes_peak = srv->arena.peak;
        if (!progress) {
            if (busy) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                srv->work_cv.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

}  // namespace pcr

extern "C" {

int pcr_server_create(pcr_ctx* ctx, uint64_t budget, int max_batch, int num_streams, pcr_server** out) {
    PCR_CHECK(ctx && out, "server_create: null argument");
    PCR_CHECK(budget >= (1ull << 20), "server_create: budget %llu below 1 MiB", (unsigned long long)budget);
    PCR_CHECK(max_batch >= 1 && max_batch <= 64, "server_create: max_batch %d (1..64)", max_batch);
    PCR_CHECK(num_streams >= 1 && num_streams <= 8, "server_create: num_streams %d (1..8)", num_streams);
    pcr_server* srv = new pcr_server();
    srv->ctx = ctx;
    srv->max_batch = max_batch;
    srv->arena.init(budget);
    srv->harena.gran = 256;
    srv->harena.init(budget / 3 < kMaxPinnedBytes ? budget / 3 : kMaxPinnedBytes);
    if (hipMalloc(&srv->d_arena, srv->arena.size) != hipSuccess ||
        hipHostMalloc(&srv->h_arena, srv->harena.size) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(srv->d_arena);
        delete srv;
        set_error("server_create: cannot allocate a %llu-byte framebuffer arena", (unsigned long long)budget);
        return PCR_E_OOM;
    }
    srv->slots.resize(num_streams);
    for (Slot& sl : srv->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            pcr_server_destroy(srv);
            set_error("server_create: cannot create the server's streams and events");
            return PCR_E_HIP;
        }
    }
    srv->worker = std::thread(worker_main, srv);
    *out = srv;
    return PCR_OK;
}

int pcr_server_add_cloud(pcr_server* srv, pcr_cloud* cloud, int* cloud_id) {
    PCR_CHECK(srv && cloud && cloud_id, "server_add_cloud: null argument");
    PCR_CHECK(cloud->d_batches, "server_add_cloud: cloud has no batched layout");
    std::unique_lock<std::mutex> lk(srv->mu);
    PCR_CHECK(srv->pending.empty(), "server_add_cloud: add clouds before submitting requests");
    for (Slot& sl : srv->slots) if (sl.busy) PCR_HIP(hipEventSynchronize(sl.done));
    if (cloud->num_batches > srv->max_nb) {
        srv->max_nb = cloud->num_batches;
        for (Slot& sl : srv->slots) {
            int rc = slot_alloc(sl, srv->max_batch, srv->max_nb);
            if (rc) return rc;
        }
    }
    srv->clouds.push_back(cloud);
    *cloud_id = (int)srv->clouds.size() - 1;
    return PCR_OK;
}

int pcr_server_submit(pcr_server* srv, const pcr_request* req, uint64_t* ticket) {
    PCR_CHECK(srv && req && ticket, "server_submit: null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    PCR_CHECK(req->cloud_id >= 0 && req->cloud_id < (int)srv->clouds.size(), "server_submit: bad cloud_id %d",
              req->cloud_id);
    PCR_CHECK(req->mode == PCR_MODE_LAS || req->mode == PCR_MODE_LAS2 || req->mode == PCR_MODE_NODES,
              "server_submit: mode %d not served (LAS, LAS2, NODES)", req->mode);
    PCR_CHECK(req->view.width > 0 && req->view.height > 0 && req->view.width <= 16384 && req->view.height <= 16384,
              "server_submit: bad view size %dx%d", req->view.width, req->view.height);
    const uint64_t bytes = srv->arena.round(fb_bytes(req->view));
    const uint64_t npix = (uint64_t)req->view.width * req->view.height;
    const uint64_t hbytes = srv->harena.round(4 * npix);
    if (bytes > srv->arena.size || hbytes > srv->harena.size) {
        srv->st.refused++;
        set_error("server_submit: a %dx%d framebuffer (%llu B, image %llu B) exceeds the %llu-byte budget "
                  "(pinned images %llu B)", req->view.width, req->view.height, (unsigned long long)bytes,
                  (unsigned long long)hbytes, (unsigned long long)srv->arena.size,
                  (unsigned long long)srv->harena.size);
        return PCR_E_BUDGET;
    }
    if (srv->worker_error) { set_error("server_submit: the server stopped after a HIP error"); return PCR_E_HIP; }
    Req r;
    r.r = *req;
    r.bytes = bytes;
    r.npix = npix;
    r.hbytes = hbytes;
    r.t_submit = now_ms();
    const uint64_t t = srv->next_ticket++;
    srv->reqs.emplace(t, r);
    srv->pending.push_back(t);
    *ticket = t;
    srv->work_cv.notify_one();
    return PCR_OK;
}

// The image is copied out with the server lock released (the worker keeps
// launching meanwhile); kTaking keeps the request and its arena range alive
// until the copy is done.
static int take_locked(pcr_server* srv, std::unique_lock<std::mutex>& lk, uint64_t ticket, uint32_t* rgba,
                       double* latency_ms) {
    auto it = srv->reqs.find(ticket);
    PCR_CHECK(it != srv->reqs.end(), "server_poll: unknown ticket %llu", (unsigned long long)ticket);
    Req& r = it->second;
    if (r.state == kCancelled) return PCR_E_CANCELLED;
    PCR_CHECK(r.state != kTaking && r.state != kHeld, "server_poll: ticket %llu is already taken",
              (unsigned long long)ticket);
    if (r.state != kDone) return PCR_E_NOT_READY;
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    r.state = kTaking;
    const unsigned char* src = srv->h_arena + r.hoff;
    const uint64_t npix = r.npix, off = r.off, bytes = r.bytes, hoff = r.hoff, hbytes = r.hbytes;
    lk.unlock();
    if (rgba) memcpy(rgba, src, 4 * npix);
    lk.lock();
    srv->arena.release(off, bytes);
    srv->harena.release(hoff, hbytes);
    srv->reqs.erase(ticket);
    srv->work_cv.notify_one();
    return PCR_OK;
}

int pcr_server_poll(pcr_server* srv, uint64_t ticket, uint32_t* rgba, double* latency_ms) {
    PCR_CHECK(srv, "server_poll: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    return take_locked(srv, lk, ticket, rgba, latency_ms);
}

int pcr_server_wait(pcr_server* srv, uint64_t ticket, uint32_t* rgba, double* latency_ms) {
    PCR_CHECK(srv, "server_wait: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        const int rc = take_locked(srv, lk, ticket, rgba, latency_ms);
        if (rc != PCR_E_NOT_READY) return rc;
        if (srv->worker_error) { set_error("server_wait: the server stopped after a HIP error"); return PCR_E_HIP; }
        srv->done_cv.wait_for(lk, std::chrono::milliseconds(2));
    }
}

int pcr_server_wait_pinned(pcr_server* srv, uint64_t ticket, const uint32_t** rgba_pinned, double* latency_ms) {
    PCR_CHECK(srv && rgba_pinned, "server_wait_pinned: null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        auto it = srv->reqs.find(ticket);
        PCR_CHECK(it != srv->reqs.end(), "server_wait_pinned: unknown ticket %llu", (unsigned long long)ticket);
        Req& r = it->second;
        if (r.state == kCancelled) return PCR_E_CANCELLED;
        PCR_CHECK(r.state != kTaking && r.state != kHeld, "server_wait_pinned: ticket %llu is already taken",
                  (unsigned long long)ticket);
        if (r.state == kDone) {
            r.state = kHeld;
            *rgba_pinned = reinterpret_cast<const uint32_t*>(srv->h_arena + r.hoff);
            if (latency_ms) *latency_ms = r.t_done - r.t_submit;
            return PCR_OK;
        }
        if (srv->worker_error) {
            set_error("server_wait_pinned: the server stopped after a HIP error");
            return PCR_E_HIP;
        }
        srv->done_cv.wait_for(lk, std::chrono::milliseconds(2));
    }
}

int pcr_server_release(pcr_server* srv, uint64_t ticket) {
    PCR_CHECK(srv, "server_release: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    PCR_CHECK(it != srv->reqs.end() && it->second.state == kHeld, "server_release: ticket %llu is not held",
              (unsigned long long)ticket);
    srv->arena.release(it->second.off, it->second.bytes);
    srv->harena.release(it->second.hoff, it->second.hbytes);
    srv->reqs.erase(it);
    srv->work_cv.notify_one();
    return PCR_OK;
}

int pcr_server_cancel(pcr_server* srv, uint64_t ticket) {
    PCR_CHECK(srv, "server_cancel: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    PCR_CHECK(it != srv->reqs.end(), "server_cancel: unknown ticket %llu", (unsigned long long)ticket);
    Req& r = it->second;
    if (r.state == kPending) {
        for (auto p = srv->pending.begin(); p != srv->pending.end(); ++p)
            if (*p == ticket) { srv->pending.erase(p); break; }
        srv->reqs.erase(it);
    } else if (r.state == kInFlight) {
        r.state = kCancelled;            // released when its launch retires
    } else if (r.state == kDone || r.state == kHeld) {
        srv->arena.release(r.off, r.bytes);
        srv->harena.release(r.hoff, r.hbytes);
        srv->reqs.erase(it);
    }   // kTaking: another call is copying the finished image out; nothing to cancel
    return PCR_OK;
}

int pcr_server_stats_get(pcr_server* srv, pcr_server_stats* out) {
    PCR_CHECK(srv && out, "server_stats: null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    *out = srv->st;
    out->fb_bytes_in_use = srv->arena.used;
    out->fb_bytes_peak = srv->arena.peak;
    return PCR_OK;
}

int pcr_server_destroy(pcr_server* srv) {
    if (!srv) return PCR_OK;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->work_cv.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    for (Slot& sl : srv->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        mv_release(sl.B);
        (void)hipHostFree(sl.h_views); (void)hipHostFree(sl.h_lod); (void)hipHostFree(sl.h_fbs);
        (void)hipHostFree(sl.h_out);
        (void)hipHostFree(sl.h_state);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    (void)hipFree(srv->d_arena);
    (void)hipHostFree(srv->h_arena);
    delete srv;
    return PCR_OK;
}

}  // extern "C"
