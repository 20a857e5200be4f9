// pcr_capi.hip — extern "C" entry points of libpcr.so (declared in include/pcr.h).
//
// The per-frame sequence of pcr_render + pcr_resolve stands in for
// ComputeLoopLas::render (modules/compute_loop_las/compute_loop_las.h:117-304):
// clear -> select (cull + precision) -> raster -> resolve, all queued on the
// context's HIP stream. Errors are status codes plus pcr_last_error(), where
// the reference silently skipped (compute_loop_las.h:123,210,241).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>
#include "pcr_generate.h"
#include "pcr_internal.h"

namespace pcr {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? PCR_E_OOM : PCR_E_HIP;
}

static bool valid_view(const pcr_view* v) {
    return v && v->width > 0 && v->height > 0 && v->width <= PCR_MAX_DIM && v->height <= PCR_MAX_DIM;
}

static int clear_fb(hipStream_t s, pcr_fb* fb) {
    const uint64_t npix = (uint64_t)fb->width * fb->height;
    if (fb->kind == PCR_FB_KEY64) return launch_fill64(s, fb->d_keys, npix, PCR_FB_CLEAR);
    int rc = launch_fill32(s, fb->d_depth, npix, PCR_DEPTH_CLEAR);
    if (rc) return rc;
    return launch_fill32(s, fb->d_acc, 4 * npix, 0u);
}

}  // namespace pcr

using namespace pcr;

extern "C" {

const char* pcr_last_error(void) { return g_err; }
int pcr_version(void) { return PCR_VERSION_MAJOR * 100 + PCR_VERSION_MINOR; }

int pcr_device_count(int* count) {
    PCR_CHECK(count, "device_count: null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
    *count = n;
    return PCR_OK;
}

int pcr_create(int device, pcr_ctx** out) {
    PCR_CHECK(out, "create: null out");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return PCR_E_NO_DEVICE;
    }
    PCR_CHECK(device >= 0 && device < n, "create: device %d out of range (%d devices)", device, n);
    PCR_HIP(hipSetDevice(device));
    pcr_ctx* c = new pcr_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    PCR_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    PCR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    PCR_HIP(hipMalloc(&c->d_state, 2 * sizeof(FrameState)));
    PCR_HIP(hipMemset(c->d_state, 0, 2 * sizeof(FrameState)));
    c->ring = new hipEvent_t[PCR_TIMING_RING][6];
    for (int f = 0; f < PCR_TIMING_RING; f++)
        for (int k = 0; k < 6; k++) PCR_HIP(hipEventCreate(&c->ring[f][k]));
    *out = c;
    return PCR_OK;
}

int pcr_destroy(pcr_ctx* ctx) {
    if (!ctx) return PCR_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int f = 0; f < PCR_TIMING_RING; f++)
        for (int k = 0; k < 6; k++) (void)hipEventDestroy(ctx->ring[f][k]);
    delete[] ctx->ring;
    (void)hipFree(ctx->d_state);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PCR_OK;
}

int pcr_synchronize(pcr_ctx* ctx) {
    PCR_CHECK(ctx, "synchronize: null ctx");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    return PCR_OK;
}

void* pcr_stream(pcr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// glm lookAtRH / perspectiveRH_NO in double (Camera.h:34-38,
// matrix_transform.inl:258-274); world = translate(origin)
// (compute_loop_las.h:170-176); the products are cast to float.
int pcr_view_look_at(const double eye[3], const double target[3], const double up[3], double fovy_deg,
                     double near_plane, double far_plane, const double origin[3], int width, int height,
                     pcr_view* out) {
    PCR_CHECK(eye && target && up && origin && out, "view_look_at: null argument");
    PCR_CHECK(width > 0 && height > 0 && width <= PCR_MAX_DIM && height <= PCR_MAX_DIM,
              "view_look_at: bad size %dx%d", width, height);
    PCR_CHECK(near_plane > 0 && far_plane > near_plane, "view_look_at: bad clip planes");
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    PCR_CHECK(fl > 0, "view_look_at: eye == target");
    for (double& x : f) x /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    PCR_CHECK(sl > 0, "view_look_at: up parallel to view direction");
    for (double& x : s) x /= sl;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double V[16] = {0}, P[16] = {0}, W[16] = {0}, VW[16], PVW[16];
    V[0] = s[0]; V[4] = s[1]; V[8] = s[2];
    V[1] = u[0]; V[5] = u[1]; V[9] = u[2];
    V[2] = -f[0]; V[6] = -f[1]; V[10] = -f[2];
    V[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    V[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    V[14] = (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
    V[15] = 1.0;
    const double th = std::tan(fovy_deg * 3.14159265358979323846 / 180.0 / 2.0);
    const double aspect = (double)width / (double)height;
    P[0] = 1.0 / (aspect * th);
    P[5] = 1.0 / th;
    P[10] = -(far_plane + near_plane) / (far_plane - near_plane);
    P[11] = -1.0;
    P[14] = -(2.0 * far_plane * near_plane) / (far_plane - near_plane);
    W[0] = W[5] = W[10] = W[15] = 1.0;
    W[12] = origin[0]; W[13] = origin[1]; W[14] = origin[2];
    auto mul = [](const double* a, const double* b, double* r) {
        for (int c = 0; c < 4; c++)
            for (int rr = 0; rr < 4; rr++) {
                double acc = 0.0;
                for (int k = 0; k < 4; k++) acc += a[k * 4 + rr] * b[c * 4 + k];
                r[c * 4 + rr] = acc;
            }
    };
    mul(V, W, VW);
    mul(P, VW, PVW);
    memset(out, 0, sizeof(*out));
    for (int i = 0; i < 16; i++) {
        out->transform[i] = (float)PVW[i];
        out->transform_frustum[i] = (float)PVW[i];
        out->view_world[i] = (float)VW[i];
        out->proj[i] = (float)P[i];
    }
    out->width = width;
    out->height = height;
    return PCR_OK;
}

int pcr_view_from_matrices(const double view[16], const double proj[16], const double origin[3], int width,
                           int height, pcr_view* out) {
    PCR_CHECK(view && proj && origin && out, "view_from_matrices: null argument");
    PCR_CHECK(width > 0 && height > 0 && width <= PCR_MAX_DIM && height <= PCR_MAX_DIM,
              "view_from_matrices: bad size %dx%d", width, height);
    double W[16] = {0}, VW[16], PVW[16];
    W[0] = W[5] = W[10] = W[15] = 1.0;
    W[12] = origin[0]; W[13] = origin[1]; W[14] = origin[2];
    auto mul = [](const double* a, const double* b, double* r) {
        for (int c = 0; c < 4; c++)
            for (int rr = 0; rr < 4; rr++) {
                double acc = 0.0;
                for (int k = 0; k < 4; k++) acc += a[k * 4 + rr] * b[c * 4 + k];
                r[c * 4 + rr] = acc;
            }
    };
    mul(view, W, VW);
    mul(proj, VW, PVW);
    memset(out, 0, sizeof(*out));
    for (int i = 0; i < 16; i++) {
        out->transform[i] = (float)PVW[i];
        out->transform_frustum[i] = (float)PVW[i];
        out->view_world[i] = (float)VW[i];
        out->proj[i] = (float)proj[i];
    }
    out->width = width;
    out->height = height;
    return PCR_OK;
}

int pcr_encode_batches_host(const float* xyz, uint64_t n, uint32_t batch_size, pcr_batch* batches,
                            uint32_t* lo, uint32_t* med, uint32_t* hi) {
    PCR_CHECK(xyz && batches && lo && med && hi, "encode_batches_host: null argument");
    PCR_CHECK(batch_size > 0 && batch_size % 4 == 0, "encode_batches_host: batch_size %u not a multiple of 4",
              batch_size);
    PCR_CHECK(n < (1ull << 32), "encode_batches_host: more than 2^32-1 points");
    encode_batches_host(xyz, n, batch_size, batches, lo, med, hi);
    return PCR_OK;
}

// ------------------------------------------------------------------ clouds
int pcr_cloud_upload_points16(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n, pcr_cloud** out) {
    PCR_CHECK(ctx && xyz && rgba && out, "upload_points16: null argument");
    PCR_CHECK(n > 0, "upload_points16: no points");
    std::vector<uint32_t> packed(4 * n);
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = 0; i < n; i++) {
        memcpy(&packed[4 * i], xyz + 3 * i, 12);
        packed[4 * i + 3] = rgba[i];
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * i + a];
            mn[a] = v < mn[a] ? v : mn[a];
            mx[a] = v > mx[a] ? v : mx[a];
        }
    }
    pcr_cloud* c = new pcr_cloud();
    c->ctx = ctx;
    c->num_points = n;
    c->points16_len = n;
    if (hipMalloc(&c->d_points16, n * 16) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        set_error("upload_points16: device allocation of %llu bytes failed", (unsigned long long)(n * 16));
        return PCR_E_OOM;
    }
    c->device_bytes = n * 16;
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = mn[a]; c->aabb_max[a] = mx[a]; }
    PCR_HIP_OR(hipMemcpy(c->d_points16, packed.data(), n * 16, hipMemcpyHostToDevice), (cloud_free(c), delete c));
    *out = c;
    return PCR_OK;
}

int pcr_scene_ground_height(const pcr_gen_cfg* cfg, float x, float y, float* height) {
    PCR_CHECK(cfg && height, "scene_ground_height: null argument");
    PCR_CHECK(cfg->scene == PCR_SCENE_TERRAIN || cfg->scene == PCR_SCENE_CITY,
              "scene_ground_height: scene %d has no ground", cfg->scene);
    *height = terrain_height(cfg->seed, cfg->amplitude, x, y);
    return PCR_OK;
}

int pcr_cloud_upload_batched(pcr_ctx* ctx, const pcr_batch* batches, uint32_t nb, const uint32_t* lo,
                             const uint32_t* med, const uint32_t* hi, const uint32_t* rgba, uint64_t npoints,
                             pcr_cloud** out) {
    PCR_CHECK(ctx && batches && lo && med && hi && rgba && out, "upload_batched: null argument");
    PCR_CHECK(nb > 0 && npoints > 0 && npoints < (1ull << 32), "upload_batched: bad sizes");
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int max_level = -1;
    for (uint32_t b = 0; b < nb; b++) {
        const pcr_batch& h = batches[b];
        PCR_CHECK((uint64_t)h.first_point + h.num_points <= npoints,
                  "upload_batched: batch %u range [%u, +%u) exceeds %llu points", b, h.first_point, h.num_points,
                  (unsigned long long)npoints);
        PCR_CHECK(h.parent < (int32_t)nb && h.parent < (int32_t)b,
                  "upload_batched: batch %u parent %d must precede it", b, h.parent);
        for (int a = 0; a < 3; a++) {
            mn[a] = h.min[a] < mn[a] ? h.min[a] : mn[a];
            mx[a] = h.max[a] > mx[a] ? h.max[a] : mx[a];
        }
        if (h.parent >= 0 || h.spacing > 0) max_level = (int)h.level > max_level ? (int)h.level : max_level;
    }
    pcr_cloud* c = new pcr_cloud();
    int rc = cloud_alloc_batched(ctx, nb, npoints, c);
    if (rc) { delete c; return rc; }
    c->batch_size = batches[0].num_points;
    c->num_levels = max_level + 1;
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = mn[a]; c->aabb_max[a] = mx[a]; }
    const auto drop = [c] { cloud_free(c); delete c; };
    PCR_HIP_OR(hipMemcpy(c->d_batches, batches, sizeof(pcr_batch) * nb, hipMemcpyHostToDevice), drop());
    PCR_HIP_OR(hipMemcpy(c->d_lo, lo, npoints * 4, hipMemcpyHostToDevice), drop());
    PCR_HIP_OR(hipMemcpy(c->d_med, med, npoints * 4, hipMemcpyHostToDevice), drop());
    PCR_HIP_OR(hipMemcpy(c->d_hi, hi, npoints * 4, hipMemcpyHostToDevice), drop());
    PCR_HIP_OR(hipMemcpy(c->d_rgba, rgba, npoints * 4, hipMemcpyHostToDevice), drop());
    PCR_HIP_OR(hipStreamSynchronize(ctx->stream), drop());
    c->h_batches.assign(batches, batches + nb);
    *out = c;
    return PCR_OK;
}

int pcr_generate_points(pcr_ctx* ctx, const pcr_gen_cfg* cfg, float* xyz_dev, uint32_t* rgba_dev) {
    PCR_CHECK(ctx, "generate_points: null ctx");
    return generate_points(ctx, cfg, xyz_dev, rgba_dev);
}

int pcr_cloud_build_device(pcr_ctx* ctx, const float* xyz_dev, const uint32_t* rgba_dev, uint64_t n,
                           const pcr_build_opts* opts, pcr_cloud** out) {
    return build_cloud_device(ctx, xyz_dev, rgba_dev, n, opts, out);
}

int pcr_cloud_generate(pcr_ctx* ctx, const pcr_gen_cfg* cfg, const pcr_build_opts* opts, pcr_cloud** out) {
    PCR_CHECK(ctx && cfg && out, "cloud_generate: null argument");
    PCR_CHECK(cfg->num_points > 0, "cloud_generate: no points");
    float* xyz = nullptr;
    uint32_t* rgba = nullptr;
    PCR_HIP(hipMalloc(&xyz, cfg->num_points * 12));
    PCR_HIP(hipMalloc(&rgba, cfg->num_points * 4));
    int rc = generate_points(ctx, cfg, xyz, rgba);
    if (rc == PCR_OK) rc = build_cloud_device(ctx, xyz, rgba, cfg->num_points, opts, out);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(xyz);
    (void)hipFree(rgba);
    return rc;
}

int pcr_cloud_info_get(const pcr_cloud* c, pcr_cloud_info* out) {
    PCR_CHECK(c && out, "cloud_info: null argument");
    memset(out, 0, sizeof(*out));
    out->num_points = c->num_points;
    out->num_batches = c->num_batches;
    out->batch_size = c->batch_size;
    out->has_points16 = c->d_points16 != nullptr;
    out->has_batches = c->d_batches != nullptr;
    out->num_levels = c->num_levels;
    for (int a = 0; a < 3; a++) { out->aabb_min[a] = c->aabb_min[a]; out->aabb_max[a] = c->aabb_max[a]; }
    out->device_bytes = c->device_bytes;
    return PCR_OK;
}

int pcr_cloud_download(pcr_cloud* c, pcr_batch* batches, uint32_t* lo, uint32_t* med, uint32_t* hi, uint32_t* rgba,
                       float* xyz16, uint32_t* rgba16) {
    PCR_CHECK(c, "cloud_download: null cloud");
    PCR_HIP(hipStreamSynchronize(c->ctx->stream));
    const uint64_t n = c->num_points;
    if (c->d_batches) {
        if (batches) PCR_HIP(hipMemcpy(batches, c->d_batches, sizeof(pcr_batch) * c->num_batches, hipMemcpyDeviceToHost));
        if (lo) PCR_HIP(hipMemcpy(lo, c->d_lo, n * 4, hipMemcpyDeviceToHost));
        if (med) PCR_HIP(hipMemcpy(med, c->d_med, n * 4, hipMemcpyDeviceToHost));
        if (hi) PCR_HIP(hipMemcpy(hi, c->d_hi, n * 4, hipMemcpyDeviceToHost));
        if (rgba) PCR_HIP(hipMemcpy(rgba, c->d_rgba, n * 4, hipMemcpyDeviceToHost));
    }
    if (c->d_points16 && (xyz16 || rgba16)) {
        std::vector<uint32_t> p(4 * c->points16_len);
        PCR_HIP(hipMemcpy(p.data(), c->d_points16, c->points16_len * 16, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < c->points16_len; i++) {
            if (xyz16) memcpy(xyz16 + 3 * i, &p[4 * i], 12);
            if (rgba16) rgba16[i] = p[4 * i + 3];
        }
    }
    return PCR_OK;
}

int pcr_cloud_destroy(pcr_cloud* c) {
    if (!c) return PCR_OK;
    (void)hipDeviceSynchronize();
    cloud_free(c);
    delete c;
    return PCR_OK;
}

// ------------------------------------------------------------ framebuffers
int pcr_fb_create(pcr_ctx* ctx, int width, int height, int kind, pcr_fb** out) {
    PCR_CHECK(ctx && out, "fb_create: null argument");
    PCR_CHECK(width > 0 && height > 0 && width <= PCR_MAX_DIM && height <= PCR_MAX_DIM &&
                  (uint64_t)width * height < (1ull << 31),
              "fb_create: bad size %dx%d", width, height);
    PCR_CHECK(kind == PCR_FB_KEY64 || kind == PCR_FB_HQS, "fb_create: bad kind %d", kind);
    const uint64_t npix = (uint64_t)width * height;
    pcr_fb* f = new pcr_fb();
    f->ctx = ctx; f->width = width; f->height = height; f->kind = kind;
    hipError_t e = hipMalloc(&f->d_rgba, npix * 4);
    if (e == hipSuccess) {
        if (kind == PCR_FB_KEY64) e = hipMalloc(&f->d_keys, npix * 8);
        else {
            e = hipMalloc(&f->d_depth, npix * 4);
            if (e == hipSuccess) e = hipMalloc(&f->d_acc, npix * 16);
        }
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(f->d_rgba); (void)hipFree(f->d_keys); (void)hipFree(f->d_depth); (void)hipFree(f->d_acc);
        delete f;
        set_error("fb_create: device allocation for %dx%d failed", width, height);
        return PCR_E_OOM;
    }
    f->bytes = npix * (kind == PCR_FB_KEY64 ? 12 : 24);
    int rc = clear_fb(ctx->stream, f);
    if (rc == PCR_OK) rc = launch_fill32(ctx->stream, f->d_rgba, npix, 0u);
    if (rc) { pcr_fb_destroy(f); return rc; }
    *out = f;
    return PCR_OK;
}

int pcr_fb_destroy(pcr_fb* f) {
    if (!f) return PCR_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(f->d_rgba); (void)hipFree(f->d_keys); (void)hipFree(f->d_depth); (void)hipFree(f->d_acc);
    delete f;
    return PCR_OK;
}

int pcr_fb_clear(pcr_ctx* ctx, pcr_fb* fb) {
    PCR_CHECK(ctx && fb, "fb_clear: null argument");
    return clear_fb(ctx->stream, fb);
}

int pcr_fb_readback(pcr_ctx* ctx, pcr_fb* fb, uint64_t* keys) {
    PCR_CHECK(ctx && fb && keys, "fb_readback: null argument");
    PCR_CHECK(fb->kind == PCR_FB_KEY64, "fb_readback: not a KEY64 framebuffer");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    PCR_HIP(hipMemcpy(keys, fb->d_keys, (uint64_t)fb->width * fb->height * 8, hipMemcpyDeviceToHost));
    return PCR_OK;
}

int pcr_fb_readback_hqs(pcr_ctx* ctx, pcr_fb* fb, uint32_t* depth, uint32_t* acc) {
    PCR_CHECK(ctx && fb, "fb_readback_hqs: null argument");
    PCR_CHECK(fb->kind == PCR_FB_HQS, "fb_readback_hqs: not an HQS framebuffer");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t npix = (uint64_t)fb->width * fb->height;
    if (depth) PCR_HIP(hipMemcpy(depth, fb->d_depth, npix * 4, hipMemcpyDeviceToHost));
    if (acc) PCR_HIP(hipMemcpy(acc, fb->d_acc, npix * 16, hipMemcpyDeviceToHost));
    return PCR_OK;
}

int pcr_fb_readback_rgba(pcr_ctx* ctx, pcr_fb* fb, uint32_t* rgba) {
    PCR_CHECK(ctx && fb && rgba, "fb_readback_rgba: null argument");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    PCR_HIP(hipMemcpy(rgba, fb->d_rgba, (uint64_t)fb->width * fb->height * 4, hipMemcpyDeviceToHost));
    return PCR_OK;
}

void* pcr_fb_device_keys(pcr_fb* fb) { return fb ? (void*)fb->d_keys : nullptr; }
void* pcr_fb_device_rgba(pcr_fb* fb) { return fb ? (void*)fb->d_rgba : nullptr; }

// ------------------------------------------------------------------ render
int pcr_render(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* view, const pcr_render_params* params, pcr_fb* fb) {
    PCR_CHECK(ctx && cloud && params && fb, "render: null argument");
    PCR_CHECK(valid_view(view), "render: bad view");
    PCR_CHECK(view->width == fb->width && view->height == fb->height,
              "render: view size %dx%d differs from framebuffer %dx%d", view->width, view->height, fb->width,
              fb->height);
    const int mode = params->mode;
    const bool hqs = mode == PCR_MODE_HQS || mode == PCR_MODE_HQS16 || mode == PCR_MODE_NODES_HQS;
    const bool pts16 = mode == PCR_MODE_PLAIN16 || mode == PCR_MODE_HQS16;
    const bool lod = mode == PCR_MODE_NODES || mode == PCR_MODE_NODES_HQS;
    PCR_CHECK(mode >= PCR_MODE_PLAIN16 && mode <= PCR_MODE_NODES_HQS, "render: unknown mode %d", mode);
    PCR_CHECK(params->time_stages >= 0 && params->time_stages <= 2, "render: time_stages %d not in 0..2",
              params->time_stages);
    PCR_CHECK(hqs == (fb->kind == PCR_FB_HQS), "render: mode %d needs a %s framebuffer", mode,
              hqs ? "HQS" : "KEY64");
    PCR_CHECK(pts16 ? cloud->d_points16 != nullptr : cloud->d_batches != nullptr,
              "render: cloud has no %s layout for mode %d", pts16 ? "16 B/pt" : "batched", mode);
    PCR_CHECK(!lod || params->lod_rule == PCR_LOD_SPACING || params->lod_rule == PCR_LOD_SIZE80,
              "render: bad lod_rule %d", params->lod_rule);
    PCR_CHECK(params->tile_px == 0 || params->tile_px == 1024 || params->tile_px == 2048 || params->tile_px == 4096,
              "render: tile_px %d not in {0,1024,2048,4096}", params->tile_px);
    PCR_CHECK(params->occlusion_skip == 0 || params->occlusion_skip == -1, "render: occlusion_skip %d not 0 or -1",
              params->occlusion_skip);
    const bool occl = params->occlusion_skip == 0;
    PCR_CHECK(params->batch_end == 0 ? params->batch_begin == 0
                                      : params->batch_begin <= params->batch_end &&
                                            params->batch_end <= (pts16 ? 0u : cloud->num_batches),
              "render: batch range [%u, %u) outside the cloud's %u batches", params->batch_begin, params->batch_end,
              pts16 ? 0u : cloud->num_batches);
    const int block = params->block_size;
    PCR_CHECK(block == 0 || block == 64 || block == 128 || block == 256 || block == 512,
              "render: block_size %d not in {64,128,256,512}", block);
    hipStream_t s = ctx->stream;
    // 1: events around every stage; 2: around the raster stage only (two
    // events a frame instead of six, each of which costs the stream a few µs)
    ctx->timing = params->time_stages;
    const bool all = ctx->timing == 1;
    const uint64_t slot = ctx->frames_timed % PCR_TIMING_RING;
    hipEvent_t* ev = ctx->ring[slot];
    if (ctx->timing) {
        ctx->ring_level[slot] = (uint8_t)ctx->timing;
        ctx->frames_timed++;
    }
    if (all) PCR_HIP(hipEventRecord(ev[0], s));
    int rc = PCR_OK;
    FrameState* st = ctx->d_state + ctx->cur;
    FrameState* st_next = ctx->d_state + (1 - ctx->cur);
    if (!ctx->state_clean[ctx->cur] && (rc = launch_frame_reset(s, st))) return rc;
    // batched modes clear the framebuffer inside the select launch; the
    // 16 B/pt modes have no select and clear with their own fill
    if (params->clear_before && pts16 && (rc = clear_fb(s, fb))) return rc;
    if (all) PCR_HIP(hipEventRecord(ev[1], s));
    ctx->last_state = ctx->cur;
    if (!pts16) {
        const uint64_t npix = (uint64_t)fb->width * fb->height;
        FillRegion f0, f1;
        if (params->clear_before) {
            if (fb->kind == PCR_FB_KEY64) {
                f0 = FillRegion{reinterpret_cast<uint32_t*>(fb->d_keys), 2 * npix, 0xFFFFFFFFu};   // PCR_FB_CLEAR
            } else {
                f0 = FillRegion{fb->d_depth, npix, PCR_DEPTH_CLEAR};
                f1 = FillRegion{fb->d_acc, 4 * npix, 0u};
            }
        }
        rc = launch_select(s, cloud, *view, *params, st, st_next, f0, f1, lod ? 80.0f : 100.0f, lod);
        if (rc) return rc;
        // this frame's state is in use; the select reset the other one
        ctx->state_clean[ctx->cur] = false;
        ctx->state_clean[1 - ctx->cur] = true;
        ctx->cur = 1 - ctx->cur;
    }
    ctx->last_nb = pts16 ? 0 : cloud->num_batches;
    if (ctx->timing) PCR_HIP(hipEventRecord(ev[2], s));
    switch (mode) {
        case PCR_MODE_PLAIN16: rc = launch_raster_points16(s, cloud, *view, fb->d_keys); break;
        case PCR_MODE_LAS: rc = launch_raster_batched(s, cloud, *view, 1, block ? block : 128, st, fb->d_keys); break;
        case PCR_MODE_LAS2:
        case PCR_MODE_NODES:
            rc = params->tile_px ? launch_raster_tiled(s, cloud, *view, params->tile_px, block, st, fb->d_keys, occl)
                                 : launch_raster_batched(s, cloud, *view, 4, block ? block : 256, st, fb->d_keys);
            break;
        case PCR_MODE_HQS:
        case PCR_MODE_NODES_HQS:
            rc = params->tile_px ? launch_hqs_tiled(s, cloud, *view, params->tile_px, block, st, fb, occl)
                                 : launch_hqs_batched(s, cloud, *view, st, fb);
            break;
        case PCR_MODE_HQS16: rc = launch_hqs_points16(s, cloud, *view, fb); break;
    }
    if (rc) return rc;
    if (ctx->timing) PCR_HIP(hipEventRecord(ev[3], s));
    return PCR_OK;
}

int pcr_resolve(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_render_params* params, pcr_fb* fb) {
    PCR_CHECK(ctx && cloud && params && fb, "resolve: null argument");
    const int mode = params->mode;
    const bool hqs = mode == PCR_MODE_HQS || mode == PCR_MODE_HQS16 || mode == PCR_MODE_NODES_HQS;
    PCR_CHECK(hqs == (fb->kind == PCR_FB_HQS), "resolve: mode %d does not match the framebuffer kind", mode);
    PCR_CHECK(params->resolve_window >= 0 && params->resolve_window <= 4, "resolve: resolve_window %d not in 0..4",
              params->resolve_window);
    PCR_CHECK(params->edl_strength == 0.0f ||
                  (params->edl_strength > 0.0f && params->edl_strength < 1.0e30f &&
                   (mode == PCR_MODE_LAS || mode == PCR_MODE_LAS2 || mode == PCR_MODE_NODES)),
              "resolve: edl_strength %g needs a finite positive value and an index mode (LAS, LAS2, NODES)",
              (double)params->edl_strength);
    PCR_CHECK(hqs || mode == PCR_MODE_PLAIN16 || cloud->d_rgba, "resolve: cloud has no colour plane");
    hipEvent_t* ev = ctx->ring[(ctx->frames_timed + PCR_TIMING_RING - 1) % PCR_TIMING_RING];
    if (ctx->timing == 1) PCR_HIP(hipEventRecord(ev[4], ctx->stream));
    int rc = launch_resolve(ctx->stream, cloud, mode, params->background, params->resolve_window,
                            params->edl_strength, fb,
                            params->debug_counters == 1 ? ctx->d_state + ctx->last_state : nullptr);
    if (rc) return rc;
    if (ctx->timing == 1) PCR_HIP(hipEventRecord(ev[5], ctx->stream));
    return PCR_OK;
}

int pcr_stats_get(pcr_ctx* ctx, pcr_stats* out) {
    PCR_CHECK(ctx && out, "stats: null argument");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    FrameState st;
    PCR_HIP(hipMemcpy(&st, ctx->d_state + ctx->last_state, sizeof(st), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->batches_total = ctx->last_nb;
    out->batches_culled = st.culled;
    for (int i = 0; i < 5; i++) out->batches_per_level[i] = st.level_count[i];
    out->points_drawn = st.points_drawn;
    out->bytes_algorithmic = st.plane_bytes + 64ull * ctx->last_nb;
    out->batches_occluded = st.occluded_batches;
    out->points_occluded = st.occluded_points;
    out->bytes_occluded = st.occluded_bytes;
    out->atomics_issued = st.atomics;
    out->pixels_visible = st.pixels_visible;
    if (ctx->timing && ctx->frames_timed) {
        hipEvent_t* ev = ctx->ring[(ctx->frames_timed - 1) % PCR_TIMING_RING];
        float ms = 0;
        if (ctx->timing == 1 && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) out->ms_clear = ms;
        if (ctx->timing == 1 && hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) out->ms_cull = ms;
        if (hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) out->ms_raster = ms;
        if (ctx->timing == 1 && hipEventElapsedTime(&ms, ev[4], ev[5]) == hipSuccess) out->ms_resolve = ms;
        (void)hipGetLastError();
    }
    ctx->last = *out;
    return PCR_OK;
}

int pcr_frame_times(pcr_ctx* ctx, int n, float* out5, int* got) {
    PCR_CHECK(ctx && out5 && got && n >= 0, "frame_times: bad argument");
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t avail = ctx->frames_timed < PCR_TIMING_RING ? ctx->frames_timed : PCR_TIMING_RING;
    if ((uint64_t)n < avail) avail = (uint64_t)n;
    auto elapsed = [](hipEvent_t a, hipEvent_t b) {
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
            (void)hipGetLastError();
            ms = -1.0f;
        }
        return ms;
    };
    for (uint64_t k = 0; k < avail; k++) {
        const uint64_t f = ctx->frames_timed - avail + k;   // frame number
        hipEvent_t* ev = ctx->ring[f % PCR_TIMING_RING];
        float* o = out5 + 5 * k;
        o[2] = elapsed(ev[2], ev[3]);
        if (ctx->ring_level[f % PCR_TIMING_RING] == 1) {
            o[0] = elapsed(ev[0], ev[1]);
            o[1] = elapsed(ev[1], ev[2]);
            o[3] = elapsed(ev[4], ev[5]);
            o[4] = elapsed(ev[0], ev[5]);
        } else {
            // raster events only: no stage times, and the frame is the period
            // from the previous frame's raster end to this one's, when the
            // ring still holds that frame
            o[0] = o[1] = o[3] = -1.0f;
            const bool prev = f > 0 && ctx->frames_timed - f < PCR_TIMING_RING;
            o[4] = prev ? elapsed(ctx->ring[(f - 1) % PCR_TIMING_RING][3], ev[3]) : -1.0f;
        }
    }
    *got = (int)avail;
    return PCR_OK;
}

int pcr_selftest_divide(pcr_ctx* ctx, const float* num_x, const float* num_y, const float* den, uint64_t n,
                        float* out2) {
    PCR_CHECK(ctx && num_x && num_y && den && out2, "selftest_divide: null argument");
    PCR_CHECK(n <= (1ull << 28), "selftest_divide: n %llu exceeds 2^28", (unsigned long long)n);
    if (n == 0) return PCR_OK;
    float* d = nullptr;
    PCR_HIP(hipMalloc(&d, n * 20));
    int rc = PCR_OK;
    hipError_t e = hipMemcpy(d, num_x, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, num_y, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 2 * n, den, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = launch_selftest_divide(ctx->stream, d, d + n, d + 2 * n, n, d + 3 * n);
        if (rc == PCR_OK) e = hipMemcpyAsync(out2, d + 3 * n, n * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (rc == PCR_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    (void)hipFree(d);
    if (rc != PCR_OK) return rc;
    PCR_HIP(e);
    return PCR_OK;
}

}  // extern "C"
