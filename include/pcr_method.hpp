// pcr_method.hpp — the reference's Method plugin interface, over libpcr's C ABI.
//
// Mirrors include/Method.h:10-24 (struct Method with name/description/group
// and per-frame update/render), include/Runtime.h:17-50 (static method
// registry, selection by name) and the headless part of Renderer::views
// (include/Renderer.h:78-82). The concrete methods stand in for
// ComputeLoopLas (modules/compute_loop_las/compute_loop_las.h), ComputeLoopLas2
// (compute_loop_las2.h), ComputeLoopLasHqs (compute_loop_las_hqs.h),
// ComputeLoopNodes (compute_loop_nodes.h) and ComputeLoopNodesHqs
// (compute_loop_nodes_hqs.h): each render() rasterises the loaded cloud for
// views[0] and resolves into that view's RGBA8 image. The two VR methods
// (compute_loop_las_hqs_vr.h, compute_loop_nodes_hqs_vr.h) render views[0]
// and views[1] together. Errors throw pcr_host::Error (the reference skipped
// silently, compute_loop_las.h:123,210).
#pragma once
#include <cstdint>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "pcr.h"

namespace pcr_host {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
    if (rc != PCR_OK) throw Error(rc, pcr_last_error());
}

// Renderer::views[i]: double view/projection matrices (column-major) and the
// resolved colour attachment, here a host RGBA8 array.
struct View {
    double view[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double proj[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    int width = 0, height = 0;
    std::vector<uint32_t> image;
};

struct Renderer {
    pcr_ctx* ctx = nullptr;
    std::vector<View> views;
};

struct Method {
    std::string name = "no name";
    std::string description = "";
    std::string group = "no group";
    virtual ~Method() = default;
    virtual void update(Renderer* renderer) = 0;
    virtual void render(Renderer* renderer) = 0;
};

struct Runtime {
    inline static std::vector<Method*> methods;
    inline static Method* selectedMethod = nullptr;
    static void addMethod(Method* m) { methods.push_back(m); }
    static void setSelectedMethod(const std::string& name) {
        for (Method* m : methods)
            if (m->name == name) selectedMethod = m;
    }
    static Method* getSelectedMethod() { return selectedMethod; }
};

// One drawable piece of a cloud: a batch range (0, 0 = all) and the origin
// its coordinates are relative to. The reference gives every LAS file its
// own translate(boxMin) (compute_loop_las.h:170-176).
struct CloudPart {
    uint32_t batch_begin = 0, batch_end = 0;
    double origin[3] = {0.0, 0.0, 0.0};
};

inline std::vector<char> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(PCR_E_INVALID, "cannot open " + path);
    return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

// LasLoaderSparse's job (LasLoaderSparse.cpp:37-212, 495-526): reads the
// files, encodes them into one cloud on the device, one part per file that
// has points (an empty file owns no batches, and its empty range would read
// as "all batches" when it starts at batch 0).
struct LasFiles {
    pcr_cloud* cloud = nullptr;
    std::vector<CloudPart> parts;
    LasFiles(pcr_ctx* ctx, const std::vector<std::string>& paths) {
        std::vector<std::vector<char>> data;
        std::vector<pcr_las_file> files;
        for (const std::string& p : paths) data.push_back(read_file(p));
        for (const auto& d : data) files.push_back(pcr_las_file{d.data(), (uint64_t)d.size()});
        std::vector<uint32_t> first(files.size() + 1);
        check(pcr_cloud_from_las(ctx, files.data(), (int)files.size(), first.data(), &cloud));
        for (size_t k = 0; k < files.size(); k++) {
            pcr_las_info info;
            check(pcr_las_parse_header(files[k].bytes, files[k].size, &info));
            if (info.num_points == 0) continue;
            CloudPart part;
            part.batch_begin = first[k];
            part.batch_end = first[k + 1];
            for (int a = 0; a < 3; a++) part.origin[a] = info.min[a];
            parts.push_back(part);
        }
    }
    ~LasFiles() { pcr_cloud_destroy(cloud); }
    LasFiles(const LasFiles&) = delete;
    LasFiles& operator=(const LasFiles&) = delete;
};

// PotreeData::create + load (PotreeData.h:144-511) for a dataset whose
// metadata fields the caller has read from metadata.json.
struct PotreeCloud {
    pcr_cloud* cloud = nullptr;
    CloudPart part;
    PotreeCloud(pcr_ctx* ctx, const pcr_potree_info& info, const std::string& dir) {
        const std::vector<char> h = read_file(dir + "/hierarchy.bin"), o = read_file(dir + "/octree.bin");
        check(pcr_cloud_from_potree(ctx, &info, h.data(), h.size(), o.data(), o.size(), &cloud));
        for (int a = 0; a < 3; a++) part.origin[a] = info.min[a];
    }
    ~PotreeCloud() { pcr_cloud_destroy(cloud); }
    PotreeCloud(const PotreeCloud&) = delete;
    PotreeCloud& operator=(const PotreeCloud&) = delete;
};

// Shared body of the batched methods: one framebuffer per method, resized
// when the view size changes (compute_loop_las.h:130-138), render of every
// part (the framebuffer is cleared before the first) + resolve + readback of
// the image each frame.
class BatchedMethod : public Method {
  public:
    BatchedMethod(pcr_ctx* ctx, pcr_cloud* cloud, std::vector<CloudPart> parts, int mode, int fb_kind)
        : ctx_(ctx), cloud_(cloud), parts_(std::move(parts)), mode_(mode), fb_kind_(fb_kind) {
        // hqs runs its depth and colour passes per render call, so several
        // parts would colour before the last part's depth is known
        if ((mode == PCR_MODE_HQS || mode == PCR_MODE_NODES_HQS) && parts_.size() > 1)
            throw Error(PCR_E_INVALID, "the hqs methods draw one part per frame");
        params_ = pcr_render_params{};
        params_.mode = mode;
        params_.frustum_culling = 1;
        params_.lod_rule = PCR_LOD_SPACING;
        params_.lod_px = 1.0f;
        params_.background = PCR_BG_LAS;
        params_.clear_before = 1;
        // the LDS-tiled kernels give the same images faster (loop_las keeps
        // its one-point-per-thread kernel)
        params_.tile_px = mode == PCR_MODE_LAS ? 0 : 2048;
    }
    ~BatchedMethod() override { pcr_fb_destroy(fb_); }
    void update(Renderer*) override {}
    void render(Renderer* r) override {
        if (!cloud_ || r->views.empty()) return;   // nothing loaded (compute_loop_las.h:123)
        View& v = r->views[0];
        if (!fb_ || fb_w_ != v.width || fb_h_ != v.height) {
            pcr_fb_destroy(fb_);
            fb_ = nullptr;
            check(pcr_fb_create(ctx_, v.width, v.height, fb_kind_, &fb_));
            fb_w_ = v.width;
            fb_h_ = v.height;
        }
        for (size_t k = 0; k < parts_.size(); k++) {
            pcr_view pv;
            check(pcr_view_from_matrices(v.view, v.proj, parts_[k].origin, v.width, v.height, &pv));
            pcr_render_params p = params_;
            p.clear_before = k == 0 ? params_.clear_before : 0;
            p.batch_begin = parts_[k].batch_begin;
            p.batch_end = parts_[k].batch_end;
            check(pcr_render(ctx_, cloud_, &pv, &p, fb_));
        }
        check(pcr_resolve(ctx_, cloud_, &params_, fb_));
        v.image.resize((size_t)v.width * v.height);
        check(pcr_fb_readback_rgba(ctx_, fb_, v.image.data()));
    }
    pcr_render_params& params() { return params_; }
    pcr_fb* framebuffer() { return fb_; }

  protected:
    pcr_ctx* ctx_;
    pcr_cloud* cloud_;
    std::vector<CloudPart> parts_;
    int mode_, fb_kind_;
    pcr_render_params params_;
    pcr_fb* fb_ = nullptr;
    int fb_w_ = 0, fb_h_ = 0;
};

inline std::vector<CloudPart> one_part(const double origin[3]) {
    std::vector<CloudPart> parts(1);
    for (int i = 0; i < 3; i++) parts[0].origin[i] = origin ? origin[i] : 0.0;
    return parts;
}

struct ComputeLoopLas : BatchedMethod {       // compute_loop_las.h:80-92
    ComputeLoopLas(pcr_ctx* c, pcr_cloud* cl, const double o[3]) : ComputeLoopLas(c, cl, one_part(o)) {}
    ComputeLoopLas(pcr_ctx* c, pcr_cloud* cl, std::vector<CloudPart> parts)
        : BatchedMethod(c, cl, std::move(parts), PCR_MODE_LAS, PCR_FB_KEY64) {
        name = "loop_las";
        description = "one workgroup per batch, frustum cull, adaptive 10/20/30-bit precision";
        group = "10-10-10 bit encoded";
    }
};
struct ComputeLoopLas2 : BatchedMethod {      // compute_loop_las2.h:80-94
    ComputeLoopLas2(pcr_ctx* c, pcr_cloud* cl, const double o[3]) : ComputeLoopLas2(c, cl, one_part(o)) {}
    ComputeLoopLas2(pcr_ctx* c, pcr_cloud* cl, std::vector<CloudPart> parts)
        : BatchedMethod(c, cl, std::move(parts), PCR_MODE_LAS2, PCR_FB_KEY64) {
        name = "loop_las_prefetch";
        description = "as loop_las, four consecutive points per thread with 128-bit loads";
        group = "10-10-10 bit encoded";
    }
};
struct ComputeLoopLasHqs : BatchedMethod {    // compute_loop_las_hqs.h:86-95
    ComputeLoopLasHqs(pcr_ctx* c, pcr_cloud* cl, const double o[3])
        : BatchedMethod(c, cl, one_part(o), PCR_MODE_HQS, PCR_FB_HQS) {
        name = "loop_las_hqs";
        description = "high-quality shading: depth pass, then averaging of points within 1% of the depth";
        group = "10-10-10 bit encoded";
    }
};
struct ComputeLoopNodes : BatchedMethod {     // compute_loop_nodes.h:74-84
    ComputeLoopNodes(pcr_ctx* c, pcr_cloud* cl, const double o[3]) : ComputeLoopNodes(c, cl, one_part(o)) {}
    ComputeLoopNodes(pcr_ctx* c, pcr_cloud* cl, std::vector<CloudPart> parts)
        : BatchedMethod(c, cl, std::move(parts), PCR_MODE_NODES, PCR_FB_KEY64) {
        name = "loop_nodes";
        description = "LOD octree nodes: frustum and screen-space error cut";
        group = "render Potree nodes";
    }
};

struct ComputeLoopNodesHqs : BatchedMethod {  // compute_loop_nodes_hqs.h:82-92
    ComputeLoopNodesHqs(pcr_ctx* c, pcr_cloud* cl, const double o[3])
        : BatchedMethod(c, cl, one_part(o), PCR_MODE_NODES_HQS, PCR_FB_HQS) {
        name = "loop_nodes_hqs";
        description = "LOD octree nodes with high-quality shading and a weighted resolve";
        group = "render Potree nodes";
    }
};

// The two-eye VR methods (compute_loop_las_hqs_vr.h, compute_loop_nodes_hqs_vr.h)
// render views[0] and views[1] with one decode of each batch: here one
// pcr_multiview plan over both views, captured in a hipGraph and rebuilt
// when a view's size changes. Empty pixels resolve to 0, as the VR resolve
// writes them (compute_loop_las_hqs_vr/resolve.cs:66-72). Controller input
// and the world_vr transform stay with the caller: fold them into the views.
class StereoHqsMethod : public Method {
  public:
    StereoHqsMethod(pcr_ctx* ctx, pcr_cloud* cloud, const double origin[3], int mode)
        : ctx_(ctx), cloud_(cloud), mode_(mode) {
        for (int i = 0; i < 3; i++) origin_[i] = origin ? origin[i] : 0.0;
        params_ = pcr_render_params{};
        params_.mode = mode;
        params_.frustum_culling = 1;
        params_.lod_rule = PCR_LOD_SPACING;
        params_.lod_px = 1.0f;
        params_.background = 0;
        params_.resolve_window = 1;
    }
    ~StereoHqsMethod() override { release(); }
    void update(Renderer*) override {}
    void render(Renderer* r) override {
        if (!cloud_ || r->views.size() < 2) return;
        pcr_view pv[2];
        for (int e = 0; e < 2; e++) {
            const View& v = r->views[e];
            check(pcr_view_from_matrices(v.view, v.proj, origin_, v.width, v.height, &pv[e]));
        }
        if (!mv_ || w_[0] != pv[0].width || h_[0] != pv[0].height || w_[1] != pv[1].width ||
            h_[1] != pv[1].height) {
            release();
            for (int e = 0; e < 2; e++) {
                check(pcr_fb_create(ctx_, pv[e].width, pv[e].height, PCR_FB_HQS, &fb_[e]));
                w_[e] = pv[e].width;
                h_[e] = pv[e].height;
            }
            pcr_tune t{256, 40, 0, 2048};
            check(pcr_multiview_create(ctx_, cloud_, pv, 2, fb_, &params_, &t, 1, &mv_));
        } else {
            check(pcr_multiview_set_views(mv_, pv, 2));
        }
        check(pcr_multiview_launch(mv_));
        for (int e = 0; e < 2; e++) {
            View& v = r->views[e];
            v.image.resize((size_t)v.width * v.height);
            check(pcr_fb_readback_rgba(ctx_, fb_[e], v.image.data()));
        }
    }
    pcr_render_params& params() { return params_; }   // takes effect when the plan is next rebuilt
    pcr_fb* framebuffer(int eye) { return fb_[eye]; }

  private:
    void release() {
        pcr_multiview_destroy(mv_);
        mv_ = nullptr;
        for (int e = 0; e < 2; e++) { pcr_fb_destroy(fb_[e]); fb_[e] = nullptr; w_[e] = h_[e] = 0; }
    }
    pcr_ctx* ctx_;
    pcr_cloud* cloud_;
    int mode_;
    double origin_[3];
    pcr_render_params params_;
    pcr_multiview* mv_ = nullptr;
    pcr_fb* fb_[2] = {nullptr, nullptr};
    int w_[2] = {0, 0}, h_[2] = {0, 0};
};

struct ComputeLoopLasHqsVR : StereoHqsMethod {     // compute_loop_las_hqs_vr.h:61-70
    ComputeLoopLasHqsVR(pcr_ctx* c, pcr_cloud* cl, const double o[3]) : StereoHqsMethod(c, cl, o, PCR_MODE_HQS) {
        name = "loop_las_hqs_vr";
        description = "hqs for two eyes: each batch is decoded once and drawn into both views";
        group = "10-10-10 bit encoded";
    }
};
struct ComputeLoopNodesHqsVR : StereoHqsMethod {   // compute_loop_nodes_hqs_vr.h:99-108
    ComputeLoopNodesHqsVR(pcr_ctx* c, pcr_cloud* cl, const double o[3])
        : StereoHqsMethod(c, cl, o, PCR_MODE_NODES_HQS) {
        name = "loop_nodes_hqs_vr";
        description = "LOD nodes with hqs for two eyes, one decode per node, weighted resolve";
        group = "render Potree nodes";
    }
};

}  // namespace pcr_host
