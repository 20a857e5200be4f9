// Drives the Method adapters of include/pcr_method.hpp the way the
// reference's main loop does (Runtime::setSelectedMethod, then update() and
// render() per frame, src/main.cpp:342,455) and checks every resolved image
// against a direct pcr_render + pcr_resolve of the same view (both eyes for
// the two VR methods).
// Prints "OK <method> <drawn pixels>" per method; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "pcr_method.hpp"

using namespace pcr_host;

static void look_at(const double e[3], const double t[3], const double up[3], double out[16]) {
    double f[3] = {t[0] - e[0], t[1] - e[1], t[2] - e[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    memset(out, 0, 16 * sizeof(double));
    out[0] = s[0]; out[4] = s[1]; out[8] = s[2];
    out[1] = u[0]; out[5] = u[1]; out[9] = u[2];
    out[2] = -f[0]; out[6] = -f[1]; out[10] = -f[2];
    out[12] = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    out[13] = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    out[14] = f[0] * e[0] + f[1] * e[1] + f[2] * e[2];
    out[15] = 1.0;
}

static void perspective(double fovy_deg, double aspect, double n, double f, double out[16]) {
    const double th = std::tan(fovy_deg * 3.14159265358979323846 / 360.0);
    memset(out, 0, 16 * sizeof(double));
    out[0] = 1.0 / (aspect * th);
    out[5] = 1.0 / th;
    out[10] = -(f + n) / (f - n);
    out[11] = -1.0;
    out[14] = -(2.0 * f * n) / (f - n);
}

// Renders `parts` of `cloud` with method `m` and directly (one pcr_render per
// part, clear before the first); returns the number of drawn pixels, or -1
// when the images differ.
static long compare_parts(pcr_ctx* ctx, Renderer& r, BatchedMethod* m, pcr_cloud* cloud,
                          const std::vector<CloudPart>& parts) {
    View& v = r.views[0];
    m->update(&r);
    m->render(&r);
    pcr_fb* fb = nullptr;
    check(pcr_fb_create(ctx, v.width, v.height, PCR_FB_KEY64, &fb));
    for (size_t k = 0; k < parts.size(); k++) {
        pcr_view pv;
        check(pcr_view_from_matrices(v.view, v.proj, parts[k].origin, v.width, v.height, &pv));
        pcr_render_params p = m->params();
        p.clear_before = k == 0;
        p.batch_begin = parts[k].batch_begin;
        p.batch_end = parts[k].batch_end;
        check(pcr_render(ctx, cloud, &pv, &p, fb));
    }
    check(pcr_resolve(ctx, cloud, &m->params(), fb));
    std::vector<uint32_t> img((size_t)v.width * v.height);
    check(pcr_fb_readback_rgba(ctx, fb, img.data()));
    pcr_fb_destroy(fb);
    long drawn = 0;
    for (uint32_t px : v.image) drawn += px != PCR_BG_LAS;
    return img == v.image ? drawn : -1;
}

// method_main --las <file>...: LAS files through LasFiles + loop_las_prefetch.
// method_main --potree <dir> <pcr_potree_info bytes>: a Potree 2.0 dataset
// through PotreeCloud + loop_nodes. The camera looks at the first part's
// origin from (150, -250, 200) m in front of it.
static int run_files(int argc, char** argv) {
    pcr_ctx* ctx = nullptr;
    check(pcr_create(0, &ctx));
    Renderer r;
    r.ctx = ctx;
    r.views.resize(1);
    View& v = r.views[0];
    v.width = 800;
    v.height = 600;
    int failures = 0;
    auto camera = [&](const double o[3]) {
        const double eye[3] = {o[0] + 150, o[1] - 250, o[2] + 200}, tgt[3] = {o[0] + 150, o[1] + 100, o[2]};
        const double up[3] = {0, 0, 1};
        look_at(eye, tgt, up, v.view);
        perspective(60.0, 800.0 / 600.0, 1.0, 5000.0, v.proj);
    };
    if (std::string(argv[1]) == "--las") {
        LasFiles las(ctx, std::vector<std::string>(argv + 2, argv + argc));
        camera(las.parts[0].origin);
        ComputeLoopLas2 m(ctx, las.cloud, las.parts);
        const long drawn = compare_parts(ctx, r, &m, las.cloud, las.parts);
        uint64_t hash = 1469598103934665603ull;   // FNV-1a over the image
        for (uint32_t px : v.image)
            for (int b = 0; b < 4; b++) hash = (hash ^ ((px >> (8 * b)) & 255u)) * 1099511628211ull;
        std::printf("%s las_files %zu %ld %016llx\n", drawn > 1000 ? "OK" : "FAIL", las.parts.size(), drawn,
                    (unsigned long long)hash);
        failures += drawn <= 1000;
    } else if (std::string(argv[1]) == "--potree" && argc == 4) {
        const std::vector<char> raw = read_file(argv[3]);
        if (raw.size() != sizeof(pcr_potree_info)) throw Error(PCR_E_INVALID, "bad pcr_potree_info file");
        pcr_potree_info info;
        std::memcpy(&info, raw.data(), sizeof(info));
        PotreeCloud pc(ctx, info, argv[2]);
        const double mid[3] = {pc.part.origin[0] + 362, pc.part.origin[1] + 512, pc.part.origin[2] + 300};
        camera(mid);
        ComputeLoopNodes m(ctx, pc.cloud, std::vector<CloudPart>{pc.part});
        const long drawn = compare_parts(ctx, r, &m, pc.cloud, std::vector<CloudPart>{pc.part});
        std::printf("%s potree %ld\n", drawn > 1000 ? "OK" : "FAIL", drawn);
        failures += drawn <= 1000;
    } else {
        throw Error(PCR_E_INVALID, "usage: method_main [--las file... | --potree dir info]");
    }
    pcr_destroy(ctx);
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc > 1) {
        try {
            return run_files(argc, argv);
        } catch (const Error& e) {
            std::printf("FAIL pcr error %d: %s\n", e.code, e.what());
            return 1;
        }
    }
    try {
        pcr_ctx* ctx = nullptr;
        check(pcr_create(0, &ctx));
        pcr_gen_cfg cfg{PCR_SCENE_TERRAIN, 0, 7, 400000, 2000.0f, 150.0f, 0.5f, 0.0f};
        pcr_build_opts sorted{10240, 1, 0}, raw{10240, 0, 1};
        pcr_cloud *cloud = nullptr, *pts = nullptr, *tree = nullptr;
        check(pcr_cloud_generate(ctx, &cfg, &sorted, &cloud));
        check(pcr_cloud_generate(ctx, &cfg, &raw, &pts));
        std::vector<float> xyz(3 * cfg.num_points);
        std::vector<uint32_t> rgba(cfg.num_points);
        check(pcr_cloud_download(pts, nullptr, nullptr, nullptr, nullptr, nullptr, xyz.data(), rgba.data()));
        pcr_octree_opts oo{3, 128.0f, 20000, 12};
        check(pcr_cloud_build_octree_host(ctx, xyz.data(), rgba.data(), cfg.num_points, &oo, &tree));

        const double origin[3] = {0, 0, 0};
        Runtime::addMethod(new ComputeLoopLas(ctx, cloud, origin));
        Runtime::addMethod(new ComputeLoopLas2(ctx, cloud, origin));
        Runtime::addMethod(new ComputeLoopLasHqs(ctx, cloud, origin));
        Runtime::addMethod(new ComputeLoopNodes(ctx, tree, origin));
        Runtime::addMethod(new ComputeLoopNodesHqs(ctx, tree, origin));
        Runtime::addMethod(new ComputeLoopLasHqsVR(ctx, cloud, origin));
        Runtime::addMethod(new ComputeLoopNodesHqsVR(ctx, tree, origin));

        Renderer r;
        r.ctx = ctx;
        r.views.resize(1);
        View& v = r.views[0];
        v.width = 800;
        v.height = 450;
        const double eye[3] = {1000, -200, 400}, tgt[3] = {1000, 1000, 50}, up[3] = {0, 0, 1};
        look_at(eye, tgt, up, v.view);
        perspective(60.0, 800.0 / 450.0, 1.0, 5000.0, v.proj);

        int failures = 0;
        for (const char* name : {"loop_las", "loop_las_prefetch", "loop_las_hqs", "loop_nodes", "loop_nodes_hqs"}) {
            Runtime::setSelectedMethod(name);
            Method* m = Runtime::getSelectedMethod();
            if (!m || m->name != name) { std::printf("FAIL %s not registered\n", name); return 1; }
            m->update(&r);
            m->render(&r);
            auto* bm = static_cast<BatchedMethod*>(m);
            // direct path, same view, fresh framebuffer
            pcr_view pv;
            check(pcr_view_from_matrices(v.view, v.proj, origin, v.width, v.height, &pv));
            pcr_fb* fb = nullptr;
            const int md = bm->params().mode;
            check(pcr_fb_create(ctx, v.width, v.height,
                                md == PCR_MODE_HQS || md == PCR_MODE_NODES_HQS ? PCR_FB_HQS : PCR_FB_KEY64, &fb));
            pcr_cloud* c = md == PCR_MODE_NODES || md == PCR_MODE_NODES_HQS ? tree : cloud;
            check(pcr_render(ctx, c, &pv, &bm->params(), fb));
            check(pcr_resolve(ctx, c, &bm->params(), fb));
            std::vector<uint32_t> img((size_t)v.width * v.height);
            check(pcr_fb_readback_rgba(ctx, fb, img.data()));
            pcr_fb_destroy(fb);
            size_t drawn = 0;
            const uint32_t bg = md == PCR_MODE_NODES_HQS ? 0u : PCR_BG_LAS;   // nodes_hqs resolves empty pixels to 0
            for (uint32_t p : v.image) drawn += p != bg;
            const bool same = img == v.image;
            std::printf("%s %s %zu\n", same && drawn > 1000 ? "OK" : "FAIL", name, drawn);
            failures += !(same && drawn > 1000);
        }
        // the two-eye methods: views[1] is a second eye 6 m to the right
        r.views.resize(2);
        r.views[1] = r.views[0];
        {
            const double eye1[3] = {1006, -200, 400}, tgt1[3] = {1006, 1000, 50};
            look_at(eye1, tgt1, up, r.views[1].view);
        }
        for (const char* name : {"loop_las_hqs_vr", "loop_nodes_hqs_vr"}) {
            Runtime::setSelectedMethod(name);
            Method* m = Runtime::getSelectedMethod();
            if (!m || m->name != name) { std::printf("FAIL %s not registered\n", name); return 1; }
            auto* sm = static_cast<StereoHqsMethod*>(m);
            const int md = sm->params().mode;
            pcr_cloud* c = md == PCR_MODE_NODES_HQS ? tree : cloud;
            for (int frame = 0; frame < 2; frame++) {   // the second frame replays the captured plan
                m->update(&r);
                m->render(&r);
                for (int e = 0; e < 2; e++) {
                    View& ve = r.views[e];
                    pcr_view pv;
                    check(pcr_view_from_matrices(ve.view, ve.proj, origin, ve.width, ve.height, &pv));
                    pcr_fb* fb = nullptr;
                    check(pcr_fb_create(ctx, ve.width, ve.height, PCR_FB_HQS, &fb));
                    pcr_render_params p = sm->params();
                    p.clear_before = 1;
                    check(pcr_render(ctx, c, &pv, &p, fb));
                    check(pcr_resolve(ctx, c, &p, fb));
                    std::vector<uint32_t> img((size_t)ve.width * ve.height);
                    check(pcr_fb_readback_rgba(ctx, fb, img.data()));
                    pcr_fb_destroy(fb);
                    size_t drawn = 0;
                    for (uint32_t px : ve.image) drawn += px != 0u;
                    const bool same = img == ve.image;
                    std::printf("%s %s frame %d eye %d %zu\n", same && drawn > 1000 ? "OK" : "FAIL", name, frame, e,
                                drawn);
                    failures += !(same && drawn > 1000);
                }
            }
            if (r.views[0].image == r.views[1].image) {
                std::printf("FAIL %s: both eyes got the same image\n", name);
                failures++;
            }
        }
        for (Method* m : Runtime::methods) delete m;
        pcr_cloud_destroy(tree);
        pcr_cloud_destroy(pts);
        pcr_cloud_destroy(cloud);
        pcr_destroy(ctx);
        return failures ? 1 : 0;
    } catch (const Error& e) {
        std::printf("FAIL pcr error %d: %s\n", e.code, e.what());
        return 1;
    }
}
