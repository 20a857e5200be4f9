// Drives the four Method adapters of include/pcr_method.hpp the way the
// reference's main loop does (Runtime::setSelectedMethod, then update() and
// render() per frame, src/main.cpp:342,455) and checks every resolved image
// against a direct pcr_render + pcr_resolve of the same view.
// Prints "OK <method> <drawn pixels>" per method; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
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

int main() {
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
        for (const char* name : {"loop_las", "loop_las_prefetch", "loop_las_hqs", "loop_nodes"}) {
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
            check(pcr_fb_create(ctx, v.width, v.height, bm->params().mode == PCR_MODE_HQS ? PCR_FB_HQS : PCR_FB_KEY64, &fb));
            pcr_cloud* c = bm->params().mode == PCR_MODE_NODES ? tree : cloud;
            check(pcr_render(ctx, c, &pv, &bm->params(), fb));
            check(pcr_resolve(ctx, c, &bm->params(), fb));
            std::vector<uint32_t> img((size_t)v.width * v.height);
            check(pcr_fb_readback_rgba(ctx, fb, img.data()));
            pcr_fb_destroy(fb);
            size_t drawn = 0;
            for (uint32_t p : v.image) drawn += p != PCR_BG_LAS;
            const bool same = img == v.image;
            std::printf("%s %s %zu\n", same && drawn > 1000 ? "OK" : "FAIL", name, drawn);
            failures += !(same && drawn > 1000);
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
