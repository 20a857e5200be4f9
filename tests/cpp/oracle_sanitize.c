/* Exercises the CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer
 * (built by tests/test_oracle_sanitize.py with gcc -fsanitize=address,undefined,
 * CPU only). Covers generation, Morton order, encoding, batched / 16 B/pt
 * rendering, hqs, LOD selection and the multithreaded renderer, including the
 * edge cases the oracle decides: zero-extent batches, partial last batches,
 * NaN coordinates and empty frames. Exit status 0 when every invariant holds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../oracle/pcr_oracle.h"

#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); return 1; } } while (0)

int main(void) {
    const uint64_t n = 20003;
    pcr_gen_cfg cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.scene = PCR_SCENE_CITY; cfg.num_buildings = 8; cfg.seed = 5; cfg.num_points = n;
    cfg.extent = 300.0f; cfg.amplitude = 30.0f; cfg.jitter = 0.5f; cfg.terrain_frac = 0.3f;
    float* xyz = malloc(12 * n);
    uint32_t* rgba = malloc(4 * n);
    orc_generate(&cfg, 0, n, xyz, rgba);
    /* a flat (zero-extent) run and a NaN point */
    for (int i = 0; i < 600; i++) xyz[3 * i + 2] = 7.0f;
    xyz[3 * 777] = NAN;

    uint32_t* perm = malloc(4 * n);
    uint64_t* keys = malloc(8 * n);
    orc_morton_order(xyz, n, perm, keys);
    for (uint64_t i = 1; i < n; i++) CHECK(keys[i] >= keys[i - 1]);
    float* sorted = malloc(12 * n);
    for (uint64_t i = 0; i < n; i++) memcpy(sorted + 3 * i, xyz + 3 * perm[i], 12);

    const uint32_t bs = 512;
    const uint32_t nb = (uint32_t)((n + bs - 1) / bs);
    pcr_batch* b = malloc(sizeof(pcr_batch) * nb);
    uint32_t *lo = malloc(4 * n), *med = malloc(4 * n), *hi = malloc(4 * n);
    orc_encode_batches(sorted, n, bs, b, lo, med, hi);
    CHECK(b[nb - 1].num_points == n - (uint64_t)(nb - 1) * bs);

    const double eye[3] = {150, -80, 90}, tgt[3] = {150, 150, 10}, up[3] = {0, 0, 1}, org[3] = {0, 0, 0};
    pcr_view v;
    orc_look_at_view(eye, tgt, up, 60.0, 1.0, 2000.0, org, 160, 90, &v);
    const uint64_t npix = 160 * 90;
    uint64_t *fb = malloc(8 * npix), *fb2 = malloc(8 * npix);
    uint8_t* levels = malloc(nb);
    orc_clear64(fb, npix);
    orc_render_batched(b, nb, lo, med, hi, &v, 1, fb, levels);
    orc_clear64(fb2, npix);
    orc_render_batched_mt(b, nb, lo, med, hi, &v, 1, fb2, 4);
    CHECK(memcmp(fb, fb2, 8 * npix) == 0);
    uint32_t* img = malloc(4 * npix);
    orc_resolve_index(fb, npix, rgba, n, 0x00443322u, img);

    orc_clear64(fb, npix);
    orc_render_points16(xyz, rgba, n, &v, fb);
    orc_resolve_color(fb, npix, 0xFF000000u, img);

    uint32_t* depth = malloc(4 * npix);
    uint32_t* acc = calloc(4 * npix, 4);
    for (uint64_t i = 0; i < npix; i++) depth[i] = 0xFFFFFFFFu;
    orc_hqs_batched(b, nb, lo, med, hi, rgba, &v, 1, depth, acc);
    orc_hqs_resolve(acc, npix, 0x00443322u, img);

    /* a view that sees nothing */
    const double away[3] = {150, 150, 5000};
    orc_look_at_view(eye, away, up, 60.0, 1.0, 10.0, org, 160, 90, &v);
    orc_clear64(fb, npix);
    orc_render_batched(b, nb, lo, med, hi, &v, 1, fb, levels);
    for (uint64_t i = 0; i < npix; i++) CHECK(fb[i] == PCR_FB_CLEAR);

    /* LOD over a two-level hierarchy made from the batches */
    for (uint32_t k = 0; k < nb; k++) { b[k].parent = k ? 0 : -1; b[k].level = k ? 1 : 0; b[k].spacing = k ? 0.5f : 1.0f; }
    uint8_t* flags = malloc(nb);
    orc_look_at_view(eye, tgt, up, 60.0, 1.0, 2000.0, org, 160, 90, &v);
    orc_lod_select(b, nb, &v, 1, PCR_LOD_SPACING, 1.0f, flags);
    for (uint32_t k = 1; k < nb; k++) CHECK(!flags[k] || flags[0]);
    orc_clear64(fb, npix);
    orc_render_nodes(b, nb, lo, med, hi, &v, 1, PCR_LOD_SIZE80, 1.0f, fb, flags);

    free(xyz); free(rgba); free(perm); free(keys); free(sorted); free(b); free(lo); free(med); free(hi);
    free(fb); free(fb2); free(levels); free(img); free(depth); free(acc); free(flags);
    printf("oracle sanitize ok\n");
    return 0;
}
