/*
 * pcr_oracle.h — CPU oracle for the point-cloud rasterizer (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference's per-frame shader logic, used only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. Nothing in the product links or calls it.
 *
 * Numerics contract (shared by the HIP kernels, restated independently there):
 *   - IEEE fp32, no contraction (-ffp-contract=off); an FMA appears only where
 *     written as fmaf(); division and sqrt are correctly rounded.
 *   - mat4 * (x,y,z,1):  r_i = fmaf(M[8+i], z, fmaf(M[4+i], y, fmaf(M[i], x, M[12+i])))
 *   - mat4 * (x,y,z,w):  r_i = fmaf(M[12+i], w, fmaf(M[8+i], z, fmaf(M[4+i], y, M[i]*x)))
 *   - length(x,y,z) = sqrtf(fmaf(z,z, fmaf(y,y, x*x)));  length(x,y) = sqrtf(fmaf(y,y, x*x))
 * Every function cites the reference file:line it follows.
 */
#ifndef PCR_ORACLE_H
#define PCR_ORACLE_H
#include <stdint.h>
#include "../include/pcr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PRNG (counter based) -------------------------------------------- */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_rand(uint64_t key, uint64_t i);
float    orc_u01(uint64_t r);

/* ---- generators -------------------------------------------------------- */
void orc_generate(const pcr_gen_cfg* cfg, uint64_t first, uint64_t count, float* xyz, uint32_t* rgba);
float orc_terrain_height(uint64_t seed, float extent, float amplitude, float x, float y);

/* ---- layout ------------------------------------------------------------ */
uint64_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);
/* 63-bit Morton keys over the cloud's bounding cube; returns the permutation
 * that sorts them (stable by index). */
void orc_morton_order(const float* xyz, uint64_t n, uint32_t* perm_out, uint64_t* keys_out);
void orc_encode_batches(const float* xyz, uint64_t n, uint32_t batch_size,
                        pcr_batch* batches, uint32_t* lo, uint32_t* med, uint32_t* hi);
void orc_quantize30(double x, double mn, double size, uint32_t* q);

/* ---- LAS ingest (LasLoaderSparse.cpp:37-212, 495-526) ------------------ */
/* 0 on success, -1 for a missing signature or a short header/record range. */
int  orc_las_parse_header(const uint8_t* bytes, uint64_t size, pcr_las_info* out);
/* One file's records -> batches of batch_size consecutive points; lo/med/hi/
 * rgba point at the file's first slot, first_point = sparse_offset + i. */
void orc_las_encode(const uint8_t* records, uint64_t n, const pcr_las_info* info, uint32_t batch_size,
                    uint64_t sparse_offset, uint32_t file_index, pcr_batch* batches,
                    uint32_t* lo, uint32_t* med, uint32_t* hi, uint32_t* rgba);

/* ---- LOD + hqs (compute_loop_nodes_hqs) --------------------------------- */
void orc_hqs_nodes(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                   const uint32_t* hi, const uint32_t* rgba, const pcr_view* v, int culling, int rule,
                   float lod_px, uint32_t* depth, uint32_t* acc);
void orc_hqs_resolve_window(const uint32_t* depth, const uint32_t* acc, int w, int h, int window, uint32_t* out);

/* ---- Potree 2.0 ingest (PotreeData.h:144-317, 439-511) ----------------- */
/* Nodes of hierarchy.bin as batches in octree.bin order, with each node's
 * double box (min xyz, max xyz) in boxes[6*i]. Returns the node count, or -1
 * for a malformed hierarchy or more than max_nodes nodes. */
int64_t orc_potree_nodes(const pcr_potree_info* info, const uint8_t* hier, uint64_t hsize,
                         pcr_batch* nodes, double* boxes, uint64_t max_nodes);
/* Encode every node's points (octree.bin records) into the planes. */
void orc_potree_encode(const pcr_potree_info* info, const uint8_t* octree, const pcr_batch* nodes,
                       const double* boxes, uint64_t nn, uint32_t* lo, uint32_t* med, uint32_t* hi,
                       uint32_t* rgba);

/* ---- per-batch stages --------------------------------------------------- */
int   orc_frustum_visible(const float tf[16], const float mn[3], const float mx[3]);
float orc_projected_px(const float mn[3], const float mx[3], float len,
                       const float vw[16], const float proj[16], int w, int h);
int   orc_precision_level(const float mn[3], const float mx[3], const pcr_view* v, float first_threshold);
void  orc_decode(int level, uint32_t lo, uint32_t med, uint32_t hi,
                 const float mn[3], const float mx[3], float out[3]);
void  orc_decode_n(int level, uint64_t n, const uint32_t* lo, const uint32_t* med, const uint32_t* hi,
                   const float mn[3], const float mx[3], float* out);
/* project a point: returns 1 and pixel index/depth bits if it lands inside */
int   orc_project(const float t[16], const float p[3], int w, int h, uint32_t* pixel, uint32_t* depth_bits);

/* ---- full frames (single thread) --------------------------------------- */
void orc_clear64(uint64_t* fb, uint64_t npix);
void orc_render_points16(const float* xyz, const uint32_t* rgba, uint64_t n,
                         const pcr_view* v, uint64_t* fb);
/* levels_out (optional): per batch level 0..4, or 255 when culled/skipped */
void orc_render_batched(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                        const uint32_t* hi, const pcr_view* v, int frustum_culling,
                        uint64_t* fb, uint8_t* levels_out);
void orc_resolve_index(const uint64_t* fb, uint64_t npix, const uint32_t* rgba, uint64_t npoints, uint32_t bg,
                       uint32_t* out);
void orc_resolve_color(const uint64_t* fb, uint64_t npix, uint32_t bg, uint32_t* out);
/* eye-dome lighting resolve (compute_loop_las/resolve.cs:148-191, switched on) */
float orc_edl_exp(float x);
void orc_resolve_edl(const uint64_t* fb, int w, int h, const uint32_t* rgba, uint64_t npoints, uint32_t bg,
                     float strength, uint32_t* out);

/* hqs: pass 1 depth, pass 2 accumulate, resolve.  src16 != 0 selects the
 * 16 B/pt point source (xyz/rgba) instead of the batched one. */
void orc_hqs_batched(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                     const uint32_t* hi, const uint32_t* rgba, const pcr_view* v,
                     int frustum_culling, uint32_t* depth, uint32_t* acc);
void orc_hqs_points16(const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_view* v,
                      uint32_t* depth, uint32_t* acc);
void orc_hqs_resolve(const uint32_t* acc, uint64_t npix, uint32_t bg, uint32_t* out);

/* LOD: node selection (flags[i]=1 if drawn) and the node render */
void orc_lod_select(const pcr_batch* nodes, uint32_t nn, const pcr_view* v, int frustum_culling,
                    int rule, float lod_px, uint8_t* flags);
void orc_render_nodes(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                      const uint32_t* hi, const pcr_view* v, int frustum_culling, int rule,
                      float lod_px, uint64_t* fb, uint8_t* flags_out);

/* ---- multithreaded frame (the timed CPU baseline) ---------------------- */
void orc_render_batched_mt(const pcr_batch* b, uint32_t nb, const uint32_t* lo, const uint32_t* med,
                           const uint32_t* hi, const pcr_view* v, int frustum_culling,
                           uint64_t* fb, int nthreads);
void orc_render_nodes_mt(const pcr_batch* nodes, uint32_t nn, const uint32_t* lo, const uint32_t* med,
                         const uint32_t* hi, const pcr_view* v, int culling, int rule, float lod_px, uint64_t* fb,
                         int nthreads);

/* ---- host camera (double precision, glm conventions) ------------------- */
void orc_look_at_view(const double eye[3], const double target[3], const double up[3],
                      double fovy_deg, double near_plane, double far_plane,
                      const double origin[3], int width, int height, pcr_view* out);

#ifdef __cplusplus
}
#endif
#endif
