/*
 * pcr.h — C ABI of the MI355X point-cloud rasterizer (libpcr.so).
 *
 * This is the drop-in boundary for the reference's per-frame Method plugin
 * (include/Method.h:10-24: `update(Renderer*)` / `render(Renderer*)` over a
 * loaded point buffer into a framebuffer). Every entry point takes plain
 * pointers and sizes; there are no GL, torch or C++ types in the signatures.
 *
 * Conventions
 *   - Every function returns PCR_OK (0) or a negative PCR_E_* code; the text
 *     of the last error on the calling thread is in pcr_last_error().
 *   - Matrices are float[16], column-major (glm layout): m[col*4 + row].
 *   - Framebuffers are row-major, pixel (x, y) at index x + y*width.
 *   - Host pointers are named *_host; device pointers returned by the library
 *     are named *_dev. Work is queued on the context's HIP stream and is
 *     asynchronous unless the function says it synchronises.
 *   - Thread safety: one context per host thread, except the pcr_server_*
 *     family, whose submit/poll calls may come from any thread.
 */
#ifndef PCR_H
#define PCR_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCR_VERSION_MAJOR 1
#define PCR_VERSION_MINOR 0

/* ---- status codes ------------------------------------------------------ */
#define PCR_OK               0
#define PCR_E_INVALID      (-1)  /* bad argument (null, size, alignment, mode) */
#define PCR_E_NO_DEVICE    (-2)  /* HIP runtime has no usable GPU            */
#define PCR_E_HIP          (-3)  /* a HIP runtime call failed                 */
#define PCR_E_OOM          (-4)  /* device allocation failed                 */
#define PCR_E_BUDGET       (-5)  /* request refused by the framebuffer budget */
#define PCR_E_NOT_READY    (-6)  /* server ticket still in flight             */
#define PCR_E_CANCELLED    (-7)  /* server ticket cancelled                   */

/* ---- fixed layout constants (modules/compute/Resources.h:4-6) --------- */
#define PCR_POINTS_PER_BATCH      10240u   /* 80 points x 128 threads       */
#define PCR_FB_CLEAR        0xFFFFFFFFFFFFFFFFull
#define PCR_DEPTH_CLEAR     0xFFFFFFFFu
#define PCR_BG_LAS          0x00443322u    /* compute_loop_las/resolve.cs:141 */
#define PCR_BG_BLACK        0xFF000000u
#define PCR_MAX_DIM         32768          /* framebuffer and view width/height limit */

/*
 * 64-byte batch / node header.
 * Replaces `struct Batch` of compute_loop_las/render.cs:22-40 (written by
 * LasLoaderSparse.cpp:110-121) and the node variant of
 * compute_loop_nodes/render.cs:26-43 (`pointOffset`@32, `level`@36).
 * min/max are the float AABB of the batch's points in cloud coordinates
 * (relative to the cloud origin). first_point may be any index (LAS files
 * concatenated in one cloud start at arbitrary offsets); the kernels read
 * 128-bit quads from the aligned word below it.
 */
typedef struct pcr_batch {
    int32_t  state;        /* @0  unused, 0                                   */
    float    min[3];       /* @4                                              */
    float    max[3];       /* @16                                             */
    uint32_t num_points;   /* @28                                             */
    uint32_t first_point;  /* @32 index of the batch's first point            */
    uint32_t level;        /* @36 octree level (LOD nodes), else file index 0 */
    int32_t  parent;       /* @40 parent node index (LOD nodes), -1 = root    */
    float    spacing;      /* @44 node point spacing (LOD nodes)              */
    uint32_t pad[4];       /* @48 .. 64                                       */
} pcr_batch;

/*
 * One camera / render target. Replaces the UBO and per-file SSBO that
 * ComputeLoopLas::render fills (compute_loop_las.h:141-199): `transform` is
 * float(proj*view*world) computed in double, `transform_frustum` the matrix
 * the frustum planes are taken from (normally equal to transform),
 * `view_world` = float(view*world) in double and `proj` the float projection,
 * both used by the precision select (render.cs:199-231).
 */
typedef struct pcr_view {
    float   transform[16];
    float   transform_frustum[16];
    float   view_world[16];
    float   proj[16];
    int32_t width;
    int32_t height;
    int32_t pad[2];
} pcr_view;

/* render modes: which reference Method the call stands in for */
typedef enum pcr_mode {
    PCR_MODE_PLAIN16  = 0,  /* 16 B/pt points, colour payload (compute_2021_earlyz)         */
    PCR_MODE_LAS      = 1,  /* batched adaptive precision, index payload (compute_loop_las) */
    PCR_MODE_LAS2     = 2,  /* same, 4-consecutive-point vector loads  (compute_loop_las2)  */
    PCR_MODE_HQS      = 3,  /* two-pass depth + accumulate            (compute_loop_las_hqs)*/
    PCR_MODE_NODES    = 4,  /* LOD octree nodes                       (compute_loop_nodes)  */
    PCR_MODE_HQS16    = 5,  /* hqs over the 16 B/pt layout                                 */
    PCR_MODE_NODES_HQS = 6  /* hqs over the LOD node cut, window-weighted resolve
                               (compute_loop_nodes_hqs)                                    */
} pcr_mode;

/* LOD criteria for PCR_MODE_NODES */
typedef enum pcr_lod_rule {
    PCR_LOD_SPACING   = 0,  /* refine a node while its projected spacing > lod_px      */
    PCR_LOD_SIZE80    = 1   /* reference: skip nodes whose projected size < 80 px      */
} pcr_lod_rule;

typedef struct pcr_render_params {
    int32_t  mode;              /* pcr_mode                                        */
    int32_t  frustum_culling;   /* uniforms.enableFrustumCulling (render.cs:250)   */
    int32_t  lod_rule;          /* pcr_lod_rule (PCR_MODE_NODES only)              */
    float    lod_px;            /* LOD screen-space error threshold in pixels      */
    uint32_t background;        /* RGBA8 written where nothing was drawn           */
    int32_t  block_size;        /* raster workgroup size: 0 = default (256)        */
    int32_t  clear_before;      /* 1: clear the framebuffer before rasterising     */
    int32_t  time_stages;       /* 1: record HIP events around each stage; 2:
                                   around the raster stage only (two events a
                                   frame instead of six; see pcr_frame_times)    */
    int32_t  tile_px;           /* LAS2/NODES/HQS/NODES_HQS: batches whose screen
                                   rectangle fits in this many pixels merge in an
                                   LDS tile first (0 = off, else 1024/2048/4096);
                                   same result. For hqs it also merges the colour
                                   sums, which is what makes dense views fast    */
    uint32_t batch_begin;       /* draw batches [batch_begin, batch_end) only;     */
    uint32_t batch_end;         /* 0, 0 = all, so an empty range cannot start at
                                   batch 0: leave such a draw out. One LAS file of
                                   a multi-file cloud is drawn with its own view
                                   (compute_loop_las.h:141-199 gives each file its
                                   own transform)                                 */
    int32_t  resolve_window;    /* PCR_MODE_NODES_HQS: neighbourhood radius of the
                                   resolve, 0..4 (0 in nodes_hqs/resolve.cs:98)    */
    float    edl_strength;      /* LAS/LAS2/NODES resolve: eye-dome lighting, the
                                   shading that compute_loop_las/resolve.cs:
                                   148-191 carries switched off (its constant is
                                   0.0005); 0 = off                              */
    int32_t  occlusion_skip;    /* tiled LAS2/NODES/HQS/NODES_HQS raster: 0 = on
                                   (default), -1 = off. A batch whose nearest
                                   possible depth lies behind every depth already
                                   drawn over its screen rectangle is skipped;
                                   the frame is the same bit for bit either way */
    int32_t  debug_counters;    /* 1: the index resolve also counts the covered
                                   pixels (pcr_stats.pixels_visible), as the
                                   reference's resolve does with debug.enabled
                                   (resolve.cs:143-145); 0 = off              */
} pcr_render_params;

/* Statistics of the last frame rendered on a context; the counts follow the
 * reference's debug counters (compute_loop_las/render.cs:73-81, read back at
 * compute_loop_las.h:260-277): #nodes processed = batches_total, #nodes
 * rendered = batches_total - batches_culled, #points processed =
 * points_drawn, #points rendered = atomics_issued, #points visible =
 * pixels_visible. */
typedef struct pcr_stats {
    uint64_t batches_total;
    uint64_t batches_culled;       /* frustum-culled or LOD-skipped               */
    uint64_t batches_per_level[5]; /* precision level histogram of drawn batches  */
    uint64_t points_drawn;         /* points in non-culled batches                */
    uint64_t bytes_algorithmic;    /* headers + planes read by the raster stage   */
    float    ms_cull;              /* stage times of the last timed frame          */
    float    ms_raster;
    float    ms_resolve;
    float    ms_clear;
    uint64_t batches_occluded;     /* tiled LAS2/NODES/HQS/NODES_HQS draws skipped
                                      because every pixel of the batch's screen
                                      rectangle already held a nearer point (hqs:
                                      both passes, a batch per pass)            */
    uint64_t points_occluded;      /* the points of those batches                 */
    uint64_t bytes_occluded;       /* the plane bytes (hqs pass 2: and colour
                                      bytes) those batches did not read; the
                                      raster's bytes are bytes_algorithmic minus
                                      these (hqs: twice the planes plus the colour
                                      of the drawn points, minus these)          */
    uint64_t atomics_issued;       /* 64-bit framebuffer atomicMins issued (LAS,
                                      LAS2, NODES): those whose compare load found
                                      a larger key                                */
    uint64_t pixels_visible;       /* pixels holding a point after the index
                                      resolve (LAS, LAS2, NODES) with
                                      params.debug_counters = 1, else 0          */
} pcr_stats;

typedef struct pcr_ctx   pcr_ctx;
typedef struct pcr_cloud pcr_cloud;
typedef struct pcr_fb    pcr_fb;

/* ---- context ---------------------------------------------------------- */
const char* pcr_last_error(void);
int pcr_version(void);                         /* major*100 + minor */
int pcr_device_count(int* count);
int pcr_create(int device, pcr_ctx** out);
int pcr_destroy(pcr_ctx* ctx);
int pcr_synchronize(pcr_ctx* ctx);
void* pcr_stream(pcr_ctx* ctx);                /* the context's hipStream_t */

/* ---- host-side helpers (no device work) ------------------------------- */
/* Build a pcr_view for a look-at perspective camera; all matrix algebra in
 * double, glm lookAtRH / perspectiveRH_NO conventions (Camera.h:34-38,
 * matrix_transform.inl:258-274), world = translate(origin) as in
 * compute_loop_las.h:170-176. fovy in degrees. */
int pcr_view_look_at(const double eye[3], const double target[3], const double up[3],
                     double fovy_deg, double near_plane, double far_plane,
                     const double origin[3], int width, int height, pcr_view* out);

/* Build a pcr_view from the renderer's double view/projection matrices
 * (Renderer.h:78-82 View::view/proj) and the cloud origin, exactly as
 * ComputeLoopLas::render fills its per-file SSBO (compute_loop_las.h:166-199):
 * transform = float(proj*view*translate(origin)) in double. */
int pcr_view_from_matrices(const double view[16], const double proj[16], const double origin[3],
                           int width, int height, pcr_view* out);

/* Encode points into the adaptive-precision batch layout on the host
 * (LasLoaderSparse.cpp:37-212). Points are taken in the given order; batches
 * of `batch_size` points; planes lo/med/hi hold bits 29..20 / 19..10 / 9..0. */
int pcr_encode_batches_host(const float* xyz_host, uint64_t n, uint32_t batch_size,
                            pcr_batch* batches_host, uint32_t* lo_host,
                            uint32_t* med_host, uint32_t* hi_host);

/* ---- clouds ------------------------------------------------------------ */
/* 16 B/pt layout: xyz float + RGBA8 (LasLoaderStandard.h:152-178). */
int pcr_cloud_upload_points16(pcr_ctx* ctx, const float* xyz_host, const uint32_t* rgba_host,
                              uint64_t n, pcr_cloud** out);
/* Adaptive-precision layout already encoded by the caller (the output of the
 * reference's LasLoaderSparse, modules/compute/LasLoaderSparse.h:46-99). */
int pcr_cloud_upload_batched(pcr_ctx* ctx, const pcr_batch* batches_host, uint32_t nbatches,
                             const uint32_t* lo_host, const uint32_t* med_host,
                             const uint32_t* hi_host, const uint32_t* rgba_host,
                             uint64_t npoints, pcr_cloud** out);

/* Synthetic scene generators, evaluated on the device from a counter-based
 * PRNG so that every point depends only on (seed, index). */
typedef enum pcr_scene {
    PCR_SCENE_CUBE    = 0,   /* xyz ~ U[0,1)^3, rgba = uint8(255*xyz), A=255       */
    PCR_SCENE_TERRAIN = 1,   /* fBm heightfield terrain                            */
    PCR_SCENE_CITY    = 2    /* box buildings on fBm terrain                       */
} pcr_scene;

typedef struct pcr_gen_cfg {
    int32_t  scene;          /* pcr_scene                                        */
    int32_t  num_buildings;  /* CITY                                             */
    uint64_t seed;
    uint64_t num_points;
    float    extent;         /* TERRAIN/CITY: side of the square area (m)        */
    float    amplitude;      /* TERRAIN/CITY: fBm amplitude (m)                  */
    float    jitter;         /* TERRAIN: +-jitter (m)                            */
    float    terrain_frac;   /* CITY: fraction of points on the terrain          */
} pcr_gen_cfg;

/* Ground height of a TERRAIN or CITY scene at (x, y): the fBm heightfield
 * the generator puts the terrain points on (before their jitter), evaluated
 * on the host with the generator's own operations. For cameras that keep a
 * height above the ground. */
int pcr_scene_ground_height(const pcr_gen_cfg* cfg, float x, float y, float* height);

/* Generate points on the device: xyz float[3n] + rgba uint32[n], unsorted. */
int pcr_generate_points(pcr_ctx* ctx, const pcr_gen_cfg* cfg, float* xyz_dev, uint32_t* rgba_dev);

/* Build options for the device-side layout builder. */
typedef struct pcr_build_opts {
    uint32_t batch_size;     /* points per batch, multiple of 4; 0 = 10240       */
    int32_t  morton_sort;    /* 1: sort by 63-bit Morton code before batching    */
    int32_t  keep_points16;  /* 1: also keep the sorted 16 B/pt copy             */
} pcr_build_opts;

/* Morton-sort, batch and encode points that are already on the device. */
int pcr_cloud_build_device(pcr_ctx* ctx, const float* xyz_dev, const uint32_t* rgba_dev,
                           uint64_t n, const pcr_build_opts* opts, pcr_cloud** out);
/* Generate + build in one call (frees the temporaries). */
int pcr_cloud_generate(pcr_ctx* ctx, const pcr_gen_cfg* cfg, const pcr_build_opts* opts,
                       pcr_cloud** out);

typedef struct pcr_cloud_info {
    uint64_t num_points;
    uint32_t num_batches;
    uint32_t batch_size;
    int32_t  has_points16;
    int32_t  has_batches;
    int32_t  num_levels;     /* LOD octree depth (0 = no octree)                 */
    float    aabb_min[3];
    float    aabb_max[3];
    uint64_t device_bytes;
} pcr_cloud_info;

int pcr_cloud_info_get(const pcr_cloud* cloud, pcr_cloud_info* out);
/* Copy the layout back to host arrays (any pointer may be NULL to skip). */
int pcr_cloud_download(pcr_cloud* cloud, pcr_batch* batches_host, uint32_t* lo_host,
                       uint32_t* med_host, uint32_t* hi_host, uint32_t* rgba_host,
                       float* xyz16_host, uint32_t* rgba16_host);
int pcr_cloud_destroy(pcr_cloud* cloud);

/* ---- LAS ingest -------------------------------------------------------- */
/* Header fields the reference reads (LasLoaderSparse.cpp:495-526). */
typedef struct pcr_las_info {
    uint64_t num_points;        /* @107 (u32) before LAS 1.4, @247 (u64) from 1.4;
                                   capped at 1e9 as the reference does           */
    uint32_t offset_to_points;  /* @96                                            */
    uint16_t record_length;     /* @105 bytes per point record                    */
    uint8_t  point_format;      /* @104 % 128 (the LAZ compression bit cleared)   */
    uint8_t  version_major;     /* @24                                            */
    uint8_t  version_minor;     /* @25                                            */
    uint8_t  rgb_offset;        /* R,G,B u16 offset in a record: 20 (format 2),
                                   28 (3), 30 (7, 8), else 0 (LasLoaderSparse.cpp:
                                   123-132 reads the first six bytes then)       */
    uint8_t  compressed;        /* 1: the point data is LAZ (LASzip VLR present,
                                   @104 bit 7 or 6 set); the record size check
                                   is then left to pcr_laz_info_get             */
    uint8_t  pad[5];
    double   scale[3];          /* @131                                           */
    double   offset[3];         /* @155                                           */
    double   min[3];            /* @187, @203, @219                               */
    double   max[3];            /* @179, @195, @211                               */
} pcr_las_info;

typedef struct pcr_las_file {
    const void* bytes;          /* the whole file: header, VLRs, point records   */
    uint64_t    size;
} pcr_las_file;

/* Parse a LAS 1.0-1.4 header (host only; no device work). */
int pcr_las_parse_header(const void* bytes, uint64_t size, pcr_las_info* out);

/* ---- LAZ (LASzip-compressed LAS) ---------------------------------------- */
/* The reference reads .laz through its vendored LASzip (libs/laszip):
 * laszip_open_reader + laszip_seek_point + laszip_read_point per point
 * (LasLoaderSparse.cpp:246-300). These replace that reader with this
 * library's own decoder of the format: compressor 2 (pointwise chunked,
 * items POINT10 / GPSTIME11 / RGB12 / BYTE version 2: point formats 0-3
 * with extra bytes) and compressor 3 (layered chunked, items POINT14 / RGB14 /
 * RGBNIR14 / BYTE14 version 3 or 4: formats 6-8 with extra bytes). Waveform
 * items (formats 4, 5, 9, 10) and the LASzip 1.x pointwise coder are refused. */
typedef struct pcr_laz_info {
    uint64_t num_points;        /* header count (legacy or LAS 1.4 field)         */
    uint32_t record_length;     /* bytes per decoded record (= the header's)      */
    uint32_t num_chunks;        /* chunks in the chunk table (0: none, decoded in
                                   order on one thread)                           */
    uint32_t chunk_size;        /* points per chunk, 0xFFFFFFFF = variable        */
    uint16_t compressor;        /* 2 pointwise chunked, 3 layered chunked         */
    uint16_t num_items;
    uint16_t item_type[8];      /* LASzip item types: 6 POINT10, 7 GPSTIME11,
                                   8 RGB12, 0 BYTE, 10 POINT14, 11 RGB14,
                                   12 RGBNIR14, 14 BYTE14                         */
    uint16_t item_size[8];
    uint16_t item_version[8];
    uint8_t  point_format;      /* @104 with the compression bits cleared         */
    uint8_t  has_rgb;           /* the records carry R,G,B                        */
    uint8_t  rgb_offset;        /* their byte offset in a record (0 if none)      */
    uint8_t  pad[5];
} pcr_laz_info;

#define PCR_LAZ_ALL_FIELDS 0u   /* decode every field of every record            */
#define PCR_LAZ_XYZ_RGB    1u   /* layered files: decode only the X/Y, Z and RGB
                                   layers; the other fields of a record keep the
                                   value of its chunk's first record (pointwise
                                   files decode everything either way)           */

/* Header, LASzip VLR and chunk table of a LAZ file (host only). */
int pcr_laz_info_get(const void* bytes, uint64_t size, pcr_laz_info* out);
/* The chunk table: chunk c holds records [chunk_first[c], chunk_first[c+1])
 * in bytes [chunk_start[c], chunk_start[c+1]) of the file (num_chunks + 1
 * entries; num_chunks = 0 when the file has no table; with capacity 0 and
 * null arrays only the count is returned). For loaders that
 * schedule their own chunk ranges, as LASreadPoint::seek finds chunks. */
int pcr_laz_chunk_table(const void* bytes, uint64_t size, uint64_t* chunk_start, uint64_t* chunk_first,
                        uint32_t capacity, uint32_t* num_chunks);
/* Decode records [first_point, first_point + num_points) of a LAZ file into
 * `records` (num_points * record_length bytes, the uncompressed LAS records)
 * on `threads` host threads (<= 0: up to 16), one chunk per thread at a time.
 * A damaged chunk fails the call with PCR_E_INVALID. Host only. */
int pcr_laz_decode(const void* bytes, uint64_t size, uint64_t first_point, uint64_t num_points, void* records,
                   uint32_t fields, int threads);
/* One cloud from nfiles LAS or LAZ files, encoded on the device as
 * LasLoaderSparse::loadLas does (LasLoaderSparse.cpp:37-212; LAZ files are
 * first decoded on the host with pcr_laz_decode, and then follow loadLaz,
 * :214-330, which differs only in colour: a LAZ file without R,G,B gets
 * colour 0, where loadLas reads the record's first six bytes): file k's points
 * follow file k-1's in the planes (the sparse offset is the running point
 * count), its batches cover 10,240 consecutive points each, positions are
 * X*scale + offset - header min in double, each batch AABB and 30-bit
 * quantisation are computed in double, the header stores the float AABB,
 * first_point = the sparse index and @36 (level) = k, and colour is
 * R|G<<8|B<<16 with 16-bit channels v < 256 ? v : v / 256 (alpha 0).
 * first_batch[k] receives file k's first batch (nfiles + 1 entries); draw
 * file k with batch_begin/batch_end = first_batch[k], first_batch[k+1] and a
 * view whose origin is file k's header min. */
int pcr_cloud_from_las(pcr_ctx* ctx, const pcr_las_file* files, int nfiles, uint32_t* first_batch,
                       pcr_cloud** out);

/* ---- Potree 2.0 ingest -------------------------------------------------- */
/* The metadata.json fields the reader uses (PotreeData.h:144-182); the caller
 * parses the JSON. Only the "DEFAULT" (uncompressed) encoding is read. */
typedef struct pcr_potree_info {
    double   scale[3];          /* "scale"                                        */
    double   offset[3];         /* "offset"                                       */
    double   min[3];            /* "boundingBox"."min"                            */
    double   max[3];            /* "boundingBox"."max"                            */
    double   spacing;           /* "spacing" (root node)                          */
    uint64_t num_points;        /* "points"                                       */
    uint64_t first_chunk_size;  /* "hierarchy"."firstChunkSize"                   */
    uint32_t bytes_per_point;   /* sum of the attribute sizes                     */
    uint32_t rgb_offset;        /* byte offset of the "rgb" attribute             */
} pcr_potree_info;

/* One LOD cloud from a Potree 2.0 hierarchy.bin and octree.bin (host bytes).
 * The hierarchy is read as PotreeData::parseHierarchy does
 * (PotreeData.h:185-248: 22-byte node records, proxy nodes of type 2 pointing
 * at further chunks, child boxes halved in float). Nodes become batches in
 * octree.bin order (PotreeData.h:276-278), with first_point = byteOffset /
 * bytes_per_point, @36 = depth, parent = the parent's batch index and spacing =
 * root spacing / 2^depth. Points are encoded on the device against their
 * node's box (PotreeData.h:439-511), with the clamps of
 * LasLoaderSparse.cpp:150-152 in place of the reference's 30-bit wrap, and
 * the 10-bit fine plane from bits 9..0. */
int pcr_cloud_from_potree(pcr_ctx* ctx, const pcr_potree_info* info, const void* hierarchy,
                          uint64_t hierarchy_size, const void* octree, uint64_t octree_size, pcr_cloud** out);

/* ---- LOD octree (Potree-style nodes) ----------------------------------- */
typedef struct pcr_octree_opts {
    uint64_t seed;            /* priority PRNG seed for the subsampling          */
    float    root_spacing_div;/* root spacing = cube side / this (128)           */
    uint32_t max_node_points; /* node chunk size limit (20000)                   */
    int32_t  max_depth;       /* deepest level                                    */
} pcr_octree_opts;

/* Build a Potree-style octree from points on the host (PotreeData.h:188-317
 * stands in as the node layout). Nodes become pcr_batch headers with
 * level/parent/spacing set; node chunks are stored in node order. A node that
 * holds more than max_node_points of the points no coarser level took keeps a
 * seeded Poisson-disk sample of them (no two closer than cube side /
 * (root_spacing_div * 2^level), maximal) and passes the rest down; a smaller
 * node keeps them all. */
int pcr_cloud_build_octree_host(pcr_ctx* ctx, const float* xyz_host, const uint32_t* rgba_host,
                                uint64_t n, const pcr_octree_opts* opts, pcr_cloud** out);
/* Same octree, byte for byte, built on the device from points already there
 * (Morton sort, per-level Poisson-disk sampling and encoding on the device; only the
 * per-chunk table visits the host). n < 2^31. */
int pcr_octree_generate(pcr_ctx* ctx, const pcr_gen_cfg* cfg, const pcr_octree_opts* opts, pcr_cloud** out);
int pcr_cloud_build_octree_device(pcr_ctx* ctx, const float* xyz_dev, const uint32_t* rgba_dev,
                                  uint64_t n, const pcr_octree_opts* opts, pcr_cloud** out);
/* Host only (no device work): the same octree as a node array plus planes, for
 * pcr_cloud_upload_batched. Call with null arrays to get *nnodes and *nslots
 * (plane words incl. 4-alignment padding), then with arrays of those sizes. */
int pcr_octree_layout_host(const float* xyz_host, const uint32_t* rgba_host, uint64_t n,
                           const pcr_octree_opts* opts, pcr_batch* nodes_host, uint32_t* nnodes,
                           uint32_t* lo_host, uint32_t* med_host, uint32_t* hi_host,
                           uint32_t* rgba_out_host, uint64_t* nslots);
/* The selected node set of the last PCR_MODE_NODES render: flags[i] = 1 if drawn. */
int pcr_lod_selection(pcr_ctx* ctx, pcr_cloud* cloud, uint8_t* flags_host, uint32_t n);

/* ---- framebuffers ------------------------------------------------------ */
typedef enum pcr_fb_kind {
    PCR_FB_KEY64 = 0,   /* uint64 depth|payload keys            */
    PCR_FB_HQS   = 1    /* uint32 depth + 4 x uint32 R,G,B,count */
} pcr_fb_kind;

int pcr_fb_create(pcr_ctx* ctx, int width, int height, int kind, pcr_fb** out);
int pcr_fb_destroy(pcr_fb* fb);
int pcr_fb_clear(pcr_ctx* ctx, pcr_fb* fb);
/* Read back (synchronises): keys uint64[w*h], or depth uint32[w*h] + acc uint32[4*w*h]. */
int pcr_fb_readback(pcr_ctx* ctx, pcr_fb* fb, uint64_t* keys_host);
int pcr_fb_readback_hqs(pcr_ctx* ctx, pcr_fb* fb, uint32_t* depth_host, uint32_t* acc_host);
int pcr_fb_readback_rgba(pcr_ctx* ctx, pcr_fb* fb, uint32_t* rgba_host);
void* pcr_fb_device_keys(pcr_fb* fb);
void* pcr_fb_device_rgba(pcr_fb* fb);

/* ---- per-frame render (Method::render, compute_loop_las.h:117-304) ------ */
/* Rasterise `cloud` for one view into `fb` (asynchronous). */
int pcr_render(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* view,
               const pcr_render_params* params, pcr_fb* fb);
/* Resolve fb to RGBA8 in the fb's own image (compute_loop_las/resolve.cs:66-195,
 * hqs/resolve.cs:62-129, compute_2021_earlyz/resolve.cs:41-78). */
int pcr_resolve(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_render_params* params, pcr_fb* fb);
int pcr_stats_get(pcr_ctx* ctx, pcr_stats* out);   /* synchronises */
/* Stage times of the last n frames rendered with time_stages != 0, oldest
 * first, five floats per frame: clear, cull, raster, resolve and the whole
 * frame from clear to the end of resolve (ms). Batched modes clear the
 * framebuffer inside the cull launch, so their clear time is about 0 and
 * their cull time includes the clear. A frame rendered with
 * time_stages = 2 has only the raster time; its clear, cull and resolve read
 * -1, and its whole-frame time is the period from the previous timed frame's
 * raster end to its own. Up to PCR_TIMING_RING frames are kept
 * (GLTimerQueries.cpp:55-153 kept 3 frames of GL timestamps for the same
 * purpose). Synchronises; *got = frames written. */
#define PCR_TIMING_RING 512
int pcr_frame_times(pcr_ctx* ctx, int n, float* out5, int* got);
/* Device check of the projection's division (no counterpart in the
 * reference; for tests): for i < n with den[i] > 0, out2[2i] and out2[2i+1]
 * receive num_x[i]/den[i] and num_y[i]/den[i] as the rasterizer computes
 * them, which must equal the IEEE quotients. Host arrays; synchronises. */
int pcr_selftest_divide(pcr_ctx* ctx, const float* num_x, const float* num_y, const float* den, uint64_t n,
                        float* out2);

/* ---- allocation-free render ---------------------------------------------- */
/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Renders one frame exactly as pcr_render (and, with
 * clear_before, the clear) does for modes LAS, LAS2 and NODES, over a batch
 * layout and a key framebuffer the caller already holds in device memory, on
 * the caller's stream of the current HIP device. Allocates nothing: the
 * frame's scratch (counters, the drawn-batch list and per-batch levels) is a
 * fixed device array of the library, one per GPU (about 1.3 MB, held until the
 * process ends), sized for PCR_STATIC_MAX_BATCHES batches; larger layouts are
 * refused with PCR_E_INVALID. Every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU race: serialise them (one stream, or events).
 *   batches_dev     nbatches headers (as pcr_cloud_upload_batched takes them)
 *   lo/med/hi_dev   position planes of plane_words words each: 16-byte
 *                   aligned, plane_words a multiple of 4 and at least every
 *                   batch's last point + 1 rounded up to 4
 *   keys_dev        view->width * view->height uint64 keys, 16-byte aligned
 *   params          time_stages must be 0; mode LAS, LAS2 or NODES
 *   stream          a hipStream_t of the current device (NULL = null stream) */
#define PCR_STATIC_MAX_BATCHES 262144u
int pcr_render_static(const pcr_batch* batches_dev, uint32_t nbatches, const uint32_t* lo_dev,
                      const uint32_t* med_dev, const uint32_t* hi_dev, uint64_t plane_words,
                      const pcr_view* view, const pcr_render_params* params, uint64_t* keys_dev,
                      void* stream);

/* ---- batched multi-view (persistent launch, hipGraph) ------------------- */
typedef struct pcr_tune {
    int32_t block_size;      /* threads per workgroup: 64,128,256,512            */
    int32_t points_per_thread;/* untiled: points per thread per work item (x4)    */
    int32_t persistent_groups;/* untiled: 0 = derived from the device             */
    int32_t tile_px;         /* 0 = untiled; 1024/2048/4096 = LDS-tiled raster,
                                one batch per item, block 128 or 256            */
} pcr_tune;

typedef struct pcr_multiview pcr_multiview;
/* Prepare one launch that renders `nviews` (1..64) views of `cloud` into `fbs`,
 * walking a (batch, chunk) work queue that reads each batch once for all the
 * views that draw it; the whole frame (clear, cull, raster, resolve) is
 * captured in a hipGraph when use_graph != 0. Modes LAS, LAS2 and NODES take
 * KEY64 framebuffers; HQS and NODES_HQS take HQS framebuffers and run both hqs
 * passes over all views, the shared-decode form of the two-eye VR methods
 * (compute_loop_las_hqs_vr/depth.cs:281-413, color.cs:294-513). Every view's
 * buffers equal a single-view pcr_render + pcr_resolve of that view. */
int pcr_multiview_create(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* views, int nviews,
                         pcr_fb* const* fbs, const pcr_render_params* params,
                         const pcr_tune* tune, int use_graph, pcr_multiview** out);
int pcr_multiview_set_views(pcr_multiview* mv, const pcr_view* views, int nviews);
int pcr_multiview_launch(pcr_multiview* mv);
/* Counts of the plan's last launch (synchronises). The algorithmic bytes are
 * the batch headers and position planes the raster needs: for the batched
 * form each kept batch once, at the most precise level any view needs; for
 * separate per-view renders the sum over the views that draw it. */
typedef struct pcr_multiview_stats {
    uint64_t batches_any_view;     /* batches drawn by at least one view        */
    uint64_t bytes_batched;        /* 64 + 4 x planes(most precise level) x n    */
    uint64_t bytes_per_view_sum;   /* sum over views of 64 + 4 x planes(level) x n */
    uint64_t points_per_view_sum;  /* points in drawn batches, summed over views */
    uint64_t points_any_view;      /* points in batches drawn by at least one view */
} pcr_multiview_stats;
int pcr_multiview_stats_get(pcr_multiview* mv, pcr_multiview_stats* out);
int pcr_multiview_destroy(pcr_multiview* mv);
/* Sweep block size / points per thread (untiled) and block size / tile size
 * (LDS-tiled) for this cloud, views and mode; writes the fastest
 * configuration and its time. hqs modes try one untiled configuration. */
int pcr_autotune(pcr_ctx* ctx, pcr_cloud* cloud, const pcr_view* views, int nviews,
                 pcr_fb* const* fbs, const pcr_render_params* params, int iters,
                 pcr_tune* best, float* best_ms);

/* ---- serving: continuous batching over device-resident clouds ----------- */
typedef struct pcr_server pcr_server;
typedef struct pcr_request {
    int32_t  cloud_id;       /* index returned by pcr_server_add_cloud          */
    int32_t  mode;           /* PCR_MODE_LAS or PCR_MODE_NODES                   */
    float    lod_px;
    int32_t  pad;
    pcr_view view;
} pcr_request;

typedef struct pcr_server_stats {
    uint64_t completed;
    uint64_t refused;
    uint64_t launches;
    uint64_t points_rendered;   /* sum over completed requests of cloud points */
    uint64_t fb_bytes_in_use;
    uint64_t fb_bytes_peak;
    uint64_t points_drawn;      /* points in the batches or LOD nodes the
                                   completed launches drew, summed over their
                                   requests                                   */
    uint64_t bytes_algorithmic; /* headers + planes those launches read (each
                                   batch once per launch, at the most precise
                                   level its requests need)                   */
} pcr_server_stats;

int pcr_server_create(pcr_ctx* ctx, uint64_t fb_budget_bytes, int max_batch, int num_streams,
                      pcr_server** out);
int pcr_server_add_cloud(pcr_server* srv, pcr_cloud* cloud, int* cloud_id);
/* Enqueue a request; returns a ticket. Refused with PCR_E_BUDGET only when the
 * request alone exceeds the budget; otherwise it waits for memory to free up. */
int pcr_server_submit(pcr_server* srv, const pcr_request* req, uint64_t* ticket);
/* Non-blocking: PCR_OK with the image copied to rgba_host (w*h uint32), or PCR_E_NOT_READY. */
int pcr_server_poll(pcr_server* srv, uint64_t ticket, uint32_t* rgba_host, double* latency_ms);
int pcr_server_wait(pcr_server* srv, uint64_t ticket, uint32_t* rgba_host, double* latency_ms);
/* Zero-copy form of pcr_server_wait (no counterpart in the reference; for
 * clients that consume images in place): waits, then returns a pointer to
 * the ticket's w*h RGBA8 image in the server's pinned result memory. The
 * image stays valid, and its framebuffer memory stays counted against the
 * budget, until pcr_server_release(ticket) (or pcr_server_cancel). */
int pcr_server_wait_pinned(pcr_server* srv, uint64_t ticket, const uint32_t** rgba_pinned, double* latency_ms);
int pcr_server_release(pcr_server* srv, uint64_t ticket);
int pcr_server_cancel(pcr_server* srv, uint64_t ticket);
int pcr_server_stats_get(pcr_server* srv, pcr_server_stats* out);
int pcr_server_destroy(pcr_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* PCR_H */
