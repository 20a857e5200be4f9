// pcr_ingest.hip — LAS and Potree 2.0 ingest on the device.
//
// Builds the batched adaptive-precision layout from uncompressed LAS point
// records the way LasLoaderSparse::loadLas does on the host
// (LasLoaderSparse.cpp:37-212, header fields :495-526):
//   k_las_batch   one workgroup per 10,240-point batch: positions
//                 X*scale + offset - header min in double, batch AABB in
//                 double (workgroup min/max), float header.
//   k_las_encode  one thread per point: 30-bit quantisation against the
//                 batch's double AABB, the three 10-bit planes, and the
//                 16-to-8-bit colour.
// Files are laid end to end in the planes (the reference's sparse buffers),
// so file k's batches start at the running point count, and @36 of each
// header holds the file index as in LasLoaderSparse.cpp:121.
//
// Potree 2.0 (PotreeData.h:144-511): the hierarchy is parsed on the host into
// node boxes, and k_potree_encode (one workgroup per node) encodes each
// node's octree.bin records against the node's box with the same per-record
// arithmetic.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>
#include "pcr_internal.h"

namespace pcr {
namespace {

struct LasParams {
    double scale[3], offset[3], min[3];
    uint32_t record_length, rgb_offset;
    uint32_t no_rgb;   // LAZ records without colour: colour 0 (LasLoaderSparse.cpp:283-285)
};

// Records are packed at any byte stride, so fields are assembled from bytes.
__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

// double(X) * scale + offset - min, evaluated left to right without
// contraction (LasLoaderSparse.cpp:97-99).
__device__ __forceinline__ double las_coord(const uint8_t* r, int a, const LasParams& p) {
    return (double)(int32_t)ld_u32(r + 4 * a) * p.scale[a] + p.offset[a] - p.min[a];
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

__global__ __launch_bounds__(256) void k_las_batch(const uint8_t* __restrict__ rec, uint64_t n, LasParams p,
                                                    uint32_t bs, uint64_t sparse, uint32_t file_index,
                                                    pcr_batch* __restrict__ out, double* __restrict__ box) {
    __shared__ double s_box[6][4];
    const uint64_t first = (uint64_t)blockIdx.x * bs;
    const uint64_t cnt = n - first < bs ? n - first : bs;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = threadIdx.x; i < cnt; i += 256) {
        const uint8_t* r = rec + (first + i) * p.record_length;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double x = las_coord(r, a, p);
            mn[a] = fmin(mn[a], x);
            mx[a] = fmax(mx[a], x);
        }
    }
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double lo = wave_min(mn[a]), hi = wave_max(mx[a]);
        if (lane == 0) { s_box[a][wave] = lo; s_box[3 + a][wave] = hi; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bmin[3], bmax[3];
        for (int a = 0; a < 3; a++) {
            bmin[a] = fmin(fmin(s_box[a][0], s_box[a][1]), fmin(s_box[a][2], s_box[a][3]));
            bmax[a] = fmax(fmax(s_box[3 + a][0], s_box[3 + a][1]), fmax(s_box[3 + a][2], s_box[3 + a][3]));
            box[6 * blockIdx.x + a] = bmin[a];
            box[6 * blockIdx.x + 3 + a] = bmax[a];
        }
        pcr_batch h;
        memset(&h, 0, sizeof(h));
        for (int a = 0; a < 3; a++) { h.min[a] = (float)bmin[a]; h.max[a] = (float)bmax[a]; }
        h.num_points = (uint32_t)cnt;
        h.first_point = (uint32_t)(sparse + first);
        h.level = file_index;
        h.parent = -1;
        out[blockIdx.x] = h;
    }
}

// LasLoaderSparse.cpp:146-152 with the zero-extent decision (size 0 -> 0).
// A position below the box (a Potree node's box is rounded to float) clamps
// to 0 rather than converting a negative double.
__device__ __forceinline__ uint32_t quantize30(double x, double mn, double size) {
    if (size == 0.0) return 0u;
    const double f = ((x - mn) / size) * 1073741824.0;
    if (!(f > 0.0)) return 0u;
    return f < 1073741823.0 ? (uint32_t)f : 1073741823u;
}

// One record -> the three plane words and the colour (LasLoaderSparse.cpp:
// 146-192, PotreeData.h:439-511); bx = box min xyz, max xyz.
__device__ __forceinline__ void encode_record(const uint8_t* r, const LasParams& p, const double* bx, uint64_t i,
                                              uint32_t* __restrict__ lo, uint32_t* __restrict__ med,
                                              uint32_t* __restrict__ hi, uint32_t* __restrict__ rgba) {
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) q[a] = quantize30(las_coord(r, a, p), bx[a], bx[3 + a] - bx[a]);
    lo[i] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
    med[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
    hi[i] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t v = p.no_rgb ? 0u : ld_u16(r + p.rgb_offset + 2 * k);
        c[k] = v < 256u ? v : v / 256u;
    }
    rgba[i] = c[0] | (c[1] << 8) | (c[2] << 16);
}

__global__ __launch_bounds__(256) void k_las_encode(const uint8_t* __restrict__ rec, uint64_t n, LasParams p,
                                                     uint32_t bs, const double* __restrict__ box,
                                                     uint32_t* __restrict__ lo, uint32_t* __restrict__ med,
                                                     uint32_t* __restrict__ hi, uint32_t* __restrict__ rgba) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        encode_record(rec + i * p.record_length, p, box + 6 * (i / bs), i, lo, med, hi, rgba);
}

// One workgroup per Potree node: its records sit at first_point.
__global__ __launch_bounds__(256) void k_potree_encode(const uint8_t* __restrict__ rec, LasParams p,
                                                        const pcr_batch* __restrict__ nodes,
                                                        const double* __restrict__ boxes, uint32_t* __restrict__ lo,
                                                        uint32_t* __restrict__ med, uint32_t* __restrict__ hi,
                                                        uint32_t* __restrict__ rgba) {
    const pcr_batch h = nodes[blockIdx.x];
    const double* bx = boxes + 6 * (uint64_t)blockIdx.x;
    for (uint32_t j = threadIdx.x; j < h.num_points; j += 256) {
        const uint64_t i = (uint64_t)h.first_point + j;
        encode_record(rec + i * p.record_length, p, bx, i, lo, med, hi, rgba);
    }
}

}  // namespace
}  // namespace pcr

using namespace pcr;

namespace {

// Decodes the first n records of a LAZ file into d_rec (device) in slabs of
// whole chunks: two pinned slabs, so that one is decoded while the other is
// copied. Returns a PCR_E_* code; a failed HIP call also sets *err.
int stage_laz(hipStream_t s, const pcr_las_file& file, const pcr_laz_info& L, uint64_t n, uint8_t* d_rec,
              hipError_t* err) {
    const uint64_t rl = L.record_length;
    uint64_t slab = 1ull << 22;   // points
    if (L.chunk_size != 0xFFFFFFFFu && L.chunk_size < slab) slab -= slab % L.chunk_size;
    slab = std::min(slab, n);
    uint8_t* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipError_t e = hipSuccess;
    int rc = PCR_OK;
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipHostMalloc(reinterpret_cast<void**>(&buf[i]), slab * rl, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
    }
    for (uint64_t s0 = 0, i = 0; s0 < n && e == hipSuccess && rc == PCR_OK; s0 += slab, i++) {
        const uint64_t cnt = std::min(slab, n - s0);
        uint8_t* b = buf[i & 1];
        if (i >= 2) e = hipEventSynchronize(done[i & 1]);   // its previous copy has finished
        if (e != hipSuccess) break;
        rc = pcr_laz_decode(file.bytes, file.size, s0, cnt, b, PCR_LAZ_XYZ_RGB, 0);
        if (rc) break;
        e = hipMemcpyAsync(d_rec + s0 * rl, b, cnt * rl, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i & 1], s);
    }
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    for (int i = 0; i < 2; i++) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (buf[i]) (void)hipHostFree(buf[i]);
    }
    *err = e;
    return e != hipSuccess ? PCR_E_HIP : rc;
}

}  // namespace

extern "C" {

int pcr_cloud_from_las(pcr_ctx* ctx, const pcr_las_file* files, int nfiles, uint32_t* first_batch,
                       pcr_cloud** out) {
    PCR_CHECK(ctx && files && out && nfiles > 0, "cloud_from_las: bad argument");
    const uint32_t bs = PCR_POINTS_PER_BATCH;
    std::vector<pcr_las_info> info(nfiles);
    std::vector<uint64_t> sparse(nfiles + 1, 0);
    std::vector<uint32_t> fb(nfiles + 1, 0);
    for (int k = 0; k < nfiles; k++) {
        const int rc = pcr_las_parse_header(files[k].bytes, files[k].size, &info[k]);
        if (rc) return rc;
        sparse[k + 1] = sparse[k] + info[k].num_points;
        fb[k + 1] = fb[k] + (uint32_t)((info[k].num_points + bs - 1) / bs);
    }
    const uint64_t n = sparse[nfiles];
    PCR_CHECK(n > 0 && n < (1ull << 32), "cloud_from_las: %llu points (1 .. 2^32-1)", (unsigned long long)n);
    pcr_cloud* c = new pcr_cloud();
    int rc = cloud_alloc_batched(ctx, fb[nfiles], n, c);
    if (rc) { delete c; return rc; }
    c->batch_size = bs;
    hipStream_t s = ctx->stream;
    uint8_t* d_rec = nullptr;
    double* d_box = nullptr;
    auto fail = [&](hipError_t e, const char* what) {
        (void)hipFree(d_rec);
        (void)hipFree(d_box);
        cloud_free(c);
        delete c;
        return hip_fail(e, what, __FILE__, __LINE__);
    };
    for (int k = 0; k < nfiles; k++) {
        const pcr_las_info& I = info[k];
        const uint64_t nk = I.num_points;
        if (nk == 0) continue;
        const uint32_t nbk = fb[k + 1] - fb[k];
        LasParams p;
        for (int a = 0; a < 3; a++) { p.scale[a] = I.scale[a]; p.offset[a] = I.offset[a]; p.min[a] = I.min[a]; }
        p.record_length = I.record_length;
        p.rgb_offset = I.rgb_offset;
        p.no_rgb = 0;
        pcr_laz_info L{};
        if (I.compressed) {
            if ((rc = pcr_laz_info_get(files[k].bytes, files[k].size, &L))) {
                cloud_free(c);
                delete c;
                return rc;
            }
            p.rgb_offset = L.rgb_offset;
            p.no_rgb = !L.has_rgb;
        }
        hipError_t e = hipMalloc(&d_rec, nk * I.record_length);
        if (e == hipSuccess) e = hipMalloc(&d_box, (size_t)nbk * 6 * sizeof(double));
        if (e == hipSuccess && !I.compressed)
            e = hipMemcpy(d_rec, static_cast<const uint8_t*>(files[k].bytes) + I.offset_to_points,
                          nk * I.record_length, hipMemcpyHostToDevice);
        if (e == hipSuccess && I.compressed) {
            rc = stage_laz(s, files[k], L, nk, d_rec, &e);
            if (rc && e == hipSuccess) {   // a decode error (its message is set)
                (void)hipFree(d_rec);
                (void)hipFree(d_box);
                cloud_free(c);
                delete c;
                return rc;
            }
        }
        if (e != hipSuccess) return fail(e, "cloud_from_las: staging the point records");
        k_las_batch<<<nbk, 256, 0, s>>>(d_rec, nk, p, bs, sparse[k], (uint32_t)k, c->d_batches + fb[k], d_box);
        const uint64_t g = (nk + 255) / 256;
        k_las_encode<<<(int)(g < 65536 ? g : 65536), 256, 0, s>>>(d_rec, nk, p, bs, d_box, c->d_lo + sparse[k],
                                                                  c->d_med + sparse[k], c->d_hi + sparse[k],
                                                                  c->d_rgba + sparse[k]);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail(e, "cloud_from_las: encoding");
        (void)hipFree(d_rec);
        (void)hipFree(d_box);
        d_rec = nullptr;
        d_box = nullptr;
    }
    c->h_batches.resize(fb[nfiles]);
    PCR_HIP(hipMemcpy(c->h_batches.data(), c->d_batches, sizeof(pcr_batch) * fb[nfiles], hipMemcpyDeviceToHost));
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = INFINITY; c->aabb_max[a] = -INFINITY; }
    for (const pcr_batch& h : c->h_batches)
        for (int a = 0; a < 3; a++) {
            c->aabb_min[a] = h.min[a] < c->aabb_min[a] ? h.min[a] : c->aabb_min[a];
            c->aabb_max[a] = h.max[a] > c->aabb_max[a] ? h.max[a] : c->aabb_max[a];
        }
    if (first_batch) memcpy(first_batch, fb.data(), sizeof(uint32_t) * (nfiles + 1));
    *out = c;
    return PCR_OK;
}

// PotreeData.h:250-278 (traverse the tree, sort the nodes by byteOffset),
// :362-380 (batch headers) and :419-511 (encode).
int pcr_cloud_from_potree(pcr_ctx* ctx, const pcr_potree_info* info, const void* hierarchy,
                          uint64_t hierarchy_size, const void* octree, uint64_t octree_size, pcr_cloud** out) {
    PCR_CHECK(ctx && info && hierarchy && octree && out, "cloud_from_potree: null argument");
    std::vector<pcr_batch> hb;
    std::vector<double> boxes;
    int max_level = 0;
    int rc = potree_layout(info, hierarchy, hierarchy_size, octree_size, hb, boxes, &max_level);
    if (rc) return rc;
    const uint32_t nn = (uint32_t)hb.size();
    const uint32_t bpp = info->bytes_per_point;
    pcr_cloud* c = new pcr_cloud();
    rc = cloud_alloc_batched(ctx, nn, info->num_points, c);
    if (rc) { delete c; return rc; }
    c->batch_size = PCR_POINTS_PER_BATCH;
    c->num_levels = max_level + 1;
    hipStream_t s = ctx->stream;
    uint8_t* d_rec = nullptr;
    double* d_box = nullptr;
    const uint64_t rec_bytes = info->num_points * bpp;
    hipError_t e = hipMalloc(&d_rec, rec_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_box, boxes.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(d_rec, octree, rec_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_box, boxes.data(), boxes.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_batches, hb.data(), sizeof(pcr_batch) * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        LasParams p;
        for (int a = 0; a < 3; a++) { p.scale[a] = info->scale[a]; p.offset[a] = info->offset[a]; p.min[a] = info->min[a]; }
        p.record_length = bpp;
        p.rgb_offset = info->rgb_offset;
        p.no_rgb = 0;
        k_potree_encode<<<nn, 256, 0, s>>>(d_rec, p, c->d_batches, d_box, c->d_lo, c->d_med, c->d_hi, c->d_rgba);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    (void)hipFree(d_rec);
    (void)hipFree(d_box);
    if (e != hipSuccess) {
        cloud_free(c);
        delete c;
        return hip_fail(e, "cloud_from_potree", __FILE__, __LINE__);
    }
    c->h_batches = hb;
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = INFINITY; c->aabb_max[a] = -INFINITY; }
    for (const pcr_batch& h : hb)
        for (int a = 0; a < 3; a++) {
            c->aabb_min[a] = h.min[a] < c->aabb_min[a] ? h.min[a] : c->aabb_min[a];
            c->aabb_max[a] = h.max[a] > c->aabb_max[a] ? h.max[a] : c->aabb_max[a];
        }
    *out = c;
    return PCR_OK;
}

}  // extern "C"
