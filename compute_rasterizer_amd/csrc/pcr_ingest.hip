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
        const uint32_t v = ld_u16(r + p.rgb_offset + 2 * k);
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

// ------------------------------------------------------- Potree 2.0 (host)
struct PotreeNode {
    double box[6] = {0, 0, 0, 0, 0, 0};
    float spacing = 0.0f;
    int level = 0, type = 0, parent = -1;
    int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    uint32_t num_points = 0;
    int64_t byte_offset = 0, byte_size = 0, hier_offset = 0, hier_size = 0;
};

// PotreeData.h:115-139: the child's box is the parent's halved in float;
// index bit 0 picks the z half, bit 1 the y half, bit 2 the x half.
void potree_child_box(const double* b, int index, double* out) {
    float mn[3], mx[3], size[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = (float)b[a];
        mx[a] = (float)b[3 + a];
        size[a] = mx[a] - mn[a];
    }
    for (int bit = 0; bit < 3; bit++) {
        const int a = 2 - bit;
        if (index & (1 << bit)) mn[a] += size[a] / 2.0f;
        else mx[a] -= size[a] / 2.0f;
    }
    for (int a = 0; a < 3; a++) { out[a] = mn[a]; out[3 + a] = mx[a]; }
}

uint32_t rd_u32h(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
int64_t rd_i64h(const uint8_t* p) { int64_t v; memcpy(&v, p, 8); return v; }

constexpr size_t kMaxPotreeNodes = 1u << 24;

// PotreeData.h:185-248: the chunk whose first record describes `top`; its
// records list the chunk's nodes breadth first, proxies (type 2) point at
// the chunks below.
bool potree_chunk(std::vector<PotreeNode>& nodes, int top, const uint8_t* h, uint64_t hsize, int depth) {
    const int64_t off = nodes[top].hier_offset, size = nodes[top].hier_size;
    if (depth > 64 || off < 0 || size <= 0 || size % 22 || (uint64_t)off + (uint64_t)size > hsize) return false;
    const int64_t nrec = size / 22;
    std::vector<int> list{top};
    for (int64_t i = 0; i < nrec; i++) {
        if (i >= (int64_t)list.size()) return false;   // fewer nodes than records
        const int cur = list[i];
        const uint8_t* r = h + off + 22 * i;
        const int type = r[0], mask = r[1];
        const uint32_t np = rd_u32h(r + 2);
        const int64_t a = rd_i64h(r + 6), b = rd_i64h(r + 14);
        if (nodes[cur].type != 2 && type == 2) {
            nodes[cur].hier_offset = a;
            nodes[cur].hier_size = b;
        } else {
            nodes[cur].byte_offset = a;
            nodes[cur].byte_size = b;
        }
        nodes[cur].num_points = np;
        nodes[cur].type = type;
        if (type == 2) continue;
        for (int k = 0; k < 8; k++) {
            if (!(mask & (1 << k))) continue;
            if (nodes.size() >= kMaxPotreeNodes) return false;
            PotreeNode ch;
            potree_child_box(nodes[cur].box, k, ch.box);
            ch.spacing = nodes[cur].spacing / 2.0f;
            ch.level = nodes[cur].level + 1;
            ch.parent = cur;
            nodes.push_back(ch);
            nodes[cur].child[k] = (int)nodes.size() - 1;
            list.push_back((int)nodes.size() - 1);
        }
    }
    if ((int64_t)list.size() > nrec) return false;      // children without records
    for (int n : list)
        if (nodes[n].type == 2 && !potree_chunk(nodes, n, h, hsize, depth + 1)) return false;
    return true;
}

uint16_t rd_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }
uint32_t rd_u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
uint64_t rd_u64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
double rd_f64(const uint8_t* p) { double v; memcpy(&v, p, 8); return v; }

}  // namespace
}  // namespace pcr

using namespace pcr;

extern "C" {

// LasLoaderSparse.cpp:495-526 (and the colour offsets of :123-132).
int pcr_las_parse_header(const void* bytes, uint64_t size, pcr_las_info* out) {
    PCR_CHECK(bytes && out, "las_parse_header: null argument");
    const uint8_t* b = static_cast<const uint8_t*>(bytes);
    memset(out, 0, sizeof(*out));
    PCR_CHECK(size >= 227 && memcmp(b, "LASF", 4) == 0, "las_parse_header: not a LAS header (%llu bytes)",
              (unsigned long long)size);
    out->version_major = b[24];
    out->version_minor = b[25];
    const bool v14 = !(out->version_major == 1 && out->version_minor < 4);
    PCR_CHECK(!v14 || size >= 255, "las_parse_header: LAS %d.%d header needs 255 bytes, got %llu",
              out->version_major, out->version_minor, (unsigned long long)size);
    const uint64_t n = v14 ? rd_u64(b + 247) : rd_u32(b + 107);
    out->num_points = n < 1000000000ull ? n : 1000000000ull;
    out->offset_to_points = rd_u32(b + 96);
    out->point_format = (uint8_t)(b[104] % 128);
    out->record_length = rd_u16(b + 105);
    const int f = out->point_format;
    out->rgb_offset = (uint8_t)(f == 2 ? 20 : f == 3 ? 28 : (f == 7 || f == 8) ? 30 : 0);
    for (int a = 0; a < 3; a++) {
        out->scale[a] = rd_f64(b + 131 + 8 * a);
        out->offset[a] = rd_f64(b + 155 + 8 * a);
        out->max[a] = rd_f64(b + 179 + 16 * a);
        out->min[a] = rd_f64(b + 187 + 16 * a);
    }
    PCR_CHECK(out->record_length >= 12 && out->record_length >= out->rgb_offset + 6,
              "las_parse_header: record length %u too short for format %d", out->record_length, f);
    PCR_CHECK((uint64_t)out->offset_to_points + out->num_points * out->record_length <= size,
              "las_parse_header: %llu records of %u B at %u exceed the %llu-byte file",
              (unsigned long long)out->num_points, out->record_length, out->offset_to_points,
              (unsigned long long)size);
    return PCR_OK;
}

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
        hipError_t e = hipMalloc(&d_rec, nk * I.record_length);
        if (e == hipSuccess) e = hipMalloc(&d_box, (size_t)nbk * 6 * sizeof(double));
        if (e == hipSuccess)
            e = hipMemcpy(d_rec, static_cast<const uint8_t*>(files[k].bytes) + I.offset_to_points,
                          nk * I.record_length, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(e, "cloud_from_las: staging the point records");
        LasParams p;
        for (int a = 0; a < 3; a++) { p.scale[a] = I.scale[a]; p.offset[a] = I.offset[a]; p.min[a] = I.min[a]; }
        p.record_length = I.record_length;
        p.rgb_offset = I.rgb_offset;
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
    const uint32_t bpp = info->bytes_per_point;
    PCR_CHECK(bpp >= 12 && info->rgb_offset + 6 <= bpp, "cloud_from_potree: bytes_per_point %u / rgb_offset %u",
              bpp, info->rgb_offset);
    PCR_CHECK(info->num_points > 0 && info->num_points < (1ull << 32),
              "cloud_from_potree: %llu points (1 .. 2^32-1)", (unsigned long long)info->num_points);
    PCR_CHECK(info->num_points * bpp <= octree_size, "cloud_from_potree: octree.bin holds %llu B, %llu points need %llu",
              (unsigned long long)octree_size, (unsigned long long)info->num_points,
              (unsigned long long)(info->num_points * bpp));
    std::vector<PotreeNode> nodes(1);
    nodes[0].type = 2;
    nodes[0].hier_size = (int64_t)info->first_chunk_size;
    nodes[0].spacing = (float)info->spacing;
    for (int a = 0; a < 3; a++) nodes[0].box[3 + a] = info->max[a] - info->min[a];
    PCR_CHECK(potree_chunk(nodes, 0, static_cast<const uint8_t*>(hierarchy), hierarchy_size, 0),
              "cloud_from_potree: malformed hierarchy.bin (%llu B, first chunk %llu B)",
              (unsigned long long)hierarchy_size, (unsigned long long)info->first_chunk_size);
    // depth-first order, then a stable sort by byteOffset (ties keep the
    // traversal order)
    std::vector<int> order;
    order.reserve(nodes.size());
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int n = stack.back();
        stack.pop_back();
        order.push_back(n);
        for (int k = 7; k >= 0; k--)
            if (nodes[n].child[k] >= 0) stack.push_back(nodes[n].child[k]);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return nodes[a].byte_offset < nodes[b].byte_offset; });
    std::vector<int> pos(nodes.size());
    for (size_t i = 0; i < order.size(); i++) pos[order[i]] = (int)i;
    const uint32_t nn = (uint32_t)order.size();
    std::vector<pcr_batch> hb(nn);
    std::vector<double> boxes(6 * (size_t)nn);
    int max_level = 0;
    for (uint32_t i = 0; i < nn; i++) {
        const PotreeNode& n = nodes[order[i]];
        PCR_CHECK(n.byte_offset >= 0 && n.byte_offset % bpp == 0 &&
                      (uint64_t)n.byte_offset / bpp + n.num_points <= info->num_points,
                  "cloud_from_potree: node %u (%u points at byte %lld) outside the %llu points", i, n.num_points,
                  (long long)n.byte_offset, (unsigned long long)info->num_points);
        pcr_batch& b = hb[i];
        memset(&b, 0, sizeof(b));
        for (int a = 0; a < 3; a++) { b.min[a] = (float)n.box[a]; b.max[a] = (float)n.box[3 + a]; }
        b.num_points = n.num_points;
        b.first_point = (uint32_t)(n.byte_offset / bpp);
        b.level = (uint32_t)n.level;
        b.parent = n.parent >= 0 ? pos[n.parent] : -1;
        b.spacing = n.spacing;
        for (int a = 0; a < 6; a++) boxes[6 * (size_t)i + a] = n.box[a];
        max_level = n.level > max_level ? n.level : max_level;
    }
    pcr_cloud* c = new pcr_cloud();
    int rc = cloud_alloc_batched(ctx, nn, info->num_points, c);
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
