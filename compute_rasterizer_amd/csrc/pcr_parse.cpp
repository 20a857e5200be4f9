// pcr_parse.cpp — host-only parsers of the ingest paths: the LAS header
// (LasLoaderSparse.cpp:495-526) and the Potree 2.0 hierarchy
// (PotreeData.h:115-248). Kept free of device code so that they build with
// the host compiler alone, under sanitizers too.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pcr_internal.h"

namespace pcr {
namespace {

uint16_t rd_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }
uint32_t rd_u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
uint64_t rd_u64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
double rd_f64(const uint8_t* p) { double v; memcpy(&v, p, 8); return v; }

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

}  // namespace

// Validates a Potree 2.0 hierarchy against its metadata and octree.bin size
// and lays its nodes out as batches (see pcr_cloud_from_potree): hb[i] and
// boxes[6 i .. 6 i + 5] (the node's double box) in octree.bin order.
int potree_layout(const pcr_potree_info* info, const void* hierarchy, uint64_t hierarchy_size, uint64_t octree_size,
                  std::vector<pcr_batch>& hb, std::vector<double>& boxes, int* max_level_out) {
    PCR_CHECK(info && hierarchy && max_level_out, "cloud_from_potree: null argument");
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
    hb.assign(nn, pcr_batch{});
    boxes.assign(6 * (size_t)nn, 0.0);
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
    *max_level_out = max_level;
    return PCR_OK;
}

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
    out->compressed = (b[104] & 0xC0) ? 1 : 0;
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
    PCR_CHECK(out->compressed || (uint64_t)out->offset_to_points + out->num_points * out->record_length <= size,
              "las_parse_header: %llu records of %u B at %u exceed the %llu-byte file",
              (unsigned long long)out->num_points, out->record_length, out->offset_to_points,
              (unsigned long long)size);
    return PCR_OK;
}

}  // extern "C"
