// pcr_octree.hip — Potree-style LOD octree in the batched plane layout.
//
// Stands in for the node layout the reference reads from Potree 2.0 files
// (modules/compute/PotreeData.h:188-317: one 64-byte header per node, points
// of a node contiguous, encoded relative to the node box as in
// LasLoaderSparse.cpp:146-182) for synthetic clouds that have no Potree file.
//
// Construction (host):
//   1. Morton keys (21 bits/axis over the bounding cube), stable LSD radix sort.
//   2. Level l: a node (Morton prefix of 3l bits) whose points not yet taken
//      by a coarser level fit in max_node_points keeps all of them and becomes
//      a leaf; a larger node keeps a seeded Poisson-disk sample of them with
//      spacing cube/(div * 2^l) (see the Poisson-disk helpers below).
//   3. Nodes are emitted level by level in Morton order (parents precede
//      children); a node with more than max_node_points points is split into
//      chunks that share the node's header fields. first_point is 4-aligned.
//   4. spacing = AABB diagonal / div / 2^level is the LOD spacing of the node.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>
#include "pcr_generate.h"
#include "pcr_internal.h"

namespace pcr {

static void radix_sort_pairs(std::vector<uint64_t>& keys, std::vector<uint32_t>& vals) {
    const size_t n = keys.size();
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> v2(n);
    for (int shift = 0; shift < 63; shift += 11) {
        size_t cnt[2048] = {0};
        for (size_t i = 0; i < n; i++) cnt[(keys[i] >> shift) & 2047]++;
        size_t sum = 0;
        for (int d = 0; d < 2048; d++) { const size_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (size_t i = 0; i < n; i++) {
            const size_t o = cnt[(keys[i] >> shift) & 2047]++;
            k2[o] = keys[i];
            v2[o] = vals[i];
        }
        keys.swap(k2);
        vals.swap(v2);
    }
}

static uint64_t spread3_host(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

static uint32_t quant30(float x, float mn, double size) {
    if (size == 0.0) return 0u;
    const uint32_t v = (uint32_t)((((double)x - (double)mn) / size) * 1073741824.0);
    return v < 1073741823u ? v : 1073741823u;
}

// ---- Poisson-disk subsampling (shared with the device builder) -----------
// A node above max_node_points at level l keeps a maximal set of its
// unassigned points with no two closer than the level's cell side
// r = cube / (div * 2^l), chosen by seeded priority. Distances are exact in
// the 21-bit Morton coordinates. The node is cut into half-cells of side r/2
// (any two points of one are closer than r, so a half-cell keeps at most one
// sample), and the half-cells are visited in 27 phases by (x, y, z) mod 3:
// half-cells of one phase are at least 3 apart, so their points are more
// than r apart and the order inside a phase does not matter (the phase
// groups of L.-Y. Wei, "Parallel Poisson disk sampling", SIGGRAPH 2008).
// A half-cell keeps its lowest-priority unassigned point that lies at least
// r from every sample already kept in the 124 half-cells around it (within
// the node), so the set is maximal: every other point is within r of a
// sample.
struct PoissonCell {
    size_t a, b;        // sorted positions [a, b) of the half-cell's points
    uint64_t prefix;    // Morton prefix of the half-cell (3t bits)
};
static inline uint32_t compact3(uint64_t x) {   // every third bit, from bit 0
    x &= 0x1249249249249249ull;
    x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x | (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x | (x >> 16)) & 0x1F00000000FFFFull;
    x = (x | (x >> 32)) & 0x1FFFFFull;
    return (uint32_t)x;
}
// cell coordinates (x, y, z) of a Morton prefix of 3t bits (key = x<<2 | y<<1 | z per bit)
static inline void morton_cell(uint64_t prefix, int t, uint32_t c[3]) {
    (void)t;
    c[0] = compact3(prefix >> 2);
    c[1] = compact3(prefix >> 1);
    c[2] = compact3(prefix);
}
static inline uint64_t morton_of(const uint32_t c[3]) {
    return (spread3_host(c[0]) << 2) | (spread3_host(c[1]) << 1) | spread3_host(c[2]);
}
// the point's 21-bit coordinates packed as x | y << 21 | z << 42
static inline uint64_t morton_point(uint64_t key) {
    return (uint64_t)compact3(key >> 2) | ((uint64_t)compact3(key >> 1) << 21) | ((uint64_t)compact3(key) << 42);
}
static inline int poisson_phase(const uint32_t c[3]) { return (int)(c[0] % 3 + 3 * (c[1] % 3) + 9 * (c[2] % 3)); }
// Samples kept in the 124 half-cells around c (same node: the top l = t-g-1
// bits of each coordinate agree), through lookup(prefix, &value).
template <class F>
static int poisson_neighbours(const uint32_t c[3], int t, int gbits, F lookup, uint64_t* out) {
    const int64_t lim = (int64_t)1 << t;
    const int sh = gbits + 1;   // half-cell bits below the node
    int n = 0;
    for (int o = 0; o < 125; o++) {
        if (o == 62) continue;
        const int64_t d[3] = {o % 5 - 2, (o / 5) % 5 - 2, o / 25 - 2};
        uint32_t nc[3];
        bool ok = true;
        for (int k = 0; k < 3; k++) {
            const int64_t v = (int64_t)c[k] + d[k];
            ok = ok && v >= 0 && v < lim && (v >> sh) == ((int64_t)c[k] >> sh);
            nc[k] = (uint32_t)v;
        }
        uint64_t v;
        if (ok && lookup(morton_of(nc), &v)) out[n++] = v;
    }
    return n;
}
static inline bool poisson_conflict(uint64_t p, const uint64_t* nb, int nn, int64_t r) {
    for (int k = 0; k < nn; k++) {
        int64_t d2 = 0;
        for (int a = 0; a < 3; a++) {
            const int64_t d = (int64_t)((p >> (21 * a)) & 0x1FFFFF) - (int64_t)((nb[k] >> (21 * a)) & 0x1FFFFF);
            d2 += d * d;
        }
        if (d2 < r * r) return true;
    }
    return false;
}

struct OctreeLayout {
    std::vector<pcr_batch> hdr;
    std::vector<uint32_t> lo, med, hi, col;
    int depth = 0;
    float mn[3], mx[3];
};

static int build_octree_layout(const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_octree_opts* o,
                               OctreeLayout& L) {
    PCR_CHECK(xyz && rgba && o, "octree: null argument");
    PCR_CHECK(n > 0 && n < (1ull << 32), "octree: point count out of range");
    const uint32_t maxpts = o->max_node_points ? o->max_node_points : 20000;
    const float div = o->root_spacing_div > 0 ? o->root_spacing_div : 128.0f;
    const int gbits = (int)std::lround(std::log2((double)div));
    PCR_CHECK(gbits >= 1 && gbits <= 12 && std::ldexp(1.0, gbits) == (double)div,
              "octree: root_spacing_div %g must be a power of two in [2, 4096]", (double)div);
    const int max_depth = std::min(o->max_depth > 0 ? o->max_depth : 12, 21 - gbits);

    // 1. AABB, cube, Morton order
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * i + a];
            mn[a] = v < mn[a] ? v : mn[a];
            mx[a] = v > mx[a] ? v : mx[a];
        }
    float side = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
    const double scale = side > 0.0f ? 2097152.0 / (double)side : 0.0;
    std::vector<uint64_t> keys(n);
    std::vector<uint32_t> idx(n);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            const double d = ((double)xyz[3 * i + a] - (double)mn[a]) * scale;
            q[a] = d >= 2097151.0 ? 2097151u : (uint32_t)d;
        }
        keys[i] = (spread3_host(q[0]) << 2) | (spread3_host(q[1]) << 1) | spread3_host(q[2]);
        idx[i] = (uint32_t)i;
    }
    radix_sort_pairs(keys, idx);

    // 2. level assignment: leaves, then Poisson-disk samples of the larger nodes
    const uint64_t pkey = stream_key(o->seed, 5);
    std::vector<int8_t> lev(n, -1);
    uint64_t remaining = n;
    int deepest = 0;
    std::unordered_map<uint64_t, uint64_t> accepted;   // half-cell prefix -> packed sample coordinates
    for (int l = 0; l <= max_depth && remaining; l++) {
        const int node_shift = 63 - 3 * l;
        bool sampling = false;
        size_t a = 0;
        while (a < n) {
            const uint64_t node = l == 0 ? 0 : keys[a] >> node_shift;
            size_t b = a;
            uint64_t un = 0;
            while (b < n && (l == 0 || (keys[b] >> node_shift) == node)) { un += lev[b] < 0; b++; }
            if (un) {
                deepest = l;
                if (un <= maxpts || l == max_depth) {
                    for (size_t i = a; i < b; i++) if (lev[i] < 0) lev[i] = (int8_t)l;
                    remaining -= un;
                } else {
                    sampling = true;
                }
            }
            a = b;
        }
        if (!sampling) continue;
        // the unassigned points left are those of nodes above max_node_points
        const int t = gbits + l + 1;                        // half-cell: 2^-t of the cube per axis
        const int64_t r = (int64_t)1 << (21 - gbits - l);   // sample spacing in 2^-21 cube units
        std::vector<PoissonCell> cells;
        for (size_t c = 0; c < n;) {
            const uint64_t pre = keys[c] >> (63 - 3 * t);
            size_t e = c;
            bool any = false;
            for (; e < n && (keys[e] >> (63 - 3 * t)) == pre; e++) any |= lev[e] < 0;
            if (any) cells.push_back(PoissonCell{c, e, pre});
            c = e;
        }
        accepted.clear();
        for (int phase = 0; phase < 27; phase++)
            for (const PoissonCell& cell : cells) {
                uint32_t cc[3];
                morton_cell(cell.prefix, t, cc);
                if (poisson_phase(cc) != phase) continue;
                uint64_t nb[124];
                const int nn = poisson_neighbours(cc, t, gbits, [&](uint64_t key, uint64_t* v) {
                    auto it = accepted.find(key);
                    if (it == accepted.end()) return false;
                    *v = it->second;
                    return true;
                }, nb);
                size_t best = SIZE_MAX;
                uint64_t bp = ~0ull;
                for (size_t i = cell.a; i < cell.b; i++) {
                    if (lev[i] >= 0) continue;
                    const uint64_t pq = morton_point(keys[i]);
                    if (poisson_conflict(pq, nb, nn, r)) continue;
                    const uint64_t pr = rnd(pkey, idx[i]);
                    if (pr < bp) { bp = pr; best = i; }   // first minimum: the lower position wins ties
                }
                if (best == SIZE_MAX) continue;
                lev[best] = (int8_t)l;
                remaining--;
                accepted[cell.prefix] = morton_point(keys[best]);
            }
    }

    // 3. nodes, level by level in Morton order
    const double diag = std::sqrt((double)(mx[0] - mn[0]) * (mx[0] - mn[0]) + (double)(mx[1] - mn[1]) * (mx[1] - mn[1]) +
                                  (double)(mx[2] - mn[2]) * (mx[2] - mn[2]));
    std::vector<pcr_batch> hdr;
    std::vector<uint32_t> order;   // sorted position of each stored point, UINT32_MAX = padding
    order.reserve(n + n / 1000 + 64);
    std::vector<std::unordered_map<uint64_t, int32_t>> first_chunk(deepest + 1);
    for (int l = 0; l <= deepest; l++) {
        const int node_shift = 63 - 3 * l;
        std::vector<uint32_t> pos;
        for (uint64_t i = 0; i < n; i++) if (lev[i] == l) pos.push_back((uint32_t)i);
        size_t a = 0;
        while (a < pos.size()) {
            const uint64_t node = l == 0 ? 0 : keys[pos[a]] >> node_shift;
            size_t b = a;
            float nmn[3] = {INFINITY, INFINITY, INFINITY}, nmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            while (b < pos.size() && (l == 0 || (keys[pos[b]] >> node_shift) == node)) {
                const float* p = xyz + 3ull * idx[pos[b]];
                for (int k = 0; k < 3; k++) {
                    nmn[k] = p[k] < nmn[k] ? p[k] : nmn[k];
                    nmx[k] = p[k] > nmx[k] ? p[k] : nmx[k];
                }
                b++;
            }
            int32_t parent = -1;
            if (l > 0) {
                auto it = first_chunk[l - 1].find(node >> 3);
                PCR_CHECK(it != first_chunk[l - 1].end(), "octree: node at level %d has no parent", l);
                parent = it->second;
            }
            first_chunk[l][node] = (int32_t)hdr.size();
            for (size_t c = a; c < b; c += maxpts) {
                const size_t e = std::min(b, c + maxpts);
                while (order.size() % 4) order.push_back(UINT32_MAX);
                pcr_batch h;
                memset(&h, 0, sizeof(h));
                for (int k = 0; k < 3; k++) { h.min[k] = nmn[k]; h.max[k] = nmx[k]; }
                h.num_points = (uint32_t)(e - c);
                h.first_point = (uint32_t)order.size();
                h.level = (uint32_t)l;
                h.parent = parent;
                h.spacing = (float)(diag / (double)div / std::ldexp(1.0, l));
                hdr.push_back(h);
                for (size_t q = c; q < e; q++) order.push_back(pos[q]);
            }
            a = b;
        }
    }
    PCR_CHECK(order.size() < (1ull << 32), "octree: too many points after padding");

    // 4. encode planes relative to each node box
    const uint64_t np = order.size();
    L.lo.assign(np, 0); L.med.assign(np, 0); L.hi.assign(np, 0); L.col.assign(np, 0);
    uint32_t *lo = L.lo.data(), *med = L.med.data(), *hi = L.hi.data(), *col = L.col.data();
    for (const pcr_batch& h : hdr) {
        double size[3];
        for (int k = 0; k < 3; k++) size[k] = (double)h.max[k] - (double)h.min[k];
        for (uint32_t q = 0; q < h.num_points; q++) {
            const uint64_t s = h.first_point + q;
            const uint32_t src = idx[order[s]];
            uint32_t qq[3];
            for (int k = 0; k < 3; k++) qq[k] = quant30(xyz[3ull * src + k], h.min[k], size[k]);
            lo[s] = ((qq[0] >> 20) & 1023u) | (((qq[1] >> 20) & 1023u) << 10) | (((qq[2] >> 20) & 1023u) << 20);
            med[s] = ((qq[0] >> 10) & 1023u) | (((qq[1] >> 10) & 1023u) << 10) | (((qq[2] >> 10) & 1023u) << 20);
            hi[s] = (qq[0] & 1023u) | ((qq[1] & 1023u) << 10) | ((qq[2] & 1023u) << 20);
            col[s] = rgba[src];
        }
    }
    L.hdr.swap(hdr);
    L.depth = deepest;
    for (int k = 0; k < 3; k++) { L.mn[k] = mn[k]; L.mx[k] = mx[k]; }
    return PCR_OK;
}

int build_octree_host(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_octree_opts* o,
                      pcr_cloud** out) {
    PCR_CHECK(ctx && out, "octree: null argument");
    OctreeLayout L;
    int rc = build_octree_layout(xyz, rgba, n, o, L);
    if (rc) return rc;
    const uint64_t np = L.lo.size();
    pcr_cloud* c = new pcr_cloud();
    rc = cloud_alloc_batched(ctx, (uint32_t)L.hdr.size(), np, c);
    if (rc) { delete c; return rc; }
    c->batch_size = o->max_node_points ? o->max_node_points : 20000;
    c->num_levels = L.depth + 1;
    for (int k = 0; k < 3; k++) { c->aabb_min[k] = L.mn[k]; c->aabb_max[k] = L.mx[k]; }
    PCR_HIP(hipMemcpy(c->d_batches, L.hdr.data(), sizeof(pcr_batch) * L.hdr.size(), hipMemcpyHostToDevice));
    PCR_HIP(hipMemcpy(c->d_lo, L.lo.data(), np * 4, hipMemcpyHostToDevice));
    PCR_HIP(hipMemcpy(c->d_med, L.med.data(), np * 4, hipMemcpyHostToDevice));
    PCR_HIP(hipMemcpy(c->d_hi, L.hi.data(), np * 4, hipMemcpyHostToDevice));
    PCR_HIP(hipMemcpy(c->d_rgba, L.col.data(), np * 4, hipMemcpyHostToDevice));
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    c->h_batches = L.hdr;
    *out = c;
    return PCR_OK;
}

}  // namespace pcr

using namespace pcr;

extern "C" {

int pcr_cloud_build_octree_host(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n,
                                const pcr_octree_opts* opts, pcr_cloud** out) {
    return build_octree_host(ctx, xyz, rgba, n, opts, out);
}

int pcr_cloud_build_octree_device(pcr_ctx* ctx, const float* xyz_dev, const uint32_t* rgba_dev, uint64_t n,
                                  const pcr_octree_opts* opts, pcr_cloud** out) {
    return build_octree_device(ctx, xyz_dev, rgba_dev, n, opts, out);
}

// Generate a synthetic scene on the device and build its octree there.
int pcr_octree_generate(pcr_ctx* ctx, const pcr_gen_cfg* cfg, const pcr_octree_opts* opts, pcr_cloud** out) {
    PCR_CHECK(ctx && cfg && opts && out, "octree_generate: null argument");
    PCR_CHECK(cfg->num_points > 0, "octree_generate: no points");
    float* xyz = nullptr;
    uint32_t* rgba = nullptr;
    PCR_HIP(hipMalloc(&xyz, cfg->num_points * 12));
    if (hipMalloc(&rgba, cfg->num_points * 4) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(xyz);
        set_error("octree_generate: device allocation failed");
        return PCR_E_OOM;
    }
    int rc = generate_points(ctx, cfg, xyz, rgba);
    if (rc == PCR_OK) rc = build_octree_device(ctx, xyz, rgba, cfg->num_points, opts, out);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(xyz);
    (void)hipFree(rgba);
    return rc;
}

int pcr_octree_layout_host(const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_octree_opts* opts,
                           pcr_batch* nodes, uint32_t* nnodes, uint32_t* lo, uint32_t* med, uint32_t* hi,
                           uint32_t* rgba_out, uint64_t* nslots) {
    PCR_CHECK(nnodes && nslots, "octree_layout_host: null size outputs");
    OctreeLayout L;
    int rc = build_octree_layout(xyz, rgba, n, opts, L);
    if (rc) return rc;
    const bool fill = nodes && lo && med && hi && rgba_out;
    if (fill) {
        PCR_CHECK(*nnodes >= L.hdr.size() && *nslots >= L.lo.size(),
                  "octree_layout_host: arrays too small (%u nodes, %llu slots needed)", (unsigned)L.hdr.size(),
                  (unsigned long long)L.lo.size());
        memcpy(nodes, L.hdr.data(), sizeof(pcr_batch) * L.hdr.size());
        memcpy(lo, L.lo.data(), 4 * L.lo.size());
        memcpy(med, L.med.data(), 4 * L.med.size());
        memcpy(hi, L.hi.data(), 4 * L.hi.size());
        memcpy(rgba_out, L.col.data(), 4 * L.col.size());
    }
    *nnodes = (uint32_t)L.hdr.size();
    *nslots = L.lo.size();
    return PCR_OK;
}

int pcr_lod_selection(pcr_ctx* ctx, pcr_cloud* cloud, uint8_t* flags, uint32_t n) {
    PCR_CHECK(ctx && cloud && flags, "lod_selection: null argument");
    PCR_CHECK(n == cloud->num_batches, "lod_selection: %u flags for %u nodes", n, cloud->num_batches);
    PCR_HIP(hipStreamSynchronize(ctx->stream));
    PCR_HIP(hipMemcpy(flags, cloud->d_levels, n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) flags[i] = flags[i] != 255;
    return PCR_OK;
}

}  // extern "C"
