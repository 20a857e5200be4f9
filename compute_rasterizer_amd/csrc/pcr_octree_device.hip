// pcr_octree_device.hip — the LOD octree of pcr_octree.hip built on the device.
//
// Same layout, same bytes as the host builder (a GPU test compares them):
//   1. Morton keys over the bounding cube + hipCUB radix sort (stable).
//   2. per level l: node and half-cell segments of the sorted keys are
//      numbered by prefix sums of "starts a new segment" flags; a node whose
//      unassigned points fit in max_node_points (or the last level) keeps them
//      all; the larger nodes keep a Poisson-disk sample: 27 phase launches, one
//      wave per half-cell of the phase, the samples kept so far in a hash table
//      keyed by half-cell (the host builder's rule and tie order).
//   3. a stable radix sort by level groups the points level by level in
//      Morton order; node and chunk segments are numbered again, the chunk
//      table goes to the host, which computes parents, spacing, 4-aligned
//      first_point and the headers (tiny: one entry per chunk);
//   4. per-chunk AABBs on the device, combined per node on the host, and a
//      final encode kernel writes the planes.
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>
#include "pcr_device.h"
#include "pcr_generate.h"
#include "pcr_internal.h"

namespace pcr {

namespace {

__device__ __forceinline__ uint64_t spread3_d(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_oct_keys(const float* __restrict__ xyz, uint64_t n, float mn0, float mn1, float mn2, double scale,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t q[3];
        const float m[3] = {mn0, mn1, mn2};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double d = ((double)xyz[3 * i + a] - (double)m[a]) * scale;
            q[a] = d >= 2097151.0 ? 2097151u : (uint32_t)d;
        }
        keys[i] = (spread3_d(q[0]) << 2) | (spread3_d(q[1]) << 1) | spread3_d(q[2]);
        vals[i] = (uint32_t)i;
    }
}

__device__ __forceinline__ uint64_t prefix(uint64_t k, int bits3) { return bits3 == 0 ? 0ull : k >> (63 - bits3); }

// flags of "first point of a node / of a half-cell" at level l (cell bits g: 3(g + l) Morton bits)
__global__ void k_oct_flags(const uint64_t* __restrict__ K, uint64_t n, int l, int g, uint32_t* __restrict__ nflag,
                            uint32_t* __restrict__ cflag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = K[i];
        const uint64_t p = i ? K[i - 1] : 0;
        nflag[i] = (i == 0 || prefix(k, 3 * l) != prefix(p, 3 * l)) ? 1u : 0u;
        cflag[i] = (i == 0 || prefix(k, 3 * (g + l)) != prefix(p, 3 * (g + l))) ? 1u : 0u;
    }
}

// one atomic per wave for a per-thread count (keeps a single counter uncontended)
__device__ __forceinline__ void wave_add(unsigned long long* p, uint32_t v) {
    uint32_t s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == 0 && s) atomicAdd(p, (unsigned long long)s);
}

__global__ void k_oct_unassigned(const int8_t* __restrict__ lev, uint64_t n, uint32_t* __restrict__ uf) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        uf[i] = lev[i] < 0 ? 1u : 0u;
}

__global__ void k_oct_nstart(const uint32_t* __restrict__ nflag, const uint32_t* __restrict__ nid, uint64_t n,
                             uint32_t* __restrict__ ns) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (nflag[i]) ns[nid[i] - 1] = (uint32_t)i;
}

// A node whose unassigned points fit in maxpts (or any node of the last
// level) keeps them all. The count is a difference of the inclusive prefix
// sum U of "unassigned" flags over the node's range [ns[node], next start).
__global__ void k_oct_leaf(const uint32_t* __restrict__ nid, const uint32_t* __restrict__ U,
                           const uint32_t* __restrict__ ns, int8_t* __restrict__ lev, uint64_t n, uint32_t maxpts,
                           int l, int last, unsigned long long* __restrict__ assigned) {
    const uint32_t nnodes = nid[n - 1];
    uint32_t mine = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (lev[i] >= 0) continue;
        const uint32_t node = nid[i] - 1, s = ns[node];
        const uint64_t e = node + 1 < nnodes ? ns[node + 1] : n;
        const uint32_t cnt = U[e - 1] - (s ? U[s - 1] : 0u);
        if (last || cnt <= maxpts) { lev[i] = (int8_t)l; mine++; }
    }
    wave_add(assigned, mine);
}

// ---- Poisson-disk subsampling (the host builder's rule, pcr_octree.hip) ---
__device__ __forceinline__ uint32_t compact3_d(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x | (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x | (x >> 16)) & 0x1F00000000FFFFull;
    x = (x | (x >> 32)) & 0x1FFFFFull;
    return (uint32_t)x;
}
__device__ __forceinline__ uint64_t point21(uint64_t key) {
    return (uint64_t)compact3_d(key >> 2) | ((uint64_t)compact3_d(key >> 1) << 21) | ((uint64_t)compact3_d(key) << 42);
}
__device__ __forceinline__ uint32_t hslot(uint64_t key, int cap_bits) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - cap_bits));
}
// half-cell starts over the sorted points, and the half-cells that still
// hold an unassigned point (nonzero = a candidate cell)
__global__ void k_ps_cells(const uint32_t* __restrict__ cid, const int8_t* __restrict__ lev, uint64_t n,
                           uint32_t* __restrict__ cstart, uint32_t* __restrict__ cand) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = cid[i] - 1;
        if (i == 0 || cid[i - 1] != cid[i]) cstart[c] = (uint32_t)i;
        if (lev[i] < 0) cand[c] = 1u;   // benign race: every writer stores 1
    }
}
__device__ __forceinline__ int cell_phase(const uint64_t* __restrict__ K, const uint32_t* __restrict__ cstart,
                                          uint64_t c, int t) {
    const uint64_t pre = K[cstart[c]] >> (63 - 3 * t);
    return (int)(compact3_d(pre >> 2) % 3 + 3 * (compact3_d(pre >> 1) % 3) + 9 * (compact3_d(pre) % 3));
}
// candidate half-cells per phase (LDS histogram, one global add per phase per block)
__global__ __launch_bounds__(256) void k_ps_hist(const uint64_t* __restrict__ K, const uint32_t* __restrict__ cstart,
                                                 const uint32_t* __restrict__ cand, uint64_t ncells, int t,
                                                 uint32_t* __restrict__ counts) {
    __shared__ uint32_t h[27];
    if (threadIdx.x < 27) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (uint64_t)gridDim.x * blockDim.x)
        if (cand[c]) atomicAdd(&h[cell_phase(K, cstart, c, t)], 1u);
    __syncthreads();
    if (threadIdx.x < 27 && h[threadIdx.x]) atomicAdd(counts + threadIdx.x, h[threadIdx.x]);
}
// the candidate half-cells listed by phase: cursor[p] starts at phase p's offset
__global__ __launch_bounds__(256) void k_ps_scatter(const uint64_t* __restrict__ K, const uint32_t* __restrict__ cstart,
                                                    const uint32_t* __restrict__ cand, uint64_t ncells, int t,
                                                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ list) {
    __shared__ uint32_t h[27], base[27];
    for (uint64_t c0 = (uint64_t)blockIdx.x * blockDim.x; c0 < ncells; c0 += (uint64_t)gridDim.x * blockDim.x) {
        if (threadIdx.x < 27) h[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t c = c0 + threadIdx.x;
        int ph = -1;
        uint32_t slot = 0;
        if (c < ncells && cand[c]) {
            ph = cell_phase(K, cstart, c, t);
            slot = atomicAdd(&h[ph], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 27 && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(cursor + threadIdx.x, h[threadIdx.x]);
        __syncthreads();
        if (ph >= 0) list[base[ph] + slot] = (uint32_t)c;
        __syncthreads();
    }
}
// One wave per half-cell of the phase: gather the samples kept in its 124
// neighbours (same node) into LDS, then every lane tests its share of the
// cell's unassigned points against them; the lowest (priority, position)
// among the free ones becomes the cell's sample.
__global__ __launch_bounds__(256) void k_ps_phase(const uint64_t* __restrict__ K, const uint32_t* __restrict__ perm,
                                                  int8_t* __restrict__ lev, const uint32_t* __restrict__ cstart,
                                                  const uint32_t* __restrict__ list, uint32_t nlist,
                                                  uint64_t ncells, uint64_t n, int l, int t, int gbits, int phase,
                                                  uint64_t pkey, unsigned long long* __restrict__ htab, int cap_bits,
                                                  unsigned long long* __restrict__ assigned) {
    __shared__ unsigned long long s_nb[4][124];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r = (int64_t)1 << (21 - gbits - l);
    const int64_t lim = (int64_t)1 << t;
    const int sh = gbits + 1;
    const uint32_t capmask = (1u << cap_bits) - 1u;
    uint32_t mine = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * 4 + w; e < nlist; e += (uint64_t)gridDim.x * 4) {
        const uint64_t c = list[e];                               // wave-uniform
        const uint64_t a = cstart[c], b = c + 1 < ncells ? cstart[c + 1] : n;
        const uint64_t pre = K[a] >> (63 - 3 * t);
        const uint32_t cc[3] = {compact3_d(pre >> 2), compact3_d(pre >> 1), compact3_d(pre)};
        // the kept samples around the cell, compacted to s_nb[w][0 .. m)
        unsigned long long found[2] = {~0ull, ~0ull};
        for (int j = (int)lane, q = 0; j < 124; j += 64, q++) {
            const int o = j < 62 ? j : j + 1;
            const int64_t d[3] = {o % 5 - 2, (o / 5) % 5 - 2, o / 25 - 2};
            bool ok = true;
            uint64_t nk = 0;
            int nph = 0, mul = 1;
            for (int k = 0; k < 3; k++) {
                const int64_t v = (int64_t)cc[k] + d[k];
                ok = ok && v >= 0 && v < lim && (v >> sh) == ((int64_t)cc[k] >> sh);
                nk |= spread3_d((uint32_t)v) << (2 - k);
                nph += mul * (int)(((v % 3) + 3) % 3);
                mul *= 3;
            }
            // only half-cells of earlier phases can hold a sample yet
            unsigned long long val = ~0ull;
            if (ok && nph < phase) {
                const unsigned long long key = nk + 1;
                for (uint32_t sl = hslot(key, cap_bits);; sl = (sl + 1) & capmask) {
                    const unsigned long long hk = htab[2 * (uint64_t)sl];
                    if (hk == key) { val = htab[2 * (uint64_t)sl + 1]; break; }
                    if (hk == 0) break;
                }
            }
            found[q] = val;
        }
        int m = 0;
        for (int q = 0; q < 2; q++) {
            const unsigned long long has = __ballot(found[q] != ~0ull);
            const int before = __popcll(has & ((1ull << lane) - 1ull));
            if (found[q] != ~0ull) s_nb[w][m + before] = found[q];
            m += __popcll(has);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        unsigned long long bp = ~0ull;
        uint32_t bi = 0xFFFFFFFFu;
        for (uint64_t i = a + lane; i < b; i += 64) {
            if (lev[i] >= 0) continue;
            const uint64_t p = point21(K[i]);
            bool free = true;
            for (int j = 0; j < m && free; j++) {
                const unsigned long long v = s_nb[w][j];
                int64_t d2 = 0;
                for (int k = 0; k < 3; k++) {
                    const int64_t d = (int64_t)((p >> (21 * k)) & 0x1FFFFF) - (int64_t)((v >> (21 * k)) & 0x1FFFFF);
                    d2 += d * d;
                }
                free = d2 >= r * r;
            }
            if (!free) continue;
            const unsigned long long pr = rnd(pkey, perm[i]);
            if (pr < bp || (pr == bp && (uint32_t)i < bi)) { bp = pr; bi = (uint32_t)i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long op = __shfl_xor(bp, off);
            const uint32_t oi = __shfl_xor(bi, off);
            if (op < bp || (op == bp && oi < bi)) { bp = op; bi = oi; }
        }
        if (lane == 0 && bi != 0xFFFFFFFFu) {
            lev[bi] = (int8_t)l;
            mine++;
            const unsigned long long key = pre + 1;
            for (uint32_t sl = hslot(key, cap_bits);; sl = (sl + 1) & capmask) {
                const unsigned long long prev = atomicCAS(htab + 2 * (uint64_t)sl, 0ull, key);
                if (prev == 0ull || prev == key) { htab[2 * (uint64_t)sl + 1] = point21(K[bi]); break; }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (lane == 0 && mine) atomicAdd(assigned, (unsigned long long)mine);
}

__global__ void k_oct_iota(uint32_t* __restrict__ v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}
__global__ void k_oct_lev_u8(const int8_t* __restrict__ lev, uint64_t n, uint8_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint8_t)lev[i];
}

// node flags over the level-grouped order (pos[j] = sorted position)
__global__ void k_oct_nodeflags(const uint32_t* __restrict__ pos, const uint64_t* __restrict__ K,
                                const int8_t* __restrict__ lev, uint64_t n, uint32_t* __restrict__ nflag) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = pos[j];
        const int L = lev[i];
        bool start = j == 0;
        if (!start) {
            const uint32_t ip = pos[j - 1];
            const int Lp = lev[ip];
            start = L != Lp || prefix(K[i], 3 * L) != prefix(K[ip], 3 * Lp);
        }
        nflag[j] = start ? 1u : 0u;
    }
}

__global__ void k_oct_nodestart(const uint32_t* __restrict__ nflag, const uint32_t* __restrict__ nid,
                                const uint32_t* __restrict__ pos, const uint64_t* __restrict__ K,
                                const int8_t* __restrict__ lev, uint64_t n, uint32_t* __restrict__ nstart,
                                uint64_t* __restrict__ nkey, uint8_t* __restrict__ nlev) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
        if (nflag[j]) {
            const uint32_t node = nid[j] - 1, i = pos[j];
            const int L = lev[i];
            nstart[node] = (uint32_t)j;
            nkey[node] = prefix(K[i], 3 * L);
            nlev[node] = (uint8_t)L;
        }
}

__global__ void k_oct_chunkflags(const uint32_t* __restrict__ nflag, const uint32_t* __restrict__ nid,
                                 const uint32_t* __restrict__ nstart, uint64_t n, uint32_t maxpts,
                                 uint32_t* __restrict__ cflag) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
        cflag[j] = (nflag[j] || ((uint32_t)j - nstart[nid[j] - 1]) % maxpts == 0) ? 1u : 0u;
}

__global__ void k_oct_chunkstart(const uint32_t* __restrict__ cflag, const uint32_t* __restrict__ cid,
                                 const uint32_t* __restrict__ nid, uint64_t n, uint32_t* __restrict__ cstart,
                                 uint32_t* __restrict__ cnode) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
        if (cflag[j]) { cstart[cid[j] - 1] = (uint32_t)j; cnode[cid[j] - 1] = nid[j] - 1; }
}

// float AABB of each chunk, one workgroup per chunk
__global__ __launch_bounds__(256) void k_oct_chunk_aabb(const float* __restrict__ xyz, const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ pos,
                                                        const uint32_t* __restrict__ cstart, uint32_t nchunks,
                                                        uint64_t n, float* __restrict__ out) {
    __shared__ float s[6][256];
    const uint32_t c = blockIdx.x;
    const uint64_t a = cstart[c], b = c + 1 < nchunks ? cstart[c + 1] : n;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t j = a + threadIdx.x; j < b; j += 256) {
        const uint64_t o = perm[pos[j]];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float v = xyz[3 * o + k];
            mn[k] = v < mn[k] ? v : mn[k];
            mx[k] = v > mx[k] ? v : mx[k];
        }
    }
    for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 3; k++) {
                const float o = s[k][threadIdx.x + off], O = s[3 + k][threadIdx.x + off];
                s[k][threadIdx.x] = o < s[k][threadIdx.x] ? o : s[k][threadIdx.x];
                s[3 + k][threadIdx.x] = O > s[3 + k][threadIdx.x] ? O : s[3 + k][threadIdx.x];
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * (uint64_t)c + threadIdx.x] = s[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t quant30_d(float x, float mn, double size) {
    if (size == 0.0) return 0u;
    const uint32_t v = (uint32_t)((((double)x - (double)mn) / size) * 1073741824.0);
    return v < 1073741823u ? v : 1073741823u;
}

__global__ void k_oct_encode(const float* __restrict__ xyz, const uint32_t* __restrict__ rgba,
                             const uint32_t* __restrict__ perm, const uint32_t* __restrict__ pos,
                             const uint32_t* __restrict__ cid, const uint32_t* __restrict__ cstart,
                             const pcr_batch* __restrict__ hdr, uint64_t n, uint32_t* __restrict__ lo,
                             uint32_t* __restrict__ med, uint32_t* __restrict__ hi, uint32_t* __restrict__ col) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = cid[j] - 1;
        const pcr_batch& h = hdr[c];
        const uint64_t slot = (uint64_t)h.first_point + ((uint32_t)j - cstart[c]);
        const uint64_t o = perm[pos[j]];
        uint32_t q[3];
#pragma unroll
        for (int k = 0; k < 3; k++) q[k] = quant30_d(xyz[3 * o + k], h.min[k], (double)h.max[k] - (double)h.min[k]);
        lo[slot] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
        med[slot] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
        hi[slot] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
        col[slot] = rgba[o];
    }
}

__global__ void k_oct_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_oct_fill64(unsigned long long* __restrict__ p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// RAII set of device buffers freed on every exit path
struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs() { for (void* p : ptrs) (void)hipFree(p); }
    template <class T>
    hipError_t alloc(T** p, uint64_t count) {
        hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back((void*)*p);
        return e;
    }
};

}  // namespace

int build_octree_device(pcr_ctx* ctx, const float* xyz, const uint32_t* rgba, uint64_t n, const pcr_octree_opts* o,
                        pcr_cloud** out) {
    PCR_CHECK(ctx && xyz && rgba && o && out, "octree_device: null argument");
    PCR_CHECK(n > 0 && n < 0x7FFFFFFFull, "octree_device: point count %llu out of range (1 .. 2^31-1)",
              (unsigned long long)n);
    const uint32_t maxpts = o->max_node_points ? o->max_node_points : 20000;
    const float div = o->root_spacing_div > 0 ? o->root_spacing_div : 128.0f;
    const int gbits = (int)std::lround(std::log2((double)div));
    PCR_CHECK(gbits >= 1 && gbits <= 12 && std::ldexp(1.0, gbits) == (double)div,
              "octree_device: root_spacing_div %g must be a power of two in [2, 4096]", (double)div);
    const int max_depth = std::min(o->max_depth > 0 ? o->max_depth : 12, 21 - gbits);
    hipStream_t s = ctx->stream;
    uint64_t gl = (n + 255) / 256;
    if (gl > (uint64_t)ctx->cu_count * 64) gl = (uint64_t)ctx->cu_count * 64;
    const int G = (int)gl;
    DevBufs B;

    // 1. AABB and Morton sort
    float mn[3], mx[3];
    const int rc0 = device_aabb(ctx, xyz, n, mn, mx);
    if (rc0) return rc0;
    float side = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
    const double scale = side > 0.0f ? 2097152.0 / (double)side : 0.0;
    uint64_t *K = nullptr, *K2 = nullptr;
    uint32_t *perm = nullptr, *v2 = nullptr;
    PCR_HIP(B.alloc(&K2, n));
    PCR_HIP(B.alloc(&K, n));
    PCR_HIP(B.alloc(&v2, n));
    PCR_HIP(B.alloc(&perm, n));
    k_oct_keys<<<G, 256, 0, s>>>(xyz, n, mn[0], mn[1], mn[2], scale, K2, v2);
    size_t tb = 0;
    PCR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, K2, K, v2, perm, (int)n, 0, 63, s));
    void* temp = nullptr;
    PCR_HIP(B.alloc((unsigned char**)&temp, tb));
    PCR_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, K2, K, v2, perm, (int)n, 0, 63, s));

    // 2. level assignment
    int8_t* lev = nullptr;
    uint32_t *nflag = nullptr, *cflag = nullptr, *nid = nullptr, *cid = nullptr, *ncount = nullptr, *cellwin = nullptr;
    unsigned long long* assigned = nullptr;
    PCR_HIP(B.alloc(&lev, n));
    PCR_HIP(B.alloc(&nflag, n));
    PCR_HIP(B.alloc(&cflag, n));
    PCR_HIP(B.alloc(&nid, n));
    PCR_HIP(B.alloc(&cid, n));
    PCR_HIP(B.alloc(&ncount, n));
    PCR_HIP(B.alloc(&cellwin, n));
    PCR_HIP(B.alloc(&assigned, 1));
    uint32_t *cellcand = nullptr, *plist = nullptr, *phase_cnt = nullptr, *phase_cur = nullptr;
    unsigned long long* htab = nullptr;   // (half-cell key + 1, packed sample) pairs
    uint64_t hcap = 0;
    PCR_HIP(B.alloc(&cellcand, n));
    PCR_HIP(B.alloc(&plist, n));
    PCR_HIP(B.alloc(&phase_cnt, 27));
    PCR_HIP(B.alloc(&phase_cur, 27));
    PCR_HIP(hipMemsetAsync(lev, 0xFF, n, s));
    PCR_HIP(hipMemsetAsync(assigned, 0, 8, s));
    size_t tscan = 0;
    PCR_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tscan, nflag, nid, (int)n, s));
    void* stemp = nullptr;
    PCR_HIP(B.alloc((unsigned char**)&stemp, tscan));
    const uint64_t pkey = stream_key(o->seed, 5);
    int deepest = 0;
    for (int l = 0; l <= max_depth; l++) {
        // cells: the half-cells of level l (at a last level past 20 - gbits only the node flags matter)
        k_oct_flags<<<G, 256, 0, s>>>(K, n, l, std::min(gbits + 1, 21 - l), nflag, cflag);
        PCR_HIP(hipcub::DeviceScan::InclusiveSum(stemp, tscan, nflag, nid, (int)n, s));
        PCR_HIP(hipcub::DeviceScan::InclusiveSum(stemp, tscan, cflag, cid, (int)n, s));
        // per-node unassigned counts as prefix-sum differences: cflag is free
        // once cid exists, cellwin holds the node starts until the leaf pass
        k_oct_unassigned<<<G, 256, 0, s>>>(lev, n, cflag);
        PCR_HIP(hipcub::DeviceScan::InclusiveSum(stemp, tscan, cflag, ncount, (int)n, s));
        k_oct_nstart<<<G, 256, 0, s>>>(nflag, nid, n, cellwin);
        k_oct_leaf<<<G, 256, 0, s>>>(nid, ncount, cellwin, lev, n, maxpts, l, l == max_depth, assigned);
        PCR_HIP(hipGetLastError());
        unsigned long long done = 0;
        PCR_HIP(hipMemcpyAsync(&done, assigned, 8, hipMemcpyDeviceToHost, s));
        PCR_HIP(hipStreamSynchronize(s));
        deepest = l;
        if (done >= n) break;
        // Poisson-disk samples of the nodes above maxpts (the rest of the points)
        const int t = gbits + l + 1;
        uint32_t ncells = 0;
        PCR_HIP(hipMemcpyAsync(&ncells, cid + n - 1, 4, hipMemcpyDeviceToHost, s));
        k_oct_fill32<<<G, 256, 0, s>>>(cellcand, n, 0u);
        PCR_HIP(hipStreamSynchronize(s));
        k_ps_cells<<<G, 256, 0, s>>>(cid, lev, n, cellwin, cellcand);
        k_oct_fill32<<<1, 64, 0, s>>>(phase_cnt, 27, 0u);
        k_ps_hist<<<G, 256, 0, s>>>(K, cellwin, cellcand, ncells, t, phase_cnt);
        uint32_t cnt[27], off[28];
        PCR_HIP(hipMemcpyAsync(cnt, phase_cnt, sizeof(cnt), hipMemcpyDeviceToHost, s));
        PCR_HIP(hipStreamSynchronize(s));
        off[0] = 0;
        for (int ph = 0; ph < 27; ph++) off[ph + 1] = off[ph] + cnt[ph];
        const uint64_t nc = off[27];
        PCR_HIP(hipMemcpyAsync(phase_cur, off, 27 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        k_ps_scatter<<<G, 256, 0, s>>>(K, cellwin, cellcand, ncells, t, phase_cur, plist);
        int cap_bits = 6;
        while ((1ull << cap_bits) < 2 * nc) cap_bits++;
        PCR_CHECK(cap_bits <= 31, "octree_device: %llu sample cells exceed the hash table", nc);
        const uint64_t cap = 1ull << cap_bits;
        if (cap > hcap) {
            PCR_HIP(B.alloc(&htab, 2 * cap));
            hcap = cap;
        }
        k_oct_fill64<<<G, 256, 0, s>>>(htab, 2 * cap, 0ull);
        for (int phase = 0; phase < 27; phase++) {
            if (!cnt[phase]) continue;
            const uint64_t gw = (cnt[phase] + 3) / 4;
            const int GP = (int)std::min<uint64_t>(gw, (uint64_t)ctx->cu_count * 32);
            k_ps_phase<<<GP, 256, 0, s>>>(K, perm, lev, cellwin, plist + off[phase], cnt[phase], ncells, n, l, t,
                                          gbits, phase, pkey, htab, cap_bits, assigned);
        }
        PCR_HIP(hipGetLastError());
        PCR_HIP(hipMemcpyAsync(&done, assigned, 8, hipMemcpyDeviceToHost, s));
        PCR_HIP(hipStreamSynchronize(s));
        if (done >= n) break;
    }

    // 3. group by level (stable), number nodes and chunks
    uint8_t *lev8 = nullptr, *lev8s = nullptr;
    uint32_t* iota = nullptr;
    uint32_t* pos = nullptr;
    PCR_HIP(B.alloc(&lev8, n));
    PCR_HIP(B.alloc(&lev8s, n));
    PCR_HIP(B.alloc(&iota, n));
    PCR_HIP(B.alloc(&pos, n));
    k_oct_lev_u8<<<G, 256, 0, s>>>(lev, n, lev8);
    k_oct_iota<<<G, 256, 0, s>>>(iota, n);
    size_t tb2 = 0;
    PCR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, lev8, lev8s, iota, pos, (int)n, 0, 5, s));
    void* temp2 = nullptr;
    PCR_HIP(B.alloc((unsigned char**)&temp2, tb2));
    PCR_HIP(hipcub::DeviceRadixSort::SortPairs(temp2, tb2, lev8, lev8s, iota, pos, (int)n, 0, 5, s));
    // reuse: nflag/nid for nodes, cflag/cid for chunks, ncount as node start
    k_oct_nodeflags<<<G, 256, 0, s>>>(pos, K, lev, n, nflag);
    PCR_HIP(hipcub::DeviceScan::InclusiveSum(stemp, tscan, nflag, nid, (int)n, s));
    uint32_t nnodes = 0;
    PCR_HIP(hipMemcpyAsync(&nnodes, nid + n - 1, 4, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipStreamSynchronize(s));
    uint64_t* nkey = nullptr;
    uint8_t* nlev = nullptr;
    PCR_HIP(B.alloc(&nkey, nnodes));
    PCR_HIP(B.alloc(&nlev, nnodes));
    uint32_t* nstart = ncount;
    k_oct_nodestart<<<G, 256, 0, s>>>(nflag, nid, pos, K, lev, n, nstart, nkey, nlev);
    k_oct_chunkflags<<<G, 256, 0, s>>>(nflag, nid, nstart, n, maxpts, cflag);
    PCR_HIP(hipcub::DeviceScan::InclusiveSum(stemp, tscan, cflag, cid, (int)n, s));
    uint32_t nchunks = 0;
    PCR_HIP(hipMemcpyAsync(&nchunks, cid + n - 1, 4, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipStreamSynchronize(s));
    uint32_t *cstart = nullptr, *cnode = nullptr;
    float* caabb = nullptr;
    PCR_HIP(B.alloc(&cstart, nchunks));
    PCR_HIP(B.alloc(&cnode, nchunks));
    PCR_HIP(B.alloc(&caabb, 6ull * nchunks));
    k_oct_chunkstart<<<G, 256, 0, s>>>(cflag, cid, nid, n, cstart, cnode);
    k_oct_chunk_aabb<<<nchunks, 256, 0, s>>>(xyz, perm, pos, cstart, nchunks, n, caabb);
    PCR_HIP(hipGetLastError());
    std::vector<uint32_t> h_cstart(nchunks), h_cnode(nchunks);
    std::vector<float> h_caabb(6ull * nchunks);
    std::vector<uint64_t> h_nkey(nnodes);
    std::vector<uint8_t> h_nlev(nnodes);
    PCR_HIP(hipMemcpyAsync(h_cstart.data(), cstart, 4ull * nchunks, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipMemcpyAsync(h_cnode.data(), cnode, 4ull * nchunks, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipMemcpyAsync(h_caabb.data(), caabb, 24ull * nchunks, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipMemcpyAsync(h_nkey.data(), nkey, 8ull * nnodes, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipMemcpyAsync(h_nlev.data(), nlev, nnodes, hipMemcpyDeviceToHost, s));
    PCR_HIP(hipStreamSynchronize(s));

    // headers on the host: node AABB over its chunks, parents, spacing, 4-aligned slots
    std::vector<float> nbox(6ull * nnodes);
    for (uint32_t k = 0; k < nnodes; k++)
        for (int a = 0; a < 3; a++) { nbox[6 * k + a] = INFINITY; nbox[6 * k + 3 + a] = -INFINITY; }
    for (uint32_t c = 0; c < nchunks; c++) {
        float* nb = &nbox[6ull * h_cnode[c]];
        for (int a = 0; a < 3; a++) {
            nb[a] = h_caabb[6ull * c + a] < nb[a] ? h_caabb[6ull * c + a] : nb[a];
            nb[3 + a] = h_caabb[6ull * c + 3 + a] > nb[3 + a] ? h_caabb[6ull * c + 3 + a] : nb[3 + a];
        }
    }
    const double diag = std::sqrt((double)(mx[0] - mn[0]) * (mx[0] - mn[0]) + (double)(mx[1] - mn[1]) * (mx[1] - mn[1]) +
                                  (double)(mx[2] - mn[2]) * (mx[2] - mn[2]));
    std::vector<std::unordered_map<uint64_t, int32_t>> first_chunk(deepest + 1);
    std::vector<pcr_batch> hdr(nchunks);
    uint64_t slot = 0;
    for (uint32_t c = 0; c < nchunks; c++) {
        const uint32_t node = h_cnode[c];
        const int L = h_nlev[node];
        const uint64_t cnt = (c + 1 < nchunks ? h_cstart[c + 1] : n) - h_cstart[c];
        const bool first_of_node = c == 0 || h_cnode[c - 1] != node;
        int32_t parent = -1;
        if (L > 0) {
            auto it = first_chunk[L - 1].find(h_nkey[node] >> 3);
            PCR_CHECK(it != first_chunk[L - 1].end(), "octree_device: node at level %d has no parent", L);
            parent = it->second;
        }
        if (first_of_node) first_chunk[L][h_nkey[node]] = (int32_t)c;
        slot = (slot + 3) & ~3ull;
        pcr_batch& h = hdr[c];
        memset(&h, 0, sizeof(h));
        for (int a = 0; a < 3; a++) { h.min[a] = nbox[6ull * node + a]; h.max[a] = nbox[6ull * node + 3 + a]; }
        h.num_points = (uint32_t)cnt;
        h.first_point = (uint32_t)slot;
        h.level = (uint32_t)L;
        h.parent = parent;
        h.spacing = (float)(diag / (double)div / std::ldexp(1.0, L));
        slot += cnt;
    }
    PCR_CHECK(slot < (1ull << 32), "octree_device: too many points after padding");

    // 4. cloud and encode
    pcr_cloud* c = new pcr_cloud();
    int rc = cloud_alloc_batched(ctx, nchunks, slot, c);
    if (rc) { delete c; return rc; }
    c->batch_size = maxpts;
    c->num_levels = deepest + 1;
    for (int a = 0; a < 3; a++) { c->aabb_min[a] = mn[a]; c->aabb_max[a] = mx[a]; }
    PCR_HIP(hipMemcpyAsync(c->d_batches, hdr.data(), sizeof(pcr_batch) * nchunks, hipMemcpyHostToDevice, s));
    PCR_HIP(hipMemsetAsync(c->d_lo, 0, 4 * c->plane_len, s));
    PCR_HIP(hipMemsetAsync(c->d_med, 0, 4 * c->plane_len, s));
    PCR_HIP(hipMemsetAsync(c->d_hi, 0, 4 * c->plane_len, s));
    PCR_HIP(hipMemsetAsync(c->d_rgba, 0, 4 * c->plane_len, s));
    k_oct_encode<<<G, 256, 0, s>>>(xyz, rgba, perm, pos, cid, cstart, c->d_batches, n, c->d_lo, c->d_med, c->d_hi,
                                   c->d_rgba);
    PCR_HIP(hipGetLastError());
    PCR_HIP(hipStreamSynchronize(s));
    c->h_batches = std::move(hdr);
    *out = c;
    return PCR_OK;
}

}  // namespace pcr
