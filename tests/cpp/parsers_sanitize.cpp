// parsers_sanitize.cpp — the library's host parsers under AddressSanitizer +
// UndefinedBehaviorSanitizer, fed mutated inputs (CPU only).
//
// Built by tests/test_parsers_sanitize.py together with the host-only
// sources compute_rasterizer_amd/csrc/pcr_parse.cpp (LAS header, Potree 2.0
// hierarchy) and pcr_laz.cpp (LAZ decoder). Each seed file is mutated many
// times (byte flips, extreme field values, truncation) with a fixed PRNG;
// every call must return cleanly, with no sanitizer report.
//
// usage: parsers_sanitize ITER las:FILE... potree:INFO_FILE:HIERARCHY_FILE:OCTREE_SIZE... laz:FILE...
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pcr_internal.h"

namespace pcr {
// normally in pcr_capi.hip: the thread's last error text
void set_error(const char* fmt, ...) {
    char buf[512];
    va_list a;
    va_start(a, fmt);
    vsnprintf(buf, sizeof(buf), fmt, a);
    va_end(a);
}
int hip_fail(hipError_t, const char*, const char*, int) { return PCR_E_HIP; }
}  // namespace pcr

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {   // splitmix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t below(uint64_t n) { return n ? next_u64() % n : 0; }

std::vector<uint8_t> read_file(const std::string& path) {
    std::vector<uint8_t> b;
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path.c_str());
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    b.resize((size_t)n);
    if (n && fread(b.data(), 1, b.size(), f) != b.size()) exit(2);
    fclose(f);
    return b;
}

// Flips bytes, writes extreme little-endian values, or truncates; `hot` bytes
// at the front (headers, tables) are hit more often.
std::vector<uint8_t> mutate(const std::vector<uint8_t>& seed, size_t hot) {
    std::vector<uint8_t> m = seed;
    if (m.empty()) return m;
    const int edits = 1 + (int)below(8);
    for (int e = 0; e < edits; e++) {
        const size_t lim = below(3) ? std::min(hot, m.size()) : m.size();
        const size_t pos = below(lim);
        switch (below(4)) {
            case 0: m[pos] ^= (uint8_t)(1u << below(8)); break;
            case 1: m[pos] = (uint8_t)next_u64(); break;
            case 2: {   // an extreme 32- or 64-bit value
                static const uint64_t ext[] = {0, 1, 0x7FFFFFFFull, 0x80000000ull, 0xFFFFFFFFull,
                                               0x7FFFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFFFFFull, 22204, 50000};
                const uint64_t v = ext[below(sizeof(ext) / sizeof(ext[0]))];
                const size_t w = below(2) ? 4 : 8;
                for (size_t k = 0; k < w && pos + k < m.size(); k++) m[pos + k] = (uint8_t)(v >> (8 * k));
                break;
            }
            default: break;
        }
    }
    if (below(5) == 0) m.resize(below(m.size() + 1));
    return m;
}

uint64_t g_calls = 0, g_ok = 0;

void las_case(const std::vector<uint8_t>& seed, int iters) {
    for (int i = 0; i < iters; i++) {
        const std::vector<uint8_t> m = mutate(seed, 375);
        pcr_las_info info;
        g_ok += pcr_las_parse_header(m.data(), m.size(), &info) == PCR_OK;
        g_calls++;
    }
}

void potree_case(const pcr_potree_info& info0, const std::vector<uint8_t>& hier, uint64_t octree_size, int iters) {
    for (int i = 0; i < iters; i++) {
        const std::vector<uint8_t> m = mutate(hier, hier.size());
        pcr_potree_info info = info0;
        if (below(4) == 0) info.first_chunk_size = below(2) ? next_u64() : below(hier.size() + 100);
        if (below(8) == 0) info.bytes_per_point = (uint32_t)below(64);
        if (below(8) == 0) info.num_points = below(2) ? next_u64() : below(info0.num_points * 2 + 2);
        std::vector<pcr_batch> hb;
        std::vector<double> boxes;
        int levels = 0;
        g_ok += pcr::potree_layout(&info, m.data(), m.size(), octree_size, hb, boxes, &levels) == PCR_OK;
        g_calls++;
    }
}

void laz_case(const std::vector<uint8_t>& seed, int iters) {
    pcr_laz_info info0;
    if (pcr_laz_info_get(seed.data(), seed.size(), &info0) != PCR_OK) {
        fprintf(stderr, "seed LAZ does not parse\n");
        exit(2);
    }
    const uint32_t off = seed.size() >= 100 ? (uint32_t)seed[96] | ((uint32_t)seed[97] << 8) |
                                                  ((uint32_t)seed[98] << 16) | ((uint32_t)seed[99] << 24)
                                            : 0;
    for (int i = 0; i < iters; i++) {
        // headers and VLR, or the first chunk and its layer sizes
        const std::vector<uint8_t> m = mutate(seed, below(2) ? (size_t)off + 512 : seed.size());
        pcr_laz_info info;
        g_calls++;
        if (pcr_laz_info_get(m.data(), m.size(), &info) != PCR_OK) continue;
        const uint64_t n = std::min<uint64_t>(info.num_points, 20000);
        if ((uint64_t)info.record_length * n > (64ull << 20)) continue;
        std::vector<uint8_t> out((size_t)info.record_length * n + 1);
        const uint64_t first = n ? below(n) : 0;
        g_ok += pcr_laz_decode(m.data(), m.size(), first, n - first, out.data(), (uint32_t)below(2),
                               1 + (int)below(3)) == PCR_OK;
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: parsers_sanitize ITER las:FILE potree:INFO:HIER:OCTREE_SIZE laz:FILE ...\n");
        return 2;
    }
    const int iters = atoi(argv[1]);
    for (int a = 2; a < argc; a++) {
        const std::string arg = argv[a];
        if (arg.rfind("las:", 0) == 0) {
            las_case(read_file(arg.substr(4)), iters);
        } else if (arg.rfind("laz:", 0) == 0) {
            laz_case(read_file(arg.substr(4)), iters);
        } else if (arg.rfind("potree:", 0) == 0) {
            const size_t p1 = arg.find(':', 7), p2 = arg.find(':', p1 + 1);
            const std::vector<uint8_t> ib = read_file(arg.substr(7, p1 - 7));
            if (ib.size() != sizeof(pcr_potree_info)) return 2;
            pcr_potree_info info;
            memcpy(&info, ib.data(), sizeof(info));
            potree_case(info, read_file(arg.substr(p1 + 1, p2 - p1 - 1)), strtoull(arg.c_str() + p2 + 1, nullptr, 10),
                        iters);
        } else {
            fprintf(stderr, "unknown argument %s\n", arg.c_str());
            return 2;
        }
    }
    printf("parsers sanitize ok: %llu calls, %llu accepted\n", (unsigned long long)g_calls,
           (unsigned long long)g_ok);
    return 0;
}
