// pcr_laz.cpp — LAZ (LASzip-compressed LAS) point decompression on the host.
//
// The reference reads .laz files through the LASzip library it vendors
// (libs/laszip, LASzip 3.x; LasLoaderSparse.cpp:214-330 calls
// laszip_open_reader / laszip_seek_point / laszip_read_point /
// laszip_get_coordinates). This file is this library's own decoder of the
// LASzip format, written from the format's definition, for the two
// compressors LASzip writes:
//
//   pointwise chunked (compressor 2), point formats 0-3 (+ extra bytes):
//     items POINT10, GPSTIME11, RGB12, BYTE, version 2. A chunk is the
//     first record raw, then one arithmetic-coded stream for the others.
//   layered chunked (compressor 3), point formats 6-8 (+ extra bytes):
//     items POINT14, RGB14, RGBNIR14, BYTE14, version 3 or 4. A chunk is the
//     first record raw, the point count, the byte size of every layer, then
//     the layers, each its own arithmetic-coded stream, with separate model
//     sets per scanner channel.
//
// Chunks are independent (every model restarts at a chunk), so decoding is
// spread over host threads, one chunk at a time per thread, through the
// chunk table at the end of the point data. Output is the uncompressed LAS
// point records, byte for byte what the LAS file held before compression.
//
// Entropy coding: adaptive binary/multi-symbol arithmetic coding with
// periodic model updates (A. Said, "Introduction to Arithmetic Coding —
// Theory and Practice", HP Labs HPL-2004-76, and his FastAC coder, which
// LASzip uses): 32-bit interval, renormalisation by bytes below 2^24,
// 13-bit binary probabilities and 15-bit symbol distributions. The constants
// and update schedules below are those of the format, since any change moves
// every decoded symbol.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "pcr_internal.h"

namespace pcr {
namespace laz {
namespace {

// ---------------------------------------------------------------- bytes
// A bounded reader: past the end it yields zeros and records the overrun,
// so a damaged file decodes to garbage but never reads out of bounds.
struct Bytes {
    const uint8_t* p = nullptr;
    const uint8_t* end = nullptr;
    bool overrun = false;
    void init(const uint8_t* b, uint64_t n) { p = b; end = b + n; overrun = false; }
    uint32_t get() {
        if (p < end) return *p++;
        overrun = true;
        return 0;
    }
    uint32_t u32() { uint32_t v = get(); v |= get() << 8; v |= get() << 16; return v | (get() << 24); }
};

// ---------------------------------------------------------------- models
constexpr uint32_t kMinLength = 0x01000000u;   // renormalise below 2^24
constexpr uint32_t kBitShift = 13;             // binary model precision
constexpr uint32_t kBitMaxCount = 1u << kBitShift;
constexpr uint32_t kSymShift = 15;             // symbol model precision
constexpr uint32_t kSymMaxCount = 1u << kSymShift;

struct BitModel {
    uint32_t p0 = 1u << (kBitShift - 1), n0 = 1, n = 2, cycle = 4, until = 4;
    void reset() { *this = BitModel(); }
    void update() {
        if ((n += cycle) > kBitMaxCount) {   // halve the counts
            n = (n + 1) >> 1;
            n0 = (n0 + 1) >> 1;
            if (n0 == n) ++n;
        }
        p0 = (n0 * (0x80000000u / n)) >> (31 - kBitShift);
        cycle = std::min((5 * cycle) >> 2, 64u);
        until = cycle;
    }
};

// Adaptive distribution over `n` symbols. Alphabets above 16 symbols keep a
// lookup table from the top bits of the scaled value to a symbol range. The
// cumulative distribution, the counts and the table share one allocation.
struct SymModel {
    uint32_t n = 0, total = 0, cycle = 0, until = 0, tbits = 0, tshift = 0;
    std::vector<uint32_t> mem;
    uint32_t* dist = nullptr;
    uint32_t* cnt = nullptr;
    uint32_t* table = nullptr;   // null for 16 symbols or fewer
    bool live = false;           // reset since the current chunk (lazily made models)

    SymModel() = default;
    SymModel(const SymModel& o) { *this = o; }
    SymModel& operator=(const SymModel& o) {   // copies re-point into their own memory
        n = o.n; total = o.total; cycle = o.cycle; until = o.until; tbits = o.tbits; tshift = o.tshift;
        mem = o.mem;
        live = o.live;
        dist = mem.empty() ? nullptr : mem.data();
        cnt = dist ? dist + n : nullptr;
        table = o.table ? dist + 2 * n : nullptr;
        return *this;
    }

    void make(uint32_t symbols) {
        n = symbols;
        uint32_t extra = 0;
        if (n > 16) {
            tbits = 3;
            while (n > (1u << (tbits + 2))) ++tbits;
            tshift = kSymShift - tbits;
            extra = (1u << tbits) + 2;
        }
        mem.assign(2 * n + extra, 0);
        dist = mem.data();
        cnt = dist + n;
        table = extra ? dist + 2 * n : nullptr;
        reset();
    }
    void reset() {
        std::fill(cnt, cnt + n, 1u);
        total = 0;
        cycle = n;
        update();
        until = cycle = (n + 6) >> 1;
        live = true;
    }
    void update() {
        if ((total += cycle) > kSymMaxCount) {
            total = 0;
            for (uint32_t k = 0; k < n; k++) total += (cnt[k] = (cnt[k] + 1) >> 1);
        }
        const uint32_t scale = 0x80000000u / total;
        uint32_t sum = 0;
        if (!table) {
            for (uint32_t k = 0; k < n; k++) {
                dist[k] = (scale * sum) >> (31 - kSymShift);
                sum += cnt[k];
            }
        } else {
            const uint32_t tsize = 1u << tbits;
            uint32_t s = 0;
            for (uint32_t k = 0; k < n; k++) {
                dist[k] = (scale * sum) >> (31 - kSymShift);
                sum += cnt[k];
                const uint32_t w = dist[k] >> tshift;
                while (s < w) table[++s] = k - 1;
            }
            table[0] = 0;
            while (s <= tsize) table[++s] = n - 1;
        }
        cycle = std::min((5 * cycle) >> 2, (n + 6) << 3);
        until = cycle;
    }
};

// ---------------------------------------------------------------- decoder
struct Decoder {
    Bytes in;
    uint32_t value = 0, length = 0xFFFFFFFFu;

    void start(const uint8_t* b, uint64_t n) {
        in.init(b, n);
        length = 0xFFFFFFFFu;
        value = in.get() << 24;
        value |= in.get() << 16;
        value |= in.get() << 8;
        value |= in.get();
    }
    void renorm() {
        do value = (value << 8) | in.get();
        while ((length <<= 8) < kMinLength);
    }
    uint32_t bit(BitModel& m) {
        const uint32_t x = m.p0 * (length >> kBitShift);
        const uint32_t b = value >= x;
        if (!b) {
            length = x;
            ++m.n0;
        } else {
            value -= x;
            length -= x;
        }
        if (length < kMinLength) renorm();
        if (--m.until == 0) m.update();
        return b;
    }
    uint32_t symbol(SymModel& m) {
        uint32_t sym, x, y = length;
        if (m.table) {
            const uint32_t dv = value / (length >>= kSymShift);
            // t <= table size (value < length keeps dv below 2^15 + 2^15 / (length >> 15)),
            // and the table holds size + 2 entries
            uint32_t t = dv >> m.tshift;
            if (t > (1u << m.tbits)) {   // value >= length: only a damaged stream
                t = 1u << m.tbits;
                in.overrun = true;
            }
            sym = m.table[t];
            uint32_t hi = m.table[t + 1] + 1;
            while (hi > sym + 1) {   // bisection inside the table's range
                const uint32_t k = (sym + hi) >> 1;
                if (m.dist[k] > dv) hi = k;
                else sym = k;
            }
            x = m.dist[sym] * length;
            if (sym != m.n - 1) y = m.dist[sym + 1] * length;
        } else {
            x = sym = 0;
            length >>= kSymShift;
            uint32_t hi = m.n, k = hi >> 1;
            do {
                const uint32_t z = length * m.dist[k];
                if (z > value) {
                    hi = k;
                    y = z;
                } else {
                    sym = k;
                    x = z;
                }
            } while ((k = (sym + hi) >> 1) != sym);
        }
        value -= x;
        length = y - x;
        if (length < kMinLength) renorm();
        ++m.cnt[sym];
        if (--m.until == 0) m.update();
        return sym;
    }
    uint32_t raw(uint32_t bits) {   // 1 <= bits <= 19: equiprobable bits
        uint32_t v = value / (length >>= bits);
        value -= length * v;
        if (length < kMinLength) renorm();
        if (v >> bits) {   // only a damaged stream
            v &= (1u << bits) - 1;
            in.overrun = true;
        }
        return v;
    }
    uint32_t bits(uint32_t n) {
        if (n > 19) {
            const uint32_t lo = raw(16);
            return (bits(n - 16) << 16) | lo;
        }
        return raw(n);
    }
    uint32_t u32() {
        const uint32_t lo = raw(16);
        return (raw(16) << 16) | lo;
    }
};

// ---------------------------------------------------------------- integers
// A value is coded as its difference (the corrector) from a prediction: k,
// the bit length of the corrector's interval, with a per-context model,
// then the position inside the interval (a model for k <= 8, the low k - 8
// bits raw otherwise). Results wrap into the value's bit range.
struct IntDecoder {
    uint32_t bits = 16, ctxs = 1, high = 8, cbits = 0, crange = 0, k = 0;
    int32_t cmin = 0;
    std::vector<SymModel> mk;   // per context: k
    BitModel m0;                // corrector 0 or 1
    std::vector<SymModel> mc;   // per k >= 1: the corrector's position

    void make(uint32_t nbits, uint32_t contexts) {
        bits = nbits;
        ctxs = contexts;
        if (bits < 32) {
            cbits = bits;
            crange = 1u << bits;
            cmin = -(int32_t)(crange / 2);
        } else {
            cbits = 32;
            crange = 0;
            cmin = INT32_MIN;
        }
        mk.assign(ctxs, SymModel());
        for (auto& m : mk) m.make(cbits + 1);
        mc.assign(cbits + 1, SymModel());
        for (uint32_t i = 1; i <= cbits; i++) mc[i].make(1u << std::min(i, high));
        reset();
    }
    void reset() {
        for (auto& m : mk) m.reset();
        m0.reset();
        for (uint32_t i = 1; i <= cbits; i++) mc[i].reset();
        k = 0;
    }
    int32_t corrector(Decoder& d, uint32_t ctx) {
        k = d.symbol(mk[ctx]);
        if (k == 0) return (int32_t)d.bit(m0);
        if (k >= 32) return cmin;
        int64_t c;
        if (k <= high) {
            c = d.symbol(mc[k]);
        } else {
            const uint32_t lowbits = k - high;
            const uint64_t h = d.symbol(mc[k]);
            c = (int64_t)((h << lowbits) | d.bits(lowbits));
        }
        // [0, 2^(k-1)) maps to [-(2^k - 1), -2^(k-1)], [2^(k-1), 2^k) to [2^(k-1) + 1, 2^k]
        if (c >= (int64_t(1) << (k - 1))) c += 1;
        else c -= (int64_t(1) << k) - 1;
        return (int32_t)c;
    }
    int32_t decode(Decoder& d, int32_t pred, uint32_t ctx = 0) {
        int64_t r = (int64_t)(int32_t)((uint32_t)pred + (uint32_t)corrector(d, ctx));
        if (crange) {   // wrap into [0, 2^bits)
            if (r < 0) r += crange;
            else if (r >= (int64_t)crange) r -= crange;
        }
        return (int32_t)r;
    }
};

// Streaming approximate median of the last differences: a sorted window of
// five that alternately drops its largest and its smallest entry.
struct Median5 {
    int32_t s[5] = {0, 0, 0, 0, 0};
    bool high = true;
    void reset() { *this = Median5(); }
    int32_t get() const { return s[2]; }
    void add(int32_t v) {
        if (high) {   // the largest leaves
            if (v < s[2]) {
                s[4] = s[3];
                s[3] = s[2];
                int i = 2;
                while (i > 0 && v < s[i - 1]) { s[i] = s[i - 1]; --i; }
                s[i] = v;
            } else {
                if (v < s[3]) { s[4] = s[3]; s[3] = v; } else s[4] = v;
                high = false;
            }
        } else {      // the smallest leaves
            if (s[2] < v) {
                s[0] = s[1];
                s[1] = s[2];
                int i = 2;
                while (i < 4 && s[i + 1] < v) { s[i] = s[i + 1]; ++i; }
                s[i] = v;
            } else {
                if (s[1] < v) { s[0] = s[1]; s[1] = v; } else s[0] = v;
                high = true;
            }
        }
    }
};

// Context tables of the format: return number r and number of returns n to
// a coarse class, for formats 0-5 (8 x 8) and 6-10 (16 x 16).
const uint8_t kReturnMap8[8][8] = {
    {15, 14, 13, 12, 11, 10, 9, 8}, {14, 0, 1, 3, 6, 10, 10, 9}, {13, 1, 2, 4, 7, 11, 11, 10},
    {12, 3, 4, 5, 8, 12, 12, 11},   {11, 6, 7, 8, 9, 13, 13, 12}, {10, 10, 11, 12, 13, 14, 14, 13},
    {9, 10, 11, 12, 13, 14, 15, 14}, {8, 9, 10, 11, 12, 13, 14, 15}};
const uint8_t kReturnMap6ctx[16][16] = {
    {0, 1, 2, 3, 4, 5, 3, 4, 4, 5, 5, 5, 5, 5, 5, 5}, {1, 0, 1, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3},
    {2, 1, 2, 4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 3, 3}, {3, 3, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4},
    {4, 3, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4}, {5, 3, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4},
    {3, 3, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4}, {4, 3, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4},
    {4, 3, 4, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4}, {5, 3, 4, 4, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4},
    {5, 3, 4, 4, 4, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4}, {5, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 4, 4, 4},
    {5, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 4, 4}, {5, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 4},
    {5, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5}, {5, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5}};
// |n - r| (formats 0-5), min(|n - r|, 7) (formats 6-10)
inline uint32_t return_level(uint32_t n, uint32_t r, uint32_t cap) {
    const uint32_t d = n > r ? n - r : r - n;
    return d < cap ? d : cap;
}
inline uint32_t fold8(int32_t v) { return (uint32_t)v & 0xFFu; }
inline int32_t clamp8(int32_t v) { return v < 0 ? 0 : v > 255 ? 255 : v; }
inline uint32_t rd16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
inline uint32_t rd32(const uint8_t* p) { return rd16(p) | (rd16(p + 2) << 16); }
inline uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }
inline void wr16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
inline void wr32(uint8_t* p, uint32_t v) { wr16(p, v); wr16(p + 2, v >> 16); }
inline void wr64(uint8_t* p, uint64_t v) { wr32(p, (uint32_t)v); wr32(p + 4, (uint32_t)(v >> 32)); }

// GPS time as a set of four interleaved sequences, each with its own last
// integer difference; multiples of that difference are cheap to code.
// `v3` selects the layered form (no "unchanged" symbol: the point layer
// says whether the time changed).
constexpr int32_t kGpsMulti = 500, kGpsMultiMinus = -10;
struct GpsTime {
    bool v3 = false;
    SymModel m_multi, m_0diff;
    IntDecoder ic;
    uint32_t last = 0, next = 0;
    uint64_t t[4] = {0, 0, 0, 0};
    int32_t diff[4] = {0, 0, 0, 0}, extreme[4] = {0, 0, 0, 0};
    uint32_t code_full() const { return v3 ? kGpsMulti - kGpsMultiMinus + 1 : kGpsMulti - kGpsMultiMinus + 2; }

    void make(bool layered) {
        v3 = layered;
        m_multi.make(kGpsMulti - kGpsMultiMinus + (v3 ? 5 : 6));
        m_0diff.make(v3 ? 5 : 6);
        ic.make(32, 9);
    }
    void reset(uint64_t t0) {
        m_multi.reset();
        m_0diff.reset();
        ic.reset();
        last = next = 0;
        for (int i = 0; i < 4; i++) { t[i] = 0; diff[i] = 0; extreme[i] = 0; }
        t[0] = t0;
    }
    void full(Decoder& d) {   // a new sequence: upper 32 bits predicted, lower 32 raw
        next = (next + 1) & 3;
        const uint32_t hi = (uint32_t)ic.decode(d, (int32_t)(uint32_t)(t[last] >> 32), 8);
        t[next] = ((uint64_t)hi << 32) | d.u32();
        last = next;
        diff[last] = 0;
        extreme[last] = 0;
    }
    void extreme_step(int32_t dt) {
        if (++extreme[last] > 3) {
            diff[last] = dt;
            extreme[last] = 0;
        }
    }
    // Returns false on a stream that keeps switching sequences (damaged).
    bool read(Decoder& d) {
        const uint32_t cf = code_full();
        for (int hop = 0; hop < 8; hop++) {
            if (diff[last] == 0) {
                const uint32_t m = d.symbol(m_0diff);
                // v2: 0 unchanged, 1 32-bit difference, 2 full, >2 switch
                // v3: 0 32-bit difference, 1 full, >1 switch
                const uint32_t m1 = v3 ? m + 1 : m;
                if (m1 == 0) return true;
                if (m1 == 1) {
                    diff[last] = ic.decode(d, 0, 0);
                    t[last] += (uint64_t)(int64_t)diff[last];
                    extreme[last] = 0;
                    return true;
                }
                if (m1 == 2) { full(d); return true; }
                last = (last + m1 - 2) & 3;
                continue;
            }
            const uint32_t m = d.symbol(m_multi);
            if (m == 1) {
                t[last] += (uint64_t)(int64_t)ic.decode(d, diff[last], 1);
                extreme[last] = 0;
                return true;
            }
            if (m < (uint32_t)(kGpsMulti - kGpsMultiMinus + 1)) {   // a multiple of the last difference
                int32_t dt;
                const auto times = [&](int32_t f) { return (int32_t)((uint32_t)f * (uint32_t)diff[last]); };
                if (m == 0) {
                    dt = ic.decode(d, 0, 7);
                    extreme_step(dt);
                } else if (m < (uint32_t)kGpsMulti) {
                    dt = ic.decode(d, times((int32_t)m), m < 10 ? 2 : 3);
                } else if (m == (uint32_t)kGpsMulti) {
                    dt = ic.decode(d, times(kGpsMulti), 4);
                    extreme_step(dt);
                } else {
                    const int32_t f = kGpsMulti - (int32_t)m;   // -1 .. -10
                    if (f > kGpsMultiMinus) {
                        dt = ic.decode(d, times(f), 5);
                    } else {
                        dt = ic.decode(d, times(kGpsMultiMinus), 6);
                        extreme_step(dt);
                    }
                }
                t[last] += (uint64_t)(int64_t)dt;
                return true;
            }
            if (m < cf) return true;   // v2: unchanged
            if (m == cf) { full(d); return true; }
            last = (last + m - cf) & 3;
        }
        return false;
    }
};

// ================================================================ formats 0-5
// POINT10 v2: the 20-byte record; changed-field mask, deltas with medians.
struct Point10 {
    Decoder* d = nullptr;
    SymModel m_changed, m_angle[2];
    SymModel m_bits[256], m_class[256], m_user[256];
    IntDecoder ic_intensity, ic_psid, ic_dx, ic_dy, ic_z;
    uint8_t last[20];
    uint16_t last_intensity[16];
    int32_t last_height[8];
    Median5 mx[16], my[16];

    void make() {
        m_changed.make(64);
        m_angle[0].make(256);
        m_angle[1].make(256);
        ic_intensity.make(16, 4);
        ic_psid.make(16, 1);
        ic_dx.make(32, 2);
        ic_dy.make(32, 22);
        ic_z.make(32, 20);
    }
    void init(const uint8_t* rec) {
        m_changed.reset();
        m_angle[0].reset();
        m_angle[1].reset();
        for (int i = 0; i < 256; i++) m_bits[i].live = m_class[i].live = m_user[i].live = false;
        ic_intensity.reset();
        ic_psid.reset();
        ic_dx.reset();
        ic_dy.reset();
        ic_z.reset();
        for (int i = 0; i < 16; i++) { mx[i].reset(); my[i].reset(); last_intensity[i] = 0; }
        for (int i = 0; i < 8; i++) last_height[i] = 0;
        memcpy(last, rec, 20);
        last[12] = last[13] = 0;   // the intensity starts from zero
    }
    static uint32_t lazy(Decoder& d, SymModel& m) {
        if (!m.live) {
            if (m.n == 0) m.make(256);
            else m.reset();
        }
        return d.symbol(m);
    }
    void read(uint8_t* out) {
        Decoder& D = *d;
        const uint32_t changed = D.symbol(m_changed);
        uint32_t r = last[14] & 7, n = (last[14] >> 3) & 7;
        if (changed) {
            if (changed & 32) last[14] = (uint8_t)lazy(D, m_bits[last[14]]);
            r = last[14] & 7;
            n = (last[14] >> 3) & 7;
            const uint32_t m = kReturnMap8[n][r];
            if (changed & 16) {
                const uint32_t v = (uint32_t)ic_intensity.decode(D, last_intensity[m], m < 3 ? m : 3);
                last_intensity[m] = (uint16_t)v;
            }
            wr16(last + 12, last_intensity[m]);
            if (changed & 8) last[15] = (uint8_t)lazy(D, m_class[last[15]]);
            if (changed & 4) {
                const uint32_t v = D.symbol(m_angle[(last[14] >> 6) & 1]);
                last[16] = (uint8_t)fold8((int32_t)v + last[16]);
            }
            if (changed & 2) last[17] = (uint8_t)lazy(D, m_user[last[17]]);
            if (changed & 1) wr16(last + 18, (uint32_t)ic_psid.decode(D, (int32_t)rd16(last + 18)));
        }
        const uint32_t m = kReturnMap8[n][r], l = return_level(n, r, 7);
        if (!changed) wr16(last + 12, last_intensity[m]);
        int32_t diff = ic_dx.decode(D, mx[m].get(), n == 1);
        wr32(last, rd32(last) + (uint32_t)diff);
        mx[m].add(diff);
        uint32_t kb = ic_dx.k;
        diff = ic_dy.decode(D, my[m].get(), (n == 1) + (kb < 20 ? kb & ~1u : 20));
        wr32(last + 4, rd32(last + 4) + (uint32_t)diff);
        my[m].add(diff);
        kb = (ic_dx.k + ic_dy.k) / 2;
        const int32_t z = ic_z.decode(D, last_height[l], (n == 1) + (kb < 18 ? kb & ~1u : 18));
        wr32(last + 8, (uint32_t)z);
        last_height[l] = z;
        memcpy(out, last, 20);
    }
};

struct Gps11 {
    Decoder* d = nullptr;
    GpsTime g;
    void make() { g.make(false); }
    void init(const uint8_t* rec) { g.reset(rd64(rec)); }
    bool read(uint8_t* out) {
        const bool ok = g.read(*d);
        wr64(out, g.t[g.last]);
        return ok;
    }
};

// RGB12 v2 / RGB14 v3 share the colour predictor: low and high bytes of
// red, then green and blue predicted from red's change.
struct RgbModels {
    SymModel used, diff[6];
    void make() {
        used.make(128);
        for (auto& m : diff) m.make(256);
    }
    void reset() {
        used.reset();
        for (auto& m : diff) m.reset();
    }
};
inline void rgb_decode(Decoder& D, RgbModels& M, const uint16_t* last, uint16_t* out) {
    const uint32_t sym = D.symbol(M.used);
    uint32_t c0 = (sym & 1) ? fold8((int32_t)D.symbol(M.diff[0]) + (last[0] & 255)) : (last[0] & 0xFF);
    c0 |= (sym & 2) ? fold8((int32_t)D.symbol(M.diff[1]) + (last[0] >> 8)) << 8 : (last[0] & 0xFF00);
    uint32_t c1, c2;
    if (sym & 64) {
        int32_t diff = (int32_t)(c0 & 0xFF) - (int32_t)(last[0] & 0xFF);
        c1 = (sym & 4) ? fold8((int32_t)D.symbol(M.diff[2]) + clamp8(diff + (last[1] & 255))) : (last[1] & 0xFF);
        if (sym & 16) {
            diff = (diff + ((int32_t)(c1 & 0xFF) - (int32_t)(last[1] & 0xFF))) / 2;
            c2 = fold8((int32_t)D.symbol(M.diff[4]) + clamp8(diff + (last[2] & 255)));
        } else {
            c2 = last[2] & 0xFF;
        }
        diff = (int32_t)(c0 >> 8) - (int32_t)(last[0] >> 8);
        c1 |= (sym & 8) ? fold8((int32_t)D.symbol(M.diff[3]) + clamp8(diff + (last[1] >> 8))) << 8
                        : (last[1] & 0xFF00);
        if (sym & 32) {
            diff = (diff + ((int32_t)(c1 >> 8) - (int32_t)(last[1] >> 8))) / 2;
            c2 |= fold8((int32_t)D.symbol(M.diff[5]) + clamp8(diff + (last[2] >> 8))) << 8;
        } else {
            c2 |= last[2] & 0xFF00;
        }
    } else {
        c1 = c2 = c0;
    }
    out[0] = (uint16_t)c0;
    out[1] = (uint16_t)c1;
    out[2] = (uint16_t)c2;
}

struct Rgb12 {
    Decoder* d = nullptr;
    RgbModels M;
    uint16_t last[3];
    void make() { M.make(); }
    void init(const uint8_t* rec) {
        M.reset();
        for (int i = 0; i < 3; i++) last[i] = (uint16_t)rd16(rec + 2 * i);
    }
    void read(uint8_t* out) {
        uint16_t c[3];
        rgb_decode(*d, M, last, c);
        for (int i = 0; i < 3; i++) { last[i] = c[i]; wr16(out + 2 * i, c[i]); }
    }
};

struct Byte10 {
    Decoder* d = nullptr;
    uint32_t n = 0;
    std::vector<SymModel> m;
    std::vector<uint8_t> last;
    void make(uint32_t count) {
        n = count;
        m.assign(n, SymModel());
        for (auto& x : m) x.make(256);
        last.assign(n, 0);
    }
    void init(const uint8_t* rec) {
        for (auto& x : m) x.reset();
        memcpy(last.data(), rec, n);
    }
    void read(uint8_t* out) {
        for (uint32_t i = 0; i < n; i++) last[i] = (uint8_t)fold8((int32_t)last[i] + (int32_t)d->symbol(m[i]));
        memcpy(out, last.data(), n);
    }
};

// ================================================================ formats 6-10
// The LAS 1.4 point fields the layered coder tracks.
struct P14 {
    int32_t x, y, z;
    uint32_t intensity, r, n, class_flags, channel, scan_dir, edge, cls, user, psid;
    int32_t angle;            // int16 scan angle
    uint64_t gps;
    bool gps_changed;         // the last point's GPS time changed
    void load(const uint8_t* b) {
        x = (int32_t)rd32(b);
        y = (int32_t)rd32(b + 4);
        z = (int32_t)rd32(b + 8);
        intensity = rd16(b + 12);
        r = b[14] & 15;
        n = b[14] >> 4;
        class_flags = b[15] & 15;
        channel = (b[15] >> 4) & 3;
        scan_dir = (b[15] >> 6) & 1;
        edge = b[15] >> 7;
        cls = b[16];
        user = b[17];
        angle = (int16_t)rd16(b + 18);
        psid = rd16(b + 20);
        gps = rd64(b + 22);
        gps_changed = false;
    }
    void store(uint8_t* b) const {
        wr32(b, (uint32_t)x);
        wr32(b + 4, (uint32_t)y);
        wr32(b + 8, (uint32_t)z);
        wr16(b + 12, intensity);
        b[14] = (uint8_t)(r | (n << 4));
        b[15] = (uint8_t)(class_flags | (channel << 4) | (scan_dir << 6) | (edge << 7));
        b[16] = (uint8_t)cls;
        b[17] = (uint8_t)user;
        wr16(b + 18, (uint32_t)angle);
        wr16(b + 20, psid);
        wr64(b + 22, gps);
    }
};

// One layer: its bytes inside the chunk and its decoder.
struct Layer {
    Decoder d;
    bool on = false;   // the layer has bytes in this chunk and is decoded
    uint32_t size = 0;
};

enum : uint32_t { kSelAll = 0, kSelXyzRgb = 1 };

struct Point14 {
    enum { XY, Z, CLS, FLAGS, INTENSITY, ANGLE, USER, PSID, GPS, NLAYERS };
    Layer L[NLAYERS];
    struct Ctx {
        bool used = false, made = false;
        P14 last;
        SymModel m_changed[8], m_channel, m_nret[16], m_rnum[16], m_rnum_same;
        IntDecoder ic_dx, ic_dy, ic_z, ic_intensity, ic_angle, ic_psid;
        SymModel m_class[64], m_flags[64], m_user[64];
        uint16_t last_intensity[8];
        int32_t last_z[8];
        Median5 mx[12], my[12];
        GpsTime g;
    };
    Ctx c[4];
    uint32_t cur = 0;
    uint32_t select = kSelAll;
    bool v4 = false;

    void create(uint32_t k, const P14& from) {
        Ctx& C = c[k];
        if (!C.made) {
            for (auto& m : C.m_changed) m.make(128);
            C.m_channel.make(3);
            C.m_rnum_same.make(13);
            C.ic_dx.make(32, 2);
            C.ic_dy.make(32, 22);
            C.ic_z.make(32, 20);
            C.ic_intensity.make(16, 4);
            C.ic_angle.make(16, 2);
            C.ic_psid.make(16, 1);
            C.g.make(true);
            C.made = true;
        }
        for (auto& m : C.m_changed) m.reset();
        C.m_channel.reset();
        C.m_rnum_same.reset();
        for (int i = 0; i < 16; i++) C.m_nret[i].live = C.m_rnum[i].live = false;
        for (int i = 0; i < 64; i++) C.m_class[i].live = C.m_flags[i].live = C.m_user[i].live = false;
        C.ic_dx.reset();
        C.ic_dy.reset();
        C.ic_z.reset();
        C.ic_intensity.reset();
        C.ic_angle.reset();
        C.ic_psid.reset();
        for (int i = 0; i < 12; i++) { C.mx[i].reset(); C.my[i].reset(); }
        for (int i = 0; i < 8; i++) { C.last_z[i] = from.z; C.last_intensity[i] = (uint16_t)from.intensity; }
        C.g.reset(from.gps);
        C.last = from;
        C.last.gps_changed = false;
        C.used = true;
    }
    static uint32_t lazy(Decoder& d, SymModel& m, uint32_t symbols) {
        if (!m.live) {
            if (m.n == 0) m.make(symbols);
            else m.reset();
        }
        return d.symbol(m);
    }
    // sizes[] holds the nine layer sizes; data the layers in order.
    void init(const uint8_t* rec, const uint32_t* sizes, const uint8_t* data) {
        uint64_t off = 0;
        for (int i = 0; i < NLAYERS; i++) {
            L[i].size = sizes[i];
            const bool wanted = select == kSelAll || i == XY || i == Z;
            L[i].on = wanted && sizes[i] > 0;
            if (i == XY || L[i].on) L[i].d.start(data + off, sizes[i]);
            off += sizes[i];
        }
        for (auto& C : c) C.used = false;
        P14 p;
        p.load(rec);
        cur = p.channel;
        create(cur, p);
    }
    bool overrun() const {
        for (int i = 0; i < NLAYERS; i++)
            if ((i == XY || L[i].on) && L[i].d.in.overrun) return true;
        return false;
    }
    // Decodes one point into out (30 bytes) and sets the context the other
    // items of the point use: version 4 hands over the scanner channel of
    // every point, version 3 only that of a point whose channel changed, and
    // context 0 otherwise.
    bool read(uint8_t* out, uint32_t* context) {
        Decoder& D = L[XY].d;
        P14* last = &c[cur].last;
        uint32_t lpr = (last->r == 1 ? 1 : 0) + (last->r >= last->n ? 2 : 0) + (last->gps_changed ? 4 : 0);
        const uint32_t changed = D.symbol(c[cur].m_changed[lpr]);
        const bool switched = changed & 64;
        if (switched) {   // another scanner channel
            const uint32_t k = (cur + D.symbol(c[cur].m_channel) + 1) % 4;
            if (!c[k].used) create(k, c[cur].last);
            cur = k;
            last = &c[cur].last;
            last->channel = k;
        }
        Ctx& C = c[cur];
        const bool psid_changed = changed & 32, gps_changed = changed & 16, angle_changed = changed & 8;
        const uint32_t last_n = last->n, last_r = last->r;
        uint32_t n = last_n, r = last_r;
        if (changed & 4) n = last->n = lazy(D, C.m_nret[last_n], 16);
        switch (changed & 3) {
            case 0: break;
            case 1: r = (last_r + 1) % 16; break;
            case 2: r = (last_r + 15) % 16; break;
            default:
                r = gps_changed ? lazy(D, C.m_rnum[last_r], 16) : (last_r + D.symbol(C.m_rnum_same) + 2) % 16;
        }
        last->r = r;
        const uint32_t m = kReturnMap6ctx[n][r], l = return_level(n, r, 7);
        const uint32_t cpr = (r == 1 ? 2 : 0) + (r >= n ? 1 : 0);
        const uint32_t mi = (m << 1) | (gps_changed ? 1 : 0);
        int32_t diff = C.ic_dx.decode(D, C.mx[mi].get(), n == 1);
        last->x = (int32_t)((uint32_t)last->x + (uint32_t)diff);
        C.mx[mi].add(diff);
        uint32_t kb = C.ic_dx.k;
        diff = C.ic_dy.decode(D, C.my[mi].get(), (n == 1) + (kb < 20 ? kb & ~1u : 20));
        last->y = (int32_t)((uint32_t)last->y + (uint32_t)diff);
        C.my[mi].add(diff);
        if (L[Z].on) {
            kb = (C.ic_dx.k + C.ic_dy.k) / 2;
            last->z = C.ic_z.decode(L[Z].d, C.last_z[l], (n == 1) + (kb < 18 ? kb & ~1u : 18));
            C.last_z[l] = last->z;
        }
        if (L[CLS].on) {
            const uint32_t ccc = ((last->cls & 0x1F) << 1) + (cpr == 3 ? 1 : 0);
            last->cls = lazy(L[CLS].d, C.m_class[ccc], 256);
        }
        if (L[FLAGS].on) {
            const uint32_t lf = (last->edge << 5) | (last->scan_dir << 4) | last->class_flags;
            const uint32_t f = lazy(L[FLAGS].d, C.m_flags[lf], 64);
            last->edge = (f >> 5) & 1;
            last->scan_dir = (f >> 4) & 1;
            last->class_flags = f & 15;
        }
        if (L[INTENSITY].on) {
            const uint32_t ix = (cpr << 1) | (gps_changed ? 1 : 0);
            const uint32_t v = (uint16_t)C.ic_intensity.decode(L[INTENSITY].d, C.last_intensity[ix], cpr);
            C.last_intensity[ix] = (uint16_t)v;
            last->intensity = v;
        }
        if (L[ANGLE].on && angle_changed)
            last->angle = (int16_t)(uint16_t)C.ic_angle.decode(L[ANGLE].d, last->angle, gps_changed ? 1 : 0);
        if (L[USER].on) last->user = lazy(L[USER].d, C.m_user[last->user / 4], 256);
        if (L[PSID].on && psid_changed) last->psid = (uint16_t)C.ic_psid.decode(L[PSID].d, (int32_t)last->psid);
        bool ok = true;
        if (L[GPS].on && gps_changed) {
            ok = C.g.read(L[GPS].d);
            last->gps = C.g.t[C.g.last];
        }
        last->store(out);
        last->gps_changed = gps_changed;
        *context = v4 || switched ? cur : 0;
        return ok;
    }
};

// RGB14 / RGBNIR14, v3 or v4. On a switch to a scanner channel already seen
// in the chunk, version 3 keeps predicting from (and updating) the previous
// channel's last colour; version 4 uses the channel's own.
struct Rgb14 {
    Layer rgb, nir;
    bool has_nir = false, v4 = false;
    uint32_t select = kSelAll;
    struct Ctx {
        bool used = false, made = false;
        RgbModels M;
        SymModel nir_used, nir_diff[2];
        uint16_t last[4];
    };
    Ctx c[4];
    uint32_t cur = 0;

    void create(uint32_t k, const uint16_t* from) {
        Ctx& C = c[k];
        if (!C.made) {
            C.M.make();
            C.nir_used.make(4);
            C.nir_diff[0].make(256);
            C.nir_diff[1].make(256);
            C.made = true;
        }
        C.M.reset();
        C.nir_used.reset();
        C.nir_diff[0].reset();
        C.nir_diff[1].reset();
        memcpy(C.last, from, sizeof(C.last));
        C.used = true;
    }
    void init(const uint8_t* rec, uint32_t context, const uint32_t* sizes, const uint8_t* data) {
        rgb.size = sizes[0];
        rgb.on = sizes[0] > 0;
        if (rgb.on) rgb.d.start(data, sizes[0]);
        nir.size = has_nir ? sizes[1] : 0;
        nir.on = has_nir && select == kSelAll && nir.size > 0;
        if (nir.on) nir.d.start(data + sizes[0], nir.size);
        for (auto& C : c) C.used = false;
        uint16_t v[4] = {0, 0, 0, 0};
        for (int i = 0; i < (has_nir ? 4 : 3); i++) v[i] = (uint16_t)rd16(rec + 2 * i);
        cur = context;
        create(cur, v);
    }
    bool overrun() const { return (rgb.on && rgb.d.in.overrun) || (nir.on && nir.d.in.overrun); }
    void read(uint8_t* out, uint32_t context) {
        uint16_t* last = c[cur].last;
        if (context != cur) {
            cur = context;
            if (!c[cur].used) {
                create(cur, last);
                last = c[cur].last;
            } else if (v4) {
                last = c[cur].last;
            }
        }
        Ctx& C = c[cur];
        uint16_t v[3];
        if (rgb.on) {
            rgb_decode(rgb.d, C.M, last, v);
            memcpy(last, v, 6);
        } else {
            memcpy(v, last, 6);
        }
        for (int i = 0; i < 3; i++) wr16(out + 2 * i, v[i]);
        if (!has_nir) return;
        uint32_t w = last[3];
        if (nir.on) {
            const uint32_t sym = nir.d.symbol(C.nir_used);
            w = (sym & 1) ? fold8((int32_t)nir.d.symbol(C.nir_diff[0]) + (last[3] & 255)) : (last[3] & 0xFF);
            w |= (sym & 2) ? fold8((int32_t)nir.d.symbol(C.nir_diff[1]) + (last[3] >> 8)) << 8 : (last[3] & 0xFF00);
            last[3] = (uint16_t)w;
        }
        wr16(out + 6, w);
    }
};

struct Byte14 {
    std::vector<Layer> L;
    bool v4 = false;
    uint32_t select = kSelAll;
    uint32_t n = 0;
    struct Ctx {
        bool used = false;
        std::vector<SymModel> m;
        std::vector<uint8_t> last;
    };
    Ctx c[4];
    uint32_t cur = 0;
    void make(uint32_t count) {
        n = count;
        L.assign(n, Layer());
    }
    void create(uint32_t k, const uint8_t* from) {
        Ctx& C = c[k];
        if (C.m.empty()) {
            C.m.assign(n, SymModel());
            for (auto& x : C.m) x.make(256);
            C.last.assign(n, 0);
        } else {
            for (auto& x : C.m) x.reset();
        }
        memcpy(C.last.data(), from, n);
        C.used = true;
    }
    void init(const uint8_t* rec, uint32_t context, const uint32_t* sizes, const uint8_t* data) {
        uint64_t off = 0;
        for (uint32_t i = 0; i < n; i++) {
            L[i].size = sizes[i];
            L[i].on = select == kSelAll && sizes[i] > 0;
            if (L[i].on) L[i].d.start(data + off, sizes[i]);
            off += sizes[i];
        }
        for (auto& C : c) C.used = false;
        cur = context;
        create(cur, rec);
    }
    bool overrun() const {
        for (const auto& l : L)
            if (l.on && l.d.in.overrun) return true;
        return false;
    }
    void read(uint8_t* out, uint32_t context) {
        uint8_t* last = c[cur].last.data();
        if (context != cur) {
            cur = context;
            if (!c[cur].used) {
                create(cur, last);
                last = c[cur].last.data();
            } else if (v4) {
                last = c[cur].last.data();
            }
        }
        Ctx& C = c[cur];
        for (uint32_t i = 0; i < n; i++) {
            if (L[i].on) last[i] = (uint8_t)fold8((int32_t)last[i] + (int32_t)L[i].d.symbol(C.m[i]));
            out[i] = last[i];
        }
    }
};

}  // namespace

// ---------------------------------------------------------------- the file
enum ItemType : uint16_t { BYTE = 0, POINT10 = 6, GPSTIME11 = 7, RGB12 = 8, POINT14 = 10, RGB14 = 11,
                           RGBNIR14 = 12, BYTE14 = 14 };

struct File {
    const uint8_t* bytes = nullptr;
    uint64_t size = 0;
    uint64_t num_points = 0;
    uint32_t offset_to_points = 0;
    uint32_t record_length = 0;
    uint8_t point_format = 0;
    uint16_t compressor = 0;
    uint32_t chunk_size = 0;
    uint16_t num_items = 0;
    uint16_t item_type[8] = {}, item_size[8] = {}, item_version[8] = {};
    // chunk c spans bytes [chunk_start[c], chunk_start[c + 1]) and points
    // [chunk_first[c], chunk_first[c + 1]); empty when the table is missing
    std::vector<uint64_t> chunk_start, chunk_first;
};

namespace {

// The chunk table: version, count, then per chunk (point count when chunks
// vary,) byte size, each predicted from the previous chunk's value.
int read_chunk_table(File& f) {
    const uint64_t p0 = f.offset_to_points;
    PCR_CHECK(p0 + 8 <= f.size, "laz: point data ends before the chunk table pointer");
    int64_t table = (int64_t)rd64(f.bytes + p0);
    const uint64_t first = p0 + 8;
    if ((uint64_t)table + 8 == first || table == 0) {   // the writer stopped before the table
        PCR_CHECK(f.chunk_size != 0xFFFFFFFFu, "laz: variable-size chunks without a chunk table");
        f.chunk_start.clear();
        f.chunk_first.clear();
        return PCR_OK;
    }
    if (table == -1) {   // written to a stream: the pointer is at the file's end
        PCR_CHECK(f.size >= 8, "laz: file too short for the chunk table pointer");
        table = (int64_t)rd64(f.bytes + f.size - 8);
    }
    PCR_CHECK(table >= (int64_t)first && (uint64_t)table + 8 <= f.size, "laz: chunk table at %lld outside the file",
              (long long)table);
    const uint8_t* t = f.bytes + table;
    PCR_CHECK(rd32(t) == 0, "laz: chunk table version %u (0 expected)", rd32(t));
    const uint32_t nchunks = rd32(t + 4);
    PCR_CHECK((uint64_t)nchunks <= f.size, "laz: %u chunks in a %llu-byte file", nchunks,
              (unsigned long long)f.size);
    f.chunk_start.assign(nchunks + 1, 0);
    f.chunk_first.assign(nchunks + 1, 0);
    f.chunk_start[0] = first;
    if (nchunks) {
        Decoder d;
        d.start(t + 8, f.size - (uint64_t)table - 8);
        IntDecoder ic;
        ic.make(32, 2);
        const bool variable = f.chunk_size == 0xFFFFFFFFu;
        uint32_t pc = 0, pb = 0;
        for (uint32_t i = 1; i <= nchunks; i++) {
            if (variable) pc = (uint32_t)ic.decode(d, (int32_t)pc, 0);
            pb = (uint32_t)ic.decode(d, (int32_t)pb, 1);
            f.chunk_first[i] = f.chunk_first[i - 1] + (variable ? pc : f.chunk_size);
            f.chunk_start[i] = f.chunk_start[i - 1] + pb;
            PCR_CHECK(f.chunk_start[i] > f.chunk_start[i - 1] && f.chunk_start[i] <= (uint64_t)table,
                      "laz: chunk %u ends at %llu, outside the point data", i - 1,
                      (unsigned long long)f.chunk_start[i]);
        }
        PCR_CHECK(!d.in.overrun, "laz: chunk table truncated");
    }
    PCR_CHECK(f.chunk_first[nchunks] >= f.num_points, "laz: %u chunks hold %llu points, the header %llu", nchunks,
              (unsigned long long)f.chunk_first[nchunks], (unsigned long long)f.num_points);
    return PCR_OK;
}

}  // namespace

int open_file(const uint8_t* b, uint64_t size, File& f) {
    PCR_CHECK(b && size >= 227 && memcmp(b, "LASF", 4) == 0, "laz: not a LAS file");
    f = File();
    f.bytes = b;
    f.size = size;
    const uint32_t header_size = rd16(b + 94);
    f.offset_to_points = rd32(b + 96);
    const uint32_t nvlr = rd32(b + 100);
    f.point_format = b[104] & 0x3F;
    f.record_length = rd16(b + 105);
    const bool v14 = !(b[24] == 1 && b[25] < 4);
    PCR_CHECK(!v14 || size >= 255, "laz: LAS 1.4 header truncated");
    f.num_points = v14 ? rd64(b + 247) : rd32(b + 107);
    if (f.num_points == 0 && v14) f.num_points = rd32(b + 107);
    PCR_CHECK(header_size >= 227 && header_size <= f.offset_to_points && f.offset_to_points <= size,
              "laz: header size %u / point offset %u outside the %llu-byte file", header_size, f.offset_to_points,
              (unsigned long long)size);
    // the LASzip VLR: user "laszip encoded", record 22204
    uint64_t p = header_size;
    bool found = false;
    for (uint32_t i = 0; i < nvlr && p + 54 <= f.offset_to_points; i++) {
        const uint32_t rec = rd16(b + p + 18), len = rd16(b + p + 20);
        PCR_CHECK(p + 54 + len <= f.offset_to_points, "laz: VLR %u runs past the point data", i);
        if (rec == 22204 && memcmp(b + p + 2, "laszip encoded", 14) == 0) {
            const uint8_t* v = b + p + 54;
            PCR_CHECK(len >= 34 && (len - 34) % 6 == 0, "laz: LASzip VLR of %u bytes", len);
            f.compressor = (uint16_t)rd16(v);
            PCR_CHECK(rd16(v + 2) == 0, "laz: coder %u (arithmetic, 0, expected)", rd16(v + 2));
            f.chunk_size = rd32(v + 12);
            f.num_items = (uint16_t)rd16(v + 32);
            PCR_CHECK(f.num_items >= 1 && f.num_items <= 8 && 34u + 6u * f.num_items == len,
                      "laz: %u items in a %u-byte VLR", f.num_items, len);
            for (int k = 0; k < f.num_items; k++) {
                f.item_type[k] = (uint16_t)rd16(v + 34 + 6 * k);
                f.item_size[k] = (uint16_t)rd16(v + 36 + 6 * k);
                f.item_version[k] = (uint16_t)rd16(v + 38 + 6 * k);
            }
            found = true;
        }
        p += 54 + len;
    }
    PCR_CHECK(found, "laz: no LASzip VLR (is the file compressed?)");
    PCR_CHECK(f.compressor == 2 || f.compressor == 3,
              "laz: compressor %u not supported (2 pointwise chunked, 3 layered chunked)", f.compressor);
    PCR_CHECK(f.chunk_size > 0, "laz: chunk size 0");
    uint32_t total = 0;
    for (int k = 0; k < f.num_items; k++) {
        const uint16_t t = f.item_type[k], s = f.item_size[k], v = f.item_version[k];
        bool ok;
        if (f.compressor == 2)
            ok = (t == POINT10 && s == 20 && k == 0) || (t == GPSTIME11 && s == 8) || (t == RGB12 && s == 6) ||
                 (t == BYTE && s >= 1);
        else
            ok = (t == POINT14 && s == 30 && k == 0) || (t == RGB14 && s == 6) || (t == RGBNIR14 && s == 8) ||
                 (t == BYTE14 && s >= 1);
        PCR_CHECK(ok, "laz: item %d (type %u, %u bytes) not supported by compressor %u", k, t, s, f.compressor);
        PCR_CHECK(f.compressor == 2 ? v == 2 : (v == 3 || v == 4),
                  "laz: item type %u version %u not supported (2 for formats 0-5, 3 or 4 for 6-10)", t, v);
        total += s;
    }
    PCR_CHECK(total == f.record_length, "laz: items add up to %u bytes, the header's records are %u", total,
              f.record_length);
    return read_chunk_table(f);
}

inline uint64_t sum_of(const uint32_t* v, uint32_t n) {
    uint64_t s = 0;
    for (uint32_t i = 0; i < n; i++) s += v[i];
    return s;
}

// One decoding worker: its item decoders are made once and restarted per chunk.
struct Worker {
    const File* f = nullptr;
    uint32_t select = kSelAll;
    // formats 0-5
    Decoder d;
    std::unique_ptr<Point10> p10;
    std::unique_ptr<Gps11> g11;
    std::unique_ptr<Rgb12> rgb12;
    std::unique_ptr<Byte10> byte10;
    // formats 6-10
    std::unique_ptr<Point14> p14;
    std::unique_ptr<Rgb14> rgb14;
    std::unique_ptr<Byte14> byte14;
    uint32_t nlayers_rgb = 0, nbytes14 = 0;
    std::vector<uint32_t> offs;   // item offsets in a record

    void make(const File& file, uint32_t sel) {
        f = &file;
        select = sel;
        uint32_t o = 0;
        offs.clear();
        for (int k = 0; k < f->num_items; k++) {
            offs.push_back(o);
            o += f->item_size[k];
            switch (f->item_type[k]) {
                case POINT10: p10.reset(new Point10()); p10->make(); p10->d = &d; break;
                case GPSTIME11: g11.reset(new Gps11()); g11->make(); g11->d = &d; break;
                case RGB12: rgb12.reset(new Rgb12()); rgb12->make(); rgb12->d = &d; break;
                case BYTE: byte10.reset(new Byte10()); byte10->make(f->item_size[k]); byte10->d = &d; break;
                case POINT14:
                    p14.reset(new Point14());
                    p14->select = sel;
                    p14->v4 = f->item_version[k] == 4;
                    break;
                case RGB14: case RGBNIR14:
                    rgb14.reset(new Rgb14());
                    rgb14->has_nir = f->item_type[k] == RGBNIR14;
                    rgb14->v4 = f->item_version[k] == 4;
                    rgb14->select = sel;
                    nlayers_rgb = rgb14->has_nir ? 2 : 1;
                    break;
                case BYTE14:
                    byte14.reset(new Byte14());
                    byte14->make(f->item_size[k]);
                    byte14->v4 = f->item_version[k] == 4;
                    byte14->select = sel;
                    nbytes14 = f->item_size[k];
                    break;
            }
        }
    }

    // Decodes the `count` points of the chunk at [b, e) into out (count
    // records); returns the bytes consumed (pointwise chunks) or 0 on error.
    uint64_t chunk(const uint8_t* b, const uint8_t* e, uint64_t count, uint8_t* out) {
        const uint32_t rl = f->record_length;
        if (count == 0) return 1;
        if ((uint64_t)(e - b) < rl) return 0;
        memcpy(out, b, rl);   // the first record is stored raw
        const uint8_t* p = b + rl;
        if (f->compressor == 2) {
            for (int k = 0; k < f->num_items; k++) {
                const uint8_t* item = out + offs[k];
                switch (f->item_type[k]) {
                    case POINT10: p10->init(item); break;
                    case GPSTIME11: g11->init(item); break;
                    case RGB12: rgb12->init(item); break;
                    case BYTE: byte10->init(item); break;
                }
            }
            d.start(p, (uint64_t)(e - p));
            bool ok = true;
            for (uint64_t i = 1; i < count; i++) {
                uint8_t* r = out + i * rl;
                for (int k = 0; k < f->num_items; k++) {
                    uint8_t* item = r + offs[k];
                    switch (f->item_type[k]) {
                        case POINT10: p10->read(item); break;
                        case GPSTIME11: ok &= g11->read(item); break;
                        case RGB12: rgb12->read(item); break;
                        case BYTE: byte10->read(item); break;
                    }
                }
            }
            if (!ok || d.in.overrun) return 0;
            return (uint64_t)(d.in.p - b);
        }
        // layered: count, the layer sizes of every item, then the layers
        const uint32_t nsizes = 9 + nlayers_rgb + nbytes14;
        if ((uint64_t)(e - p) < 4ull * (1 + nsizes)) return 0;
        const uint32_t stored = rd32(p);
        if (stored != count) return 0;
        p += 4;
        std::vector<uint32_t> sizes(nsizes);
        uint64_t sum = 0;
        for (uint32_t i = 0; i < nsizes; i++) { sizes[i] = rd32(p + 4 * i); sum += sizes[i]; }
        p += 4ull * nsizes;
        if (sum > (uint64_t)(e - p)) return 0;
        uint32_t ctx = 0;
        const uint8_t* data = p;
        p14->init(out + offs[0], sizes.data(), data);
        data += sum_of(sizes.data(), 9);
        ctx = p14->cur;
        for (int k = 1; k < f->num_items; k++) {
            if (f->item_type[k] == RGB14 || f->item_type[k] == RGBNIR14) {
                rgb14->init(out + offs[k], ctx, sizes.data() + 9, data);
                data += sum_of(sizes.data() + 9, nlayers_rgb);
            } else {
                byte14->init(out + offs[k], ctx, sizes.data() + 9 + nlayers_rgb, data);
                data += sum_of(sizes.data() + 9 + nlayers_rgb, nbytes14);
            }
        }
        bool ok = true;
        for (uint64_t i = 1; i < count; i++) {
            uint8_t* r = out + i * rl;
            ok &= p14->read(r + offs[0], &ctx);
            for (int k = 1; k < f->num_items; k++) {
                if (f->item_type[k] == BYTE14) byte14->read(r + offs[k], ctx);
                else rgb14->read(r + offs[k], ctx);
            }
        }
        if (!ok || p14->overrun() || (rgb14 && rgb14->overrun()) || (byte14 && byte14->overrun())) return 0;
        return (uint64_t)(p + sum - b);
    }
};

}  // namespace laz
}  // namespace pcr

using namespace pcr;

namespace {

uint32_t item_offset(const laz::File& f, uint16_t type, uint16_t type2 = 0xFFFF) {
    uint32_t o = 0;
    for (int k = 0; k < f.num_items; k++) {
        if (f.item_type[k] == type || f.item_type[k] == type2) return o;
        o += f.item_size[k];
    }
    return 0xFFFFFFFFu;
}

}  // namespace

extern "C" {

int pcr_laz_info_get(const void* bytes, uint64_t size, pcr_laz_info* out) {
    PCR_CHECK(bytes && out, "laz_info_get: null argument");
    laz::File f;
    const int rc = laz::open_file(static_cast<const uint8_t*>(bytes), size, f);
    if (rc) return rc;
    memset(out, 0, sizeof(*out));
    out->num_points = f.num_points;
    out->record_length = f.record_length;
    out->num_chunks = f.chunk_start.empty() ? 0 : (uint32_t)(f.chunk_start.size() - 1);
    out->chunk_size = f.chunk_size;
    out->compressor = f.compressor;
    out->num_items = f.num_items;
    for (int k = 0; k < f.num_items; k++) {
        out->item_type[k] = f.item_type[k];
        out->item_size[k] = f.item_size[k];
        out->item_version[k] = f.item_version[k];
    }
    out->point_format = f.point_format;
    uint32_t ro = item_offset(f, laz::RGB12, laz::RGB14);
    if (ro == 0xFFFFFFFFu) ro = item_offset(f, laz::RGBNIR14);
    out->has_rgb = ro != 0xFFFFFFFFu;
    out->rgb_offset = out->has_rgb ? (uint8_t)ro : 0;
    return PCR_OK;
}

int pcr_laz_chunk_table(const void* bytes, uint64_t size, uint64_t* chunk_start, uint64_t* chunk_first,
                        uint32_t capacity, uint32_t* num_chunks) {
    PCR_CHECK(bytes && num_chunks, "laz_chunk_table: null argument");
    laz::File f;
    try {
        const int rc = laz::open_file(static_cast<const uint8_t*>(bytes), size, f);
        if (rc) return rc;
    } catch (const std::exception& e) {
        set_error("laz_chunk_table: %s", e.what());
        return PCR_E_INVALID;
    }
    const uint32_t n = f.chunk_start.empty() ? 0 : (uint32_t)(f.chunk_start.size() - 1);
    *num_chunks = n;
    if (n == 0 || (capacity == 0 && !chunk_start && !chunk_first)) return PCR_OK;   // the count only
    PCR_CHECK(capacity >= n + 1 && chunk_start && chunk_first, "laz_chunk_table: %u chunks need %u entries, got %u",
              n, n + 1, capacity);
    for (uint32_t c = 0; c <= n; c++) {
        chunk_start[c] = f.chunk_start[c];
        chunk_first[c] = std::min(f.chunk_first[c], f.num_points);
    }
    return PCR_OK;
}

static int laz_decode(const void* bytes, uint64_t size, uint64_t first_point, uint64_t num_points, void* records,
                      uint32_t fields, int threads) {
    PCR_CHECK(bytes && (records || num_points == 0), "laz_decode: null argument");
    PCR_CHECK(fields == PCR_LAZ_ALL_FIELDS || fields == PCR_LAZ_XYZ_RGB, "laz_decode: fields %u not 0 or 1", fields);
    laz::File f;
    int rc = laz::open_file(static_cast<const uint8_t*>(bytes), size, f);
    if (rc) return rc;
    PCR_CHECK(first_point <= f.num_points && num_points <= f.num_points - first_point,
              "laz_decode: points [%llu, %llu) outside the file's %llu", (unsigned long long)first_point,
              (unsigned long long)(first_point + num_points), (unsigned long long)f.num_points);
    if (num_points == 0) return PCR_OK;
    const uint64_t rl = f.record_length, end_point = first_point + num_points;
    uint8_t* out = static_cast<uint8_t*>(records);
    const uint8_t* base = f.bytes;

    if (f.chunk_start.empty()) {   // no chunk table: chunk after chunk, in order
        laz::Worker w;
        w.make(f, fields);
        std::vector<uint8_t> tmp(rl * f.chunk_size > (1ull << 31) ? 0 : rl * (uint64_t)f.chunk_size);
        PCR_CHECK(!tmp.empty(), "laz_decode: chunk of %u records too large", f.chunk_size);
        uint64_t pos = f.offset_to_points + 8, p0 = 0;
        for (uint32_t c = 0; p0 < end_point; c++) {
            const uint64_t cnt = std::min<uint64_t>(f.chunk_size, f.num_points - p0);
            const uint64_t used = w.chunk(base + pos, base + f.size, cnt, tmp.data());
            PCR_CHECK(used > 0, "laz_decode: chunk %u (points %llu..) is damaged", c, (unsigned long long)p0);
            const uint64_t a = std::max(p0, first_point), b = std::min(p0 + cnt, end_point);
            if (a < b) memcpy(out + (a - first_point) * rl, tmp.data() + (a - p0) * rl, (b - a) * rl);
            pos += used;
            p0 += cnt;
        }
        return PCR_OK;
    }

    // chunk table: the chunks that overlap the range, spread over threads
    const uint32_t nchunks = (uint32_t)(f.chunk_start.size() - 1);
    uint32_t c0 = 0;
    while (c0 < nchunks && f.chunk_first[c0 + 1] <= first_point) c0++;
    uint32_t c1 = c0;
    while (c1 < nchunks && f.chunk_first[c1] < end_point) c1++;
    int nt = threads > 0 ? threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::max(1, std::min<int>({nt, 64, (int)(c1 - c0)}));
    std::atomic<uint32_t> next{c0};
    std::atomic<int64_t> bad{-1};
    auto work = [&]() {
        laz::Worker w;
        w.make(f, fields);
        std::vector<uint8_t> tmp;
        for (uint32_t c; (c = next.fetch_add(1)) < c1 && bad.load() < 0;) {
            const uint64_t p0 = f.chunk_first[c];
            const uint64_t cnt = std::min(f.chunk_first[c + 1], f.num_points) - std::min(p0, f.num_points);
            const uint64_t a = std::max(p0, first_point), b = std::min(p0 + cnt, end_point);
            uint8_t* dst;
            if (a == p0 && b == p0 + cnt) {
                dst = out + (p0 - first_point) * rl;   // the whole chunk is wanted
            } else {
                tmp.resize(cnt * rl);
                dst = tmp.data();
            }
            const uint64_t used = w.chunk(base + f.chunk_start[c], base + f.chunk_start[c + 1], cnt, dst);
            if (used == 0) {
                bad.store(c);
                return;
            }
            if (dst == tmp.data() && a < b) memcpy(out + (a - first_point) * rl, tmp.data() + (a - p0) * rl, (b - a) * rl);
        }
    };
    if (nt == 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; t++) pool.emplace_back(work);
        for (auto& t : pool) t.join();
    }
    PCR_CHECK(bad.load() < 0, "laz_decode: chunk %lld is damaged", (long long)bad.load());
    return PCR_OK;
}

int pcr_laz_decode(const void* bytes, uint64_t size, uint64_t first_point, uint64_t num_points, void* records,
                   uint32_t fields, int threads) {
    try {
        return laz_decode(bytes, size, first_point, num_points, records, fields, threads);
    } catch (const std::exception& e) {   // allocation or thread start failed
        set_error("laz_decode: %s", e.what());
        return PCR_E_INVALID;
    }
}

}  // extern "C"
