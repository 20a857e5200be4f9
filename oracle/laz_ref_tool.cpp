// laz_ref_tool.cpp — test infrastructure, not part of the library.
//
// Drives the LASzip library that the reference vendors (libs/laszip/src,
// compiled from its own sources into oracle/_ref by `make -C oracle ref`)
// to make the LAZ fixtures of the tests and to decode LAZ files with the
// reference's own reader:
//
//   laz_ref_tool compress IN.las OUT.laz [CHUNK] [v4]
//       CHUNK: points per chunk (default 50000) or "variable:N" for
//       variable-size chunks closed every N points; "v4" marks the LAS 1.4
//       items as version 4 (LASzip 3.2+'s context-switch fix).
//   laz_ref_tool decompress IN.laz OUT.records
//       writes the decoded point records (num_points * record_length bytes).
//
// Only the tests' fixture script and the oracle's own checks run it.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bytestreamin_file.hpp"
#include "bytestreamout_file.hpp"
#include "lasreadpoint.hpp"
#include "laswriteitemraw.hpp"
#include "laswritepoint.hpp"
#include "laszip.hpp"

namespace {

std::vector<unsigned char> read_file(const char* path) {
    std::vector<unsigned char> b;
    FILE* f = fopen(path, "rb");
    if (!f) return b;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    b.resize(n > 0 ? (size_t)n : 0);
    if (n > 0 && fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
    fclose(f);
    return b;
}
template <class T> T get(const std::vector<unsigned char>& b, size_t off) {
    T v;
    memcpy(&v, b.data() + off, sizeof(T));
    return v;
}
template <class T> void put(std::vector<unsigned char>& b, size_t off, T v) { memcpy(b.data() + off, &v, sizeof(T)); }

int fail(const char* what) {
    fprintf(stderr, "laz_ref_tool: %s\n", what);
    return 1;
}

uint64_t point_count(const std::vector<unsigned char>& b) {
    const bool v14 = !(b[24] == 1 && b[25] < 4);
    uint64_t n = v14 ? get<uint64_t>(b, 247) : get<uint32_t>(b, 107);
    if (n == 0) n = get<uint32_t>(b, 107);
    return n;
}

// A LAS 1.4 record (30 bytes) as LASzip's in-memory point, which its
// POINT14 coders read: the legacy fields plus the extended ones.
void to_internal14(const unsigned char* r, unsigned char* item) {
    memset(item, 0, 64);
    LAStempWritePoint10* p = reinterpret_cast<LAStempWritePoint10*>(item);
    memcpy(&p->X, r, 4);
    memcpy(&p->Y, r + 4, 4);
    memcpy(&p->Z, r + 8, 4);
    memcpy(&p->intensity, r + 12, 2);
    const unsigned rn = r[14] & 15, nr = r[14] >> 4, flags = r[15] & 15, ch = (r[15] >> 4) & 3;
    p->return_number = rn > 7 ? 7 : rn;
    p->number_of_returns = nr > 7 ? 7 : nr;
    p->scan_direction_flag = (r[15] >> 6) & 1;
    p->edge_of_flight_line = r[15] >> 7;
    p->classification = (unsigned char)((r[16] < 32 ? r[16] : 0) | ((flags & 7) << 5));
    p->user_data = r[17];
    memcpy(&p->point_source_ID, r + 20, 2);
    p->extended_point_type = 1;
    p->extended_scanner_channel = ch;
    p->extended_classification_flags = flags;
    p->extended_classification = r[16];
    p->extended_return_number = rn;
    p->extended_number_of_returns = nr;
    memcpy(&p->extended_scan_angle, r + 18, 2);
    memcpy(&p->gps_time, r + 22, 8);
}

int compress(const char* in, const char* out, const char* chunk_arg, bool v4) {
    std::vector<unsigned char> b = read_file(in);
    if (b.size() < 227 || memcmp(b.data(), "LASF", 4) != 0) return fail("input is not a LAS file");
    const uint32_t offset = get<uint32_t>(b, 96);
    const unsigned char format = b[104];
    const uint16_t rl = get<uint16_t>(b, 105);
    const uint64_t n = point_count(b);
    if (format & 0xC0) return fail("input is already compressed");
    if ((uint64_t)offset + n * rl > b.size()) return fail("input records exceed the file");

    LASzip zip;
    const unsigned short comp = format >= 6 ? LASZIP_COMPRESSOR_LAYERED_CHUNKED : LASZIP_COMPRESSOR_CHUNKED;
    if (!zip.setup(format, rl, comp)) return fail("LASzip::setup refused the point format");
    unsigned every = 0;
    if (chunk_arg) {
        if (strncmp(chunk_arg, "variable:", 9) == 0) {
            every = (unsigned)atoi(chunk_arg + 9);
            if (!zip.set_chunk_size(0xFFFFFFFFu)) return fail("set_chunk_size");
        } else if (!zip.set_chunk_size((unsigned)atoi(chunk_arg))) {
            return fail("set_chunk_size");
        }
    }
    if (v4)
        for (int i = 0; i < zip.num_items; i++)
            if (zip.items[i].type == LASitem::POINT14 || zip.items[i].type == LASitem::RGB14 ||
                zip.items[i].type == LASitem::RGBNIR14 || zip.items[i].type == LASitem::BYTE14)
                zip.items[i].version = 4;
    unsigned char* vlr = nullptr;
    int vlr_n = 0;
    if (!zip.pack(vlr, vlr_n)) return fail("LASzip::pack");

    // header + existing VLRs, then the LASzip VLR
    std::vector<unsigned char> head(b.begin(), b.begin() + offset);
    const size_t at = head.size();
    head.resize(at + 54 + vlr_n, 0);
    put<uint16_t>(head, at + 18, 22204);
    memcpy(head.data() + at + 2, "laszip encoded", 14);
    put<uint16_t>(head, at + 20, (uint16_t)vlr_n);
    memcpy(head.data() + at + 22, "LASzip fixture", 14);
    memcpy(head.data() + at + 54, vlr, vlr_n);   // vlr stays owned by zip
    put<uint32_t>(head, 96, (uint32_t)head.size());
    put<uint32_t>(head, 100, get<uint32_t>(head, 100) + 1);
    head[104] = (unsigned char)(format | 128);

    FILE* f = fopen(out, "wb");
    if (!f) return fail("cannot open the output");
    fwrite(head.data(), 1, head.size(), f);
    ByteStreamOutFileLE stream(f);
    LASwritePoint w;
    if (!w.setup(zip.num_items, zip.items, &zip) || !w.init(&stream)) return fail("LASwritePoint setup");
    std::vector<unsigned char> items(64 * zip.num_items);
    std::vector<unsigned char*> ptr(zip.num_items);
    for (uint64_t i = 0; i < n; i++) {
        const unsigned char* r = b.data() + offset + i * rl;
        size_t o = 0;
        for (int k = 0; k < zip.num_items; k++) {
            unsigned char* it = items.data() + 64 * k;
            if (zip.items[k].type == LASitem::POINT14) to_internal14(r + o, it);
            else memcpy(it, r + o, zip.items[k].size);
            ptr[k] = it;
            o += zip.items[k].size;
        }
        if (!w.write(ptr.data())) return fail("LASwritePoint::write");
        if (every && (i + 1) % every == 0 && i + 1 < n && !w.chunk()) return fail("LASwritePoint::chunk");
    }
    if (!w.done()) return fail("LASwritePoint::done");
    fclose(f);
    return 0;
}

int decompress(const char* in, const char* out) {
    std::vector<unsigned char> b = read_file(in);
    if (b.size() < 227 || memcmp(b.data(), "LASF", 4) != 0) return fail("input is not a LAS file");
    const uint16_t header_size = get<uint16_t>(b, 94);
    const uint32_t offset = get<uint32_t>(b, 96), nvlr = get<uint32_t>(b, 100);
    const uint16_t rl = get<uint16_t>(b, 105);
    const uint64_t n = point_count(b);
    LASzip zip;
    bool found = false;
    size_t p = header_size;
    for (uint32_t i = 0; i < nvlr && p + 54 <= offset; i++) {
        const uint16_t rec = get<uint16_t>(b, p + 18), len = get<uint16_t>(b, p + 20);
        if (rec == 22204 && memcmp(b.data() + p + 2, "laszip encoded", 14) == 0) {
            if (!zip.unpack(b.data() + p + 54, len)) return fail("LASzip::unpack");
            found = true;
        }
        p += 54 + len;
    }
    if (!found) return fail("no LASzip VLR");
    FILE* f = fopen(in, "rb");
    if (!f) return fail("cannot reopen the input");
    fseek(f, offset, SEEK_SET);
    ByteStreamInFileLE stream(f);
    LASreadPoint r;
    if (!r.setup(zip.num_items, zip.items, &zip) || !r.init(&stream)) return fail("LASreadPoint setup");
    // write the records back with LASzip's own raw item writers
    LASzip raw;
    if (!raw.setup(zip.num_items, zip.items, LASZIP_COMPRESSOR_NONE)) return fail("raw setup");
    for (int k = 0; k < raw.num_items; k++) raw.items[k].version = 0;
    FILE* o = fopen(out, "wb");
    if (!o) return fail("cannot open the output");
    ByteStreamOutFileLE ostream(o);
    LASwritePoint w;
    if (!w.setup(raw.num_items, raw.items, &raw) || !w.init(&ostream)) return fail("raw writer setup");
    std::vector<unsigned char> items(128 * zip.num_items);
    std::vector<unsigned char*> ptr(zip.num_items);
    for (int k = 0; k < zip.num_items; k++) ptr[k] = items.data() + 128 * k;
    unsigned total = 0;
    for (int k = 0; k < zip.num_items; k++) total += zip.items[k].size;
    if (total != rl) return fail("items do not add up to the record length");
    for (uint64_t i = 0; i < n; i++) {
        if (!r.read(ptr.data())) return fail("LASreadPoint::read");
        // the decoded LAS 1.4 point is an extended one (as laszip_read_point marks it)
        for (int k = 0; k < zip.num_items; k++)
            if (zip.items[k].type == LASitem::POINT14)
                reinterpret_cast<LAStempWritePoint10*>(ptr[k])->extended_point_type = 1;
        if (!w.write(ptr.data())) return fail("raw write");
    }
    r.done();
    w.done();
    fclose(o);
    fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 4 && strcmp(argv[1], "compress") == 0)
        return compress(argv[2], argv[3], argc >= 5 ? argv[4] : nullptr, argc >= 6 && strcmp(argv[5], "v4") == 0);
    if (argc == 4 && strcmp(argv[1], "decompress") == 0) return decompress(argv[2], argv[3]);
    fprintf(stderr, "usage: laz_ref_tool compress IN.las OUT.laz [CHUNK|variable:N] [v4]\n"
                    "       laz_ref_tool decompress IN.laz OUT.records\n");
    return 2;
}
