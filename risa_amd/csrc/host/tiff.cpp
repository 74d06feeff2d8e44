// Uncompressed single-channel float32 TIFF writer for the output side
// (OfflineSaver::writeToTiff / writeTiffSequence, OfflineSaver.cu:135-178),
// written from the TIFF 6.0 and BigTIFF file layouts.  Each page carries the
// tags the reference sets -- width, length, 32 bits per sample, no
// compression, min-is-black, bilevel threshholding, one sample per pixel,
// software "RISA", IEEE floating point samples -- plus the strip tags (one
// strip per page) and contiguous planar configuration.  Pages are chained
// IFDs: a multi-page file is the reference's image sequence.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "risa_hip.h"
#include "common.h"

namespace risa {
namespace {

class Out {
public:
    Out(bool big_endian) : be_(big_endian) {}
    void u8(uint8_t v) { b_.push_back(v); }
    void put(uint64_t v, int bytes)
    {
        for (int i = 0; i < bytes; i++) {
            const int shift = be_ ? 8 * (bytes - 1 - i) : 8 * i;
            b_.push_back((uint8_t)(v >> shift));
        }
    }
    void u16(uint16_t v) { put(v, 2); }
    void u32(uint32_t v) { put(v, 4); }
    void u64(uint64_t v) { put(v, 8); }
    void f32(float v)
    {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        put(u, 4);
    }
    void pad_to(size_t align)
    {
        while (b_.size() % align) b_.push_back(0);
    }
    size_t size() const { return b_.size(); }
    void patch(size_t at, uint64_t v, int bytes)
    {
        for (int i = 0; i < bytes; i++) {
            const int shift = be_ ? 8 * (bytes - 1 - i) : 8 * i;
            b_[at + i] = (uint8_t)(v >> shift);
        }
    }
    const std::vector<uint8_t>& bytes() const { return b_; }

private:
    bool be_;
    std::vector<uint8_t> b_;
};

enum : uint16_t { T_ASCII = 2, T_SHORT = 3, T_LONG = 4, T_LONG8 = 16 };

struct Entry {
    uint16_t tag, type;
    uint64_t count, value;   // value: the number itself, or a file offset
    bool inline_bytes;       // ASCII stored in the value field
    char text[8];
};

}  // namespace

void tiff_write_f32(const std::string& path, const float* images, int count, int width, int height, int format)
{
    if (!images || count < 1 || width < 1 || height < 1) fail(RISA_ERR_ARG, "bad TIFF arguments");
    const bool big = format == RISA_TIFF_BIGTIFF;
    const bool be = format == RISA_TIFF_CLASSIC_BE;
    if (format != RISA_TIFF_BIGTIFF && format != RISA_TIFF_CLASSIC_BE && format != RISA_TIFF_CLASSIC_LE)
        fail(RISA_ERR_ARG, "unknown TIFF format %d", format);
    const uint64_t page = (uint64_t)width * height * 4;
    if (!big && (page + 512) * (uint64_t)count >= (1ull << 32))
        fail(RISA_ERR_UNSUPPORTED, "classic TIFF limited to 4 GiB: use the BigTIFF format");
    Out o(be);
    // header
    if (be) {
        o.u8('M');
        o.u8('M');
    } else {
        o.u8('I');
        o.u8('I');
    }
    size_t first_ifd_at;
    if (big) {
        o.u16(43);
        o.u16(8);
        o.u16(0);
        first_ifd_at = o.size();
        o.u64(0);
    } else {
        o.u16(42);
        first_ifd_at = o.size();
        o.u32(0);
    }
    size_t link_at = first_ifd_at;   // where the offset of the next IFD goes
    const int off_bytes = big ? 8 : 4;
    for (int k = 0; k < count; k++) {
        // pixel data of page k, rows top to bottom (TIFFWriteScanline order)
        o.pad_to(8);
        const uint64_t data_at = o.size();
        const float* img = images + (size_t)k * width * height;
        for (size_t i = 0; i < (size_t)width * height; i++) o.f32(img[i]);
        std::vector<Entry> e = {
            {256, T_LONG, 1, (uint64_t)width, false, {}},
            {257, T_LONG, 1, (uint64_t)height, false, {}},
            {258, T_SHORT, 1, 32, false, {}},
            {259, T_SHORT, 1, 1, false, {}},   // no compression
            {262, T_SHORT, 1, 1, false, {}},   // min-is-black
            {263, T_SHORT, 1, 1, false, {}},   // bilevel threshholding
            {273, big ? T_LONG8 : T_LONG, 1, data_at, false, {}},
            {277, T_SHORT, 1, 1, false, {}},
            {278, T_LONG, 1, (uint64_t)height, false, {}},
            {279, big ? T_LONG8 : T_LONG, 1, page, false, {}},
            {284, T_SHORT, 1, 1, false, {}},   // contiguous
            {305, T_ASCII, 5, 0, true, {'R', 'I', 'S', 'A', 0}},
            {339, T_SHORT, 1, 3, false, {}},   // IEEE floating point
        };
        // the software string fits the BigTIFF value field; classic TIFF
        // stores it before the IFD
        uint64_t text_at = 0;
        if (!big) {
            text_at = o.size();
            for (int i = 0; i < 5; i++) o.u8((uint8_t)e[11].text[i]);
        }
        o.pad_to(2);
        if (big) o.pad_to(8);
        const uint64_t ifd_at = o.size();
        o.patch(link_at, ifd_at, off_bytes);
        if (big) o.u64(e.size());
        else o.u16((uint16_t)e.size());
        for (const Entry& x : e) {
            o.u16(x.tag);
            o.u16(x.type);
            if (big) o.u64(x.count);
            else o.u32((uint32_t)x.count);
            const int field = big ? 8 : 4;
            if (x.type == T_ASCII) {
                if (big) {
                    for (int i = 0; i < 8; i++) o.u8(i < 5 ? (uint8_t)x.text[i] : 0);
                } else {
                    o.u32((uint32_t)text_at);
                }
            } else if (x.type == T_SHORT) {
                o.u16((uint16_t)x.value);   // left-justified in the field
                for (int i = 2; i < field; i++) o.u8(0);
            } else if (x.type == T_LONG) {
                o.u32((uint32_t)x.value);
                for (int i = 4; i < field; i++) o.u8(0);
            } else {
                o.u64(x.value);
            }
        }
        link_at = o.size();
        if (big) o.u64(0);
        else o.u32(0);
    }
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) fail(RISA_ERR_ARG, "could not open %s for writing", path.c_str());
    const size_t n = std::fwrite(o.bytes().data(), 1, o.size(), f);
    const int rc = std::fclose(f);
    if (n != o.size() || rc != 0) fail(RISA_ERR_ARG, "could not write %s", path.c_str());
}

}  // namespace risa

extern "C" int risa_tiff_write_f32(const char* path, const float* images, int count, int width, int height, int format)
{
    return risa::guarded("risa_tiff_write_f32", [&] {
        if (!path) risa::fail(RISA_ERR_ARG, "null path");
        risa::tiff_write_f32(path, images, count, width, height, format);
    });
}
