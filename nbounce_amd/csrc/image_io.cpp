// Image files of the RGBA32F accumulation target (the output texture, pathtracer.rs:192-206) for
// inspection: Portable Float Map (RGB, little-endian) and OpenEXR 2 (single-part scanline,
// uncompressed, channels A B G R as FLOAT or HALF).  Input rows are in target order: row 0 is the
// bottom of the view (the camera maps pixel row 0 to ndc y = -1, pathtracing.wgsl:29).  PFM
// stores rows bottom to top, so rows go out as given; EXR scanline 0 is the top of the image, so
// its scanline y holds target row height - 1 - y.
#include "nbounce.h"
#include "nb_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace nb {
namespace {

// IEEE binary32 -> binary16, round to nearest even; overflow to infinity, NaN kept quiet.
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xFFu;
    uint32_t mant = x & 0x7FFFFFu;
    if (exp == 0xFFu) return static_cast<uint16_t>(sign | 0x7C00u | (mant ? 0x200u | (mant >> 13) : 0u));
    const int e = static_cast<int>(exp) - 127 + 15;
    if (e >= 31) return static_cast<uint16_t>(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return static_cast<uint16_t>(sign);               // below half the smallest subnormal
        mant |= 0x800000u;                                              // implicit bit
        const uint32_t shift = static_cast<uint32_t>(14 - e);           // 24-bit significand -> subnormal
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) ++h;
        return static_cast<uint16_t>(sign | h);
    }
    uint32_t h = (static_cast<uint32_t>(e) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;           // may carry into the exponent
    return static_cast<uint16_t>(sign | h);
}

struct Out {
    std::vector<uint8_t> b;
    void bytes(const void *p, size_t n) {
        const uint8_t *q = static_cast<const uint8_t *>(p);
        b.insert(b.end(), q, q + n);
    }
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) { bytes(&v, 4); }           // little-endian hosts only (x86-64)
    void u64(uint64_t v) { bytes(&v, 8); }
    void f32(float v) { bytes(&v, 4); }
    void str(const char *s) { bytes(s, std::strlen(s) + 1); }
    void attr(const char *name, const char *type, int32_t size) {
        str(name);
        str(type);
        i32(size);
    }
};

nb_status write_file(const char *path, const std::vector<uint8_t> &data) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(NB_ERR_IO, "cannot open %s for writing", path);
    const size_t n = data.empty() ? 0 : std::fwrite(data.data(), 1, data.size(), f);
    const bool ok = n == data.size() && std::fclose(f) == 0;
    if (!ok) return set_error(NB_ERR_IO, "cannot write %s", path);
    return NB_OK;
}

nb_status check_image(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba) return set_error(NB_ERR_INVALID_ARGUMENT, "null path or pixels");
    if (width == 0 || height == 0 || width > 65536 || height > 65536)
        return set_error(NB_ERR_INVALID_ARGUMENT, "image %ux%u: sides must be in [1, 65536]", width, height);
    return NB_OK;
}

}  // namespace
}  // namespace nb

using namespace nb;

extern "C" {

nb_status nb_image_write_pfm(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    nb_status st = check_image(path, rgba, width, height);
    if (st != NB_OK) return st;
    try {
        Out o;
        char head[64];
        int n = std::snprintf(head, sizeof(head), "PF\n%u %u\n-1.0\n", width, height);
        o.bytes(head, static_cast<size_t>(n));
        o.b.reserve(o.b.size() + static_cast<size_t>(width) * height * 12);
        for (size_t i = 0; i < static_cast<size_t>(width) * height; ++i) o.bytes(rgba + 4 * i, 12);
        return write_file(path, o.b);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory writing %s", path);
    }
}

nb_status nb_image_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height, uint32_t flags) {
    nb_status st = check_image(path, rgba, width, height);
    if (st != NB_OK) return st;
    if (flags & ~static_cast<uint32_t>(NB_EXR_HALF)) return set_error(NB_ERR_INVALID_ARGUMENT, "unknown EXR flags %#x", flags);
    const bool half = flags & NB_EXR_HALF;
    try {
        Out o;
        o.i32(20000630);                       // magic 76 2f 31 01
        o.i32(2);                              // version 2, single-part scanline, short names
        const char *names[4] = {"A", "B", "G", "R"};   // channel list sorted by name
        o.attr("channels", "chlist", 4 * (2 + 16) + 1);
        for (const char *c : names) {
            o.str(c);
            o.i32(half ? 1 : 2);               // pixel type: HALF 1, FLOAT 2
            o.u8(0);                           // pLinear
            o.u8(0); o.u8(0); o.u8(0);         // reserved
            o.i32(1);                          // xSampling
            o.i32(1);                          // ySampling
        }
        o.u8(0);
        o.attr("compression", "compression", 1);
        o.u8(0);                               // NO_COMPRESSION
        for (const char *win : {"dataWindow", "displayWindow"}) {
            o.attr(win, "box2i", 16);
            o.i32(0); o.i32(0);
            o.i32(static_cast<int32_t>(width - 1)); o.i32(static_cast<int32_t>(height - 1));
        }
        o.attr("lineOrder", "lineOrder", 1);
        o.u8(0);                               // INCREASING_Y
        o.attr("pixelAspectRatio", "float", 4);
        o.f32(1.0f);
        o.attr("screenWindowCenter", "v2f", 8);
        o.f32(0.0f); o.f32(0.0f);
        o.attr("screenWindowWidth", "float", 4);
        o.f32(1.0f);
        o.u8(0);                               // end of header

        const size_t comp = half ? 2 : 4;
        const size_t line_data = static_cast<size_t>(width) * 4 * comp;
        const size_t table = o.b.size();
        const size_t first = table + 8 * static_cast<size_t>(height);
        for (uint32_t y = 0; y < height; ++y) o.u64(first + static_cast<uint64_t>(y) * (8 + line_data));
        o.b.reserve(first + static_cast<size_t>(height) * (8 + line_data));
        const int channel_of[4] = {3, 2, 1, 0};     // A B G R -> rgba component
        for (uint32_t y = 0; y < height; ++y) {
            o.i32(static_cast<int32_t>(y));
            o.i32(static_cast<int32_t>(line_data));
            const float *row = rgba + static_cast<size_t>(height - 1 - y) * width * 4;
            for (int c = 0; c < 4; ++c)
                for (uint32_t x = 0; x < width; ++x) {
                    const float v = row[4 * x + channel_of[c]];
                    if (half) {
                        uint16_t h = to_half(v);
                        o.bytes(&h, 2);
                    } else {
                        o.f32(v);
                    }
                }
        }
        return write_file(path, o.b);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory writing %s", path);
    }
}

}  // extern "C"
