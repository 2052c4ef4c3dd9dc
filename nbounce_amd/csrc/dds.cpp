// Environment maps from DDS files: the reference's EnvMap::load (src/pathtracing/envmap.rs:12) reads
// a BC6H_UF16 cube DDS with the ddsfile crate and hands the compressed blocks to the GPU, whose
// texture unit decodes them (src/common/texture.rs:89-141; the shader samples LOD 0 only,
// pathtracing.wgsl:168).  Here the file is parsed and mip 0 of the six faces decoded on the host
// to RGBA32F (a BC6H texel is a half float, exactly representable), then uploaded as any cube.
//
// Formats: BC6H_UF16 and BC6H_SF16 (the D3D11 / Vulkan BC6H format), R16G16B16A16_FLOAT and
// R32G32B32A32_FLOAT, through the DX10 header or the legacy D3DFMT codes 113 / 116.  Cube faces
// are stored face-major (+X, -X, +Y, -Y, +Z, -Z), each with its mip chain; like the reference
// (which forces the array size to 6) a file is read as six faces whatever its array size says.
//
// The BC6H decoder follows the published format: a 128-bit block, a mode in its first 2 or 5
// bits, two-region modes 1-10 (3-bit indices, 32 partitions) and one-region modes 11-14 (4-bit
// indices), endpoints stored as a base plus sign-extended deltas in the transformed modes,
// unquantised to 16 bits, interpolated with 6-bit weights and scaled by 31/64 (unsigned) or
// 31/32 (signed) into half-float bits.  Reserved modes decode to zero.  A BC6H encoder (mode 11
// only) writes files for round trips.
#include "nbounce.h"
#include "nb_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

using namespace nb;

namespace {

// ---- BC6H --------------------------------------------------------------------------------------

// Endpoint fields of a block, in the order a layout names them.
enum Field : uint8_t { RW, GW, BW, RX, GX, BX, RY, GY, BY, RZ, GZ, BZ, PART, NFIELDS };

struct Mode {
    uint8_t value;          // the mode bits' value (2-bit modes 0 and 1, else 5 bits)
    bool two_regions, transformed;
    uint8_t epb;            // endpoint precision in bits
    uint8_t delta[3];       // delta bits per channel (transformed modes)
    const char *layout;     // the fields after the mode bits, lowest block bit first
};

// Field segments "name hi:lo" (or "name bit"): the block's next bits hold bit lo, then lo+1 ...
// up to hi (or, written lo > hi, bit lo, lo-1, ... down to hi: the reversed segments of modes 13
// and 14).
const Mode kModes[14] = {
    {0x00, true, true, 10, {5, 5, 5},
     "gy4 by4 bz4 rw9:0 gw9:0 bw9:0 rx4:0 gz4 gy3:0 gx4:0 bz0 gz3:0 bx4:0 bz1 by3:0 ry4:0 bz2 rz4:0 bz3 d4:0"},
    {0x01, true, true, 7, {6, 6, 6},
     "gy5 gz4 gz5 rw6:0 bz0 bz1 by4 gw6:0 by5 bz2 gy4 bw6:0 bz3 bz5 bz4 rx5:0 gy3:0 gx5:0 gz3:0 bx5:0 by3:0 ry5:0 "
     "rz5:0 d4:0"},
    {0x02, true, true, 11, {5, 4, 4},
     "rw9:0 gw9:0 bw9:0 rx4:0 rw10 gy3:0 gx3:0 gw10 bz0 gz3:0 bx3:0 bw10 bz1 by3:0 ry4:0 bz2 rz4:0 bz3 d4:0"},
    {0x06, true, true, 11, {4, 5, 4},
     "rw9:0 gw9:0 bw9:0 rx3:0 rw10 gz4 gy3:0 gx4:0 gw10 gz3:0 bx3:0 bw10 bz1 by3:0 ry3:0 bz0 bz2 rz3:0 gy4 bz3 d4:0"},
    {0x0A, true, true, 11, {4, 4, 5},
     "rw9:0 gw9:0 bw9:0 rx3:0 rw10 by4 gy3:0 gx3:0 gw10 bz0 gz3:0 bx4:0 bw10 by3:0 ry3:0 bz1 bz2 rz3:0 bz4 bz3 d4:0"},
    {0x0E, true, true, 9, {5, 5, 5},
     "rw8:0 by4 gw8:0 gy4 bw8:0 bz4 rx4:0 gz4 gy3:0 gx4:0 bz0 gz3:0 bx4:0 bz1 by3:0 ry4:0 bz2 rz4:0 bz3 d4:0"},
    {0x12, true, true, 8, {6, 5, 5},
     "rw7:0 gz4 by4 gw7:0 bz2 gy4 bw7:0 bz3 bz4 rx5:0 gy3:0 gx4:0 bz0 gz3:0 bx4:0 bz1 by3:0 ry5:0 rz5:0 d4:0"},
    {0x16, true, true, 8, {5, 6, 5},
     "rw7:0 bz0 by4 gw7:0 gy5 gy4 bw7:0 gz5 bz4 rx4:0 gz4 gy3:0 gx5:0 gz3:0 bx4:0 bz1 by3:0 ry4:0 bz2 rz4:0 bz3 d4:0"},
    {0x1A, true, true, 8, {5, 5, 6},
     "rw7:0 bz1 by4 gw7:0 by5 gy4 bw7:0 bz5 bz4 rx4:0 gz4 gy3:0 gx4:0 bz0 gz3:0 bx5:0 by3:0 ry4:0 bz2 rz4:0 bz3 d4:0"},
    {0x1E, true, false, 6, {6, 6, 6},
     "rw5:0 gz4 bz0 bz1 by4 gw5:0 gy5 by5 bz2 gy4 bw5:0 gz5 bz3 bz5 bz4 rx5:0 gy3:0 gx5:0 gz3:0 bx5:0 by3:0 ry5:0 "
     "rz5:0 d4:0"},
    {0x03, false, false, 10, {10, 10, 10}, "rw9:0 gw9:0 bw9:0 rx9:0 gx9:0 bx9:0"},
    {0x07, false, true, 11, {9, 9, 9}, "rw9:0 gw9:0 bw9:0 rx8:0 rw10 gx8:0 gw10 bx8:0 bw10"},
    {0x0B, false, true, 12, {8, 8, 8}, "rw9:0 gw9:0 bw9:0 rx7:0 rw10:11 gx7:0 gw10:11 bx7:0 bw10:11"},
    {0x0F, false, true, 16, {4, 4, 4}, "rw9:0 gw9:0 bw9:0 rx3:0 rw10:15 gx3:0 gw10:15 bx3:0 bw10:15"},
};

// One bit of a field: block bits are assigned to (field, bit) in layout order.
struct BitRef { uint8_t field, bit; };
struct ModeBits { BitRef bits[128]; uint8_t n; };

// Expands a layout string into per-bit field references.
ModeBits expand(const char *layout) {
    ModeBits m{};
    const char *p = layout;
    while (*p) {
        while (*p == ' ') ++p;
        if (!*p) break;
        char a = p[0], b = p[1];
        p += 2;
        int field;
        if (a == 'd') { field = PART; --p; b = 0; }
        else {
            const int ch = a == 'r' ? 0 : (a == 'g' ? 1 : 2);
            const int ep = b == 'w' ? 0 : (b == 'x' ? 1 : (b == 'y' ? 2 : 3));
            field = ep * 3 + ch;
        }
        int hi = 0;
        while (*p >= '0' && *p <= '9') hi = hi * 10 + (*p++ - '0');
        int lo = hi;
        if (*p == ':') {
            ++p;
            lo = 0;
            while (*p >= '0' && *p <= '9') lo = lo * 10 + (*p++ - '0');
        }
        const int step = lo <= hi ? 1 : -1;
        for (int bit = lo;; bit += step) {
            m.bits[m.n++] = BitRef{static_cast<uint8_t>(field), static_cast<uint8_t>(bit)};
            if (bit == hi) break;
        }
    }
    return m;
}

const ModeBits &mode_bits(int k) {
    static ModeBits table[14];
    static bool init = [] {
        for (int i = 0; i < 14; ++i) table[i] = expand(kModes[i].layout);
        return true;
    }();
    (void)init;
    return table[k];
}

// Two-region partitions: bit i of kPartition[p] is the region of texel i (row-major 4x4), and the
// second region's anchor texel (its index has one bit fewer).
const uint16_t kPartition[32] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8, 0xFF00, 0xFFF0, 0xF000,
    0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110, 0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C,
};
const uint8_t kAnchor2[32] = {
    15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15,
    15, 2, 8, 2, 2, 8, 8, 15, 2, 8, 2, 2, 8, 8, 2, 2,
};
const int kWeights3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
const int kWeights4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

struct BitReader {
    const uint8_t *b;
    uint32_t pos = 0;
    uint32_t get(uint32_t n) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < n; ++i, ++pos) v |= static_cast<uint32_t>((b[pos >> 3] >> (pos & 7)) & 1u) << i;
        return v;
    }
};

int32_t sext(int32_t v, int bits) {
    const int32_t m = 1 << (bits - 1);
    v &= (1 << bits) - 1;
    return (v ^ m) - m;
}

int32_t unquantize(int32_t c, int epb, bool is_signed) {
    if (!is_signed) {
        if (epb >= 15) return c;
        if (c == 0) return 0;
        if (c == (1 << epb) - 1) return 0xFFFF;
        return ((c << 16) + 0x8000) >> epb;
    }
    if (epb >= 16) return c;
    const bool neg = c < 0;
    if (neg) c = -c;
    int32_t u;
    if (c == 0) u = 0;
    else if (c >= (1 << (epb - 1)) - 1) u = 0x7FFF;
    else u = ((c << 15) + 0x4000) >> (epb - 1);
    return neg ? -u : u;
}

uint16_t finish(int32_t v, bool is_signed) {
    if (!is_signed) return static_cast<uint16_t>((v * 31) >> 6);
    const bool neg = v < 0;
    const int32_t m = ((neg ? -v : v) * 31) >> 5;
    return static_cast<uint16_t>((neg ? 0x8000 : 0) | m);
}

// Decodes one block into 16 texels of half-float RGB (row-major).
void decode_block(const uint8_t *blk, bool is_signed, uint16_t out[48]) {
    BitReader r{blk};
    uint32_t mv = r.get(2);
    if (mv >= 2) mv |= r.get(3) << 2;
    int k = -1;
    for (int i = 0; i < 14; ++i)
        if (kModes[i].value == mv) { k = i; break; }
    if (k < 0) { std::memset(out, 0, 96); return; }      // reserved mode
    const Mode &M = kModes[k];
    const ModeBits &mb = mode_bits(k);
    int32_t f[NFIELDS] = {};
    for (uint32_t i = 0; i < mb.n; ++i) f[mb.bits[i].field] |= static_cast<int32_t>(r.get(1)) << mb.bits[i].bit;
    const int n_ep = M.two_regions ? 4 : 2;
    int32_t e[4][3];
    for (int ep = 0; ep < n_ep; ++ep)
        for (int ch = 0; ch < 3; ++ch) e[ep][ch] = f[ep * 3 + ch];
    const int32_t mask = (1 << M.epb) - 1;
    if (M.transformed) {
        for (int ch = 0; ch < 3; ++ch) {
            if (is_signed) e[0][ch] = sext(e[0][ch], M.epb);
            for (int ep = 1; ep < n_ep; ++ep) {
                int32_t v = (e[0][ch] + sext(e[ep][ch], M.delta[ch])) & mask;
                e[ep][ch] = is_signed ? sext(v, M.epb) : v;
            }
        }
    } else if (is_signed) {
        for (int ep = 0; ep < n_ep; ++ep)
            for (int ch = 0; ch < 3; ++ch) e[ep][ch] = sext(e[ep][ch], M.epb);
    }
    for (int ep = 0; ep < n_ep; ++ep)
        for (int ch = 0; ch < 3; ++ch) e[ep][ch] = unquantize(e[ep][ch], M.epb, is_signed);
    const uint32_t part = static_cast<uint32_t>(f[PART]);
    for (int t = 0; t < 16; ++t) {
        int region = 0, w;
        if (M.two_regions) {
            region = (kPartition[part] >> t) & 1;
            const bool anchor = t == 0 || t == kAnchor2[part];
            w = kWeights3[r.get(anchor ? 2 : 3)];
        } else {
            w = kWeights4[r.get(t == 0 ? 3 : 4)];
        }
        for (int ch = 0; ch < 3; ++ch) {
            const int32_t a = e[2 * region][ch], b = e[2 * region + 1][ch];
            out[3 * t + ch] = finish((a * (64 - w) + b * w + 32) >> 6, is_signed);
        }
    }
}

// Encodes 16 texels (RGB, non-negative floats) as one mode-11 block (one region, 10-bit
// endpoints, 4-bit indices): per-channel minimum and maximum as the endpoints, each texel on the
// nearest of the 16 interpolated colours along them.
void encode_block_mode11(const float *rgb, uint8_t out[16]) {
    auto to_q = [](float x) -> int32_t {   // value -> 10-bit endpoint whose decode is nearest
        if (!(x > 0.0f)) return 0;
        const float h = x > 65504.0f ? 65504.0f : x;
        // half bits of h, then invert finish (x 64/31) and unquantize (x 1024/65536)
        _Float16 hh = static_cast<_Float16>(h);
        uint16_t bits;
        std::memcpy(&bits, &hh, 2);
        const double u = bits * 64.0 / 31.0;
        long q = std::lround(u * 1023.0 / 65535.0);
        return static_cast<int32_t>(q < 0 ? 0 : (q > 1023 ? 1023 : q));
    };
    int32_t lo[3], hi[3];
    for (int ch = 0; ch < 3; ++ch) {
        float mn = rgb[ch], mx = rgb[ch];
        for (int t = 1; t < 16; ++t) { mn = std::fmin(mn, rgb[3 * t + ch]); mx = std::fmax(mx, rgb[3 * t + ch]); }
        lo[ch] = to_q(mn);
        hi[ch] = to_q(mx);
    }
    // decoded endpoint colours and the 16 palette entries as floats
    float pal[16][3];
    for (int ch = 0; ch < 3; ++ch) {
        const int32_t a = unquantize(lo[ch], 10, false), b = unquantize(hi[ch], 10, false);
        for (int i = 0; i < 16; ++i) {
            const uint16_t h = finish((a * (64 - kWeights4[i]) + b * kWeights4[i] + 32) >> 6, false);
            _Float16 hv;
            std::memcpy(&hv, &h, 2);
            pal[i][ch] = static_cast<float>(hv);
        }
    }
    int idx[16];
    for (int t = 0; t < 16; ++t) {
        int best = 0;
        double bd = 1e300;
        for (int i = 0; i < 16; ++i) {
            double d = 0;
            for (int ch = 0; ch < 3; ++ch) { const double z = pal[i][ch] - rgb[3 * t + ch]; d += z * z; }
            if (d < bd) { bd = d; best = i; }
        }
        idx[t] = best;
    }
    // texel 0's index has 3 bits: with its top bit set, swap the endpoints and mirror the indices
    if (idx[0] & 8) {
        for (int ch = 0; ch < 3; ++ch) std::swap(lo[ch], hi[ch]);
        for (int t = 0; t < 16; ++t) idx[t] = 15 - idx[t];
    }
    std::memset(out, 0, 16);
    uint32_t pos = 0;
    auto put = [&](uint32_t v, uint32_t n) {
        for (uint32_t i = 0; i < n; ++i, ++pos) out[pos >> 3] |= static_cast<uint8_t>(((v >> i) & 1u) << (pos & 7));
    };
    put(0x03, 5);
    for (int ch = 0; ch < 3; ++ch) put(static_cast<uint32_t>(lo[ch]), 10);
    for (int ch = 0; ch < 3; ++ch) put(static_cast<uint32_t>(hi[ch]), 10);
    for (int t = 0; t < 16; ++t) put(static_cast<uint32_t>(idx[t]), t == 0 ? 3 : 4);
}

float half_to_float(uint16_t h) {
    _Float16 v;
    std::memcpy(&v, &h, 2);
    return static_cast<float>(v);
}

// ---- DDS -----------------------------------------------------------------------------------------

constexpr uint32_t kMagic = 0x20534444u;        // "DDS "
constexpr uint32_t kFourCC_DX10 = 0x30315844u;  // "DX10"
constexpr uint32_t kDdpfFourCC = 0x4u;
constexpr uint32_t kDxgiRgba32f = 2, kDxgiRgba16f = 10, kDxgiBc6hUf16 = 95, kDxgiBc6hSf16 = 96;
constexpr uint32_t kD3dRgba16f = 113, kD3dRgba32f = 116;

uint32_t rd32(const uint8_t *p) { uint32_t v; std::memcpy(&v, p, 4); return v; }

enum class Fmt { RGBA32F, RGBA16F, BC6H_UF16, BC6H_SF16 };

// Bytes of one mip level of one face.
size_t level_bytes(Fmt f, uint32_t w, uint32_t h) {
    if (f == Fmt::BC6H_UF16 || f == Fmt::BC6H_SF16) return static_cast<size_t>((w + 3) / 4) * ((h + 3) / 4) * 16;
    return static_cast<size_t>(w) * h * (f == Fmt::RGBA32F ? 16 : 8);
}

// Parses a cube DDS and decodes mip 0 of the six faces into RGBA32F (6 x n x n x 4).
nb_status decode_cube(const uint8_t *d, size_t size, std::vector<float> &out, uint32_t &n) {
    if (size < 128 || rd32(d) != kMagic || rd32(d + 4) != 124)
        return set_error(NB_ERR_FORMAT, "not a DDS file (magic or header size)");
    const uint32_t height = rd32(d + 12), width = rd32(d + 16);
    uint32_t mips = rd32(d + 28);
    if (mips == 0) mips = 1;
    const uint32_t pf_flags = rd32(d + 80), fourcc = rd32(d + 84);
    size_t off = 128;
    Fmt fmt;
    if (!(pf_flags & kDdpfFourCC)) return set_error(NB_ERR_FORMAT, "DDS pixel format without a FourCC is not supported");
    if (fourcc == kFourCC_DX10) {
        if (size < 148) return set_error(NB_ERR_FORMAT, "truncated DX10 header");
        const uint32_t dxgi = rd32(d + 128);
        off = 148;
        if (dxgi == kDxgiBc6hUf16) fmt = Fmt::BC6H_UF16;
        else if (dxgi == kDxgiBc6hSf16) fmt = Fmt::BC6H_SF16;
        else if (dxgi == kDxgiRgba16f) fmt = Fmt::RGBA16F;
        else if (dxgi == kDxgiRgba32f) fmt = Fmt::RGBA32F;
        else return set_error(NB_ERR_FORMAT, "unsupported DXGI format %u (BC6H_UF16, BC6H_SF16, RGBA16F, RGBA32F)", dxgi);
    } else if (fourcc == kD3dRgba16f) {
        fmt = Fmt::RGBA16F;
    } else if (fourcc == kD3dRgba32f) {
        fmt = Fmt::RGBA32F;
    } else {
        return set_error(NB_ERR_FORMAT, "unsupported DDS FourCC %#x", fourcc);
    }
    if (width == 0 || width != height || width > 16384)
        return set_error(NB_ERR_FORMAT, "cube faces must be square and at most 16384 texels (%ux%u)", width, height);
    if (mips > 15) return set_error(NB_ERR_FORMAT, "%u mip levels", mips);
    size_t face_bytes = 0;
    for (uint32_t m = 0; m < mips; ++m)
        face_bytes += level_bytes(fmt, std::max(1u, width >> m), std::max(1u, height >> m));
    if (size < off + 6 * face_bytes)
        return set_error(NB_ERR_FORMAT, "DDS data holds %zu bytes, six faces need %zu", size - off, 6 * face_bytes);
    n = width;
    out.assign(static_cast<size_t>(6) * n * n * 4, 1.0f);
    for (int f = 0; f < 6; ++f) {
        const uint8_t *src = d + off + f * face_bytes;
        float *dst = out.data() + static_cast<size_t>(f) * n * n * 4;
        if (fmt == Fmt::RGBA32F) {
            std::memcpy(dst, src, static_cast<size_t>(n) * n * 16);
        } else if (fmt == Fmt::RGBA16F) {
            for (size_t i = 0; i < static_cast<size_t>(n) * n * 4; ++i) {
                uint16_t h;
                std::memcpy(&h, src + 2 * i, 2);
                dst[i] = half_to_float(h);
            }
        } else {
            const bool sg = fmt == Fmt::BC6H_SF16;
            const uint32_t bw = (n + 3) / 4, bh = (n + 3) / 4;
            uint16_t texels[48];
            for (uint32_t by = 0; by < bh; ++by)
                for (uint32_t bx = 0; bx < bw; ++bx) {
                    decode_block(src + 16 * (static_cast<size_t>(by) * bw + bx), sg, texels);
                    for (uint32_t t = 0; t < 16; ++t) {
                        const uint32_t x = 4 * bx + (t & 3), y = 4 * by + (t >> 2);
                        if (x >= n || y >= n) continue;
                        float *px = dst + 4 * (static_cast<size_t>(y) * n + x);
                        for (int ch = 0; ch < 3; ++ch) px[ch] = half_to_float(texels[3 * t + ch]);
                        px[3] = 1.0f;
                    }
                }
        }
    }
    return NB_OK;
}

nb_status read_file(const char *path, std::vector<uint8_t> &bytes) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(NB_ERR_IO, "cannot open %s", path);
    std::fseek(fp, 0, SEEK_END);
    const long len = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (len < 0) { std::fclose(fp); return set_error(NB_ERR_IO, "cannot size %s", path); }
    bytes.resize(static_cast<size_t>(len));
    const size_t got = len ? std::fread(bytes.data(), 1, bytes.size(), fp) : 0;
    std::fclose(fp);
    if (got != bytes.size()) return set_error(NB_ERR_IO, "short read of %s", path);
    return NB_OK;
}

nb_status copy_out(const std::vector<float> &faces, uint32_t n, float *rgba, uint64_t cap, uint32_t *face_size) {
    *face_size = n;
    if (!rgba) return NB_OK;
    if (cap < faces.size()) return set_error(NB_ERR_INVALID_ARGUMENT, "output holds %llu floats, the cube needs %zu",
                                             static_cast<unsigned long long>(cap), faces.size());
    std::memcpy(rgba, faces.data(), faces.size() * 4);
    return NB_OK;
}

}  // namespace

extern "C" {

nb_status nb_dds_read_cube_memory(const void *data, size_t size, float *rgba, uint64_t cap_floats, uint32_t *face_size) {
    if (!data || !face_size) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    try {
        std::vector<float> faces;
        uint32_t n = 0;
        nb_status st = decode_cube(static_cast<const uint8_t *>(data), size, faces, n);
        if (st != NB_OK) return st;
        return copy_out(faces, n, rgba, cap_floats, face_size);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed decoding a DDS cube");
    }
}

nb_status nb_dds_read_cube(const char *path, float *rgba, uint64_t cap_floats, uint32_t *face_size) {
    if (!path || !face_size) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    try {
        std::vector<uint8_t> bytes;
        nb_status st = read_file(path, bytes);
        if (st != NB_OK) return st;
        return nb_dds_read_cube_memory(bytes.data(), bytes.size(), rgba, cap_floats, face_size);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed reading %s", path);
    }
}

nb_status nb_envmap_load(nb_context *ctx, const char *path, nb_envmap **out) {
    if (!ctx || !path || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    try {
        std::vector<uint8_t> bytes;
        nb_status st = read_file(path, bytes);
        if (st != NB_OK) return st;
        std::vector<float> faces;
        uint32_t n = 0;
        st = decode_cube(bytes.data(), bytes.size(), faces, n);
        if (st != NB_OK) return st;
        return nb_envmap_create_cube(ctx, faces.data(), n, out);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed loading %s", path);
    }
}

nb_status nb_dds_write_cube(const char *path, const float *rgba, uint32_t face_size, uint32_t format) {
    if (!path || !rgba) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (face_size == 0 || face_size > 16384) return set_error(NB_ERR_INVALID_ARGUMENT, "face size %u", face_size);
    uint32_t dxgi;
    if (format == NB_DDS_RGBA32F) dxgi = kDxgiRgba32f;
    else if (format == NB_DDS_RGBA16F) dxgi = kDxgiRgba16f;
    else if (format == NB_DDS_BC6H_UF16) dxgi = kDxgiBc6hUf16;
    else return set_error(NB_ERR_INVALID_ARGUMENT, "unknown DDS format %u", format);
    try {
        const uint32_t n = face_size;
        const Fmt fmt = format == NB_DDS_RGBA32F ? Fmt::RGBA32F : (format == NB_DDS_RGBA16F ? Fmt::RGBA16F : Fmt::BC6H_UF16);
        std::vector<uint8_t> file(148, 0);
        auto w32 = [&](size_t at, uint32_t v) { std::memcpy(file.data() + at, &v, 4); };
        w32(0, kMagic);
        w32(4, 124);
        w32(8, 0x1u | 0x2u | 0x4u | 0x1000u | (fmt == Fmt::BC6H_UF16 ? 0x80000u : 0x8u));   // caps, height, width, pixel format, size
        w32(12, n);
        w32(16, n);
        w32(20, fmt == Fmt::BC6H_UF16 ? static_cast<uint32_t>(level_bytes(fmt, n, n)) : n * (fmt == Fmt::RGBA32F ? 16 : 8));
        w32(28, 1);                       // one mip level
        w32(76, 32);                      // pixel format size
        w32(80, kDdpfFourCC);
        w32(84, kFourCC_DX10);
        w32(108, 0x1000u | 0x8u);         // DDSCAPS_TEXTURE | DDSCAPS_COMPLEX
        w32(112, 0x200u | 0xFC00u);       // DDSCAPS2_CUBEMAP with all six faces
        w32(128, dxgi);
        w32(132, 3);                      // texture 2D
        w32(136, 0x4u);                   // TEXTURECUBE
        w32(140, 1);                      // one cube
        const size_t fb = level_bytes(fmt, n, n);
        file.resize(148 + 6 * fb);
        for (int f = 0; f < 6; ++f) {
            const float *src = rgba + static_cast<size_t>(f) * n * n * 4;
            uint8_t *dst = file.data() + 148 + f * fb;
            if (fmt == Fmt::RGBA32F) {
                std::memcpy(dst, src, fb);
            } else if (fmt == Fmt::RGBA16F) {
                for (size_t i = 0; i < static_cast<size_t>(n) * n * 4; ++i) {
                    _Float16 h = static_cast<_Float16>(src[i]);
                    std::memcpy(dst + 2 * i, &h, 2);
                }
            } else {
                const uint32_t bw = (n + 3) / 4;
                float blk[48];
                for (uint32_t by = 0; by < bw; ++by)
                    for (uint32_t bx = 0; bx < bw; ++bx) {
                        for (uint32_t t = 0; t < 16; ++t) {
                            const uint32_t x = std::min(4 * bx + (t & 3), n - 1), y = std::min(4 * by + (t >> 2), n - 1);
                            for (int ch = 0; ch < 3; ++ch) blk[3 * t + ch] = src[4 * (static_cast<size_t>(y) * n + x) + ch];
                        }
                        encode_block_mode11(blk, dst + 16 * (static_cast<size_t>(by) * bw + bx));
                    }
            }
        }
        FILE *fp = std::fopen(path, "wb");
        if (!fp) return set_error(NB_ERR_IO, "cannot create %s", path);
        const size_t put = std::fwrite(file.data(), 1, file.size(), fp);
        const int closed = std::fclose(fp);
        if (put != file.size() || closed != 0) return set_error(NB_ERR_IO, "short write of %s", path);
        return NB_OK;
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed writing %s", path);
    }
}

nb_status nb_bc6h_decode_block(const uint8_t *block, uint32_t is_signed, uint16_t *rgb_half) {
    if (!block || !rgb_half) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    decode_block(block, is_signed != 0, rgb_half);
    return NB_OK;
}

}  // extern "C"
