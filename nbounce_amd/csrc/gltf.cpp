// glTF 2.0 ingestion (Scene::parse_gltf, src/pathtracing/scene.rs:108-217) into a host scene of
// the reference's Vertex / index / Primitive arrays, plus the C ABI of the host scene.
//
// Semantics kept from the reference:
//   * meshes are visited in document order, their primitives in order; each primitive appends
//     its vertices (POSITION, NORMAL, TEXCOORD_0 or (0, 0), TANGENT, zipped to the shortest
//     attribute) and its indices offset by the first vertex (scene.rs:137-180);
//   * a primitive that is not a triangle list, or lacks positions, normals, tangents or indices,
//     fails the whole parse with the reference's error text (scene.rs:69-98, 141-176);
//   * every node with a mesh (document order, not the scene graph) appends one Primitive per
//     mesh primitive with the node's own local matrix only - parents are ignored (scene.rs:186);
//   * the material gives colour = (emissive, 1) when emissive_factor != 0 else base colour,
//     roughness, metallic, emissive flag 0/1; glTF defaults when absent (scene.rs:188-205).
// A TRS node's matrix is T * R * S with the quaternion-to-matrix and column products of the
// `gltf` 1.4.1 crate's math module (the expressions below).
//
// Differences, on purpose: accessors whose component type does not match the attribute are
// rejected (the reference reinterprets the bytes in release builds), out-of-range views are an
// error rather than a panic, and a failed parse leaves the host scene unchanged.  With
// NB_GLTF_SHARE_GEOMETRY, mesh primitives that name the same accessors reuse the first copy's
// vertices and index range instead of appending a duplicate, so NB_BUILD_SHARE_BLAS can build
// each geometry once.
#include "nbounce.h"
#include "nb_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

struct nb_host_scene {
    std::vector<nb_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<nb_primitive> prims;
};

namespace nb {
namespace {

// ---- JSON -----------------------------------------------------------------------------------

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> members;

    const Json *get(const char *key) const {
        if (kind != Object) return nullptr;
        for (const auto &m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct ParseError {
    std::string what;
};

[[noreturn]] void fail(const std::string &what) { throw ParseError{what}; }

class JsonReader {
public:
    JsonReader(const char *p, size_t n) : p_(p), end_(p + n) {}

    Json document() {
        Json v = value(0);
        skip_ws();
        if (p_ != end_) fail("trailing characters after the JSON document");
        return v;
    }

private:
    const char *p_, *end_;

    void skip_ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    char peek() {
        skip_ws();
        if (p_ == end_) fail("unexpected end of JSON");
        return *p_;
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "' in JSON");
        ++p_;
    }
    bool literal(const char *word) {
        size_t n = std::strlen(word);
        if (static_cast<size_t>(end_ - p_) >= n && std::memcmp(p_, word, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Json value(int depth) {
        if (depth > 64) fail("JSON nested too deeply");
        Json v;
        char c = peek();
        if (c == '{') {
            ++p_;
            v.kind = Json::Object;
            if (peek() == '}') { ++p_; return v; }
            for (;;) {
                if (peek() != '"') fail("expected a member name in JSON");
                std::string key = string_body();
                expect(':');
                v.members.emplace_back(std::move(key), value(depth + 1));
                char d = peek();
                ++p_;
                if (d == '}') break;
                if (d != ',') fail("expected ',' or '}' in JSON");
            }
        } else if (c == '[') {
            ++p_;
            v.kind = Json::Array;
            if (peek() == ']') { ++p_; return v; }
            for (;;) {
                v.items.push_back(value(depth + 1));
                char d = peek();
                ++p_;
                if (d == ']') break;
                if (d != ',') fail("expected ',' or ']' in JSON");
            }
        } else if (c == '"') {
            v.kind = Json::String;
            v.str = string_body();
        } else if (literal("true")) {
            v.kind = Json::Bool;
            v.b = true;
        } else if (literal("false")) {
            v.kind = Json::Bool;
        } else if (literal("null")) {
            v.kind = Json::Null;
        } else {
            v.kind = Json::Number;
            v.num = number();
        }
        return v;
    }

    double number() {
        const char *s = p_;
        if (p_ < end_ && *p_ == '-') ++p_;
        auto digits = [&] {
            const char *d = p_;
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
            return p_ > d;
        };
        if (!digits()) fail("malformed JSON number");
        if (p_ < end_ && *p_ == '.') {
            ++p_;
            if (!digits()) fail("malformed JSON number");
        }
        if (p_ < end_ && (*p_ == 'e' || *p_ == 'E')) {
            ++p_;
            if (p_ < end_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (!digits()) fail("malformed JSON number");
        }
        std::string text(s, p_);
        return std::strtod(text.c_str(), nullptr);   // correctly rounded to binary64
    }

    static void put_utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }

    uint32_t hex4() {
        if (end_ - p_ < 4) fail("truncated \\u escape in JSON");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            char h = *p_++;
            v <<= 4;
            if (h >= '0' && h <= '9') v |= h - '0';
            else if (h >= 'a' && h <= 'f') v |= h - 'a' + 10;
            else if (h >= 'A' && h <= 'F') v |= h - 'A' + 10;
            else fail("bad \\u escape in JSON");
        }
        return v;
    }

    std::string string_body() {
        ++p_;  // opening quote
        std::string out;
        for (;;) {
            if (p_ == end_) fail("unterminated JSON string");
            char c = *p_++;
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ == end_) fail("unterminated JSON string");
            char e = *p_++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                        p_ += 2;
                        uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair in JSON");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape in JSON string");
            }
        }
        return out;
    }
};

// ---- document access --------------------------------------------------------------------------

const Json &member_array(const Json &doc, const char *key) {
    static const Json empty = [] { Json j; j.kind = Json::Array; return j; }();
    const Json *a = doc.get(key);
    if (!a) return empty;
    if (a->kind != Json::Array) fail(std::string("'") + key + "' is not an array");
    return *a;
}

const Json *opt_member(const Json &obj, const char *key, Json::Kind kind, const char *what) {
    const Json *v = obj.get(key);
    if (v && v->kind != kind) fail(std::string(what) + ": '" + key + "' has the wrong type");
    return v;
}

uint64_t as_index(const Json &v, const char *what) {
    if (v.kind != Json::Number || !(v.num >= 0.0) || v.num != std::floor(v.num) || v.num > 4294967295.0)
        fail(std::string(what) + " is not a valid index");
    return static_cast<uint64_t>(v.num);
}

uint64_t opt_index(const Json &obj, const char *key, uint64_t dflt, const char *what) {
    const Json *v = obj.get(key);
    return v ? as_index(*v, what) : dflt;
}

float as_f32(const Json &v, const char *what) {
    if (v.kind != Json::Number) fail(std::string(what) + " is not a number");
    return static_cast<float>(v.num);
}

void read_floats(const Json *arr, float *out, size_t n, const char *what) {
    if (!arr) return;
    if (arr->kind != Json::Array || arr->items.size() != n) fail(std::string(what) + " must hold " + std::to_string(n) + " numbers");
    for (size_t i = 0; i < n; ++i) out[i] = as_f32(arr->items[i], what);
}

// ---- buffers ----------------------------------------------------------------------------------

int b64_value(unsigned char c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
}

std::vector<uint8_t> base64_decode(const std::string &s, size_t from) {
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        unsigned char c = static_cast<unsigned char>(s[i]);
        if (c == '=') break;
        int v = b64_value(c);
        if (v < 0) fail("bad base64 in a data URI");
        acc = (acc << 6) | static_cast<uint32_t>(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(static_cast<uint8_t>(acc >> bits));
        }
    }
    return out;
}

std::string percent_decode(const std::string &s) {
    std::string out;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size()) {
            char hex[3] = {s[i + 1], s[i + 2], 0};
            char *end = nullptr;
            long v = std::strtol(hex, &end, 16);
            if (end == hex + 2) {
                out += static_cast<char>(v);
                i += 2;
                continue;
            }
        }
        out += s[i];
    }
    return out;
}

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) { std::fclose(f); return false; }
    out.resize(static_cast<size_t>(n));
    size_t got = n ? std::fread(out.data(), 1, out.size(), f) : 0;
    std::fclose(f);
    return got == out.size();
}

struct Document {
    Json json;
    std::vector<std::vector<uint8_t>> buffers;
};

uint32_t le32(const uint8_t *p) { return p[0] | (p[1] << 8) | (p[2] << 16) | (static_cast<uint32_t>(p[3]) << 24); }

// gltf::import: a .glb container (12-B header, JSON chunk, optional BIN chunk) or a JSON
// document, with buffers from the BIN chunk, data URIs or files beside the document (base_dir;
// nullptr: a document with no directory, whose file URIs are an error).
Document load_document(const uint8_t *data, size_t size, const std::string *base_dir) {
    Document doc;
    const uint8_t *bin = nullptr;
    size_t bin_size = 0;
    bool glb = size >= 4 && std::memcmp(data, "glTF", 4) == 0;
    if (glb) {
        if (size < 20) fail("truncated .glb header");
        uint32_t version = le32(data + 4), length = le32(data + 8);
        if (version != 2) fail("unsupported .glb version " + std::to_string(version));
        if (length > size) fail(".glb length exceeds the data");
        uint32_t json_len = le32(data + 12), json_type = le32(data + 16);
        if (json_type != 0x4E4F534Au) fail(".glb first chunk is not JSON");
        if (static_cast<uint64_t>(json_len) + 20 > length) fail(".glb JSON chunk exceeds the file");
        doc.json = JsonReader(reinterpret_cast<const char *>(data + 20), json_len).document();
        size_t off = 20 + static_cast<size_t>(json_len);
        if (off + 8 <= length) {
            uint32_t bl = le32(data + off), bt = le32(data + off + 4);
            if (bt == 0x004E4942u) {
                if (off + 8 + static_cast<uint64_t>(bl) > length) fail(".glb BIN chunk exceeds the file");
                bin = data + off + 8;
                bin_size = bl;
            }
        }
    } else {
        doc.json = JsonReader(reinterpret_cast<const char *>(data), size).document();
    }
    if (doc.json.kind != Json::Object) fail("glTF root is not an object");
    const Json *asset = doc.json.get("asset");
    if (!asset || asset->kind != Json::Object) fail("glTF 'asset' is missing");
    const Json &bufs = member_array(doc.json, "buffers");
    for (size_t i = 0; i < bufs.items.size(); ++i) {
        const Json &b = bufs.items[i];
        uint64_t len = opt_index(b, "byteLength", ~0ull, "buffer byteLength");
        if (len == ~0ull) fail("buffer " + std::to_string(i) + " has no byteLength");
        const Json *uri = opt_member(b, "uri", Json::String, "buffer");
        std::vector<uint8_t> bytes;
        if (!uri) {
            if (!bin || i != 0) fail("buffer " + std::to_string(i) + " has no uri and no .glb BIN chunk");
            bytes.assign(bin, bin + bin_size);
        } else if (uri->str.compare(0, 5, "data:") == 0) {
            size_t comma = uri->str.find(',');
            if (comma == std::string::npos || uri->str.find(";base64") == std::string::npos || uri->str.find(";base64") > comma)
                fail("buffer " + std::to_string(i) + ": only base64 data URIs are supported");
            bytes = base64_decode(uri->str, comma + 1);
        } else {
            if (uri->str.find("://") != std::string::npos) fail("buffer " + std::to_string(i) + ": external URI scheme not supported");
            if (!base_dir) fail("buffer " + std::to_string(i) + ": a file URI needs the document's directory");
            const std::string path = *base_dir + "/" + percent_decode(uri->str);
            if (!read_file(path, bytes)) fail("cannot read buffer file " + path);
        }
        if (bytes.size() < len) fail("buffer " + std::to_string(i) + " is shorter than its byteLength");
        doc.buffers.push_back(std::move(bytes));
    }
    return doc;
}

// ---- accessors --------------------------------------------------------------------------------

enum : uint32_t { kI8 = 5120, kU8 = 5121, kI16 = 5122, kU16 = 5123, kU32 = 5125, kF32 = 5126 };
// The most bytes a sparse accessor without a bufferView may take (memory that no buffer of the
// document backs, so the document's size does not bound it).
constexpr size_t kMaxSparseOnlyBytes = size_t(1) << 28;

uint32_t component_size(uint32_t t) {
    switch (t) {
        case kI8: case kU8: return 1;
        case kI16: case kU16: return 2;
        case kU32: case kF32: return 4;
        default: fail("unknown accessor componentType " + std::to_string(t));
    }
}

uint32_t type_components(const std::string &t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    fail("unknown accessor type " + t);
}

// An accessor's elements as raw component values (one double per component), after sparse
// substitution.  Returns false when the accessor has neither a buffer view nor sparse values
// (the reference's reader then reports the attribute as missing).
struct Accessor {
    uint32_t comp_type = 0, comps = 0;
    size_t count = 0;
    bool normalized = false;
    std::vector<uint8_t> bytes;   // count * comps * component_size, tightly packed
};

const uint8_t *view_bytes(const Document &doc, uint64_t view_index, uint64_t offset, size_t elem, size_t count,
                          size_t *stride_out) {
    const Json &views = member_array(doc.json, "bufferViews");
    if (view_index >= views.items.size()) fail("bufferView index out of range");
    const Json &v = views.items[view_index];
    uint64_t buf = opt_index(v, "buffer", ~0ull, "bufferView buffer");
    if (buf >= doc.buffers.size()) fail("bufferView buffer index out of range");
    uint64_t voff = opt_index(v, "byteOffset", 0, "bufferView byteOffset");
    uint64_t vlen = opt_index(v, "byteLength", ~0ull, "bufferView byteLength");
    if (vlen == ~0ull) fail("bufferView has no byteLength");
    uint64_t stride = opt_index(v, "byteStride", 0, "bufferView byteStride");
    if (stride == 0) stride = elem;
    if (stride < elem) fail("bufferView byteStride smaller than the element");
    if (voff + vlen > doc.buffers[buf].size()) fail("bufferView exceeds its buffer");
    if (count > 0 && offset + stride * (count - 1) + elem > vlen) fail("accessor exceeds its bufferView");
    *stride_out = static_cast<size_t>(stride);
    return doc.buffers[buf].data() + voff + offset;
}

bool read_accessor(const Document &doc, uint64_t index, Accessor &a) {
    const Json &accs = member_array(doc.json, "accessors");
    if (index >= accs.items.size()) fail("accessor index out of range");
    const Json &j = accs.items[index];
    a.comp_type = static_cast<uint32_t>(opt_index(j, "componentType", 0, "accessor componentType"));
    const Json *type = opt_member(j, "type", Json::String, "accessor");
    if (!type) fail("accessor has no type");
    a.comps = type_components(type->str);
    a.count = static_cast<size_t>(opt_index(j, "count", ~0ull, "accessor count"));
    if (a.count == static_cast<size_t>(~0ull)) fail("accessor has no count");
    const Json *norm = opt_member(j, "normalized", Json::Bool, "accessor");
    a.normalized = norm && norm->b;
    const size_t cs = component_size(a.comp_type), elem = cs * a.comps;
    const Json *view = j.get("bufferView");
    const Json *sparse = opt_member(j, "sparse", Json::Object, "accessor");
    if (!view && !sparse) return false;
    // the view's bounds are checked before anything is allocated, so a count in the JSON cannot
    // ask for more memory than the document's buffers hold; a sparse accessor without a view
    // (all zeros but its overrides) is capped instead
    size_t stride = 0;
    const uint8_t *src = nullptr;
    if (view)
        src = view_bytes(doc, as_index(*view, "accessor bufferView"), opt_index(j, "byteOffset", 0, "accessor byteOffset"),
                         elem, a.count, &stride);
    else if (a.count * elem > kMaxSparseOnlyBytes)
        fail("sparse accessor without a bufferView takes more than 256 MiB");
    a.bytes.assign(a.count * elem, 0);
    for (size_t i = 0; src && i < a.count; ++i) std::memcpy(a.bytes.data() + i * elem, src + i * stride, elem);
    if (sparse) {
        size_t n = static_cast<size_t>(opt_index(*sparse, "count", 0, "sparse count"));
        const Json *si = opt_member(*sparse, "indices", Json::Object, "sparse");
        const Json *sv = opt_member(*sparse, "values", Json::Object, "sparse");
        if (!si || !sv) fail("sparse accessor without indices or values");
        uint32_t it = static_cast<uint32_t>(opt_index(*si, "componentType", 0, "sparse indices componentType"));
        if (it != kU8 && it != kU16 && it != kU32) fail("sparse indices must be unsigned integers");
        size_t istride, vstride;
        const uint8_t *ip = view_bytes(doc, opt_index(*si, "bufferView", ~0ull, "sparse indices bufferView"),
                                       opt_index(*si, "byteOffset", 0, "sparse byteOffset"), component_size(it), n, &istride);
        const uint8_t *vp = view_bytes(doc, opt_index(*sv, "bufferView", ~0ull, "sparse values bufferView"),
                                       opt_index(*sv, "byteOffset", 0, "sparse byteOffset"), elem, n, &vstride);
        istride = component_size(it);
        for (size_t k = 0; k < n; ++k) {
            const uint8_t *q = ip + k * istride;
            uint32_t target = it == kU8 ? q[0] : it == kU16 ? static_cast<uint32_t>(q[0] | (q[1] << 8)) : le32(q);
            if (target >= a.count) fail("sparse index out of range");
            std::memcpy(a.bytes.data() + target * elem, vp + k * elem, elem);
        }
    }
    return true;
}

float f32_at(const Accessor &a, size_t i) {
    float f;
    std::memcpy(&f, a.bytes.data() + 4 * i, 4);
    return f;
}

// An attribute the reference reads as [f32; comps] (positions, normals, tangents).
bool read_f32_attribute(const Document &doc, const Json &attrs, const char *name, uint32_t comps, Accessor &a) {
    const Json *idx = attrs.get(name);
    if (!idx) return false;
    if (!read_accessor(doc, as_index(*idx, name), a)) return false;
    if (a.comp_type != kF32 || a.comps != comps)
        fail(std::string(name) + " accessor must be float32 with " + std::to_string(comps) + " components");
    return true;
}

// ---- transforms (gltf 1.4.1 crate math: Transform::matrix, Matrix4 products) -------------------

struct M4 {
    float c[4][4];   // c[column][row]
};

M4 mat_mul(const M4 &a, const M4 &r) {
    // Column j of a * r = a.c0 * r.cj.x + a.c1 * r.cj.y + a.c2 * r.cj.z + a.c3 * r.cj.w, summed
    // left to right.
    M4 o;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            float s = a.c[0][i] * r.c[j][0];
            s = s + a.c[1][i] * r.c[j][1];
            s = s + a.c[2][i] * r.c[j][2];
            s = s + a.c[3][i] * r.c[j][3];
            o.c[j][i] = s;
        }
    return o;
}

M4 trs_matrix(const float t[3], const float q[4], const float s[3]) {
    M4 T{}, R{}, S{};
    for (int i = 0; i < 4; ++i) T.c[i][i] = R.c[i][i] = S.c[i][i] = 1.0f;
    T.c[3][0] = t[0]; T.c[3][1] = t[1]; T.c[3][2] = t[2];
    S.c[0][0] = s[0]; S.c[1][1] = s[1]; S.c[2][2] = s[2];
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx2 = x2 * x, xy2 = x2 * y, xz2 = x2 * z;
    const float yy2 = y2 * y, yz2 = y2 * z, zz2 = z2 * z;
    const float sy2 = y2 * w, sz2 = z2 * w, sx2 = x2 * w;
    R.c[0][0] = 1.0f - yy2 - zz2; R.c[0][1] = xy2 + sz2;        R.c[0][2] = xz2 - sy2;
    R.c[1][0] = xy2 - sz2;        R.c[1][1] = 1.0f - xx2 - zz2; R.c[1][2] = yz2 + sx2;
    R.c[2][0] = xz2 + sy2;        R.c[2][1] = yz2 - sx2;        R.c[2][2] = 1.0f - xx2 - yy2;
    return mat_mul(mat_mul(T, R), S);
}

void node_matrix(const Json &node, float out[16]) {
    if (const Json *m = node.get("matrix")) {
        read_floats(m, out, 16, "node matrix");
        return;
    }
    float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
    read_floats(node.get("translation"), t, 3, "node translation");
    read_floats(node.get("rotation"), q, 4, "node rotation");
    read_floats(node.get("scale"), s, 3, "node scale");
    M4 m = trs_matrix(t, q, s);
    std::memcpy(out, m.c, sizeof(m.c));
}

// ---- Scene::parse_gltf ----------------------------------------------------------------------

struct MeshError {
    const char *what;
};

void parse(const Document &doc, uint32_t flags, nb_host_scene &scene) {
    const Json &meshes = member_array(doc.json, "meshes");
    const Json &materials = member_array(doc.json, "materials");
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> ranges(meshes.items.size());
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t, uint64_t>, std::pair<uint32_t, uint32_t>> shared;

    for (size_t mi = 0; mi < meshes.items.size(); ++mi) {
        const Json &prims = member_array(meshes.items[mi], "primitives");
        for (const Json &pj : prims.items) {
            if (opt_index(pj, "mode", 4, "primitive mode") != 4) throw MeshError{"Not a triangle list"};
            const Json *attrs = opt_member(pj, "attributes", Json::Object, "primitive");
            if (!attrs) fail("primitive has no attributes");
            const Json *tc_idx = attrs->get("TEXCOORD_0");
            const Json *ind_idx = pj.get("indices");
            if (flags & NB_GLTF_SHARE_GEOMETRY) {
                auto id = [&](const char *n) { const Json *v = attrs->get(n); return v ? as_index(*v, n) : ~0ull; };
                auto key = std::make_tuple(id("POSITION"), id("NORMAL"), id("TEXCOORD_0"), id("TANGENT"),
                                           ind_idx ? as_index(*ind_idx, "indices") : ~0ull);
                auto hit = shared.find(key);
                if (hit != shared.end()) {
                    ranges[mi].push_back(hit->second);
                    continue;
                }
            }
            Accessor pos, nrm, tc, tan, ind;
            if (!read_f32_attribute(doc, *attrs, "POSITION", 3, pos)) throw MeshError{"Missing positions"};
            if (!read_f32_attribute(doc, *attrs, "NORMAL", 3, nrm)) throw MeshError{"Missing normals"};
            bool have_tc = tc_idx && read_accessor(doc, as_index(*tc_idx, "TEXCOORD_0"), tc);
            if (have_tc) {
                if (tc.comps != 2) fail("TEXCOORD_0 accessor must be VEC2");
                if (tc.comp_type == kU8 || tc.comp_type == kU16) {
                    if (!tc.normalized) fail("integer TEXCOORD_0 must be normalized");
                } else if (tc.comp_type != kF32) {
                    fail("TEXCOORD_0 accessor must be float32, unorm8 or unorm16");
                }
            }
            if (!read_f32_attribute(doc, *attrs, "TANGENT", 4, tan)) throw MeshError{"Tangent generation is not yet supported"};

            const uint32_t start_vertex = static_cast<uint32_t>(scene.vertices.size());
            const uint32_t start_index = static_cast<uint32_t>(scene.indices.size());
            size_t n = std::min(std::min(pos.count, nrm.count), tan.count);
            if (have_tc) n = std::min(n, tc.count);
            for (size_t k = 0; k < n; ++k) {
                nb_vertex v;
                for (int c = 0; c < 3; ++c) v.position[c] = f32_at(pos, 3 * k + c);
                for (int c = 0; c < 3; ++c) v.normal[c] = f32_at(nrm, 3 * k + c);
                for (int c = 0; c < 4; ++c) v.tangent[c] = f32_at(tan, 4 * k + c);
                float uv[2] = {0.0f, 0.0f};
                if (have_tc) {
                    for (int c = 0; c < 2; ++c) {
                        size_t e = 2 * k + c;
                        if (tc.comp_type == kF32) uv[c] = f32_at(tc, e);
                        else if (tc.comp_type == kU8) uv[c] = static_cast<float>(tc.bytes[e]) / 255.0f;
                        else uv[c] = static_cast<float>(tc.bytes[2 * e] | (tc.bytes[2 * e + 1] << 8)) / 65535.0f;
                    }
                }
                v.u = uv[0];
                v.v = uv[1];
                scene.vertices.push_back(v);
            }
            if (!ind_idx || !read_accessor(doc, as_index(*ind_idx, "indices"), ind)) throw MeshError{"Missing indices"};
            if (ind.comps != 1 || (ind.comp_type != kU8 && ind.comp_type != kU16 && ind.comp_type != kU32))
                fail("indices accessor must be an unsigned integer scalar");
            for (size_t k = 0; k < ind.count; ++k) {
                uint32_t i = ind.comp_type == kU8    ? ind.bytes[k]
                             : ind.comp_type == kU16 ? static_cast<uint32_t>(ind.bytes[2 * k] | (ind.bytes[2 * k + 1] << 8))
                                                     : le32(ind.bytes.data() + 4 * k);
                scene.indices.push_back(i + start_vertex);
            }
            auto range = std::make_pair(start_index, static_cast<uint32_t>(scene.indices.size()));
            ranges[mi].push_back(range);
            if (flags & NB_GLTF_SHARE_GEOMETRY) {
                auto id = [&](const char *nm) { const Json *v = attrs->get(nm); return v ? as_index(*v, nm) : ~0ull; };
                shared[std::make_tuple(id("POSITION"), id("NORMAL"), id("TEXCOORD_0"), id("TANGENT"),
                                       as_index(*ind_idx, "indices"))] = range;
            }
        }
    }

    for (const Json &node : member_array(doc.json, "nodes").items) {
        const Json *mesh = node.get("mesh");
        if (!mesh) continue;
        uint64_t mi = as_index(*mesh, "node mesh");
        if (mi >= meshes.items.size()) fail("node mesh index out of range");
        nb_primitive base;
        node_matrix(node, base.local_to_world);
        const Json &prims = member_array(meshes.items[mi], "primitives");
        for (size_t pi = 0; pi < prims.items.size(); ++pi) {
            const Json &pj = prims.items[pi];
            float base_color[4] = {1, 1, 1, 1}, emissive[3] = {0, 0, 0};
            float roughness = 1.0f, metallic = 1.0f;
            if (const Json *mat = pj.get("material")) {
                uint64_t k = as_index(*mat, "primitive material");
                if (k >= materials.items.size()) fail("material index out of range");
                const Json &m = materials.items[k];
                read_floats(m.get("emissiveFactor"), emissive, 3, "emissiveFactor");
                if (const Json *pbr = opt_member(m, "pbrMetallicRoughness", Json::Object, "material")) {
                    read_floats(pbr->get("baseColorFactor"), base_color, 4, "baseColorFactor");
                    if (const Json *r = pbr->get("roughnessFactor")) roughness = as_f32(*r, "roughnessFactor");
                    if (const Json *r = pbr->get("metallicFactor")) metallic = as_f32(*r, "metallicFactor");
                }
            }
            const bool is_emissive = !(emissive[0] == 0.0f && emissive[1] == 0.0f && emissive[2] == 0.0f);
            nb_primitive p = base;
            if (is_emissive) {
                p.color[0] = emissive[0]; p.color[1] = emissive[1]; p.color[2] = emissive[2]; p.color[3] = 1.0f;
            } else {
                std::memcpy(p.color, base_color, sizeof(base_color));
            }
            p.roughness = roughness;
            p.metallic = metallic;
            p.emissive = is_emissive ? 1.0f : 0.0f;
            p.index_start = ranges[mi][pi].first;
            p.index_end = ranges[mi][pi].second;
            scene.prims.push_back(p);
        }
    }
}

nb_status parse_into(nb_host_scene *scene, const uint8_t *data, size_t size, const std::string *base_dir, uint32_t flags) {
    if (flags & ~static_cast<uint32_t>(NB_GLTF_SHARE_GEOMETRY)) return set_error(NB_ERR_INVALID_ARGUMENT, "unknown glTF flags %#x", flags);
    nb_host_scene work = *scene;
    try {
        Document doc = load_document(data, size, base_dir);
        parse(doc, flags, work);
    } catch (const MeshError &e) {
        return set_error(NB_ERR_FORMAT, "%s", e.what);
    } catch (const ParseError &e) {
        return set_error(NB_ERR_FORMAT, "Gltf error: %s", e.what.c_str());
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory while parsing glTF");
    }
    if (work.vertices.size() > 0xFFFFFFFFull || work.indices.size() > 0xFFFFFFFFull)
        return set_error(NB_ERR_UNSUPPORTED, "scene exceeds 2^32 vertices or indices");
    *scene = std::move(work);
    return NB_OK;
}

}  // namespace
}  // namespace nb

using namespace nb;

extern "C" {

nb_status nb_host_scene_create(nb_host_scene **out) {
    if (!out) return set_error(NB_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = new (std::nothrow) nb_host_scene();
    return *out ? NB_OK : set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory");
}

nb_status nb_host_scene_destroy(nb_host_scene *scene) {
    delete scene;
    return NB_OK;
}

nb_status nb_host_scene_parse_gltf(nb_host_scene *scene, const char *path, uint32_t flags) {
    if (!scene || !path) return set_error(NB_ERR_INVALID_ARGUMENT, "null scene or path");
    std::vector<uint8_t> bytes;
    try {
        if (!read_file(path, bytes)) return set_error(NB_ERR_IO, "cannot read %s", path);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory reading %s", path);
    }
    std::string dir(path);
    size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash == 0 ? 1 : slash);
    return parse_into(scene, bytes.data(), bytes.size(), &dir, flags);
}

nb_status nb_host_scene_parse_gltf_memory(nb_host_scene *scene, const void *data, size_t size, const char *base_dir,
                                          uint32_t flags) {
    if (!scene || (!data && size)) return set_error(NB_ERR_INVALID_ARGUMENT, "null scene or data");
    if (!base_dir) return parse_into(scene, static_cast<const uint8_t *>(data), size, nullptr, flags);
    const std::string dir(base_dir);
    return parse_into(scene, static_cast<const uint8_t *>(data), size, &dir, flags);
}

nb_status nb_host_scene_counts(const nb_host_scene *scene, uint32_t *n_vertices, uint32_t *n_indices, uint32_t *n_prims) {
    if (!scene || !n_vertices || !n_indices || !n_prims) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *n_vertices = static_cast<uint32_t>(scene->vertices.size());
    *n_indices = static_cast<uint32_t>(scene->indices.size());
    *n_prims = static_cast<uint32_t>(scene->prims.size());
    return NB_OK;
}

nb_status nb_host_scene_copy(const nb_host_scene *scene, nb_vertex *vertices, uint32_t *indices, nb_primitive *prims) {
    if (!scene) return set_error(NB_ERR_INVALID_ARGUMENT, "null scene");
    if ((!vertices && !scene->vertices.empty()) || (!indices && !scene->indices.empty()) || (!prims && !scene->prims.empty()))
        return set_error(NB_ERR_INVALID_ARGUMENT, "null output array");
    if (!scene->vertices.empty()) std::memcpy(vertices, scene->vertices.data(), scene->vertices.size() * sizeof(nb_vertex));
    if (!scene->indices.empty()) std::memcpy(indices, scene->indices.data(), scene->indices.size() * sizeof(uint32_t));
    if (!scene->prims.empty()) std::memcpy(prims, scene->prims.data(), scene->prims.size() * sizeof(nb_primitive));
    return NB_OK;
}

}  // extern "C"
