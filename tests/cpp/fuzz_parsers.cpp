// Mutation fuzzing of the library's readers of untrusted files: the glTF / GLB reader
// (nb_host_scene_parse_gltf_memory, gltf.cpp) and the cube DDS reader (nb_dds_read_cube_memory,
// dds.cpp).  This program is built from the readers' own sources with AddressSanitizer and
// UndefinedBehaviorSanitizer, so an out-of-bounds access, a use of freed memory, a signed
// overflow or a leak ends it with a report and a non-zero status.
//
// Every seed file is mutated `iterations` times, deterministically from `seed`.  Each mutant must
// come back as a status; a glTF mutant that parses must leave a scene whose arrays copy out,
// whose primitives' index ranges lie inside its index array and which parses with geometry
// sharing too, and a DDS mutant that decodes must fill exactly 6 x n x n RGBA texels.  A parsed
// scene then goes through the host scene builder (scene_build.cpp), whose trees must obey the
// rules the scene upload checks.  No mutant
// may make a reader ask for more than 1 GiB of memory.  (Index
// values are not checked against the vertex count by the reader, as in the reference; the scene
// builder rejects out-of-range ones.)  GLB seeds are mutated as a JSON chunk plus a BIN chunk and put back
// together with consistent lengths most of the time, so that mutants get past the container
// checks into the accessor and material code; .gltf seeds are mutated as text; DDS seeds get
// interesting values in their header fields.
//
// usage: fuzz_parsers <iterations> <seed> <file.glb | file.gltf | file.dds>...
// Prints one line per seed: the mutants tried, how many parsed, and the statuses of the rest.
#include "nbounce.h"
#include "nb_internal.h"

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace nb {
// The readers report errors through this; the library's own version lives with the GPU code.
nb_status set_error(nb_status st, const char *fmt, ...) {
    static thread_local char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return st;
}
}  // namespace nb

// An allocation above 1 GiB stops the program with an out-of-memory report instead of taking the
// machine's memory: no mutant of a few KB may ask for that much.
extern "C" const char *__asan_default_options() { return "max_allocation_size_mb=1024"; }

// nb_envmap_load (dds.cpp) uploads through this; no mutant reaches it here.
extern "C" nb_status nb_envmap_create_cube(nb_context *, const float *, uint32_t, nb_envmap **) {
    return NB_ERR_UNSUPPORTED;
}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {   // xorshift64*
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
    uint32_t below(uint32_t n) { return n ? static_cast<uint32_t>(next() % n) : 0; }
};

using Bytes = std::vector<uint8_t>;

const uint32_t kU32s[] = {0u, 1u, 2u, 3u, 4u, 6u, 7u, 8u, 15u, 16u, 17u, 31u, 32u, 64u, 113u, 116u, 255u, 256u,
                          0x7FFFu, 0x8000u, 0xFFFFu, 0x10000u, 16384u, 16385u, 0x7FFFFFFFu, 0x80000000u,
                          0xFFFFFFFEu, 0xFFFFFFFFu, 0x30315844u, 0x4E4F534Au, 0x004E4942u};
const char *const kNumbers[] = {"0", "1", "-1", "2", "3", "4", "5", "6", "7", "8", "16", "255", "256", "65535",
                                "65536", "4294967295", "4294967296", "18446744073709551616", "1e308", "-1e308",
                                "1e-320", "0.5", "-0", "268435456", "268435457", "5120", "5121", "5122", "5123",
                                "5125", "5126", "34962", "34963", "99999999999999999999", "3.4e38", "1e39"};
const char *const kTokens[] = {"{}", "[]", "null", "true", "\"\"", "\"VEC3\"", "\"MAT4\"", "\"SCALAR\"",
                               "[[[[[[[[", "\"data:application/octet-stream;base64,AAAA\"", "\"x.bin\"",
                               "\"data:,\"", "{\"count\":1}", "-", "1e", "\"\\u0000\"", "\"\\ud800\""};

void flip_bits(Bytes &b, Rng &r) {
    if (b.empty()) return;
    for (uint32_t k = 1 + r.below(8); k--;) b[r.below(b.size())] ^= static_cast<uint8_t>(1u << r.below(8));
}

void set_u32(Bytes &b, Rng &r, size_t limit) {
    if (b.size() < 4) return;
    const size_t span = std::min(limit, b.size() - 3);
    const size_t at = (r.below(2) ? 4 * r.below(static_cast<uint32_t>((span + 3) / 4)) : r.below(static_cast<uint32_t>(span)));
    if (at + 4 > b.size()) return;
    uint32_t v = kU32s[r.below(sizeof kU32s / sizeof kU32s[0])];
    if (r.below(4) == 0) v = static_cast<uint32_t>(b.size() + r.below(3) - 1);
    std::memcpy(b.data() + at, &v, 4);
}

void splice(Bytes &b, Rng &r) {
    if (b.empty()) return;
    const size_t at = r.below(b.size()), n = std::min<size_t>(1 + r.below(64), b.size() - at);
    switch (r.below(3)) {
        case 0: b.erase(b.begin() + at, b.begin() + at + n); break;
        case 1: { Bytes dup(b.begin() + at, b.begin() + at + n); b.insert(b.begin() + r.below(b.size()), dup.begin(), dup.end()); break; }
        default:   // truncate anywhere, or by a few bytes at the end
            b.resize(r.below(2) ? r.below(b.size() + 1) : b.size() - std::min<size_t>(b.size(), 1 + r.below(64)));
            break;
    }
}

// Replaces a number, a string or a punctuation run of JSON text with another token.
void mutate_json(std::string &js, Rng &r) {
    if (js.empty()) return;
    const uint32_t kind = r.below(6);
    size_t at = r.below(js.size());
    if (kind <= 2) {   // a number
        size_t tries = 0;
        while (tries++ < js.size() && !(js[at] >= '0' && js[at] <= '9')) at = (at + 1) % js.size();
        if (!(js[at] >= '0' && js[at] <= '9')) return;
        size_t s = at, e = at;
        while (s > 0 && (std::strchr("0123456789.eE+-", js[s - 1]))) --s;
        while (e < js.size() && std::strchr("0123456789.eE+-", js[e])) ++e;
        js.replace(s, e - s, kNumbers[r.below(sizeof kNumbers / sizeof kNumbers[0])]);
    } else if (kind == 3) {   // a value after a colon
        size_t tries = 0;
        while (tries++ < js.size() && js[at] != ':') at = (at + 1) % js.size();
        if (js[at] != ':') return;
        size_t e = at + 1;
        int depth = 0;
        bool in_str = false;
        for (; e < js.size(); ++e) {
            const char c = js[e];
            if (in_str) { if (c == '\\') ++e; else if (c == '"') in_str = false; continue; }
            if (c == '"') in_str = true;
            else if (c == '{' || c == '[') ++depth;
            else if (c == '}' || c == ']') { if (depth-- == 0) break; }
            else if (c == ',' && depth == 0) break;
        }
        js.replace(at + 1, e - at - 1, kTokens[r.below(sizeof kTokens / sizeof kTokens[0])]);
    } else if (kind == 4) {   // delete or duplicate a short run
        const size_t n = std::min<size_t>(1 + r.below(24), js.size() - at);
        if (r.below(2)) js.erase(at, n);
        else js.insert(r.below(js.size()), js.substr(at, n));
    } else {   // one character
        js[at] = "{}[],:\"\\0-.e "[r.below(13)];
    }
}

struct Glb {
    std::string json;
    Bytes bin;
    bool has_bin = false;
};

bool split_glb(const Bytes &d, Glb &g) {
    if (d.size() < 20 || std::memcmp(d.data(), "glTF", 4) != 0) return false;
    uint32_t jl;
    std::memcpy(&jl, d.data() + 12, 4);
    if (20ull + jl > d.size()) return false;
    g.json.assign(reinterpret_cast<const char *>(d.data() + 20), jl);
    const size_t off = 20 + jl;
    if (off + 8 <= d.size()) {
        uint32_t bl;
        std::memcpy(&bl, d.data() + off, 4);
        if (off + 8 + static_cast<uint64_t>(bl) <= d.size()) {
            g.bin.assign(d.begin() + off + 8, d.begin() + off + 8 + bl);
            g.has_bin = true;
        }
    }
    return true;
}

Bytes join_glb(const Glb &g) {
    std::string js = g.json;
    while (js.size() % 4) js += ' ';
    Bytes out(20);
    const uint32_t total = static_cast<uint32_t>(20 + js.size() + (g.has_bin ? 8 + g.bin.size() : 0));
    const uint32_t head[5] = {0x46546C67u, 2u, total, static_cast<uint32_t>(js.size()), 0x4E4F534Au};
    std::memcpy(out.data(), head, 20);
    out.insert(out.end(), js.begin(), js.end());
    if (g.has_bin) {
        const uint32_t bh[2] = {static_cast<uint32_t>(g.bin.size()), 0x004E4942u};
        out.insert(out.end(), reinterpret_cast<const uint8_t *>(bh), reinterpret_cast<const uint8_t *>(bh) + 8);
        out.insert(out.end(), g.bin.begin(), g.bin.end());
    }
    return out;
}

bool ends_with(const std::string &s, const char *suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

[[noreturn]] void fail_build(const char *what, uint32_t a, uint32_t b) {
    std::printf("FAIL: scene build: %s (%u, %u)\n", what, a, b);
    std::exit(1);
}

// The tree rules nb_scene_create enforces: children after their parent and inside the array,
// leaf ranges non-empty inside [0, n_prims), at most 32 levels.
void check_built_tree(const std::vector<nb_bvh_node> &t, uint32_t n, uint32_t n_prims) {
    std::vector<uint32_t> depth(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const nb_bvh_node &d = t[i];
        if (d.end > 0) {
            if (d.start >= d.end || d.end > n_prims) fail_build("leaf range outside the primitives", i, d.end);
        } else {
            if (d.start <= i || d.start >= n - 1) fail_build("children not after their parent", i, d.start);
            if (depth[i] + 1 > 32) fail_build("tree deeper than 32 levels", i, depth[i]);
            depth[d.start] = std::max(depth[d.start], depth[i] + 1);
            depth[d.start + 1] = std::max(depth[d.start + 1], depth[i] + 1);
        }
    }
}

// A parsed scene through the host builder (scene_build.cpp, SceneBuffers::from_scene's build),
// with and without BLAS sharing: it must return a status, and a built scene must obey the tree
// rules, point its instances at BLAS nodes and keep the index values (a permutation).
void check_build(const std::vector<nb_vertex> &v, const std::vector<uint32_t> &idx, const std::vector<nb_primitive> &p) {
    if (p.empty()) return;
    for (uint32_t flags : {0u, static_cast<uint32_t>(NB_BUILD_SHARE_BLAS)}) {
        uint32_t bcap = 0, tcap = 0;
        if (nb::build_scene_capacity(static_cast<uint32_t>(idx.size()), p.data(), static_cast<uint32_t>(p.size()), flags,
                                     &bcap, &tcap) != NB_OK)
            continue;
        if (bcap > (1u << 15)) continue;   // large scenes: their builds are tested elsewhere, and slow here
        std::vector<uint32_t> ix(idx);
        std::vector<nb_bvh_node> blas(bcap), tlas(tcap);
        std::vector<nb_instance> inst(p.size());
        uint32_t nb_ = 0, nt = 0;
        if (nb::build_scene(v.data(), static_cast<uint32_t>(v.size()), ix.data(), static_cast<uint32_t>(ix.size()), p.data(),
                            static_cast<uint32_t>(p.size()), flags, blas.data(), bcap, &nb_, tlas.data(), tcap, &nt,
                            inst.data()) != NB_OK)
            continue;
        if (nb_ > bcap || nt > tcap) fail_build("more nodes than the capacity", nb_, nt);
        check_built_tree(blas, nb_, static_cast<uint32_t>(ix.size() / 3));
        check_built_tree(tlas, nt, static_cast<uint32_t>(p.size()));
        for (const nb_instance &q : inst)
            if (q.node >= nb_) fail_build("instance BLAS root outside the BLAS", q.node, nb_);
        std::vector<uint32_t> a(idx), b(ix);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        if (a != b) fail_build("index values changed", 0, 0);
    }
}

// One glTF mutant through the reader; returns its status.  A parse that succeeds is checked.
nb_status try_gltf(const Bytes &data) {
    nb_host_scene *sc = nullptr;
    if (nb_host_scene_create(&sc) != NB_OK) std::abort();
    nb_status st = nb_host_scene_parse_gltf_memory(sc, data.data(), data.size(), nullptr, 0);
    if (st == NB_OK) {
        uint32_t nv = 0, ni = 0, np = 0;
        if (nb_host_scene_counts(sc, &nv, &ni, &np) != NB_OK) std::abort();
        std::vector<nb_vertex> v(nv);
        std::vector<uint32_t> idx(ni);
        std::vector<nb_primitive> p(np);
        if (nb_host_scene_copy(sc, v.data(), idx.data(), p.data()) != NB_OK) std::abort();
        for (const nb_primitive &q : p)
            if (q.index_start > q.index_end || q.index_end > ni) {
                std::printf("FAIL: primitive range [%u, %u) outside %u indices\n", q.index_start, q.index_end, ni);
                std::exit(1);
            }
        check_build(v, idx, p);
        // the same document with geometry sharing must parse too
        nb_host_scene *sh = nullptr;
        if (nb_host_scene_create(&sh) != NB_OK) std::abort();
        if (nb_host_scene_parse_gltf_memory(sh, data.data(), data.size(), nullptr, NB_GLTF_SHARE_GEOMETRY) != NB_OK) {
            std::printf("FAIL: parses alone but not with geometry sharing\n");
            std::exit(1);
        }
        nb_host_scene_destroy(sh);
    }
    nb_host_scene_destroy(sc);
    return st;
}

nb_status try_dds(const Bytes &data) {
    uint32_t n = 0;
    nb_status st = nb_dds_read_cube_memory(data.data(), data.size(), nullptr, 0, &n);
    if (st != NB_OK) return st;
    std::vector<float> px(static_cast<size_t>(6) * n * n * 4 + 4, -7.0f);
    uint32_t n2 = 0;
    st = nb_dds_read_cube_memory(data.data(), data.size(), px.data(), px.size() - 4, &n2);
    if (st != NB_OK || n2 != n) {
        std::printf("FAIL: size query and decode disagree (%u vs %u, status %d)\n", n, n2, st);
        std::exit(1);
    }
    for (size_t i = px.size() - 4; i < px.size(); ++i)
        if (px[i] != -7.0f) {
            std::printf("FAIL: decode wrote past 6 x %u^2 texels\n", n);
            std::exit(1);
        }
    return NB_OK;
}

Bytes read_all(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::printf("cannot open %s\n", path); std::exit(2); }
    Bytes b;
    uint8_t buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + got);
    std::fclose(f);
    return b;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4) {
        std::printf("usage: %s <iterations> <seed> <file.glb|file.gltf|file.dds>...\n", argv[0]);
        return 2;
    }
    const uint32_t iterations = static_cast<uint32_t>(std::strtoul(argv[1], nullptr, 10));
    const uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    for (int a = 3; a < argc; ++a) {
        const std::string path = argv[a];
        const Bytes orig = read_all(argv[a]);
        const bool dds = ends_with(path, ".dds"), glb_seed = ends_with(path, ".glb");
        Rng r{(seed + 1) * 0x9E3779B97F4A7C15ULL ^ static_cast<uint64_t>(a) * 0xD1B54A32D192ED03ULL};
        Glb g;
        if (glb_seed && !split_glb(orig, g)) { std::printf("%s: not a .glb\n", argv[a]); return 2; }
        // the seed itself must be read
        if ((dds ? try_dds(orig) : try_gltf(orig)) != NB_OK) { std::printf("%s: the seed fails to parse\n", argv[a]); return 2; }
        uint32_t ok = 0;
        std::map<int, uint32_t> errors;
        for (uint32_t it = 0; it < iterations; ++it) {
            Bytes m;
            if (dds) {
                m = orig;
                for (uint32_t k = 1 + r.below(3); k--;) {
                    switch (r.below(4)) {
                        case 0: case 1: set_u32(m, r, 152); break;
                        case 2: flip_bits(m, r); break;
                        default: splice(m, r); break;
                    }
                }
            } else if (glb_seed) {
                Glb h = g;
                for (uint32_t k = 1 + r.below(3); k--;) {
                    if (r.below(4) || !h.has_bin) mutate_json(h.json, r);
                    else { Bytes &bb = h.bin; if (r.below(2)) flip_bits(bb, r); else set_u32(bb, r, bb.size()); }
                }
                m = join_glb(h);
                if (r.below(5) == 0) {   // the container itself
                    if (r.below(2)) set_u32(m, r, 24); else splice(m, r);
                }
            } else {
                std::string js(orig.begin(), orig.end());
                for (uint32_t k = 1 + r.below(3); k--;) mutate_json(js, r);
                m.assign(js.begin(), js.end());
                if (r.below(8) == 0) flip_bits(m, r);
            }
            // an allocation of exactly the mutant's size, so a read past its end meets the
            // sanitizer's red zone (a vector's spare capacity would hide it)
            const Bytes exact(m.begin(), m.end());
            const nb_status st = dds ? try_dds(exact) : try_gltf(exact);
            if (st == NB_OK) ++ok;
            else ++errors[static_cast<int>(st)];
        }
        std::printf("%s: %u mutants, %u parsed", argv[a], iterations, ok);
        for (const auto &e : errors) std::printf(", status %d x%u", e.first, e.second);
        std::printf("\n");
    }
    return 0;
}
