// Host-side scene build: the SAH BVH builder of src/pathtracing/bvh.rs and the flattening of
// SceneBuffers::from_scene (src/pathtracing/scene.rs:280-307), bit-exact with the reference.
//
// Bit-exactness rests on evaluating every float expression in binary32, in the reference's
// order, without FMA contraction (this file is compiled with -ffp-contract=off):
//   - triangle centroid (v0 + v1 + v2) / 3.0 by division (bvh.rs:97);
//   - surface area ex*ey + ex*ez + ey*ez summed left to right (bvh.rs:69,145);
//   - bin index floor((c - min) / step) converted with Rust's saturating `as u32` (bvh.rs:268);
//   - NaN-ignoring component min / max (glam Vec3::min / max over f32::min / max);
//   - children appended as a (left, left + 1) pair and the LIFO stack pushing left then right,
//     so the right subtree is numbered first (bvh.rs:172-186).
//
// Design: primitives are held structure-of-arrays (lo / hi / centroid / payload) so the
// partition and binning loops stream; BLAS builds over disjoint index ranges run on separate
// threads and are relocated into the shared node array in primitive order afterwards.
#include "nbounce.h"
#include "nb_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace nb {
namespace {

struct V3 { float x, y, z; };
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float at(const V3 &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// f32::min / f32::max: the non-NaN operand when exactly one is NaN.
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline float fmx(float a, float b) { return std::fmax(a, b); }
inline V3 vmin(V3 a, V3 b) { return v3(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return v3(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
inline bool finite3(V3 e) { return std::isfinite(e.x) && std::isfinite(e.y) && std::isfinite(e.z); }

constexpr float kInf = INFINITY;
constexpr uint32_t kMaxDepth = 32;   // bvh.rs:153
constexpr int kBins = 16;            // bvh.rs:154

inline float sah_cost(V3 lo, V3 hi, uint32_t count) {  // bvh.rs:65-74, 142-150
    V3 e = hi - lo;
    if (!finite3(e)) return kInf;
    float area = e.x * e.y + e.x * e.z + e.y * e.z;
    return static_cast<float>(count) * area;
}

inline uint32_t as_u32_saturating(float x) {            // Rust `f32 as u32`
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return static_cast<uint32_t>(x);
}

struct Box {
    V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    uint32_t n = 0;
    void add(V3 l, V3 h) { lo = vmin(lo, l); hi = vmax(hi, h); ++n; }
    void add(const Box &b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); n += b.n; }
    float cost() const { return sah_cost(lo, hi, n); }
};

// Primitive set, structure of arrays.  payload: 3 vertex indices per triangle, or the original
// instance number for instances.
struct Prims {
    std::vector<V3> lo, hi, ctr;
    std::vector<uint32_t> payload;
    int width = 1;  // payload words per primitive
    void swap(uint32_t a, uint32_t b) {
        if (a == b) return;
        std::swap(lo[a], lo[b]); std::swap(hi[a], hi[b]); std::swap(ctr[a], ctr[b]);
        for (int k = 0; k < width; ++k) std::swap(payload[a * width + k], payload[b * width + k]);
    }
};

struct Split { int axis; float mid; };

class Builder {
public:
    explicit Builder(Prims &p) : p_(p) {}

    // BVHTree::append (bvh.rs:162-190) into `out` with node indices relative to `base`.
    uint32_t append(std::vector<nb_bvh_node> &out, uint32_t base, uint32_t start, uint32_t end) {
        uint32_t root = base + static_cast<uint32_t>(out.size());
        out.push_back(leaf_over(start, end));
        stack_.clear();
        stack_.emplace_back(0u, root);
        while (!stack_.empty()) {
            auto [depth, idx] = stack_.back();
            stack_.pop_back();
            if (depth >= kMaxDepth) continue;
            nb_bvh_node cur = out[idx - base];
            nb_bvh_node l, r;
            if (!split_node(cur, l, r)) continue;
            uint32_t li = base + static_cast<uint32_t>(out.size());
            out[idx - base].end = 0;        // make_inner (bvh.rs:60-63)
            out[idx - base].start = li;
            out.push_back(l);
            out.push_back(r);
            stack_.emplace_back(depth + 1, li);
            stack_.emplace_back(depth + 1, li + 1);
        }
        return root;
    }

private:
    static nb_bvh_node node_of(V3 lo, V3 hi, uint32_t s, uint32_t e) {
        nb_bvh_node n;
        n.min[0] = lo.x; n.min[1] = lo.y; n.min[2] = lo.z; n.start = s;
        n.max[0] = hi.x; n.max[1] = hi.y; n.max[2] = hi.z; n.end = e;
        return n;
    }
    static V3 lo_of(const nb_bvh_node &n) { return v3(n.min[0], n.min[1], n.min[2]); }
    static V3 hi_of(const nb_bvh_node &n) { return v3(n.max[0], n.max[1], n.max[2]); }
    static float cost_of(const nb_bvh_node &n) { return sah_cost(lo_of(n), hi_of(n), n.end - n.start); }

    nb_bvh_node leaf_over(uint32_t s, uint32_t e) const {  // BVHNode::new_leaf, bvh.rs:27-39
        V3 lo = p_.lo[s], hi = p_.hi[s];
        for (uint32_t i = s + 1; i < e; ++i) { lo = vmin(lo, p_.lo[i]); hi = vmax(hi, p_.hi[i]); }
        return node_of(lo, hi, s, e);
    }

    bool split_node(const nb_bvh_node &par, nb_bvh_node &l, nb_bvh_node &r) {  // bvh.rs:208-230
        uint32_t count = par.end - par.start;
        if (count <= 1) return false;
        if (count == 2) {
            l = leaf_over(par.start, par.start + 1);
            r = leaf_over(par.start + 1, par.start + 2);
            return cost_of(l) + cost_of(r) < cost_of(par);
        }
        Split s;
        bool ok = count <= 11 ? exhaustive(par, s) : binned(par, s);
        return ok && partition(par, s, l, r);
    }

    // bvh.rs:232-258: every primitive centroid on every axis as a candidate plane.
    bool exhaustive(const nb_bvh_node &par, Split &out) const {
        float best = cost_of(par);
        bool found = false;
        for (int axis = 0; axis < 3; ++axis) {
            for (uint32_t i = par.start; i < par.end; ++i) {
                float mid = at(p_.ctr[i], axis);
                Box left, right;
                for (uint32_t j = par.start; j < par.end; ++j)
                    (at(p_.ctr[j], axis) < mid ? left : right).add(p_.lo[j], p_.hi[j]);
                float c = left.cost() + right.cost();
                if (c < best) { best = c; out = Split{axis, mid}; found = true; }
            }
        }
        return found;
    }

    // bvh.rs:260-296: 16 bins per axis, candidate planes between bins.
    bool binned(const nb_bvh_node &par, Split &out) const {
        Box bins[3][kBins];
        V3 plo = lo_of(par), ext = hi_of(par) - plo;
        const float step[3] = {ext.x / 16.0f, ext.y / 16.0f, ext.z / 16.0f};
        const float base[3] = {plo.x, plo.y, plo.z};
        for (uint32_t i = par.start; i < par.end; ++i) {
            V3 c = p_.ctr[i];
            const float cc[3] = {c.x, c.y, c.z};
            for (int a = 0; a < 3; ++a) {
                uint32_t b = as_u32_saturating(std::floor((cc[a] - base[a]) / step[a]));
                bins[a][b < 15u ? b : 15u].add(p_.lo[i], p_.hi[i]);
            }
        }
        float best = cost_of(par);
        bool found = false;
        for (int a = 0; a < 3; ++a) {
            // suffix unions: right side of plane i is bins i+1..15
            Box suffix[kBins + 1];
            for (int j = kBins - 1; j >= 0; --j) { suffix[j] = suffix[j + 1]; suffix[j].add(bins[a][j]); }
            Box left;
            for (int i = 0; i < kBins - 1; ++i) {
                left.add(bins[a][i]);
                float c = left.cost() + suffix[i + 1].cost();
                if (c < best) {
                    best = c;
                    out = Split{a, base[a] + step[a] * static_cast<float>(i + 1)};
                    found = true;
                }
            }
        }
        return found;
    }

    // bvh.rs:307-330: left-stable partition by swapping into the growing left block.
    bool partition(const nb_bvh_node &par, Split s, nb_bvh_node &l, nb_bvh_node &r) {
        Box left, right;
        for (uint32_t i = par.start; i < par.end; ++i) {
            V3 lo = p_.lo[i], hi = p_.hi[i];
            if (at(p_.ctr[i], s.axis) < s.mid) {
                left.add(lo, hi);
                p_.swap(i, par.start + left.n - 1);
            } else {
                right.add(lo, hi);
            }
        }
        if (left.n == 0 || right.n == 0) return false;
        l = node_of(left.lo, left.hi, par.start, par.start + left.n);
        r = node_of(right.lo, right.hi, par.start + left.n, par.start + left.n + right.n);
        return true;
    }

    Prims &p_;
    std::vector<std::pair<uint32_t, uint32_t>> stack_;
};

// glam Mat4::inverse (cofactors in glm's arrangement), used for Instance::world_to_local.
void mat4_inverse(const float *m, float *out) {
    auto M = [m](int c, int r) { return m[c * 4 + r]; };
    float c00 = M(2,2) * M(3,3) - M(3,2) * M(2,3), c02 = M(1,2) * M(3,3) - M(3,2) * M(1,3);
    float c03 = M(1,2) * M(2,3) - M(2,2) * M(1,3), c04 = M(2,1) * M(3,3) - M(3,1) * M(2,3);
    float c06 = M(1,1) * M(3,3) - M(3,1) * M(1,3), c07 = M(1,1) * M(2,3) - M(2,1) * M(1,3);
    float c08 = M(2,1) * M(3,2) - M(3,1) * M(2,2), c10 = M(1,1) * M(3,2) - M(3,1) * M(1,2);
    float c11 = M(1,1) * M(2,2) - M(2,1) * M(1,2), c12 = M(2,0) * M(3,3) - M(3,0) * M(2,3);
    float c14 = M(1,0) * M(3,3) - M(3,0) * M(1,3), c15 = M(1,0) * M(2,3) - M(2,0) * M(1,3);
    float c16 = M(2,0) * M(3,2) - M(3,0) * M(2,2), c18 = M(1,0) * M(3,2) - M(3,0) * M(1,2);
    float c19 = M(1,0) * M(2,2) - M(2,0) * M(1,2), c20 = M(2,0) * M(3,1) - M(3,0) * M(2,1);
    float c22 = M(1,0) * M(3,1) - M(3,0) * M(1,1), c23 = M(1,0) * M(2,1) - M(2,0) * M(1,1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float a0[4] = {M(1,0), M(0,0), M(0,0), M(0,0)}, a1[4] = {M(1,1), M(0,1), M(0,1), M(0,1)};
    const float a2[4] = {M(1,2), M(0,2), M(0,2), M(0,2)}, a3[4] = {M(1,3), M(0,3), M(0,3), M(0,3)};
    const float sa[4] = {1.0f, -1.0f, 1.0f, -1.0f}, sb[4] = {-1.0f, 1.0f, -1.0f, 1.0f};
    float inv[16];
    for (int i = 0; i < 4; ++i) {
        inv[0 * 4 + i] = sa[i] * ((a1[i] * f0[i] - a2[i] * f1[i]) + a3[i] * f2[i]);
        inv[1 * 4 + i] = sb[i] * ((a0[i] * f0[i] - a2[i] * f3[i]) + a3[i] * f4[i]);
        inv[2 * 4 + i] = sa[i] * ((a0[i] * f1[i] - a1[i] * f3[i]) + a3[i] * f5[i]);
        inv[3 * 4 + i] = sb[i] * ((a0[i] * f2[i] - a1[i] * f4[i]) + a2[i] * f5[i]);
    }
    float d0 = M(0,0) * inv[0], d1 = M(0,1) * inv[4], d2 = M(0,2) * inv[8], d3 = M(0,3) * inv[12];
    float rcp = 1.0f / ((d0 + d2) + (d1 + d3));
    for (int k = 0; k < 16; ++k) out[k] = inv[k] * rcp;
}

// glam Mat4::transform_point3: ((c0*x + c1*y) + c2*z) + c3
V3 transform_point(const float *m, V3 p) {
    return v3(((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12],
              ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13],
              ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14]);
}

uint64_t tris_of(const nb_primitive &p) { return p.index_end / 3 - p.index_start / 3; }

}  // namespace

// ---------------------------------------------------------------------------------------
nb_status build_scene_capacity(uint32_t n_indices, const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                               uint32_t *blas_cap, uint32_t *tlas_cap) {
    uint64_t total = 0;
    std::map<std::pair<uint32_t, uint32_t>, int> seen;
    for (uint32_t k = 0; k < n_prims; ++k) {
        uint32_t ts = prims[k].index_start / 3, te = prims[k].index_end / 3;
        if (te <= ts || te > n_indices / 3) return set_error(NB_ERR_INVALID_ARGUMENT, "primitive %u has an empty or out-of-range index range", k);
        if ((flags & NB_BUILD_SHARE_BLAS) && !seen.emplace(std::make_pair(ts, te), 1).second) continue;
        total += 2 * tris_of(prims[k]) - 1;
    }
    if (total > 0xFFFFFFFFull) return set_error(NB_ERR_UNSUPPORTED, "BLAS node count exceeds 2^32");
    *blas_cap = static_cast<uint32_t>(total);
    *tlas_cap = n_prims ? 2 * n_prims - 1 : 0;
    return NB_OK;
}

nb_status build_scene(const nb_vertex *vertices, uint32_t n_vertices, uint32_t *indices, uint32_t n_indices,
                      const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                      nb_bvh_node *blas_out, uint32_t blas_cap, uint32_t *n_blas,
                      nb_bvh_node *tlas_out, uint32_t tlas_cap, uint32_t *n_tlas, nb_instance *inst_out) {
    if (n_prims == 0) return set_error(NB_ERR_INVALID_ARGUMENT, "scene has no primitives");
    uint32_t need_blas = 0, need_tlas = 0;
    nb_status st = build_scene_capacity(n_indices, prims, n_prims, flags, &need_blas, &need_tlas);
    if (st != NB_OK) return st;
    if (blas_cap < need_blas || tlas_cap < need_tlas)
        return set_error(NB_ERR_INVALID_ARGUMENT, "output capacity too small (need %u BLAS, %u TLAS nodes)", need_blas, need_tlas);
    const uint32_t n_tris = n_indices / 3;
    for (uint32_t i = 0; i < n_tris * 3; ++i)
        if (indices[i] >= n_vertices) return set_error(NB_ERR_INVALID_ARGUMENT, "index %u out of range", i);

    // build_triangle_cache (bvh.rs:90-104)
    Prims tris;
    tris.width = 3;
    tris.lo.resize(n_tris); tris.hi.resize(n_tris); tris.ctr.resize(n_tris);
    tris.payload.assign(indices, indices + 3 * static_cast<size_t>(n_tris));
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float *p0 = vertices[indices[3 * t]].position, *p1 = vertices[indices[3 * t + 1]].position,
                    *p2 = vertices[indices[3 * t + 2]].position;
        V3 a = v3(p0[0], p0[1], p0[2]), b = v3(p1[0], p1[1], p1[2]), c = v3(p2[0], p2[1], p2[2]);
        V3 s = (a + b) + c;
        tris.ctr[t] = v3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f);
        tris.lo[t] = vmin(vmin(a, b), c);
        tris.hi[t] = vmax(vmax(a, b), c);
    }

    // Chains of primitives per distinct triangle range; chains over disjoint ranges build in
    // parallel.  Overlapping but unequal ranges make builds order-dependent: build serially.
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> chains;
    for (uint32_t k = 0; k < n_prims; ++k) chains[{prims[k].index_start / 3, prims[k].index_end / 3}].push_back(k);
    bool disjoint = true;
    uint32_t prev_end = 0;
    for (auto &kv : chains) { if (kv.first.first < prev_end) disjoint = false; prev_end = std::max(prev_end, kv.first.second); }
    if (!disjoint && (flags & NB_BUILD_SHARE_BLAS))
        return set_error(NB_ERR_UNSUPPORTED, "NB_BUILD_SHARE_BLAS needs index ranges that are equal or disjoint");

    std::vector<std::vector<nb_bvh_node>> local(n_prims);   // per primitive, indices relative to 0
    std::vector<int64_t> shared_with(n_prims, -1);
    auto build_chain = [&](const std::vector<uint32_t> &chain) {
        Builder b(tris);
        for (size_t c = 0; c < chain.size(); ++c) {
            uint32_t k = chain[c];
            if (c > 0 && (flags & NB_BUILD_SHARE_BLAS)) { shared_with[k] = chain[0]; continue; }
            b.append(local[k], 0, prims[k].index_start / 3, prims[k].index_end / 3);
        }
    };
    if (disjoint) {
        std::vector<const std::vector<uint32_t> *> work;
        for (auto &kv : chains) work.push_back(&kv.second);
        std::sort(work.begin(), work.end(), [&](auto *a, auto *b) {
            return tris_of(prims[a->front()]) * a->size() > tris_of(prims[b->front()]) * b->size();
        });
        unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
        nt = std::min<unsigned>(nt, static_cast<unsigned>(work.size()));
        std::atomic_size_t next{0};
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t)
            pool.emplace_back([&] { for (size_t i; (i = next.fetch_add(1)) < work.size();) build_chain(*work[i]); });
        for (auto &th : pool) th.join();
    } else {
        Builder b(tris);
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> first;
        for (uint32_t k = 0; k < n_prims; ++k) {
            auto key = std::make_pair(prims[k].index_start / 3, prims[k].index_end / 3);
            auto it = first.find(key);
            if (it != first.end() && (flags & NB_BUILD_SHARE_BLAS)) { shared_with[k] = it->second; continue; }
            first.emplace(key, k);
            b.append(local[k], 0, key.first, key.second);
        }
    }

    // Relocate into the shared node array in primitive order (scene.rs:284-288).
    std::vector<uint32_t> root(n_prims);
    uint32_t off = 0;
    for (uint32_t k = 0; k < n_prims; ++k) {
        if (shared_with[k] >= 0) { root[k] = root[shared_with[k]]; continue; }
        const uint32_t base = off;
        root[k] = base;
        for (const nb_bvh_node &n : local[k]) {
            nb_bvh_node m = n;
            if (m.end == 0) m.start += base;
            blas_out[off++] = m;
        }
        std::vector<nb_bvh_node>().swap(local[k]);
    }
    *n_blas = off;

    // Instances with world bounds (scene.rs:239-260, 290-298) and the TLAS (scene.rs:301-302).
    std::vector<nb_instance> inst(n_prims);
    Prims ib;
    ib.width = 1;
    ib.lo.resize(n_prims); ib.hi.resize(n_prims); ib.ctr.resize(n_prims); ib.payload.resize(n_prims);
    for (uint32_t k = 0; k < n_prims; ++k) {
        const nb_primitive &p = prims[k];
        nb_instance &in = inst[k];
        mat4_inverse(p.local_to_world, in.world_to_local);
        std::memcpy(in.local_to_world, p.local_to_world, sizeof(in.local_to_world));
        std::memcpy(in.color, p.color, sizeof(in.color));
        in.roughness = p.roughness; in.metallic = p.metallic; in.emissive = p.emissive;
        in.node = root[k];
        const nb_bvh_node &r = blas_out[root[k]];
        V3 lo = v3(kInf, kInf, kInf), hi = v3(-kInf, -kInf, -kInf);
        for (int c = 0; c < 8; ++c) {
            V3 corner = v3((c & 1) ? r.max[0] : r.min[0], (c & 2) ? r.max[1] : r.min[1], (c & 4) ? r.max[2] : r.min[2]);
            V3 w = transform_point(p.local_to_world, corner);
            lo = vmin(lo, w); hi = vmax(hi, w);
        }
        V3 s = lo + hi;
        ib.lo[k] = lo; ib.hi[k] = hi; ib.ctr[k] = v3(s.x * 0.5f, s.y * 0.5f, s.z * 0.5f);
        ib.payload[k] = k;
    }
    std::vector<nb_bvh_node> tlas;
    tlas.reserve(2 * n_prims);
    Builder(ib).append(tlas, 0, 0, n_prims);
    std::copy(tlas.begin(), tlas.end(), tlas_out);
    *n_tlas = static_cast<uint32_t>(tlas.size());
    for (uint32_t k = 0; k < n_prims; ++k) inst_out[k] = inst[ib.payload[k]];

    // flatten_triangle_list (bvh.rs:106-110)
    std::copy(tris.payload.begin(), tris.payload.end(), indices);
    return NB_OK;
}

void camera_look_at(const float *eye, const float *target, const float *up, float fov_y, float aspect,
                    float z_near, nb_camera *out) {
    auto sub = [](V3 a, V3 b) { return a - b; };
    auto dot = [](V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    auto cross = [](V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); };
    auto norm = [&](V3 a) { float s = 1.0f / std::sqrt(dot(a, a)); return v3(a.x * s, a.y * s, a.z * s); };
    V3 e = v3(eye[0], eye[1], eye[2]), t = v3(target[0], target[1], target[2]), u0 = v3(up[0], up[1], up[2]);
    V3 f = norm(sub(t, e));             // glam look_at_rh -> look_to_rh(eye, center - eye, up)
    V3 s = norm(cross(f, u0));
    V3 u = cross(s, f);
    const float view[16] = {s.x, u.x, -f.x, 0.0f, s.y, u.y, -f.y, 0.0f, s.z, u.z, -f.z, 0.0f,
                            -dot(e, s), -dot(e, u), dot(e, f), 1.0f};
    float fl = 1.0f / std::tan(0.5f * fov_y);   // perspective_infinite_rh
    const float proj[16] = {fl / aspect, 0, 0, 0, 0, fl, 0, 0, 0, 0, -1.0f, -1.0f, 0, 0, -z_near, 0};
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            out->world_to_clip[j * 4 + r] = ((proj[0 * 4 + r] * view[j * 4 + 0] + proj[1 * 4 + r] * view[j * 4 + 1]) +
                                             proj[2 * 4 + r] * view[j * 4 + 2]) + proj[3 * 4 + r] * view[j * 4 + 3];
    mat4_inverse(out->world_to_clip, out->clip_to_world);
}

}  // namespace nb
