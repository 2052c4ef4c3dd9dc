// C ABI of libnbounce: contexts, device scenes, Sobol table, environment cube, render targets
// and render calls (declared in include/nbounce.h).
#include "nbounce.h"
#include "nb_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace nb {

static thread_local std::string g_error;

nb_status set_error(nb_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return st;
}

#define NB_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_error(e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "%s: %s", \
                             #call, hipGetErrorString(e_));                                            \
    } while (0)

}  // namespace nb

using namespace nb;

// ---- scratch of nb_primary_hits_noalloc ------------------------------------------------------
// A file-scope device array, so that nb_primary_hits_noalloc allocates nothing: each GPU holds
// its own copy for the life of the process, 5 words per pixel for up to kHitsScratchPixels
// pixels (40 MiB).  In host code `hits_scratch` only names the symbol; its device address on a
// GPU comes from hipGetSymbolAddress with that GPU current, taken once per device when the first
// context on it is created and kept in g_hits_scratch.
constexpr uint32_t kHitsScratchPixels = 1u << 21;
static __device__ __align__(256) uint32_t hits_scratch[5u * kHitsScratchPixels];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_hits_scratch_mutex;
uint32_t *g_hits_scratch[kMaxDevices] = {};

// Device address of hits_scratch on `device` (the current device), looked up once per device.
nb_status hits_scratch_address(int device, uint32_t **out) {
    if (device < 0 || device >= kMaxDevices) return set_error(NB_ERR_UNSUPPORTED, "device %d: at most %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_hits_scratch_mutex);
    if (!g_hits_scratch[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(hits_scratch));
        if (e != hipSuccess || !ptr)
            return set_error(NB_ERR_HIP, "scratch symbol address on device %d: %s", device, hipGetErrorString(e));
        g_hits_scratch[device] = static_cast<uint32_t *>(ptr);
    }
    *out = g_hits_scratch[device];
    return NB_OK;
}

nb_status use_device(const nb_context *ctx) {
    NB_HIP(hipSetDevice(ctx->device));
    return NB_OK;
}

#define NB_TRY(expr)                   \
    do {                               \
        nb_status s_ = (expr);         \
        if (s_ != NB_OK) return s_;    \
    } while (0)

// ---- scene re-layout ------------------------------------------------------------------------

struct Relayout {
    std::vector<float4> pairs, tris, inst, ishade, vshade;
    std::vector<uint32_t> tri_leaf, inst_leaf;
    uint32_t tlas_root = kNoRecord;
};

// Validates one tree (children after parents, leaves inside [0, n_prims), depth <= 32).
nb_status check_tree(const nb_bvh_node *nodes, uint32_t n, uint32_t n_prims, const char *name) {
    std::vector<uint8_t> depth(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const nb_bvh_node &d = nodes[i];
        if (d.end > 0) {
            if (d.start >= d.end || d.end > n_prims)
                return set_error(NB_ERR_BAD_LAYOUT, "%s node %u: leaf range [%u, %u) outside [0, %u)", name, i, d.start, d.end, n_prims);
            if (d.start > kLeafStartMask)
                return set_error(NB_ERR_UNSUPPORTED, "%s node %u: leaf start %u exceeds 2^26", name, i, d.start);
        } else {
            if (d.start <= i || d.start >= n - 1)   // (d.start + 1 would wrap for start = 2^32 - 1)
                return set_error(NB_ERR_BAD_LAYOUT, "%s node %u: children %u, %u not after it", name, i, d.start, d.start + 1);
            uint8_t cd = static_cast<uint8_t>(depth[i] + 1);
            if (cd > kMaxTreeDepth)
                return set_error(NB_ERR_BAD_LAYOUT, "%s deeper than %u levels", name, kMaxTreeDepth);
            if (cd > depth[d.start]) depth[d.start] = cd;
            if (cd > depth[d.start + 1]) depth[d.start + 1] = cd;
        }
    }
    return NB_OK;
}

nb_status child_ref(const nb_bvh_node &n, const std::vector<uint32_t> &pid, uint32_t idx, std::vector<uint32_t> &side,
                    uint32_t *out) {
    if (n.end == 0) { *out = pid[idx]; return NB_OK; }
    uint32_t count = n.end - n.start;
    uint32_t field = count <= kLeafMaxInline ? count : 0u;
    if (field == 0) {
        if (side[n.start] != 0 && side[n.start] != count)
            return set_error(NB_ERR_UNSUPPORTED, "two leaves of more than %u primitives start at %u with different counts",
                             kLeafMaxInline, n.start);
        side[n.start] = count;
    }
    *out = kLeafBit | (field << kLeafStartBits) | n.start;
    return NB_OK;
}

// One child-pair record (4 float4): boxes of l and r (r == nullptr: an all-NaN box) and refs.
void make_pair(const nb_bvh_node &l, uint32_t lref, const nb_bvh_node *r, uint32_t rref, uint32_t is_root, float4 *q) {
    const float nan = std::nanf("");
    float rmin[3] = {nan, nan, nan}, rmax[3] = {nan, nan, nan};
    if (r) { std::memcpy(rmin, r->min, 12); std::memcpy(rmax, r->max, 12); }
    q[0] = make_float4(l.min[0], l.min[1], l.min[2], l.max[0]);
    q[1] = make_float4(l.max[1], l.max[2], rmin[0], rmin[1]);
    q[2] = make_float4(rmin[2], rmax[0], rmax[1], rmax[2]);
    uint32_t w[4] = {lref, rref, is_root, 0u};
    std::memcpy(&q[3], w, 16);
}

// ---- TLAS refit ------------------------------------------------------------------------------
// The reference bounds an instance by the 8 transformed corners of its BLAS root box
// (scene.rs:239-260).  For a rotated mesh that box is loose (a rotated cube's corners reach up
// to sqrt(3) times further along an axis than its contents), so rays enter instances and TLAS
// nodes they miss.  The kernels traverse the same TLAS topology with boxes refitted bottom-up to
// each instance's transformed vertices, computed in double precision and padded outward by
// 2^-14 of the coordinates' magnitude (well above the float error of the ray transform and the
// slab test, so every ray that hits an instance's geometry still enters it at a distance no
// greater than its hit).  The closest hit is the same; only which nodes are visited on the way
// changes.  A bound that is not finite keeps the reference box.  NB_TLAS_REFIT=0 keeps the
// reference boxes, read at scene creation.
bool tlas_refit_enabled() {
    const char *e = std::getenv("NB_TLAS_REFIT");
    return !(e && std::strcmp(e, "0") == 0);
}

nb_status tlas_refit(const nb_bvh_node *blas, const nb_bvh_node *tlas, uint32_t n_tlas, const nb_instance *inst,
                     uint32_t n_inst, const nb_vertex *vert, const uint32_t *idx, std::vector<nb_bvh_node> &out) {
    out.assign(tlas, tlas + n_tlas);
    if (!tlas_refit_enabled() || n_tlas == 0) return NB_OK;
    try {
        // the unique vertices each instanced BLAS root covers (its leaves' triangle range), kept
        // per range: the reference's one-BLAS-per-instance layout has many roots over one range
        std::map<uint32_t, uint64_t> range_of;                     // root -> lo << 32 | hi
        std::map<uint64_t, std::vector<uint32_t>> verts_of;        // range -> vertex ids
        std::vector<uint32_t> stack;
        for (uint32_t j = 0; j < n_inst; ++j) {
            const uint32_t root = inst[j].node;
            if (range_of.count(root)) continue;
            uint32_t lo = 0xFFFFFFFFu, hi = 0;
            stack.assign(1, root);
            while (!stack.empty()) {
                const nb_bvh_node &n = blas[stack.back()];
                stack.pop_back();
                if (n.end > 0) {
                    lo = std::min(lo, n.start);
                    hi = std::max(hi, n.end);
                } else {
                    stack.push_back(n.start);
                    stack.push_back(n.start + 1);
                }
            }
            const uint64_t key = hi > lo ? (static_cast<uint64_t>(lo) << 32) | hi : 0u;
            range_of[root] = key;
            if (verts_of.count(key)) continue;
            std::vector<uint32_t> &v = verts_of[key];
            if (hi > lo) {
                v.assign(idx + 3 * static_cast<size_t>(lo), idx + 3 * static_cast<size_t>(hi));
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
            }
        }
        // each instance's padded world bounds (min xyz, max xyz); NaN when not finite
        std::vector<std::array<float, 6>> ib(n_inst);
        for (uint32_t j = 0; j < n_inst; ++j) {
            const float *m = inst[j].local_to_world;   // column-major
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t vi : verts_of[range_of[inst[j].node]]) {
                const float *p = vert[vi].position;
                for (int a = 0; a < 3; ++a) {
                    const double w = static_cast<double>(m[a]) * p[0] + static_cast<double>(m[4 + a]) * p[1] +
                                     static_cast<double>(m[8 + a]) * p[2] + static_cast<double>(m[12 + a]);
                    lo[a] = std::min(lo[a], w);
                    hi[a] = std::max(hi[a], w);
                }
            }
            bool ok = true;
            for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(lo[a]) && std::isfinite(hi[a]);
            for (int a = 0; a < 3; ++a) {
                if (!ok) { ib[j][a] = ib[j][3 + a] = NAN; continue; }
                const double pad = std::ldexp(std::max({std::fabs(lo[a]), std::fabs(hi[a]), hi[a] - lo[a]}), -14);
                ib[j][a] = std::nextafter(static_cast<float>(lo[a] - pad), -INFINITY);
                ib[j][3 + a] = std::nextafter(static_cast<float>(hi[a] + pad), INFINITY);
            }
        }
        // bottom-up: children follow their parents in the node array
        for (uint32_t i = n_tlas; i-- > 0;) {
            nb_bvh_node &n = out[i];
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            bool ok = true;
            auto add = [&](const float *bmin, const float *bmax) {
                for (int a = 0; a < 3; ++a) {
                    ok = ok && std::isfinite(bmin[a]) && std::isfinite(bmax[a]);
                    lo[a] = std::min(lo[a], bmin[a]);
                    hi[a] = std::max(hi[a], bmax[a]);
                }
            };
            if (n.end > 0) {
                for (uint32_t j = n.start; j < n.end; ++j) add(ib[j].data(), ib[j].data() + 3);
            } else {
                add(out[n.start].min, out[n.start].max);
                add(out[n.start + 1].min, out[n.start + 1].max);
            }
            if (!ok) continue;   // keeps the reference box
            std::memcpy(n.min, lo, sizeof(lo));
            std::memcpy(n.max, hi, sizeof(hi));
        }
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    return NB_OK;
}

nb_status relayout(const nb_bvh_node *blas, uint32_t n_blas, const nb_bvh_node *tlas, uint32_t n_tlas,
                   const nb_instance *inst, uint32_t n_inst, const nb_vertex *vert, uint32_t n_vert,
                   const uint32_t *idx, uint32_t n_idx, Relayout &R) {
    const uint32_t n_tris = n_idx / 3;
    for (uint32_t i = 0; i < n_tris * 3; ++i)
        if (idx[i] >= n_vert) return set_error(NB_ERR_BAD_LAYOUT, "index %u = %u out of range (%u vertices)", i, idx[i], n_vert);
    NB_TRY(check_tree(blas, n_blas, n_tris, "BLAS"));
    NB_TRY(check_tree(tlas, n_tlas, n_inst, "TLAS"));
    if (n_tlas > 0 && n_inst == 0) return set_error(NB_ERR_BAD_LAYOUT, "TLAS without instances");
    if (n_inst > kLeafStartMask) return set_error(NB_ERR_UNSUPPORTED, "more than 2^26 instances");
    // the kernels address triangles (48 B each) with 32-bit byte offsets
    if (static_cast<uint64_t>(n_tris) * 48u > 0xFFFFFFFFull) return set_error(NB_ERR_UNSUPPORTED, "more than 89478485 triangles");
    for (uint32_t j = 0; j < n_inst; ++j)
        if (inst[j].node >= n_blas) return set_error(NB_ERR_BAD_LAYOUT, "instance %u: BLAS root %u out of range", j, inst[j].node);
    // the TLAS the kernels traverse: the given topology with boxes refitted to the instances'
    // transformed geometry (tlas_refit)
    std::vector<nb_bvh_node> tl;
    NB_TRY(tlas_refit(blas, tlas, n_tlas, inst, n_inst, vert, idx, tl));
    tlas = tl.data();

    // Record ids in depth-first preorder of each tree (left child first), each tree preceded by
    // its root record: a record's left child is the next record, so descents stay in one cache
    // line half of the time.  BLAS trees in order of first use by an instance, then the TLAS.
    // Nodes no instance reaches get no record.
    std::vector<uint32_t> pid_b(n_blas, kNoRecord), pid_t(n_tlas, kNoRecord);
    struct Rec { uint8_t kind; uint32_t node; };     // 0 BLAS inner, 1 BLAS root, 2 TLAS inner, 3 TLAS root
    std::vector<Rec> recs;
    std::map<uint32_t, uint32_t> root_rec;
    std::vector<uint32_t> stack;
    auto dfs = [&](const nb_bvh_node *nodes, uint32_t root, std::vector<uint32_t> &pid, uint8_t kind) {
        if (nodes[root].end != 0) return;
        stack.assign(1, root);
        while (!stack.empty()) {
            const uint32_t i = stack.back();
            stack.pop_back();
            if (pid[i] != kNoRecord) continue;   // a node reached twice keeps its first record
            pid[i] = static_cast<uint32_t>(recs.size());
            recs.push_back({kind, i});
            const uint32_t l = nodes[i].start, r = l + 1;
            if (nodes[r].end == 0) stack.push_back(r);
            if (nodes[l].end == 0) stack.push_back(l);
        }
    };
    for (uint32_t j = 0; j < n_inst; ++j) {
        const uint32_t root = inst[j].node;
        if (root_rec.count(root)) continue;
        root_rec[root] = static_cast<uint32_t>(recs.size());
        recs.push_back({1, root});
        if (pid_b[root] == kNoRecord) dfs(blas, root, pid_b, 0);
    }
    if (n_tlas > 0) {
        R.tlas_root = static_cast<uint32_t>(recs.size());
        recs.push_back({3, 0u});
        dfs(tlas, 0, pid_t, 2);
    }

    // the kernels address records with 32-bit byte offsets (64 B each)
    if (recs.size() > (size_t{1} << 26)) return set_error(NB_ERR_UNSUPPORTED, "more than 2^26 node records");
    R.tri_leaf.assign(n_tris ? n_tris : 1, 0u);
    R.inst_leaf.assign(n_inst ? n_inst : 1, 0u);
    auto pair_of = [&](const nb_bvh_node *nodes, std::vector<uint32_t> &pid, std::vector<uint32_t> &side, uint32_t i,
                       float4 *q) -> nb_status {
        const uint32_t l = nodes[i].start, r = l + 1;
        uint32_t lr, rr;
        NB_TRY(child_ref(nodes[l], pid, l, side, &lr));
        NB_TRY(child_ref(nodes[r], pid, r, side, &rr));
        make_pair(nodes[l], lr, &nodes[r], rr, 0u, q);
        return NB_OK;
    };
    R.pairs.assign(kRecordFloat4 * recs.size(), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (size_t k = 0; k < recs.size(); ++k) {
        const Rec &rc = recs[k];
        const bool tl = rc.kind >= 2;
        const nb_bvh_node *nodes = tl ? tlas : blas;
        std::vector<uint32_t> &pid = tl ? pid_t : pid_b;
        std::vector<uint32_t> &side = tl ? R.inst_leaf : R.tri_leaf;
        float4 *q = R.pairs.data() + kRecordFloat4 * k;
        if (rc.kind == 1 || rc.kind == 3) {
            uint32_t ref;
            NB_TRY(child_ref(nodes[rc.node], pid, rc.node, side, &ref));
            make_pair(nodes[rc.node], ref, nullptr, 0u, 1u, q);
        } else {
            NB_TRY(pair_of(nodes, pid, side, rc.node, q));
        }
    }

    R.tris.resize(3 * static_cast<size_t>(n_tris));
    for (uint32_t t = 0; t < n_tris; ++t) {
        const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
        const float *p0 = vert[i0].position, *p1 = vert[i1].position, *p2 = vert[i2].position;
        float b0, b1, b2;
        std::memcpy(&b0, &i0, 4); std::memcpy(&b1, &i1, 4); std::memcpy(&b2, &i2, 4);
        R.tris[3 * t + 0] = make_float4(p0[0], p0[1], p0[2], b0);
        R.tris[3 * t + 1] = make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], b1);
        R.tris[3 * t + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], b2);
    }
    R.inst.resize(4 * static_cast<size_t>(n_inst));
    R.ishade.resize(6 * static_cast<size_t>(n_inst));
    for (uint32_t j = 0; j < n_inst; ++j) {
        const float *w = inst[j].world_to_local, *l = inst[j].local_to_world;
        for (int r = 0; r < 3; ++r) R.inst[4 * j + r] = make_float4(w[0 + r], w[4 + r], w[8 + r], w[12 + r]);
        float root_bits;
        uint32_t rid = root_rec[inst[j].node];
        std::memcpy(&root_bits, &rid, 4);
        R.inst[4 * j + 3] = make_float4(root_bits, 0.0f, 0.0f, 0.0f);
        R.ishade[6 * j + 0] = make_float4(w[0], w[1], w[2], inst[j].roughness);
        R.ishade[6 * j + 1] = make_float4(w[4], w[5], w[6], inst[j].metallic);
        R.ishade[6 * j + 2] = make_float4(w[8], w[9], w[10], inst[j].emissive);
        R.ishade[6 * j + 3] = make_float4(l[0], l[1], l[2], inst[j].color[0]);
        R.ishade[6 * j + 4] = make_float4(l[4], l[5], l[6], inst[j].color[1]);
        R.ishade[6 * j + 5] = make_float4(l[8], l[9], l[10], inst[j].color[2]);
    }
    R.vshade.resize(2 * static_cast<size_t>(n_vert));
    for (uint32_t v = 0; v < n_vert; ++v) {
        const nb_vertex &x = vert[v];
        R.vshade[2 * v] = make_float4(x.normal[0], x.normal[1], x.normal[2], 0.0f);
        R.vshade[2 * v + 1] = make_float4(x.tangent[0], x.tangent[1], x.tangent[2], x.tangent[3]);
    }
    return NB_OK;
}

// ---- cube borders ---------------------------------------------------------------------------
// Vulkan cube faces: major axis (index, sign), then the (axis, sign) giving sc and tc.
struct FaceAxes { int ma; float ms; int sa; float ss; int ta; float ts; };
constexpr FaceAxes kFaces[6] = {
    {0, 1, 2, -1, 1, -1}, {0, -1, 2, 1, 1, -1}, {1, 1, 0, 1, 2, 1},
    {1, -1, 0, 1, 2, -1}, {2, 1, 0, 1, 1, -1}, {2, -1, 0, -1, 1, -1},
};

// The texel across the edge of face f for (i, j) with exactly one coordinate one step outside.
void across_edge(int f, int i, int j, int n, int &g, int &gi, int &gj) {
    const FaceAxes &F = kFaces[f];
    const bool s_out = i < 0 || i >= n;
    const float side = s_out ? (i < 0 ? -1.0f : 1.0f) : (j < 0 ? -1.0f : 1.0f);
    const int out_axis = s_out ? F.sa : F.ta;
    const float out_sign = side * (s_out ? F.ss : F.ts);     // direction component on the out axis
    const int along_axis = s_out ? F.ta : F.sa;
    const float along_sign = s_out ? F.ts : F.ss;
    const int along = s_out ? j : i;
    g = out_axis * 2 + (out_sign > 0.0f ? 0 : 1);
    const FaceAxes &G = kFaces[g];
    // On g, the axis that was f's major axis sits at the edge (value F.ms); the other keeps `along`.
    if (G.sa == F.ma) {
        gi = G.ss * F.ms > 0.0f ? n - 1 : 0;
        gj = G.ts * along_sign > 0.0f ? along : n - 1 - along;
    } else {
        gj = G.ts * F.ms > 0.0f ? n - 1 : 0;
        gi = G.ss * along_sign > 0.0f ? along : n - 1 - along;
    }
    (void)along_axis;
}

float4 face_texel(const float *tex, int n, int f, int i, int j) {
    const float *t = tex + 4 * ((static_cast<size_t>(f) * n + j) * n + i);
    return make_float4(t[0], t[1], t[2], t[3]);
}

// Padded faces: (n+2)^2 texels each; border texels hold the seamless neighbours, corners the
// average of the face's corner texel and its two neighbours across the edges.
std::vector<float4> pad_cube(const float *tex, int n) {
    const int pn = n + 2;
    std::vector<float4> out(6 * static_cast<size_t>(pn) * pn);
    for (int f = 0; f < 6; ++f)
        for (int pj = 0; pj < pn; ++pj)
            for (int pi = 0; pi < pn; ++pi) {
                const int i = pi - 1, j = pj - 1;
                const bool oi = i < 0 || i >= n, oj = j < 0 || j >= n;
                float4 v;
                if (!oi && !oj) {
                    v = face_texel(tex, n, f, i, j);
                } else if (oi && oj) {
                    const int ci = i < 0 ? 0 : n - 1, cj = j < 0 ? 0 : n - 1;
                    int g1, i1, j1, g2, i2, j2;
                    across_edge(f, i, cj, n, g1, i1, j1);
                    across_edge(f, ci, j, n, g2, i2, j2);
                    float4 a = face_texel(tex, n, f, ci, cj), b = face_texel(tex, n, g1, i1, j1),
                           c = face_texel(tex, n, g2, i2, j2);
                    v = make_float4(((a.x + b.x) + c.x) / 3.0f, ((a.y + b.y) + c.y) / 3.0f, ((a.z + b.z) + c.z) / 3.0f,
                                    ((a.w + b.w) + c.w) / 3.0f);
                } else {
                    int g, gi, gj;
                    across_edge(f, i, j, n, g, gi, gj);
                    v = face_texel(tex, n, g, gi, gj);
                }
                out[(static_cast<size_t>(f) * pn + pj) * pn + pi] = v;
            }
    return out;
}

nb_status check_render_args(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                            const nb_globals *g, const nb_envmap *env) {
    if (!ctx || !scene || !target || !camera || !g || !env)
        return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx || target->ctx != ctx || env->ctx != ctx)
        return set_error(NB_ERR_INVALID_ARGUMENT, "scene, target and environment must belong to the context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (g->bounces > kMaxBounceLimit)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bounces %u exceeds %u", g->bounces, kMaxBounceLimit);
    return NB_OK;
}

RenderParams make_params(nb_context *ctx, const nb_scene *scene, nb_target *t, const nb_camera *cam,
                         const nb_globals *g, const nb_envmap *env) {
    RenderParams p{};
    p.scene = scene->dev;
    p.env = dev_env(env);
    p.target = t->pixels;
    p.width = t->width;
    p.height = t->height;
    std::memcpy(p.c2w, cam->clip_to_world, sizeof(p.c2w));
    p.sobol = ctx->sobol;
    p.sobol_entries = ctx->sobol_entries;
    p.sobol_stride = ctx->sobol_stride;
    p.first_sample = g->sample;
    p.n_frames = 1;
    p.first_weight = g->weight;
    p.weights_from_sample = 0;
    p.bounces = g->bounces;
    p.contribution_factor = g->contribution_factor;
    p.counters = ctx->counting ? ctx->counters : nullptr;
    p.tiles_x = t->width / 8;
    p.n_tiles = (t->width / 8) * (t->height / 8);
    p.tiles_per_xcd = (p.n_tiles + 7) / 8;
    p.hits_out = nullptr;
    p.work = ctx->work;
    p.rr_first = ctx->rr_first;
    return p;
}

// ---- sample-parallel frames ----------------------------------------------------------------
// nb_render_frames keeps a pixel's frames in one lane (the mix chain runs in order).  A target of
// fewer pixels than the GPU holds lanes (4096 resident waves x 64 at 4 waves per SIMD) leaves
// most of them idle, so such calls render one path per (frame, pixel) with the serving kernels
// instead: the frames' paths in 16-path units (serve_pt_kernel, one request, one pass), each
// path's radiance in a result slot, then the ordered accumulation mixes each pixel's samples in
// frame order with weight 1/sample.  Every sample and the mix chain are the same, so the image
// is the same bit for bit.  NB_SAMPLE_PARALLEL=0 turns this off, NB_SAMPLE_PARALLEL_PIXELS
// moves the size limit.
constexpr uint32_t kSpPathsPerLaunch = 1u << 22;   // result slots: 64 MiB

uint32_t sample_parallel_pixels() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_SAMPLE_PARALLEL");
        if (e && std::atoi(e) == 0) return 0u;
        const char *n = std::getenv("NB_SAMPLE_PARALLEL_PIXELS");
        return n ? static_cast<uint32_t>(std::strtoul(n, nullptr, 10)) : (1u << 19);
    }();
    return v;
}

struct SpDesc {
    ServeStep step;
    ServePass pass;
    ServeGroup group;
    ServeRequestDev request;
    ServeSceneDev scene;
};

nb_status render_frames_sample_parallel(nb_context *ctx, const nb_scene *scene, nb_target *t, const nb_camera *cam,
                                        const nb_globals *first, uint32_t n_frames, const nb_envmap *env) {
    const uint32_t pixels = t->width * t->height;
    if (!ctx->sp_block) {
        const size_t desc = (sizeof(SpDesc) + 255) / 256 * 256;
        const size_t bytes = desc + kWorkStride * sizeof(uint32_t) + static_cast<size_t>(kSpPathsPerLaunch) * 16;
        NB_HIP(hipMalloc(&ctx->sp_block, bytes));
        ctx->sp_results = reinterpret_cast<float4 *>(static_cast<char *>(ctx->sp_block) + desc + kWorkStride * sizeof(uint32_t));
        ctx->sp_max_paths = kSpPathsPerLaunch;
    }
    SpDesc *d = static_cast<SpDesc *>(ctx->sp_block);
    uint32_t *work = reinterpret_cast<uint32_t *>(static_cast<char *>(ctx->sp_block) + (sizeof(SpDesc) + 255) / 256 * 256);
    const uint32_t per_launch = std::max(1u, ctx->sp_max_paths / pixels);
    for (uint32_t done = 0; done < n_frames;) {
        const uint32_t count = std::min(per_launch, n_frames - done);
        SpDesc h{};
        h.scene = ServeSceneDev{scene->dev, dev_env(env)};
        std::memcpy(h.request.c2w, cam->clip_to_world, sizeof(h.request.c2w));
        h.request.target = t->pixels;
        h.request.width = t->width;
        h.request.height = t->height;
        h.request.scene = 0;
        h.request.bounces = first->bounces;
        h.request.contribution_factor = first->contribution_factor;
        h.request.rr_first = ctx->rr_first;
        h.pass = ServePass{0u, first->sample + done, count, pixels, 0u, 0u, 0u, 0u};
        h.group = ServeGroup{0u, 0u, count * (pixels / 16u), 0u, 1u, {0u, 0u, 0u}};
        ServeStep &st = h.step;
        st.scenes = &d->scene;
        st.requests = &d->request;
        st.passes = &d->pass;
        st.groups = &d->group;
        st.n_passes = 1;
        st.n_groups = 1;
        st.n_units = count * (pixels / 16u);
        st.n_pixels = pixels;
        st.sobol = ctx->sobol;
        st.sobol_entries = ctx->sobol_entries;
        st.sobol_stride = ctx->sobol_stride;
        st.results = ctx->sp_results;
        st.work = work;
        st.counters = ctx->counting ? ctx->counters : nullptr;
        st.regen_min_idle = 24;
        st.inner_quorum = 28;
        // 2-unit grabs (32 paths): a 256x256, 16-frame launch holds about 4 paths per lane, and
        // a wave's last grab sets its tail (Mrays/s: 1 unit 2157, 2 units 2222-2227, 4 units
        // 2182-2212, 8 units 2045-2082; refills after 16 or 32 waiting lanes instead of 24: 2155,
        // 2134)
        st.grab_units = 2;
        NB_HIP(hipMemcpyAsync(d, &h, sizeof(SpDesc), hipMemcpyHostToDevice, ctx->stream));
        NB_TRY(serve_launch_reset(&d->step, ctx->stream));
        NB_TRY(serve_launch_step(&d->step, ctx->stream, 4u));
        int cus = 0;
        NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        NB_TRY(serve_launch_accumulate(&d->step, static_cast<uint32_t>(cus > 0 ? cus : 64) * 8u, ctx->stream));
        done += count;
    }
    return NB_OK;
}

// ---- pipelined nb_render -----------------------------------------------------------------------
// One nb_render launch leaves CUs idle while its last, longest paths finish.  Consecutive frames'
// paths are independent; only the mix into the target depends on the previous frame.  So each
// frame's path kernel runs on one of two side streams and writes the frame's samples to one of
// the target's two sample buffers, and a mix kernel on the context stream applies the
// accumulation step in frame order.  A path kernel waits for the context stream as it was just
// before the previous frame's mix was enqueued: that includes every earlier mix (the last reader
// of the sample buffer it writes), not the previous frame's paths, so the two frames' paths run
// side by side.  After a counter reset it waits for all of the context stream instead.  The
// work-unit order's cost halving and rebuilds on the context stream may overlap a path kernel
// (order.hip: atomic costs, two order buffers).  The image is the same bit for bit (the mix is
// the same binary32 expression).  NB_PIPELINE=0 renders on the context stream, one launch after
// another.
bool pipeline_enabled() {
    const char *e = std::getenv("NB_PIPELINE");
    return !(e && std::strcmp(e, "0") == 0);
}

nb_status pipe_setup(nb_context *ctx, nb_target *t) {
    if (!ctx->pipe_gate) {
        for (int i = 0; i < 2; ++i) {
            NB_HIP(hipStreamCreateWithFlags(&ctx->pipe_stream[i], hipStreamNonBlocking));
            NB_HIP(hipMalloc(&ctx->pipe_work[i], kWorkBytes));
            NB_HIP(hipMemset(ctx->pipe_work[i], 0, kWorkBytes));
        }
        NB_HIP(hipEventCreateWithFlags(&ctx->pipe_done, hipEventDisableTiming));
        NB_HIP(hipEventCreateWithFlags(&ctx->pipe_gate, hipEventDisableTiming));
    }
    if (!t->samples[0]) {
        const size_t bytes = static_cast<size_t>(t->width) * t->height * 16;
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipMalloc(&t->samples[i], bytes);
            if (e != hipSuccess) {
                hipFree(t->samples[0]);
                t->samples[0] = t->samples[1] = nullptr;
                return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP,
                                 "sample buffers: %s", hipGetErrorString(e));
            }
        }
    }
    return NB_OK;
}

nb_status render_pipelined(nb_context *ctx, const nb_scene *scene, nb_target *t, const nb_camera *cam,
                           const nb_globals *g, RenderParams &p) {
    NB_TRY(pipe_setup(ctx, t));
    // the gate this frame's path kernel waits for: the one recorded before the previous mix, or,
    // after a counter reset (or on the first frame), all of the context stream
    const bool refresh = !ctx->pipe_gate_valid || ctx->pipe_fence;
    NB_TRY(unit_order_prepare(t, scene, cam, g, p, ctx->stream, refresh ? ctx->order_sorts : ctx->pipe_gate_sorts));
    const uint32_t slot = static_cast<uint32_t>(ctx->pipe_seq & 1u);
    hipStream_t side = ctx->pipe_stream[slot];
    if (refresh) {
        NB_HIP(hipEventRecord(ctx->pipe_gate, ctx->stream));
        ctx->pipe_gate_sorts = ctx->order_sorts;
        ctx->pipe_gate_valid = true;
        ctx->pipe_fence = false;
    }
    NB_HIP(hipStreamWaitEvent(side, ctx->pipe_gate, 0));
    p.sample_out = t->samples[slot];
    p.work = ctx->pipe_work[slot];
    if (ctx->render_timing) NB_HIP(hipEventRecord(ctx->ev_start, side));
    NB_TRY(launch_render(p, side));
    if (ctx->render_timing) {
        NB_HIP(hipEventRecord(ctx->ev_stop, side));
        ctx->timed = true;
    }
    NB_HIP(hipEventRecord(ctx->pipe_done, side));
    NB_HIP(hipEventRecord(ctx->pipe_gate, ctx->stream));   // before this frame's mix
    ctx->pipe_gate_sorts = ctx->order_sorts;
    NB_HIP(hipStreamWaitEvent(ctx->stream, ctx->pipe_done, 0));
    NB_TRY(launch_mix_samples(t->pixels, p.sample_out, t->width * t->height, p.first_weight, ctx->stream));
    ++ctx->pipe_seq;
    return unit_order_update(t, p, 1u, ctx->stream);
}

nb_status timed_frames(nb_context *ctx, const RenderParams &p) {
    // event records between launches cost the stream about 0.5 % at 1080p, so they are made only
    // when the caller asked for per-render times
    if (!ctx->render_timing) return launch_render(p, ctx->stream);
    NB_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    NB_TRY(launch_render(p, ctx->stream));
    NB_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    ctx->timed = true;
    return NB_OK;
}

}  // namespace

static void release_context(nb_context *ctx);

// Waits for all of the context's work: its stream and both side streams of the pipelined frames.
// Each side-stream path kernel is normally followed by its mix on the context stream, which
// waits for it; the side streams are synchronised as well so that nothing a path kernel reads is
// freed under it even after an error left a frame without its mix.
static hipError_t ctx_quiesce(nb_context *ctx) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2; ++i)
        if (ctx->pipe_stream[i]) {
            const hipError_t f = hipStreamSynchronize(ctx->pipe_stream[i]);
            if (e == hipSuccess) e = f;
        }
    return e;
}

extern "C" {

const char *nb_last_error(void) { return g_error.c_str(); }

uint32_t nb_version(void) { return 100; }

nb_status nb_context_create(int device, nb_context **out) {
    if (!out) return set_error(NB_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int n = 0;
    NB_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(NB_ERR_INVALID_ARGUMENT, "device %d not present (%d devices)", device, n);
    NB_HIP(hipSetDevice(device));
    nb_context *c = new nb_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // The per-render events only time the launch (nb_last_render_ms): without the system-scope
    // fence their records do not write back and invalidate the caches between launches.
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_start, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_stop, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_region_start);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_region_stop);
    if (e == hipSuccess) e = hipMalloc(&c->counters, kCounterBytes);
    if (e == hipSuccess) e = hipMemset(c->counters, 0, kCounterBytes);
    if (e == hipSuccess) e = hipMalloc(&c->work, kWorkBytes);
    if (e == hipSuccess) e = hipMemset(c->work, 0, kWorkBytes);
    if (e != hipSuccess) {
        c->closing = true;
        release_context(c);
        return set_error(NB_ERR_HIP, "context creation: %s", hipGetErrorString(e));
    }
    // the device address of nb_primary_hits_noalloc's scratch, with this device current
    if (nb_status st = hits_scratch_address(device, &c->hits_scratch); st != NB_OK) {
        c->closing = true;
        release_context(c);
        return st;
    }
    *out = c;
    return NB_OK;
}

static void release_context(nb_context *ctx) {
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2; ++i) {
        if (ctx->pipe_stream[i]) {
            hipStreamSynchronize(ctx->pipe_stream[i]);
            hipStreamDestroy(ctx->pipe_stream[i]);
        }
        hipFree(ctx->pipe_work[i]);
    }
    if (ctx->pipe_gate) hipEventDestroy(ctx->pipe_gate);
    if (ctx->pipe_done) hipEventDestroy(ctx->pipe_done);
    hipFree(ctx->sobol);
    hipFree(ctx->counters);
    hipFree(ctx->work);
    hipFree(ctx->sp_block);
    if (ctx->ev_start) hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_region_start) hipEventDestroy(ctx->ev_region_start);
    if (ctx->ev_region_stop) hipEventDestroy(ctx->ev_region_stop);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

static void child_released(nb_context *ctx) {
    if (--ctx->children == 0 && ctx->closing) release_context(ctx);
}


nb_status nb_context_destroy(nb_context *ctx) {
    if (!ctx) return NB_OK;
    if (ctx->closing) return NB_OK;
    ctx->closing = true;
    if (ctx->children == 0) release_context(ctx);
    return NB_OK;
}

nb_status nb_context_synchronize(nb_context *ctx) {
    if (!ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "null context");
    NB_TRY(use_device(ctx));
    NB_HIP(ctx_quiesce(ctx));
    return NB_OK;
}

nb_status nb_build_scene_capacity(uint32_t n_indices, const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                                  uint32_t *blas_cap, uint32_t *tlas_cap) {
    if (!prims || !blas_cap || !tlas_cap) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    return build_scene_capacity(n_indices, prims, n_prims, flags, blas_cap, tlas_cap);
}

nb_status nb_build_scene(const nb_vertex *vertices, uint32_t n_vertices, uint32_t *indices, uint32_t n_indices,
                         const nb_primitive *prims, uint32_t n_prims, uint32_t flags, nb_bvh_node *blas,
                         uint32_t blas_cap, uint32_t *n_blas, nb_bvh_node *tlas, uint32_t tlas_cap, uint32_t *n_tlas,
                         nb_instance *instances) {
    if (!vertices || !indices || !prims || !blas || !n_blas || !tlas || !n_tlas || !instances)
        return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    try {
        return build_scene(vertices, n_vertices, indices, n_indices, prims, n_prims, flags, blas, blas_cap, n_blas,
                           tlas, tlas_cap, n_tlas, instances);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed in scene build");
    }
}

nb_status nb_scene_create(nb_context *ctx, const nb_bvh_node *blas, uint32_t n_blas, const nb_bvh_node *tlas,
                          uint32_t n_tlas, const nb_instance *instances, uint32_t n_instances, const nb_vertex *vertices,
                          uint32_t n_vertices, const uint32_t *indices, uint32_t n_indices, nb_scene **out) {
    if (!ctx || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if ((n_blas && !blas) || (n_tlas && !tlas) || (n_instances && !instances) || (n_vertices && !vertices) ||
        (n_indices && !indices))
        return set_error(NB_ERR_INVALID_ARGUMENT, "null array with nonzero length");
    NB_TRY(use_device(ctx));
    Relayout R;
    try {
        NB_TRY(relayout(blas, n_blas, tlas, n_tlas, instances, n_instances, vertices, n_vertices, indices, n_indices, R));
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed in scene upload");
    }
    auto bytes_of = [](size_t n) { return (n * 4 + 255) / 256 * 256; };
    const size_t sizes[7] = {R.pairs.size() * 16, R.tris.size() * 16, R.inst.size() * 16, R.ishade.size() * 16,
                             R.vshade.size() * 16, R.tri_leaf.size() * 4, R.inst_leaf.size() * 4};
    size_t total = 0, off[7];
    for (int k = 0; k < 7; ++k) { off[k] = total; total += bytes_of((sizes[k] + 3) / 4); }
    nb_scene *s = new nb_scene();
    s->ctx = ctx;
    hipError_t e = hipMalloc(&s->block, total ? total : 256);
    if (e != hipSuccess) {
        delete s;
        return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "scene allocation (%zu B): %s",
                         total, hipGetErrorString(e));
    }
    s->bytes = total;
    char *b = static_cast<char *>(s->block);
    const void *src[7] = {R.pairs.data(), R.tris.data(), R.inst.data(), R.ishade.data(), R.vshade.data(),
                          R.tri_leaf.data(), R.inst_leaf.data()};
    for (int k = 0; k < 7; ++k) {
        if (!sizes[k]) continue;
        e = hipMemcpy(b + off[k], src[k], sizes[k], hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(s->block);
            delete s;
            return set_error(NB_ERR_HIP, "scene upload: %s", hipGetErrorString(e));
        }
    }
    s->dev.pairs = reinterpret_cast<const float4 *>(b + off[0]);
    s->dev.tris = reinterpret_cast<const float4 *>(b + off[1]);
    s->dev.inst = reinterpret_cast<const float4 *>(b + off[2]);
    s->dev.ishade = reinterpret_cast<const float4 *>(b + off[3]);
    s->dev.vshade = reinterpret_cast<const float4 *>(b + off[4]);
    s->dev.tri_leaf_count = reinterpret_cast<const uint32_t *>(b + off[5]);
    s->dev.inst_leaf_count = reinterpret_cast<const uint32_t *>(b + off[6]);
    s->dev.tlas_root = R.tlas_root;
    s->dev.n_pairs = static_cast<uint32_t>(R.pairs.size() / kRecordFloat4);
    ++ctx->children;
    *out = s;
    return NB_OK;
}

nb_status nb_scene_destroy(nb_scene *scene) {
    if (!scene) return NB_OK;
    hipSetDevice(scene->ctx->device);
    ctx_quiesce(scene->ctx);
    hipFree(scene->block);
    nb_context *ctx = scene->ctx;
    delete scene;
    child_released(ctx);
    return NB_OK;
}

nb_status nb_scene_device_bytes(const nb_scene *scene, uint64_t *bytes) {
    if (!scene || !bytes) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = scene->bytes;
    return NB_OK;
}

nb_status nb_sobol_generate(uint32_t rows, uint32_t dims, uint32_t seed, float *out) {
    if (!out) return set_error(NB_ERR_INVALID_ARGUMENT, "null output");
    return sobol_generate(rows, dims, seed, out);
}

nb_status nb_sobol_set(nb_context *ctx, const float *table, uint32_t rows, uint32_t dims, uint32_t stride) {
    if (!ctx || !table || rows == 0 || dims == 0 || stride == 0)
        return set_error(NB_ERR_INVALID_ARGUMENT, "null table or zero size");
    uint64_t entries = static_cast<uint64_t>(rows) * dims;
    if (entries > 0x7FFFFFFFull) return set_error(NB_ERR_INVALID_ARGUMENT, "table too large");
    NB_TRY(use_device(ctx));
    NB_HIP(ctx_quiesce(ctx));
    float4 *d = nullptr;
    NB_HIP(hipMalloc(&d, entries * 16));
    hipError_t e = hipMemcpy(d, table, entries * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) { hipFree(d); return set_error(NB_ERR_HIP, "table upload: %s", hipGetErrorString(e)); }
    hipFree(ctx->sobol);
    ctx->sobol = d;
    ctx->sobol_entries = static_cast<uint32_t>(entries);
    ctx->sobol_stride = stride;
    return NB_OK;
}

nb_status nb_envmap_create_cube(nb_context *ctx, const float *texels, uint32_t face_size, nb_envmap **out) {
    if (!ctx || !texels || !out || face_size == 0 || face_size > 16384)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bad environment arguments");
    *out = nullptr;
    NB_TRY(use_device(ctx));
    std::vector<float4> padded = pad_cube(texels, static_cast<int>(face_size));
    nb_envmap *m = new nb_envmap();
    m->ctx = ctx;
    m->n = face_size;
    hipError_t e = hipMalloc(&m->texels, padded.size() * 16);
    if (e == hipSuccess) e = hipMemcpy(m->texels, padded.data(), padded.size() * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(m->texels);
        delete m;
        return set_error(NB_ERR_HIP, "environment upload: %s", hipGetErrorString(e));
    }
    ++ctx->children;
    *out = m;
    return NB_OK;
}

nb_status nb_envmap_create_latlong(nb_context *ctx, const float *texels, uint32_t width, uint32_t height,
                                   nb_envmap **out) {
    if (!ctx || !texels || !out || width == 0 || height == 0 || width > 65536 || height > 32768)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bad lat-long environment arguments");
    *out = nullptr;
    NB_TRY(use_device(ctx));
    // one border texel on every side: the columns wrap across the u seam, the rows repeat the
    // pole rows, so the lookup's four taps never leave the padded map; a header texel ahead of
    // them holds the size
    const uint32_t pw = width + 2, ph = height + 2;
    std::vector<float4> padded;
    try {
        padded.resize(static_cast<size_t>(pw) * ph + 1);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    const uint32_t header[4] = {width, height, 0u, 0u};
    std::memcpy(&padded[0], header, sizeof(header));
    for (uint32_t j = 0; j < ph; ++j) {
        const uint32_t sj = j == 0 ? 0u : (j == ph - 1 ? height - 1 : j - 1);
        for (uint32_t i = 0; i < pw; ++i) {
            const uint32_t si = i == 0 ? width - 1 : (i == pw - 1 ? 0u : i - 1);
            const float *t = texels + (static_cast<size_t>(sj) * width + si) * 4;
            padded[1 + static_cast<size_t>(j) * pw + i] = make_float4(t[0], t[1], t[2], t[3]);
        }
    }
    nb_envmap *m = new nb_envmap();
    m->ctx = ctx;
    m->n = 0;
    m->width = width;
    m->height = height;
    hipError_t e = hipMalloc(&m->texels, padded.size() * 16);
    if (e == hipSuccess) e = hipMemcpy(m->texels, padded.data(), padded.size() * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(m->texels);
        delete m;
        return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "environment upload: %s",
                         hipGetErrorString(e));
    }
    ++ctx->children;
    *out = m;
    return NB_OK;
}

nb_status nb_render_options_set(nb_context *ctx, const nb_render_options *options) {
    if (!ctx || !options) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    // (beyond the largest bounce limit a render accepts, it means no roulette at all)
    ctx->rr_first = std::min(options->roulette_first_bounce, kMaxBounceLimit);
    return NB_OK;
}

nb_status nb_envmap_destroy(nb_envmap *env) {
    if (!env) return NB_OK;
    hipSetDevice(env->ctx->device);
    ctx_quiesce(env->ctx);
    hipFree(env->texels);
    nb_context *ctx = env->ctx;
    delete env;
    child_released(ctx);
    return NB_OK;
}

nb_status nb_camera_look_at(const float eye[3], const float target[3], const float up[3], float fov_y, float aspect,
                            float z_near, nb_camera *out) {
    if (!eye || !target || !up || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    camera_look_at(eye, target, up, fov_y, aspect, z_near, out);
    return NB_OK;
}

nb_status nb_target_create(nb_context *ctx, uint32_t width, uint32_t height, nb_target **out) {
    if (!ctx || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0 || width % 8 || height % 8 || width > 32768 || height > 32768)
        return set_error(NB_ERR_INVALID_ARGUMENT, "target %ux%u: sides must be nonzero multiples of 8", width, height);
    NB_TRY(use_device(ctx));
    nb_target *t = new nb_target();
    t->ctx = ctx;
    t->width = width;
    t->height = height;
    const size_t bytes = static_cast<size_t>(width) * height * 16;
    hipError_t e = hipMalloc(&t->pixels, bytes);
    if (e == hipSuccess) e = hipMemset(t->pixels, 0, bytes);
    if (e != hipSuccess) {
        hipFree(t->pixels);
        delete t;
        return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "target allocation: %s",
                         hipGetErrorString(e));
    }
    if (nb_status st = unit_order_alloc(t); st != NB_OK) {
        unit_order_free(t);
        hipFree(t->pixels);
        delete t;
        return st;
    }
    ++ctx->children;
    *out = t;
    return NB_OK;
}

nb_status nb_target_destroy(nb_target *t) {
    if (!t) return NB_OK;
    hipSetDevice(t->ctx->device);
    ctx_quiesce(t->ctx);
    for (int i = 0; i < 2; ++i) hipFree(t->samples[i]);
    hipFree(t->pixels);
    unit_order_free(t);
    nb_context *ctx = t->ctx;
    delete t;
    child_released(ctx);
    return NB_OK;
}

nb_status nb_target_read(nb_target *t, float *rgba) {
    if (!t || !rgba) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    NB_TRY(use_device(t->ctx));
    NB_HIP(hipStreamSynchronize(t->ctx->stream));
    NB_HIP(hipMemcpy(rgba, t->pixels, static_cast<size_t>(t->width) * t->height * 16, hipMemcpyDeviceToHost));
    return NB_OK;
}

nb_status nb_target_write(nb_target *t, const float *rgba) {
    if (!t || !rgba) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    NB_TRY(use_device(t->ctx));
    NB_HIP(hipStreamSynchronize(t->ctx->stream));
    NB_HIP(hipMemcpy(t->pixels, rgba, static_cast<size_t>(t->width) * t->height * 16, hipMemcpyHostToDevice));
    return NB_OK;
}

nb_status nb_target_clear(nb_target *t) {
    if (!t) return set_error(NB_ERR_INVALID_ARGUMENT, "null target");
    NB_TRY(use_device(t->ctx));
    NB_HIP(hipMemsetAsync(t->pixels, 0, static_cast<size_t>(t->width) * t->height * 16, t->ctx->stream));
    return NB_OK;
}

nb_status nb_render(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                    const nb_globals *globals, const nb_envmap *env) {
    NB_TRY(check_render_args(ctx, scene, target, camera, globals, env));
    NB_TRY(use_device(ctx));
    RenderParams p = make_params(ctx, scene, target, camera, globals, env);
    if (pipeline_enabled()) return render_pipelined(ctx, scene, target, camera, globals, p);
    NB_TRY(unit_order_prepare(target, scene, camera, globals, p, ctx->stream, UINT64_MAX));
    NB_TRY(timed_frames(ctx, p));
    return unit_order_update(target, p, 1u, ctx->stream);
}

nb_status nb_render_frames(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                           const nb_globals *first, uint32_t n_frames, const nb_envmap *env) {
    NB_TRY(check_render_args(ctx, scene, target, camera, first, env));
    if (n_frames == 0) return NB_OK;
    if (first->sample == 0 || first->sample > 0xFFFFFFFFu - n_frames)
        return set_error(NB_ERR_INVALID_ARGUMENT, "frames must have 1-based sample numbers");
    NB_TRY(use_device(ctx));
    const uint32_t pixels = target->width * target->height;
    if (n_frames >= 2 && pixels <= sample_parallel_pixels() && pixels <= kSpPathsPerLaunch) {
        if (!ctx->render_timing) return render_frames_sample_parallel(ctx, scene, target, camera, first, n_frames, env);
        NB_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        NB_TRY(render_frames_sample_parallel(ctx, scene, target, camera, first, n_frames, env));
        NB_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        ctx->timed = true;
        return NB_OK;
    }
    RenderParams p = make_params(ctx, scene, target, camera, first, env);
    p.n_frames = n_frames;
    p.weights_from_sample = 1;
    NB_TRY(unit_order_prepare(target, scene, camera, first, p, ctx->stream, UINT64_MAX));
    NB_TRY(timed_frames(ctx, p));
    return unit_order_update(target, p, n_frames, ctx->stream);
}

nb_status nb_render_timing(nb_context *ctx, uint32_t on) {
    if (!ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "null context");
    ctx->render_timing = on != 0;
    if (!ctx->render_timing) ctx->timed = false;
    return NB_OK;
}

nb_status nb_last_render_ms(nb_context *ctx, float *ms) {
    if (!ctx || !ms) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->timed) return set_error(NB_ERR_INVALID_ARGUMENT, "no timed render (nb_render_timing(ctx, 1) before rendering)");
    NB_TRY(use_device(ctx));
    NB_HIP(hipEventSynchronize(ctx->ev_stop));
    NB_HIP(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return NB_OK;
}

nb_status nb_stream_timer_start(nb_context *ctx) {
    if (!ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "null context");
    NB_TRY(use_device(ctx));
    NB_HIP(hipEventRecord(ctx->ev_region_start, ctx->stream));
    ctx->region_started = true;
    return NB_OK;
}

nb_status nb_stream_timer_stop(nb_context *ctx, float *ms) {
    if (!ctx || !ms) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->region_started) return set_error(NB_ERR_INVALID_ARGUMENT, "timer not started");
    NB_TRY(use_device(ctx));
    NB_HIP(hipEventRecord(ctx->ev_region_stop, ctx->stream));
    NB_HIP(hipEventSynchronize(ctx->ev_region_stop));
    NB_HIP(hipEventElapsedTime(ms, ctx->ev_region_start, ctx->ev_region_stop));
    ctx->region_started = false;
    return NB_OK;
}

nb_status nb_counters_get(nb_context *ctx, nb_counters *out) {
    if (!ctx || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    NB_TRY(use_device(ctx));
    NB_HIP(ctx_quiesce(ctx));
    std::vector<unsigned long long> v(kCounterBytes / sizeof(unsigned long long));
    unsigned long long sum[kCounterWords] = {};
    NB_HIP(hipMemcpy(v.data(), ctx->counters, kCounterBytes, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < kCounterSlots; ++k)
        for (uint32_t i = 0; i < kCounterWords; ++i) sum[i] += v[k * kCounterStride + i];
    out->rays = sum[0]; out->n_aabb = sum[1]; out->n_tri = sum[2]; out->n_inst = sum[3]; out->samples = sum[4];
    out->hits = sum[5];
    return NB_OK;
}

nb_status nb_counters_enable(nb_context *ctx, uint32_t on) {
    if (!ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "null context");
    ctx->counting = on != 0;
    return NB_OK;
}

nb_status nb_counters_reset(nb_context *ctx) {
    if (!ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "null context");
    NB_TRY(use_device(ctx));
    NB_HIP(hipMemsetAsync(ctx->counters, 0, kCounterBytes, ctx->stream));
    ctx->pipe_fence = true;   // the next pipelined path kernel counts after the reset
    return NB_OK;
}

nb_status nb_primary_hits(nb_context *ctx, const nb_scene *scene, const nb_camera *camera, uint32_t width,
                          uint32_t height, uint32_t sample, uint32_t *out) {
    if (!ctx || !scene || !camera || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "scene belongs to another context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (width == 0 || height == 0 || width % 8 || height % 8)
        return set_error(NB_ERR_INVALID_ARGUMENT, "sides must be nonzero multiples of 8");
    NB_TRY(use_device(ctx));
    const size_t bytes = static_cast<size_t>(width) * height * 5 * 4;
    uint32_t *d = nullptr;
    NB_HIP(hipMalloc(&d, bytes));
    nb_target tmp;
    tmp.ctx = ctx;
    tmp.width = width;
    tmp.height = height;
    nb_globals g{sample, 0.0f, 0u, 0.0f};
    RenderParams p = make_params(ctx, scene, &tmp, camera, &g, nullptr);
    p.hits_out = d;
    p.counters = nullptr;
    nb_status st = launch_primary_hits(p, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (st == NB_OK && e == hipSuccess) e = hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost);
    hipFree(d);
    if (st != NB_OK) return st;
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "primary hits: %s", hipGetErrorString(e));
    return NB_OK;
}

nb_status nb_primary_hits_noalloc(nb_context *ctx, const nb_scene *scene, const nb_camera *camera, uint32_t width,
                                  uint32_t height, uint32_t sample, uint32_t *out) {
    if (!ctx || !scene || !camera || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "scene belongs to another context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (width == 0 || height == 0 || width % 8 || height % 8)
        return set_error(NB_ERR_INVALID_ARGUMENT, "sides must be nonzero multiples of 8");
    if (static_cast<uint64_t>(width) * height > kHitsScratchPixels)
        return set_error(NB_ERR_INVALID_ARGUMENT, "%ux%u exceeds the %u-pixel scratch of nb_primary_hits_noalloc",
                         width, height, kHitsScratchPixels);
    NB_TRY(use_device(ctx));
    // every pixel's five words are written by the kernel, so the scratch needs no clearing
    nb_target tmp;
    tmp.ctx = ctx;
    tmp.width = width;
    tmp.height = height;
    nb_globals g{sample, 0.0f, 0u, 0.0f};
    RenderParams p = make_params(ctx, scene, &tmp, camera, &g, nullptr);
    p.hits_out = ctx->hits_scratch;
    p.counters = nullptr;
    NB_TRY(launch_primary_hits(p, ctx->stream));
    NB_HIP(hipMemcpyAsync(out, ctx->hits_scratch, static_cast<size_t>(width) * height * 5 * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
    NB_HIP(hipStreamSynchronize(ctx->stream));
    return NB_OK;
}

}  // extern "C"

namespace nb {
void context_add_child(nb_context *ctx) { ++ctx->children; }
void context_release_child(nb_context *ctx) { child_released(ctx); }
}  // namespace nb
