// Device-side arithmetic of the path tracer shared by the progressive megakernel (trace.hip) and
// the serving kernels (serve.hip): binary32 math in the reference's order, the two-level BVH
// traversal with an LDS short stack, the cube lookup and one bounce of the integrator.
#pragma once

#include "nb_internal.h"

namespace nb {
namespace dev {


constexpr float kNoHit = 3.40282346638528859812e+38f;  // raytracing_sw.wgsl:1-2
constexpr float kEps = 0.00000001f;                    // raytracing_sw.wgsl:3
constexpr float kBias = 0.1f;                          // raytracing_sw.wgsl:4
constexpr float kPi = 3.14159265359f;                  // common.wgsl:1
constexpr float kTwoPi = 2.0f * kPi;

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ F3 scale(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ F3 divs(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ F3 rcp3(F3 a) { return f3(1.0f / a.x, 1.0f / a.y, 1.0f / a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 normalize(F3 v) {
    float s = 1.0f / sqrtf(dot(v, v));
    return f3(v.x * s, v.y * s, v.z * s);
}
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float luminance(F3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float pow5(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x; }
__device__ __forceinline__ float pow8(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x4; }
__device__ __forceinline__ F3 xyz(float4 q) { return f3(q.x, q.y, q.z); }

// sin / cos on binary32: quadrant q = rint(x * 2/pi), r = x - q*pi/2 with pi/2 split in three
// parts, then the Cephes minimax polynomials on [-pi/4, pi/4].
__device__ __forceinline__ float reduce_quadrant(float x, int &q) {
    float fq = rintf(x * 0.636619772367581343f);
    q = static_cast<int>(fq);
    float r = x - fq * 1.5703125f;
    r = r - fq * 4.837512969970703125e-4f;
    r = r - fq * 7.54978995489188216e-8f;
    return r;
}
__device__ __forceinline__ float poly_sin(float r) {
    float z = r * r;
    float p = -1.9515295891e-4f;
    p = p * z + 8.3321608736e-3f;
    p = p * z + -1.6666654611e-1f;
    return r + r * z * p;
}
__device__ __forceinline__ float poly_cos(float r) {
    float z = r * r;
    float p = 2.443315711809948e-5f;
    p = p * z + -1.388731625493765e-3f;
    p = p * z + 4.166664568298827e-2f;
    return (1.0f - 0.5f * z) + z * z * p;
}
__device__ __forceinline__ void sin_cos(float x, float &s, float &c) {
    int q;
    float r = reduce_quadrant(x, q);
    float ps = poly_sin(r), pc = poly_cos(r);
    switch (q & 3) {
    case 0: s = ps; c = pc; break;
    case 1: s = pc; c = -ps; break;
    case 2: s = -ps; c = -pc; break;
    default: s = -pc; c = ps; break;
    }
}

// PCG4D, common.wgsl:27-41
__device__ __forceinline__ float4 hash4f(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    uint32_t x = a * 1664525u + 1013904223u, y = b * 1664525u + 1013904223u;
    uint32_t z = c * 1664525u + 1013904223u, w = d * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16;
    x += y * w; y += z * x; z += x * y; w += y * z;
    return make_float4(__uint_as_float((x >> 9) | 0x3F800000u) - 1.0f, __uint_as_float((y >> 9) | 0x3F800000u) - 1.0f,
                       __uint_as_float((z >> 9) | 0x3F800000u) - 1.0f, __uint_as_float((w >> 9) | 0x3F800000u) - 1.0f);
}

struct Ray { F3 o, d, inv; };

// Slab test, raytracing_sw.wgsl:81-89 (NaN-ignoring min / max: v_min_f32 / v_max_f32).
__device__ __forceinline__ float slab(const Ray &r, float lx, float ly, float lz, float hx, float hy, float hz) {
    float ax = (lx - r.o.x) * r.inv.x, ay = (ly - r.o.y) * r.inv.y, az = (lz - r.o.z) * r.inv.z;
    float bx = (hx - r.o.x) * r.inv.x, by = (hy - r.o.y) * r.inv.y, bz = (hz - r.o.z) * r.inv.z;
    float tn = fmaxf(fminf(ax, bx), fmaxf(fminf(ay, by), fminf(az, bz)));
    float tf = fminf(fmaxf(ax, bx), fminf(fmaxf(ay, by), fmaxf(az, bz)));
    return (tn <= tf && tf >= 0.0f) ? tn : kNoHit;
}

struct LaneCounters { uint32_t rays, aabb, tri, inst, hits; };

// Adds a wave's counters (and `samples` finished samples) to the counter slot of the XCD the
// wave runs on (workgroup b runs on XCD b % 8), so the atomics of the XCDs do not contend for
// one address.  Waves with nothing to add issue no atomics.  Every lane of the wave calls this.
__device__ __forceinline__ void flush_counters(unsigned long long *counters, const LaneCounters &c,
                                               uint32_t samples) {
    auto sum = [](uint32_t v) { for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64); return v; };
    const uint32_t r = sum(c.rays), a = sum(c.aabb), t = sum(c.tri), i = sum(c.inst), s = sum(samples), h = sum(c.hits);
    if ((threadIdx.x & 63u) == 0 && counters && (r | a | t | i | s | h)) {
        unsigned long long *slot = counters + kCounterStride * (blockIdx.x % kCounterSlots);
        if (r) atomicAdd(slot + 0, static_cast<unsigned long long>(r));
        if (a) atomicAdd(slot + 1, static_cast<unsigned long long>(a));
        if (t) atomicAdd(slot + 2, static_cast<unsigned long long>(t));
        if (i) atomicAdd(slot + 3, static_cast<unsigned long long>(i));
        if (s) atomicAdd(slot + 4, static_cast<unsigned long long>(s));
        if (h) atomicAdd(slot + 5, static_cast<unsigned long long>(h));
    }
}

struct Hit { float t, u, v; uint32_t tri, inst; };

// LDS (address space 3) view, so stack accesses compile to ds_read_b64 / ds_write_b64.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

// Traversal stack: top K entries per lane in LDS, older entries in a private array owned by the
// kernel (declared beside the Stack, so the Stack's own fields stay in registers: a struct that
// holds a dynamically indexed array is placed in scratch memory as a whole).  An entry is one
// 64-bit word, distance bits high and reference low; a lane's slot s sits at s * 64 + lane of the
// wave's LDS array, so the 64 lanes of a push or pop touch consecutive words.
typedef unsigned long long SpillArray[kStackEntries];

template <int K>
struct Stack {
    lds_u64 *ent;                 // slot s of this lane at ent[s * 64]
    unsigned long long *spill;    // kStackEntries entries
    uint32_t sp, lo;

    __device__ __forceinline__ void init(unsigned long long *lds_base, uint32_t lane, unsigned long long *spill_array) {
        ent = (lds_u64 *)(lds_base + lane);
        spill = spill_array;
        sp = 0;
        lo = 0;
    }
    __device__ __forceinline__ void push(uint32_t r, float t) {
        if (sp >= kStackEntries) return;  // unreachable for trees of depth <= 32 (checked at upload)
        if (sp - lo == static_cast<uint32_t>(K)) {
            spill[lo] = ent[(lo & (K - 1)) * 64];
            ++lo;
        }
        ent[(sp & (K - 1)) * 64] = (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | r;
        ++sp;
    }
    __device__ __forceinline__ void pop(uint32_t &r, float &t) {
        --sp;
        unsigned long long e;
        if (sp < lo) {
            e = spill[sp];
            lo = sp;
        } else {
            e = ent[(sp & (K - 1)) * 64];
        }
        r = static_cast<uint32_t>(e);
        t = __uint_as_float(static_cast<uint32_t>(e >> 32));
    }
};

// Element at byte offset `off` of a device array, in 32-bit arithmetic: node records, triangles
// and instance records each stay below 4 GiB (checked at upload), and the load takes the array's
// scalar base plus one 32-bit vector offset.
__device__ __forceinline__ const float4 *at_byte(const float4 *base, uint32_t off) {
    return reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + off);
}

// A 16-B load through the global address space.  The scene arrays are global memory; where their
// pointers come from memory rather than from kernel arguments (the serving kernel reads them
// from the step description) the compiler cannot tell, and would use flat loads with 64-bit
// vector addresses.
typedef float g_v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_global(const float4 *p) {
    const g_v4f v = *(const __attribute__((address_space(1))) g_v4f *)(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ld_global_u4(const float4 *p) {
    const g_v4f v = *(const __attribute__((address_space(1))) g_v4f *)(p);
    return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w));
}

__device__ __forceinline__ uint32_t leaf_count(uint32_t ref, const uint32_t *side) {
    uint32_t c = (ref >> kLeafStartBits) & kLeafMaxInline;
    return c != 0 ? c : *(const __attribute__((address_space(1))) uint32_t *)(side + (ref & kLeafStartMask));
}

// Two-level BVH traversal in resumable form: one step (an inner node, a TLAS leaf's instance
// entry or a BLAS leaf's triangles, then the pop) per trace_step call, so that a wave can run
// the traversal of some lanes while other lanes shade their finished hits.
//
// TLAS and BLAS share one stack; entering an instance records the stack height at which its BLAS
// ends.  Ties follow the reference: the right child is the near one when the entry distances
// are equal, and a later triangle or instance replaces the hit only if strictly nearer.  The
// current closest distance is handed into each BLAS (the reference starts every BLAS search
// from infinity, raytracing_sw.wgsl:255): this prunes boxes and triangles that could not win,
// so the selected hit is the same.
struct TraceState {
    Ray ray;                    // the ray in the space of the tree being walked
    Hit hit;
    uint32_t cur, blas_base, inst_cur, inst_end;
    bool have;
};

template <int K>
__device__ __forceinline__ void trace_begin(const DevScene &S, const Ray &world, TraceState &ts, Stack<K> &st) {
    ts.hit.t = kNoHit;
    ts.hit.tri = 0; ts.hit.inst = 0; ts.hit.u = 0.0f; ts.hit.v = 0.0f;
    ts.ray = world;
    st.sp = 0; st.lo = 0;
    ts.blas_base = kNoRecord; ts.inst_cur = 0; ts.inst_end = 0;
    ts.cur = S.tlas_root;
    ts.have = S.tlas_root != kNoRecord;
}

// The inner-node part of a traversal step at record ts.cur: test both children, push the far one
// if it is nearer than the current hit, continue with the near one (have = true) or not.
template <int K>
__device__ __forceinline__ void visit_inner(const DevScene &S, TraceState &ts, Stack<K> &st, LaneCounters &c) {
    const float4 *rec = at_byte(S.pairs, ts.cur * (kRecordFloat4 * 16u));
    const float4 q0 = ld_global(rec), q1 = ld_global(rec + 1), q2 = ld_global(rec + 2);
    const uint4 q3 = ld_global_u4(rec + 3);
    float tl = slab(ts.ray, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y);
    float tr = slab(ts.ray, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w);
    c.aabb += q3.z ? 1u : 2u;
    // ties: the right child is the near one (raytracing_sw.wgsl:229-234)
    bool left_near = tl < tr;
    uint32_t nref = left_near ? q3.x : q3.y, fref = left_near ? q3.y : q3.x;
    float nt = left_near ? tl : tr, ft = left_near ? tr : tl;
    if (ft < ts.hit.t) st.push(fref, ft);
    ts.have = nt < ts.hit.t;
    if (ts.have) ts.cur = nref;
}

// True when the next action of a lane is an inner-node visit or a pop that neither leaves a
// BLAS nor ends the trace: the cheap, common steps that the persistent kernels' inner loop runs.
template <int K>
__device__ __forceinline__ bool inner_pending(const TraceState &ts, const Stack<K> &st) {
    if (ts.have) return !(ts.cur & kLeafBit);
    return st.sp != 0 && !(ts.blas_base != kNoRecord && st.sp == ts.blas_base);
}

// A plain pop (when the lane has no node) followed by the visit of the popped node if it is an
// inner node still nearer than the hit, or the visit alone (requires inner_pending).  The same
// operations in the same order as trace_step; a pop and the visit it leads to share one
// iteration, so lanes that pop and lanes that descend run the visit code together.
template <int K>
__device__ __forceinline__ void inner_step(const DevScene &S, TraceState &ts, Stack<K> &st, LaneCounters &c) {
    if (!ts.have) {
        float t;
        st.pop(ts.cur, t);
        ts.have = t < ts.hit.t;   // skip entries farther than the current hit
        if (!ts.have || (ts.cur & kLeafBit)) return;
    }
    visit_inner<K>(S, ts, st, c);
}

// One traversal step; true when the closest hit of `world` is final in ts.hit.
template <int K>
__device__ __forceinline__ bool trace_step(const DevScene &S, const Ray &world, TraceState &ts, Stack<K> &st,
                                           LaneCounters &c) {
    // Lanes of a wave take different cases; the ray changes of all of them (instance entry from a
    // TLAS leaf, entry of a leaf's next instance, return to world space at a BLAS exit) share one
    // transform and one 1/d at the end, instead of one copy per case.
    bool enter = false, leave = false, done = false;
    if (ts.have) {
        if (!(ts.cur & kLeafBit)) {
            visit_inner<K>(S, ts, st, c);
        } else if (ts.blas_base == kNoRecord) {
            // TLAS leaf: instances [start, start + count), each searched in order
            ts.inst_cur = ts.cur & kLeafStartMask;
            ts.inst_end = ts.inst_cur + leaf_count(ts.cur, S.inst_leaf_count);
            enter = true;
        } else {
            const uint32_t start = ts.cur & kLeafStartMask, end = start + leaf_count(ts.cur, S.tri_leaf_count);
            for (uint32_t j = start; j < end; ++j) {
                const float4 *tr = at_byte(S.tris, j * 48u);
                const float4 a = ld_global(tr), b = ld_global(tr + 1), e = ld_global(tr + 2);
                ++c.tri;
                // Möller–Trumbore with back-face culling, raytracing_sw.wgsl:51-78
                F3 e1 = xyz(b), e2 = xyz(e);
                F3 h = cross(ts.ray.d, e2);
                float det = dot(e1, h);
                if (det < kEps) continue;
                float inv_det = 1.0f / det;
                F3 s = sub(ts.ray.o, xyz(a));
                float u = inv_det * dot(s, h);
                if (u < 0.0f || u > 1.0f) continue;
                F3 q = cross(s, e1);
                float v = inv_det * dot(ts.ray.d, q);
                if (v < 0.0f || u + v > 1.0f) continue;
                float t = inv_det * dot(e2, q);
                if (t < kBias) continue;
                if (t < ts.hit.t) { ts.hit.t = t; ts.hit.u = u; ts.hit.v = v; ts.hit.tri = j; ts.hit.inst = ts.inst_cur; }
            }
            ts.have = false;
        }
    }
    if (!ts.have && !enter) {
        // pop the next entry; at the bottom of a BLAS go to the leaf's next instance, or back to
        // world space after its last one
        if (ts.blas_base != kNoRecord && st.sp == ts.blas_base) {
            if (++ts.inst_cur < ts.inst_end) {
                enter = true;
            } else {
                ts.blas_base = kNoRecord;
                leave = true;
            }
        }
        if (!enter) {
            if (st.sp == 0) {
                done = true;
            } else {
                float t;
                st.pop(ts.cur, t);
                ts.have = t < ts.hit.t;   // skip entries farther than the current hit
            }
        }
    }
    if (enter || leave) {
        // back in world space, 1/d is recomputed (the same bits) so callers need not keep it
        F3 o = world.o, d = world.d;
        if (enter) {
            // (W2L * (o, 1)).xyz and mat3(W2L) * d of instance ts.inst_cur, raytracing_sw.wgsl:203-205
            const float4 *rec = at_byte(S.inst, ts.inst_cur * 64u);
            const float4 r0 = ld_global(rec), r1 = ld_global(rec + 1), r2 = ld_global(rec + 2);
            const uint32_t root = ld_global_u4(rec + 3).x;
            o = f3(((r0.x * world.o.x + r0.y * world.o.y) + r0.z * world.o.z) + r0.w * 1.0f,
                   ((r1.x * world.o.x + r1.y * world.o.y) + r1.z * world.o.z) + r1.w * 1.0f,
                   ((r2.x * world.o.x + r2.y * world.o.y) + r2.z * world.o.z) + r2.w * 1.0f);
            d = f3((r0.x * world.d.x + r0.y * world.d.y) + r0.z * world.d.z,
                   (r1.x * world.d.x + r1.y * world.d.y) + r1.z * world.d.z,
                   (r2.x * world.d.x + r2.y * world.d.y) + r2.z * world.d.z);
            ++c.inst;
            ts.blas_base = st.sp;
            ts.cur = root;
            ts.have = true;
        }
        ts.ray.o = o;
        ts.ray.d = d;
        ts.ray.inv = rcp3(d);
    }
    return done;
}

// Lane states of the persistent kernels.
enum LaneMode : uint32_t { kIdle = 0, kTracing = 1, kReady = 2 };

// Closest hit of `world` over the two-level BVH (intersect_TLAS / intersect_BLAS,
// raytracing_sw.wgsl:176-319).
template <int K>
__device__ __forceinline__ void trace(const DevScene &S, const Ray &world, Hit &hit, Stack<K> &st, LaneCounters &c) {
    TraceState ts{};
    trace_begin<K>(S, world, ts, st);
    while (!trace_step<K>(S, world, ts, st, c)) {}
    hit = ts.hit;
}

// atan on binary32 by Cephes' atanf (S. L. Moshier): |x| above tan(3pi/8) maps to pi/2 +
// atan(-1/|x|), above tan(pi/8) to pi/4 + atan((|x|-1)/(|x|+1)), then z + z^3 P(z^2) with
// Cephes' degree-3 P; the sign of x is restored at the end.
__device__ __forceinline__ float atan_f(float x) {
    const float ax = fabsf(x);
    float y0 = 0.0f, z = ax;
    if (ax > 2.414213562373095f) {
        y0 = 1.5707963267948966f;
        z = -1.0f / ax;
    } else if (ax > 0.4142135623730950f) {
        y0 = 0.7853981633974483f;
        z = (ax - 1.0f) / (ax + 1.0f);
    }
    const float z2 = z * z;
    const float q = (((8.05374449538e-2f * z2 - 1.38776856032e-1f) * z2 + 1.99777106478e-1f) * z2 - 3.33329491539e-1f) * z2;
    const float r = y0 + (q * z + z);
    return x < 0.0f ? -r : r;
}

// atan2 on binary32 from atan_f: x = 0 gives +-pi/2 (0 when y is 0 too); x < 0 adds pi for
// y >= 0 (so both signed zeros of y give +pi) and subtracts it for y < 0.
__device__ __forceinline__ float atan2_f(float y, float x) {
    if (x == 0.0f) return y > 0.0f ? 1.5707963267948966f : (y < 0.0f ? -1.5707963267948966f : 0.0f);
    const float a = atan_f(y / x);
    if (x > 0.0f) return a;
    return y >= 0.0f ? a + 3.14159265358979323846f : a - 3.14159265358979323846f;
}

// Equirectangular lookup (an environment form the reference does not have; it reads only the
// cube): y is up, u = atan2(d.z, d.x) / 2pi + 1/2 (+x at u = 1/2, +z at 3/4, -z at 1/4) and
// v = atan2(|d.xz|, d.y) / pi (+y at the top row), bilinear at LOD 0 over the padded map, so
// the four taps wrap across the u seam and repeat the pole rows.
__device__ F3 latlong_sample(const DevEnv &E, F3 d) {
    const float r = sqrtf(d.x * d.x + d.z * d.z);
    const float u = atan2_f(d.z, d.x) * 0.15915494309189535f + 0.5f;
    const float v = atan2_f(r, d.y) * 0.31830988618379067f;
    const uint4 dims = *reinterpret_cast<const uint4 *>(E.texels);   // header texel: width, height
    const float w = static_cast<float>(dims.x), h = static_cast<float>(dims.y);
    const float x = u * w - 0.5f, y = v * h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    if (!(fx >= -1.0f)) fx = -1.0f;
    if (fx > w - 1.0f) fx = w - 1.0f;
    if (!(fy >= -1.0f)) fy = -1.0f;
    if (fy > h - 1.0f) fy = h - 1.0f;
    const float a = x - fx, b = y - fy;
    const uint32_t pw = dims.x + 2;
    const int i0 = static_cast<int>(fx) + 1, j0 = static_cast<int>(fy) + 1;
    const float4 *base = E.texels + 1 + static_cast<size_t>(j0) * pw + i0;
    const float4 t00 = base[0], t10 = base[1], t01 = base[pw], t11 = base[pw + 1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return f3(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
              ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
              ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z);
}

// Bilinear cube lookup at LOD 0 with Vulkan face selection; faces carry a one-texel border
// holding the seamless-edge neighbours, so the four taps never leave the padded face.  A map of
// face size 0 is a lat-long map.
__device__ F3 env_sample(const DevEnv &E, F3 d) {
    if (E.n == 0) return latlong_sample(E, d);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int f;
    float ma, sc, tc;
    if (ax >= ay && ax >= az) {
        f = d.x >= 0.0f ? 0 : 1; ma = ax;
        sc = f == 0 ? -1.0f * d.z : 1.0f * d.z; tc = -1.0f * d.y;
    } else if (ay >= az) {
        f = d.y >= 0.0f ? 2 : 3; ma = ay;
        sc = 1.0f * d.x; tc = f == 2 ? 1.0f * d.z : -1.0f * d.z;
    } else {
        f = d.z >= 0.0f ? 4 : 5; ma = az;
        sc = f == 4 ? 1.0f * d.x : -1.0f * d.x; tc = -1.0f * d.y;
    }
    float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    float n = static_cast<float>(E.n);
    float u = s * n - 0.5f, v = t * n - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    if (!(fu >= -1.0f)) fu = -1.0f;
    if (fu > n - 1.0f) fu = n - 1.0f;
    if (!(fv >= -1.0f)) fv = -1.0f;
    if (fv > n - 1.0f) fv = n - 1.0f;
    float a = u - fu, b = v - fv;
    const uint32_t pn = E.n + 2;
    const int i0 = static_cast<int>(fu) + 1, j0 = static_cast<int>(fv) + 1;
    const float4 *base = E.texels + (static_cast<size_t>(f) * pn + j0) * pn + i0;
    float4 t00 = base[0], t10 = base[1], t01 = base[pn], t11 = base[pn + 1];
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return f3(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
              ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
              ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z);
}

// Sobol entry `entry` rotated by the pixel's shift, pathtracing.wgsl:151-159; entries past the
// end of the table read as zero.
struct SobolView { const float4 *table; uint32_t entries, stride; };

__device__ __forceinline__ float4 sobol_fetch(const SobolView &t, uint32_t entry, const float4 &shift) {
    float4 s = entry < t.entries ? t.table[entry] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return make_float4(fract(s.x + shift.x), fract(s.y + shift.y), fract(s.z + shift.z), fract(s.w + shift.w));
}

// generate_ray, pathtracing.wgsl:27-41 (m = clip_to_world, column-major; any pointer to 16 floats)
template <typename M>
__device__ __forceinline__ Ray camera_ray(M m, uint32_t width, uint32_t height, uint32_t x, uint32_t y, float jx,
                                          float jy) {
    float uvx = (2.0f * (static_cast<float>(x) + jx)) / static_cast<float>(width) - 1.0f;
    float uvy = (2.0f * (static_cast<float>(y) + jy)) / static_cast<float>(height) - 1.0f;
    F3 pos = f3(m[12] / m[15], m[13] / m[15], m[14] / m[15]);
    float cx = -uvx, cy = -uvy;
    float wx = ((m[0] * cx + m[4] * cy) + m[8] * -1.0f) + m[12] * 1.0f;
    float wy = ((m[1] * cx + m[5] * cy) + m[9] * -1.0f) + m[13] * 1.0f;
    float wz = ((m[2] * cx + m[6] * cy) + m[10] * -1.0f) + m[14] * 1.0f;
    float ww = ((m[3] * cx + m[7] * cy) + m[11] * -1.0f) + m[15] * 1.0f;
    Ray r;
    r.o = pos;
    r.d = sub(pos, f3(wx / ww, wy / ww, wz / ww));
    r.inv = rcp3(r.d);
    return r;
}

// sample_vndf_iso, pathtracing.wgsl:68-98
__device__ F3 sample_vndf_iso(float r0, float r1, F3 wi, float alpha, F3 n) {
    if (alpha == 0.0f) return n;
    F3 wi_z = scale(dot(wi, n), n);
    F3 wi_xy = sub(wi, wi_z);
    F3 wi_std = normalize(sub(wi_z, scale(alpha, wi_xy)));
    float wi_std_z = dot(wi_std, n);
    float phi = (2.0f * r0 - 1.0f) * kPi;
    float z = (1.0f - r1) * (1.0f + wi_std_z) - wi_std_z;
    float sin_theta = sqrtf(fminf(fmaxf(1.0f - z * z, 0.0f), 1.0f));
    float sp, cp;
    sin_cos(phi, sp, cp);
    F3 c_std = f3(sin_theta * cp, sin_theta * sp, z);
    F3 wr = add(n, f3(0.0f, 0.0f, 1.0f));
    float safe_wrz = fmaxf(wr.z, 1e-6f);
    float dwc = dot(wr, c_std);
    F3 c = sub(divs(scale(dwc, wr), safe_wrz), c_std);
    F3 wm_std = add(c, wi_std);
    F3 wm_std_z = scale(dot(n, wm_std), n);
    F3 wm_std_xy = sub(wm_std_z, wm_std);
    return normalize(add(wm_std_z, scale(alpha, wm_std_xy)));
}

__device__ __forceinline__ F3 schlick(float hdotv, F3 r0) {
    float p = pow5(1.0f - hdotv);
    return f3(r0.x + (1.0f - r0.x) * p, r0.y + (1.0f - r0.y) * p, r0.z + (1.0f - r0.z) * p);
}

__device__ __forceinline__ float lambda_tr(float ndotv, float alpha2) {
    float c2 = ndotv * ndotv;
    float s2 = 1.0f - c2;
    float t2 = s2 / c2;
    return (-1.0f + sqrtf(1.0f + alpha2 * t2)) / 2.0f;
}

// Path parameters that stay fixed along a path.
struct PathConst {
    DevScene scene;
    DevEnv env;
    SobolView sobol;
    uint32_t sample;
    // the bounce limit in bits 0..15 and, above it, the first bounce with Russian roulette (the
    // reference: 0); packed so that the capped kernels hold one register less
    uint32_t bounces;
    float contribution_factor;
    float4 shift;
};

enum StepResult { kContinue = 0, kDone = 1 };

// Shading point of a non-emissive hit, as the radiance cache sees it.
struct Vertex {
    F3 pos, wo, n, albedo;
    float roughness, metallic;
    float lobe_alpha;       // alpha of the sampled lobe: the GGX alpha if specular, 1 if diffuse
};

// Closest hit of `ray`, counted as one intersect_scene call.
template <int K>
__device__ __forceinline__ void trace_counted(const PathConst &pc, const Ray &ray, Hit &h, Stack<K> &st,
                                              LaneCounters &c) {
    trace<K>(pc.scene, ray, h, st, c);
    ++c.rays;
    if (h.t != kNoHit) ++c.hits;
}

// Shading of hit h of `ray` (the body of the bounce loop after intersect_scene,
// pathtracing.wgsl:172-238).  kDone with `out` for emitters and Russian-roulette deaths,
// otherwise the next ray and throughput.  With kRecord, `vtx` receives the shading point.
template <bool kRecord>
__device__ StepResult shade(const PathConst &pc, uint32_t bounce, const Hit &h, Ray &ray, F3 &thr, F3 &out,
                            Vertex &vtx) {
    const DevScene &S = pc.scene;
    const float4 *sh = S.ishade + 6 * static_cast<size_t>(h.inst);
    const F3 albedo = f3(sh[3].w, sh[4].w, sh[5].w);
    if (sh[2].w > 0.0f) { out = mul(thr, albedo); return kDone; }
    const float4 s0 = sh[0], s1 = sh[1], s2 = sh[2];
    const float4 *tr = S.tris + 3 * static_cast<size_t>(h.tri);
    const uint32_t i0 = __float_as_uint(tr[0].w), i1 = __float_as_uint(tr[1].w), i2 = __float_as_uint(tr[2].w);
    const float4 n0 = S.vshade[2 * i0], n1 = S.vshade[2 * i1], n2 = S.vshade[2 * i2];
    const float bx = 1.0f - h.u - h.v, by = h.u, bz = h.v;
    F3 pos = add(ray.o, scale(h.t, ray.d));
    F3 ln = add(add(scale(bx, xyz(n0)), scale(by, xyz(n1))), scale(bz, xyz(n2)));
    F3 nrm = f3((s0.x * ln.x + s0.y * ln.y) + s0.z * ln.z, (s1.x * ln.x + s1.y * ln.y) + s1.z * ln.z,
                (s2.x * ln.x + s2.y * ln.y) + s2.z * ln.z);
    float roughness = s0.w, metallic = s1.w;

    float alpha = roughness * roughness;
    float alpha2 = alpha * alpha;
    F3 n = normalize(nrm);
    if (kRecord) {
        vtx.pos = pos;
        vtx.wo = normalize(f3(-ray.d.x, -ray.d.y, -ray.d.z));
        vtx.n = n;
        vtx.albedo = albedo;
        vtx.roughness = roughness;
        vtx.metallic = metallic;
    }
    uint32_t e0 = pc.sample * pc.sobol.stride + 1u + bounce * 2u;
    float4 sob0 = sobol_fetch(pc.sobol, e0, pc.shift), sob1 = sobol_fetch(pc.sobol, e0 + 1u, pc.shift);
    F3 wo = normalize(f3(-ray.d.x, -ray.d.y, -ray.d.z));
    float cos_o = dot(wo, n);
    F3 wi;
    F3 f0 = f3(0.04f * (1.0f - metallic) + albedo.x * metallic, 0.04f * (1.0f - metallic) + albedo.y * metallic,
               0.04f * (1.0f - metallic) + albedo.z * metallic);
    float spec_w = luminance(schlick(dot(wo, n), f0));
    float diff_w = (1.0f - metallic) * luminance(albedo);
    float p_spec = spec_w / (spec_w + diff_w);
    float p_diff = 1.0f - p_spec;
    if (sob0.x < p_spec) {
        F3 wm = sample_vndf_iso(sob0.y, sob0.z, wo, alpha, n);
        F3 mwo = f3(-wo.x, -wo.y, -wo.z);
        float k2 = 2.0f * dot(wm, mwo);
        wi = sub(mwo, scale(k2, wm));
        float cos_d = dot(wo, wm);
        float cos_i = dot(wi, n);
        F3 F = schlick(cos_d, f0);
        float ll = lambda_tr(cos_i, alpha2), lv = lambda_tr(cos_o, alpha2);
        float den = (1.0f + ll) + lv;
        F3 spec = f3((F.x * (1.0f + lv)) / den, (F.y * (1.0f + lv)) / den, (F.z * (1.0f + lv)) / den);
        thr = mul(thr, divs(spec, p_spec));
        if (kRecord) vtx.lobe_alpha = alpha;
    } else {
        if (kRecord) vtx.lobe_alpha = 1.0f;
        // the tangent is read here, where only the diffuse lobe needs it, so that its vertex and
        // instance data are not held in registers through the rest of the shading
        const float4 t0 = S.vshade[2 * i0 + 1], t1 = S.vshade[2 * i1 + 1], t2 = S.vshade[2 * i2 + 1];
        const float4 s3 = sh[3], s4 = sh[4], s5 = sh[5];
        F3 lt = f3((t0.x * bx + t1.x * by) + t2.x * bz, (t0.y * bx + t1.y * by) + t2.y * bz,
                   (t0.z * bx + t1.z * by) + t2.z * bz);
        F3 tan = f3((s3.x * lt.x + s4.x * lt.y) + s5.x * lt.z, (s3.y * lt.x + s4.y * lt.y) + s5.y * lt.z,
                    (s3.z * lt.x + s4.z * lt.y) + s5.z * lt.z);
        F3 nt = normalize(sub(tan, scale(dot(tan, n), n)));
        F3 b = cross(n, nt);
        float phi = kTwoPi * sob1.y;
        float st_ = sqrtf(1.0f - sob1.z), ct = sqrtf(sob1.z);
        float sp, cp;
        sin_cos(phi, sp, cp);
        F3 cs = f3(cp * st_, sp * st_, ct);
        wi = add(add(scale(cs.x, nt), scale(cs.y, b)), scale(cs.z, n));
        F3 wm = normalize(add(wi, wo));
        float cos_d = dot(wi, wm);
        float cos_i = dot(wi, n);
        float fd90 = 0.5f + (2.0f * alpha) * (cos_d * cos_d);
        float resp = (1.0f + (fd90 - 1.0f) * pow5(1.0f - cos_i)) * (1.0f + (fd90 - 1.0f) * pow5(1.0f - cos_o));
        float om = 1.0f - metallic;
        F3 diff = f3((om * albedo.x) * resp, (om * albedo.y) * resp, (om * albedo.z) * resp);
        thr = mul(thr, divs(diff, p_diff));
    }
    // Russian roulette, pathtracing.wgsl:227-236, from bounce bounces >> 16 on (the reference: 0;
    // before it every path continues with its throughput as it is)
    if (bounce >= (pc.bounces >> 16)) {
        float p_cont = fminf(1.0f - pow8(static_cast<float>(bounce) / static_cast<float>(pc.bounces & 0xFFFFu)), 1.0f);
        p_cont *= fminf(luminance(thr) * pc.contribution_factor, 1.0f);
        if (sob0.z < p_cont) {
            thr = divs(thr, p_cont);
        } else {
            out = f3(0.0f, 0.0f, 0.0f);
            return kDone;
        }
    }
    ray.o = pos;
    ray.d = wi;
    ray.inv = rcp3(wi);
    return kContinue;
}

// One iteration of the bounce loop of sample_rendering_eq (pathtracing.wgsl:164-239) with
// intersect_scene (raytracing_sw.wgsl:138-173) inlined.  Either finishes the path (kDone,
// radiance in `out`) or updates `ray` and `thr` for the next bounce (kContinue).
template <int K, bool kRecord = false>
__device__ __forceinline__ StepResult bounce_step(const PathConst &pc, uint32_t bounce, Ray &ray, F3 &thr, F3 &out,
                                                  Stack<K> &st, LaneCounters &c, Vertex *vtx = nullptr) {
    Hit h{};
    trace_counted<K>(pc, ray, h, st, c);
    if (h.t == kNoHit) { out = mul(thr, env_sample(pc.env, ray.d)); return kDone; }
    Vertex dummy{};
    return shade<kRecord>(pc, bounce, h, ray, thr, out, kRecord ? *vtx : dummy);
}

// One whole path (sample_rendering_eq, pathtracing.wgsl:161-241).
template <int K>
__device__ F3 sample_path(const PathConst &pc, Ray ray, Stack<K> &st, LaneCounters &c) {
    F3 thr = f3(1.0f, 1.0f, 1.0f), out = f3(0.0f, 0.0f, 0.0f);
    for (uint32_t bounce = 0; bounce <= (pc.bounces & 0xFFFFu); ++bounce)
        if (bounce_step<K>(pc, bounce, ray, thr, out, st, c) == kDone) return out;
    return f3(0.0f, 0.0f, 0.0f);
}

}  // namespace dev
}  // namespace nb
