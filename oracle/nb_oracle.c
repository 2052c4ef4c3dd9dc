/*
 * nb_oracle.c — CPU restatement of nBounce's path-tracing path (see nb_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Compiled with -ffp-contract=off and without fast-math: every expression below is
 * evaluated in IEEE binary32 in the order written, which is the order the reference's
 * source states it (left to right, vector ops per component).
 */
#include "nb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NO_HIT 3.40282346638528859812e+38f /* 0x1.fffffep+127f, raytracing_sw.wgsl:1-2 */
#define EPS 0.00000001f                    /* raytracing_sw.wgsl:3 */
#define BIAS 0.1f                          /* raytracing_sw.wgsl:4 */
#define STACK_SIZE 64 /* raytracing_sw.wgsl:5-6 holds 32; a depth-32 tree can need 33, so hold 64 */
#define PI_F 3.14159265359f                /* common.wgsl:1 */
#define TWO_PI_F (2.0f * PI_F)             /* common.wgsl:2 */

typedef struct { float x, y, z; } v3;

static inline v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float *p) { return mk3(p[0], p[1], p[2]); }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 div3(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl3(float s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline v3 divs3(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* normalize(v) = v * (1 / sqrt(dot(v, v))): one correctly rounded sqrt and divide. */
static inline v3 normalize3(v3 v) {
    float s = 1.0f / sqrtf(dot3(v, v));
    return mk3(v.x * s, v.y * s, v.z * s);
}
static inline float fract1(float x) { return x - floorf(x); }
/* luminance, common.wgsl:44-46 */
static inline float luminance(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
/* NaN-ignoring min/max (f32::min / f32::max and WGSL min/max lowered to minNum/maxNum). */
static inline float fmin_(float a, float b) { return fminf(a, b); }
static inline float fmax_(float a, float b) { return fmaxf(a, b); }
static inline v3 vmin3(v3 a, v3 b) { return mk3(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
static inline v3 vmax3(v3 a, v3 b) { return mk3(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }
static inline float comp3(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
/* pow with small integer exponents as products (WGSL pow is exp2(y*log2(x)) on most
 * drivers and undefined for x < 0; products agree with it to a few ulp for x >= 0). */
static inline float pow2f(float x) { return x * x; }
static inline float pow5f(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x; }
static inline float pow8f(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x4; }

/* ------------------------------------------------------------------------------------------
 * PCG4D hash, common.wgsl:27-41 (Jarzynski & Olano 2020)
 * ------------------------------------------------------------------------------------------ */
void orc_hash4u(const uint32_t s[4], uint32_t out[4]) {
    uint32_t x = s[0] * 1664525u + 1013904223u;
    uint32_t y = s[1] * 1664525u + 1013904223u;
    uint32_t z = s[2] * 1664525u + 1013904223u;
    uint32_t w = s[3] * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16;
    x += y * w; y += z * x; z += x * y; w += y * z;
    out[0] = x; out[1] = y; out[2] = z; out[3] = w;
}

static inline float map1f(uint32_t x) {
    union { uint32_t u; float f; } c;
    c.u = (x >> 9) | 0x3F800000u;
    return c.f - 1.0f;
}

void orc_map4f(const uint32_t x[4], float out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = map1f(x[i]);
}

/* ------------------------------------------------------------------------------------------
 * sin / cos: Cody-Waite reduction by pi/2 (three-part constant) and the Cephes single
 * precision minimax polynomials (S. Moshier, Cephes Math Library) on [-pi/4, pi/4].
 * WGSL leaves sin/cos accuracy to the implementation; this one is deterministic.
 * ------------------------------------------------------------------------------------------ */
static float reduce_half_pi(float x, int *quadrant) {
    float q = rintf(x * 0.636619772367581343f);
    *quadrant = (int)q;
    float r = x - q * 1.5703125f;
    r = r - q * 4.837512969970703125e-4f;
    r = r - q * 7.54978995489188216e-8f;
    return r;
}
static float sin_poly(float r) {
    float z = r * r;
    float p = -1.9515295891e-4f;
    p = p * z + 8.3321608736e-3f;
    p = p * z + -1.6666654611e-1f;
    return r + r * z * p;
}
static float cos_poly(float r) {
    float z = r * r;
    float p = 2.443315711809948e-5f;
    p = p * z + -1.388731625493765e-3f;
    p = p * z + 4.166664568298827e-2f;
    return (1.0f - 0.5f * z) + z * z * p;
}
float orc_sin(float x) {
    int q;
    float r = reduce_half_pi(x, &q);
    switch (q & 3) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
    }
}
float orc_cos(float x) {
    int q;
    float r = reduce_half_pi(x, &q);
    switch (q & 3) {
    case 0: return cos_poly(r);
    case 1: return -sin_poly(r);
    case 2: return -cos_poly(r);
    default: return sin_poly(r);
    }
}

/* ------------------------------------------------------------------------------------------
 * atan / atan2, for the lat-long environment form (no counterpart in the reference, which
 * samples only the cube): Cephes atanf (S. Moshier).  |x| > tan(3pi/8) becomes pi/2 +
 * atan(-1/|x|), |x| > tan(pi/8) becomes pi/4 + atan((|x|-1)/(|x|+1)); then the degree-3
 * minimax polynomial in z^2 gives z + z^3 P(z^2).  atan2 adds +-pi for x < 0 (+pi when y >= 0,
 * signed zeros included) and returns +-pi/2 or 0 for x == 0.
 * ------------------------------------------------------------------------------------------ */
float orc_atan(float x) {
    float ax = fabsf(x), base = 0.0f, z = ax;
    if (ax > 2.414213562373095f) {
        base = 1.5707963267948966f;
        z = -1.0f / ax;
    } else if (ax > 0.4142135623730950f) {
        base = 0.7853981633974483f;
        z = (ax - 1.0f) / (ax + 1.0f);
    }
    float zz = z * z;
    float poly = 8.05374449538e-2f;
    poly = poly * zz - 1.38776856032e-1f;
    poly = poly * zz + 1.99777106478e-1f;
    poly = poly * zz - 3.33329491539e-1f;
    poly = poly * zz;
    float r = base + (poly * z + z);
    return x < 0.0f ? -r : r;
}
float orc_atan2(float y, float x) {
    if (x == 0.0f) {
        if (y > 0.0f) return 1.5707963267948966f;
        if (y < 0.0f) return -1.5707963267948966f;
        return 0.0f;
    }
    float a = orc_atan(y / x);
    if (x > 0.0f) return a;
    if (y >= 0.0f) return a + 3.14159265358979323846f;
    return a - 3.14159265358979323846f;
}

/* ------------------------------------------------------------------------------------------
 * Mat4 (column-major, m[c*4 + r]) as glam 0.29 computes it.
 * ------------------------------------------------------------------------------------------ */
/* Cofactor inverse in glm's arrangement (glam Mat4::inverse); determinant as (x+z)+(y+w),
 * then every entry times 1/det. */
void orc_mat4_inverse(const float m[16], float out[16]) {
#define M(c, r) m[(c) * 4 + (r)]
    float c00 = M(2,2) * M(3,3) - M(3,2) * M(2,3);
    float c02 = M(1,2) * M(3,3) - M(3,2) * M(1,3);
    float c03 = M(1,2) * M(2,3) - M(2,2) * M(1,3);
    float c04 = M(2,1) * M(3,3) - M(3,1) * M(2,3);
    float c06 = M(1,1) * M(3,3) - M(3,1) * M(1,3);
    float c07 = M(1,1) * M(2,3) - M(2,1) * M(1,3);
    float c08 = M(2,1) * M(3,2) - M(3,1) * M(2,2);
    float c10 = M(1,1) * M(3,2) - M(3,1) * M(1,2);
    float c11 = M(1,1) * M(2,2) - M(2,1) * M(1,2);
    float c12 = M(2,0) * M(3,3) - M(3,0) * M(2,3);
    float c14 = M(1,0) * M(3,3) - M(3,0) * M(1,3);
    float c15 = M(1,0) * M(2,3) - M(2,0) * M(1,3);
    float c16 = M(2,0) * M(3,2) - M(3,0) * M(2,2);
    float c18 = M(1,0) * M(3,2) - M(3,0) * M(1,2);
    float c19 = M(1,0) * M(2,2) - M(2,0) * M(1,2);
    float c20 = M(2,0) * M(3,1) - M(3,0) * M(2,1);
    float c22 = M(1,0) * M(3,1) - M(3,0) * M(1,1);
    float c23 = M(1,0) * M(2,1) - M(2,0) * M(1,1);
    float fac0[4] = {c00, c00, c02, c03};
    float fac1[4] = {c04, c04, c06, c07};
    float fac2[4] = {c08, c08, c10, c11};
    float fac3[4] = {c12, c12, c14, c15};
    float fac4[4] = {c16, c16, c18, c19};
    float fac5[4] = {c20, c20, c22, c23};
    float vec0[4] = {M(1,0), M(0,0), M(0,0), M(0,0)};
    float vec1[4] = {M(1,1), M(0,1), M(0,1), M(0,1)};
    float vec2[4] = {M(1,2), M(0,2), M(0,2), M(0,2)};
    float vec3[4] = {M(1,3), M(0,3), M(0,3), M(0,3)};
    const float sign_a[4] = {1.0f, -1.0f, 1.0f, -1.0f};
    const float sign_b[4] = {-1.0f, 1.0f, -1.0f, 1.0f};
    float inv[4][4];
    for (int i = 0; i < 4; ++i) {
        inv[0][i] = sign_a[i] * ((vec1[i] * fac0[i] - vec2[i] * fac1[i]) + vec3[i] * fac2[i]);
        inv[1][i] = sign_b[i] * ((vec0[i] * fac0[i] - vec2[i] * fac3[i]) + vec3[i] * fac4[i]);
        inv[2][i] = sign_a[i] * ((vec0[i] * fac1[i] - vec1[i] * fac3[i]) + vec3[i] * fac5[i]);
        inv[3][i] = sign_b[i] * ((vec0[i] * fac2[i] - vec1[i] * fac4[i]) + vec2[i] * fac5[i]);
    }
    float d0 = M(0,0) * inv[0][0];
    float d1 = M(0,1) * inv[1][0];
    float d2 = M(0,2) * inv[2][0];
    float d3 = M(0,3) * inv[3][0];
    float det = (d0 + d2) + (d1 + d3);
    float rcp = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[c * 4 + r] = inv[c][r] * rcp;
#undef M
}

/* out = a * b; column j = ((a.c0*b[j].x + a.c1*b[j].y) + a.c2*b[j].z) + a.c3*b[j].w */
void orc_mat4_mul(const float a[16], const float b[16], float out[16]) {
    float tmp[16];
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            tmp[j * 4 + r] = ((a[0 * 4 + r] * b[j * 4 + 0] + a[1 * 4 + r] * b[j * 4 + 1]) +
                              a[2 * 4 + r] * b[j * 4 + 2]) + a[3 * 4 + r] * b[j * 4 + 3];
    memcpy(out, tmp, sizeof(tmp));
}

/* glam transform_point3: ((c0*x + c1*y) + c2*z) + c3 */
static v3 transform_point3(const float m[16], v3 p) {
    return mk3(((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12],
               ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13],
               ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14]);
}

/* camera.rs:121-142: world_to_clip = perspective_infinite_rh(fov, aspect, near) * look_at_rh(eye, target, up) */
void orc_camera_compute(const float eye_[3], const float target_[3], const float up_[3],
                        float fov_y, float aspect, float z_near, orc_camera *out) {
    v3 eye = ld3(eye_), target = ld3(target_), up = ld3(up_);
    v3 f = normalize3(sub3(target, eye));
    v3 s = normalize3(cross3(f, up));
    v3 u = cross3(s, f);
    float view[16] = {s.x, u.x, -f.x, 0.0f,
                      s.y, u.y, -f.y, 0.0f,
                      s.z, u.z, -f.z, 0.0f,
                      -dot3(eye, s), -dot3(eye, u), dot3(eye, f), 1.0f};
    float fl = 1.0f / tanf(0.5f * fov_y);
    float proj[16] = {fl / aspect, 0.0f, 0.0f, 0.0f,
                      0.0f, fl, 0.0f, 0.0f,
                      0.0f, 0.0f, -1.0f, -1.0f,
                      0.0f, 0.0f, -z_near, 0.0f};
    orc_mat4_mul(proj, view, out->world_to_clip);
    orc_mat4_inverse(out->world_to_clip, out->clip_to_world);
}

/* ------------------------------------------------------------------------------------------
 * SAH BVH builder, bvh.rs:15-330, and the scene flattening of scene.rs:239-305.
 * ------------------------------------------------------------------------------------------ */
typedef struct { v3 center, min, max; uint32_t idx[3]; } tri_t;          /* bvh.rs:77-88 */
typedef struct { v3 min, max; } prim_bounds;

typedef struct {
    int kind;             /* 0 = triangles, 1 = instances */
    tri_t *tris;
    struct inst_wb *insts;
} prims_t;

typedef struct inst_wb { orc_instance inst; v3 wmin, wmax; } inst_wb;     /* scene.rs:232-236 */

static inline v3 p_min(const prims_t *p, uint32_t i) { return p->kind == 0 ? p->tris[i].min : p->insts[i].wmin; }
static inline v3 p_max(const prims_t *p, uint32_t i) { return p->kind == 0 ? p->tris[i].max : p->insts[i].wmax; }
/* BVHPrimitive::center default = (min + max) * 0.5 (bvh.rs:10-12); triangles cache theirs (bvh.rs:87) */
static inline v3 p_center(const prims_t *p, uint32_t i) {
    if (p->kind == 0) return p->tris[i].center;
    v3 s = add3(p->insts[i].wmin, p->insts[i].wmax);
    return mk3(s.x * 0.5f, s.y * 0.5f, s.z * 0.5f);
}
static void p_swap(prims_t *p, uint32_t a, uint32_t b) {
    if (a == b) return;
    if (p->kind == 0) { tri_t t = p->tris[a]; p->tris[a] = p->tris[b]; p->tris[b] = t; }
    else { inst_wb t = p->insts[a]; p->insts[a] = p->insts[b]; p->insts[b] = t; }
}

typedef struct { v3 min; uint32_t start; v3 max; uint32_t end; } node_t;
typedef struct { v3 min, max; uint32_t count; } bin_t;                      /* bvh.rs:112-127 */

static inline bin_t bin_empty(void) {
    bin_t b; b.min = mk3(INFINITY, INFINITY, INFINITY); b.max = mk3(-INFINITY, -INFINITY, -INFINITY); b.count = 0;
    return b;
}
static inline void bin_include(bin_t *b, v3 mn, v3 mx) { b->min = vmin3(b->min, mn); b->max = vmax3(b->max, mx); b->count += 1; }
static inline void bin_include_bin(bin_t *b, const bin_t *o) { b->min = vmin3(b->min, o->min); b->max = vmax3(b->max, o->max); b->count += o->count; }
static inline int finite3(v3 e) { return isfinite(e.x) && isfinite(e.y) && isfinite(e.z); }
/* Bin::cost / BVHNode::cost, bvh.rs:65-74,142-150 */
static inline float box_cost(v3 mn, v3 mx, uint32_t count) {
    v3 e = sub3(mx, mn);
    if (!finite3(e)) return INFINITY;
    float area = e.x * e.y + e.x * e.z + e.y * e.z;
    return (float)count * area;
}
static inline float node_cost(const node_t *n) { return box_cost(n->min, n->max, n->end - n->start); }

static node_t new_leaf(const prims_t *p, uint32_t start, uint32_t end) {  /* bvh.rs:27-39 */
    node_t n;
    n.min = p_min(p, start); n.max = p_max(p, start);
    for (uint32_t i = start + 1; i < end; ++i) { n.min = vmin3(n.min, p_min(p, i)); n.max = vmax3(n.max, p_max(p, i)); }
    n.start = start; n.end = end;
    return n;
}

/* Rust `f32 as u32`: NaN and negatives saturate to 0, too large to u32::MAX. */
static inline uint32_t f32_to_u32_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

typedef struct { int axis; float mid; } split_t;

/* bvh.rs:232-258 */
static int find_best_split(const prims_t *p, const node_t *parent, split_t *out) {
    float best = node_cost(parent);
    int found = 0;
    for (int axis = 0; axis < 3; ++axis) {
        for (uint32_t i = parent->start; i < parent->end; ++i) {
            float mid = comp3(p_center(p, i), axis);
            bin_t left = bin_empty(), right = bin_empty();
            for (uint32_t j = parent->start; j < parent->end; ++j) {
                if (comp3(p_center(p, j), axis) < mid) bin_include(&left, p_min(p, j), p_max(p, j));
                else bin_include(&right, p_min(p, j), p_max(p, j));
            }
            float cost = box_cost(left.min, left.max, left.count) + box_cost(right.min, right.max, right.count);
            if (cost < best) { best = cost; out->axis = axis; out->mid = mid; found = 1; }
        }
    }
    return found;
}

/* bvh.rs:260-296 */
static int approximate_best_split(const prims_t *p, const node_t *parent, split_t *out) {
    bin_t bins[48];
    for (int i = 0; i < 48; ++i) bins[i] = bin_empty();
    v3 ext = sub3(parent->max, parent->min);
    v3 step = mk3(ext.x / 16.0f, ext.y / 16.0f, ext.z / 16.0f);
    for (uint32_t i = parent->start; i < parent->end; ++i) {
        v3 rel = div3(sub3(p_center(p, i), parent->min), step);
        uint32_t bx = f32_to_u32_sat(floorf(rel.x)); if (bx > 15u) bx = 15u;
        uint32_t by = f32_to_u32_sat(floorf(rel.y)); if (by > 15u) by = 15u;
        uint32_t bz = f32_to_u32_sat(floorf(rel.z)); if (bz > 15u) bz = 15u;
        v3 mn = p_min(p, i), mx = p_max(p, i);
        bin_include(&bins[bx], mn, mx);
        bin_include(&bins[16 + by], mn, mx);
        bin_include(&bins[32 + bz], mn, mx);
    }
    float best = node_cost(parent);
    int found = 0;
    for (int axis = 0; axis < 3; ++axis) {
        bin_t left = bin_empty();
        for (int i = 0; i < 15; ++i) {
            bin_include_bin(&left, &bins[axis * 16 + i]);
            bin_t right = bin_empty();
            for (int j = i + 1; j < 16; ++j) bin_include_bin(&right, &bins[axis * 16 + j]);
            float cost = box_cost(left.min, left.max, left.count) + box_cost(right.min, right.max, right.count);
            if (cost < best) {
                best = cost;
                out->axis = axis;
                out->mid = comp3(parent->min, axis) + comp3(step, axis) * (float)(i + 1);
                found = 1;
            }
        }
    }
    return found;
}

/* bvh.rs:307-330: left-stable in-place partition by swap */
static int do_split(prims_t *p, const node_t *parent, split_t s, node_t *l, node_t *r) {
    bin_t left = bin_empty(), right = bin_empty();
    for (uint32_t i = parent->start; i < parent->end; ++i) {
        v3 mn = p_min(p, i), mx = p_max(p, i);
        float c = comp3(p_center(p, i), s.axis);
        if (c < s.mid) {
            bin_include(&left, mn, mx);
            p_swap(p, i, parent->start + left.count - 1);
        } else {
            bin_include(&right, mn, mx);
        }
    }
    if (left.count == 0 || right.count == 0) return 0;
    l->min = left.min; l->max = left.max; l->start = parent->start; l->end = parent->start + left.count;
    r->min = right.min; r->max = right.max; r->start = l->end; r->end = l->end + right.count;
    return 1;
}

/* bvh.rs:208-230 */
static int split_node(prims_t *p, const node_t *parent, node_t *l, node_t *r) {
    uint32_t count = parent->end - parent->start;
    if (count <= 1) return 0;
    if (count == 2) {
        *l = new_leaf(p, parent->start, parent->start + 1);
        *r = new_leaf(p, parent->start + 1, parent->start + 2);
        return node_cost(l) + node_cost(r) < node_cost(parent);
    }
    split_t s;
    if (count <= 11) { if (!find_best_split(p, parent, &s)) return 0; }
    else { if (!approximate_best_split(p, parent, &s)) return 0; }
    return do_split(p, parent, s, l, r);
}

typedef struct { node_t *nodes; uint32_t n, cap; int overflow; } tree_t;

static void tree_push(tree_t *t, node_t n) {
    if (t->n >= t->cap) { t->overflow = 1; return; }
    t->nodes[t->n++] = n;
}

/* BVHTree::append, bvh.rs:162-190: LIFO stack, children appended as (L, L+1), R popped first. */
static uint32_t tree_append(tree_t *t, prims_t *p, uint32_t start, uint32_t end) {
    uint32_t parent_index = t->n;
    tree_push(t, new_leaf(p, start, end));
    if (t->overflow) return parent_index;
    uint32_t cap = 64, sp = 0;
    uint32_t (*stack)[2] = malloc(sizeof(*stack) * cap);
    stack[sp][0] = 0; stack[sp][1] = parent_index; ++sp;
    while (sp > 0) {
        --sp;
        uint32_t depth = stack[sp][0], idx = stack[sp][1];
        if (depth >= 32u) continue;                      /* MAX_DEPTH, bvh.rs:153,173 */
        node_t l, r;
        node_t cur = t->nodes[idx];
        if (split_node(p, &cur, &l, &r)) {
            uint32_t li = t->n;
            t->nodes[idx].end = 0; t->nodes[idx].start = li;   /* make_inner, bvh.rs:60-63 */
            tree_push(t, l); tree_push(t, r);
            if (t->overflow) break;
            if (sp + 2 > cap) { cap *= 2; stack = realloc(stack, sizeof(*stack) * cap); }
            stack[sp][0] = depth + 1; stack[sp][1] = li; ++sp;
            stack[sp][0] = depth + 1; stack[sp][1] = li + 1; ++sp;
        }
    }
    free(stack);
    return parent_index;
}

static void store_node(const node_t *n, orc_node *o) {
    o->min[0] = n->min.x; o->min[1] = n->min.y; o->min[2] = n->min.z; o->start = n->start;
    o->max[0] = n->max.x; o->max[1] = n->max.y; o->max[2] = n->max.z; o->end = n->end;
}

/* build_triangle_cache, bvh.rs:90-104 */
static tri_t *triangle_cache(const orc_vertex *v, const uint32_t *idx, uint32_t n_idx) {
    uint32_t nt = n_idx / 3;
    tri_t *t = malloc(sizeof(tri_t) * (nt ? nt : 1));
    for (uint32_t i = 0; i < nt; ++i) {
        v3 a = ld3(v[idx[3 * i]].position), b = ld3(v[idx[3 * i + 1]].position), c = ld3(v[idx[3 * i + 2]].position);
        v3 s = add3(add3(a, b), c);
        t[i].center = mk3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f);
        t[i].min = vmin3(vmin3(a, b), c);
        t[i].max = vmax3(vmax3(a, b), c);
        t[i].idx[0] = idx[3 * i]; t[i].idx[1] = idx[3 * i + 1]; t[i].idx[2] = idx[3 * i + 2];
    }
    return t;
}

int orc_build_blas(const orc_vertex *vertices, uint32_t n_vertices, uint32_t *indices, uint32_t n_indices,
                   uint32_t tri_start, uint32_t tri_end, orc_node *nodes_out, uint32_t cap, uint32_t *n_nodes) {
    (void)n_vertices;
    if (tri_end <= tri_start || tri_end > n_indices / 3) return -1;
    tri_t *tris = triangle_cache(vertices, indices, n_indices);
    prims_t p = {0, tris, NULL};
    tree_t t = {malloc(sizeof(node_t) * cap), 0, cap, 0};
    tree_append(&t, &p, tri_start, tri_end);
    int rc = t.overflow ? -2 : 0;
    for (uint32_t i = 0; i < t.n; ++i) store_node(&t.nodes[i], &nodes_out[i]);
    *n_nodes = t.n;
    for (uint32_t i = 0; i < n_indices / 3; ++i)
        for (int k = 0; k < 3; ++k) indices[3 * i + k] = tris[i].idx[k];
    free(t.nodes); free(tris);
    return rc;
}

/* SceneBuffers::from_scene, scene.rs:280-307 */
int orc_build_scene(const orc_vertex *vertices, uint32_t n_vertices, uint32_t *indices, uint32_t n_indices,
                    const orc_primitive *prims, uint32_t n_prims,
                    orc_node *blas_out, uint32_t blas_cap, uint32_t *n_blas,
                    orc_node *tlas_out, uint32_t tlas_cap, uint32_t *n_tlas,
                    orc_instance *instances_out) {
    (void)n_vertices;
    if (n_prims == 0) return -1;
    tri_t *tris = triangle_cache(vertices, indices, n_indices);
    prims_t pt = {0, tris, NULL};
    tree_t blas = {malloc(sizeof(node_t) * (blas_cap ? blas_cap : 1)), 0, blas_cap, 0};
    inst_wb *insts = malloc(sizeof(inst_wb) * n_prims);
    for (uint32_t k = 0; k < n_prims; ++k) {
        const orc_primitive *pr = &prims[k];
        uint32_t ts = pr->index_start / 3, te = pr->index_end / 3;
        if (te <= ts || te > n_indices / 3) { free(tris); free(blas.nodes); free(insts); return -1; }
        uint32_t root = tree_append(&blas, &pt, ts, te);
        if (blas.overflow) { free(tris); free(blas.nodes); free(insts); return -2; }
        v3 lmin = blas.nodes[root].min, lmax = blas.nodes[root].max;
        orc_instance *in = &insts[k].inst;
        orc_mat4_inverse(pr->local_to_world, in->world_to_local);
        memcpy(in->local_to_world, pr->local_to_world, 64);
        memcpy(in->color, pr->color, 16);
        in->roughness = pr->roughness; in->metallic = pr->metallic; in->emissive = pr->emissive;
        in->node = root;
        /* InstanceWithBounds::approximate_from_instance, scene.rs:239-260 */
        v3 wmin = mk3(INFINITY, INFINITY, INFINITY), wmax = mk3(-INFINITY, -INFINITY, -INFINITY);
        for (int i = 0; i < 8; ++i) {
            v3 l = mk3((i & 1) == 0 ? lmin.x : lmax.x, (i & 2) == 0 ? lmin.y : lmax.y, (i & 4) == 0 ? lmin.z : lmax.z);
            v3 w = transform_point3(pr->local_to_world, l);
            wmin = vmin3(wmin, w); wmax = vmax3(wmax, w);
        }
        insts[k].wmin = wmin; insts[k].wmax = wmax;
    }
    prims_t pi = {1, NULL, insts};
    tree_t tlas = {malloc(sizeof(node_t) * (tlas_cap ? tlas_cap : 1)), 0, tlas_cap, 0};
    tree_append(&tlas, &pi, 0, n_prims);
    if (tlas.overflow) { free(tris); free(blas.nodes); free(insts); free(tlas.nodes); return -2; }
    for (uint32_t i = 0; i < blas.n; ++i) store_node(&blas.nodes[i], &blas_out[i]);
    for (uint32_t i = 0; i < tlas.n; ++i) store_node(&tlas.nodes[i], &tlas_out[i]);
    *n_blas = blas.n; *n_tlas = tlas.n;
    for (uint32_t k = 0; k < n_prims; ++k) instances_out[k] = insts[k].inst;
    for (uint32_t i = 0; i < n_indices / 3; ++i)                           /* flatten_triangle_list, bvh.rs:106-110 */
        for (int k = 0; k < 3; ++k) indices[3 * i + k] = tris[i].idx[k];
    free(tris); free(blas.nodes); free(insts); free(tlas.nodes);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Intersection, raytracing_sw.wgsl
 * ------------------------------------------------------------------------------------------ */
typedef struct { v3 o, d, inv; } ray_t;

/* Möller–Trumbore with back-face culling, raytracing_sw.wgsl:51-78 */
static v3 intersect_triangle(const ray_t *ray, v3 v0, v3 v1, v3 v2) {
    v3 nohit = mk3(NO_HIT, NO_HIT, NO_HIT);
    v3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    v3 h = cross3(ray->d, e2);
    float det = dot3(e1, h);
    if (det < EPS) return nohit;
    float inv_det = 1.0f / det;
    v3 s = sub3(ray->o, v0);
    float u = inv_det * dot3(s, h);
    if (u < 0.0f || u > 1.0f) return nohit;
    v3 q = cross3(s, e1);
    float v = inv_det * dot3(ray->d, q);
    if (v < 0.0f || u + v > 1.0f) return nohit;
    float t = inv_det * dot3(e2, q);
    if (t < BIAS) return nohit;
    return mk3(t, u, v);
}

/* slab test, raytracing_sw.wgsl:81-89 */
static float intersect_aabb(const ray_t *ray, v3 mn, v3 mx) {
    v3 tmin = mul3(sub3(mn, ray->o), ray->inv);
    v3 tmax = mul3(sub3(mx, ray->o), ray->inv);
    v3 t1 = vmin3(tmin, tmax), t2 = vmax3(tmin, tmax);
    float tn = fmax_(t1.x, fmax_(t1.y, t1.z));
    float tf = fmin_(t2.x, fmin_(t2.y, t2.z));
    return (tn <= tf && tf >= 0.0f) ? tn : NO_HIT;
}

void orc_intersect_triangle(const float o[3], const float d[3], const float v0[3], const float v1[3],
                            const float v2[3], float out[3]) {
    ray_t r; r.o = ld3(o); r.d = ld3(d); r.inv = div3(mk3(1.0f, 1.0f, 1.0f), r.d);
    v3 t = intersect_triangle(&r, ld3(v0), ld3(v1), ld3(v2));
    out[0] = t.x; out[1] = t.y; out[2] = t.z;
}

float orc_intersect_aabb(const float o[3], const float inv_d[3], const float bmin[3], const float bmax[3]) {
    ray_t r; r.o = ld3(o); r.inv = ld3(inv_d); r.d = mk3(0, 0, 0);
    return intersect_aabb(&r, ld3(bmin), ld3(bmax));
}

static inline v3 node_min(const orc_node *n) { return ld3(n->min); }
static inline v3 node_max(const orc_node *n) { return ld3(n->max); }

typedef struct { uint32_t index; float dist; } stack_entry;

/* intersect_BLAS, raytracing_sw.wgsl:252-319 */
static orc_raw_hit intersect_blas(const orc_scene *sc, const ray_t *ray, uint32_t index_top) {
    stack_entry stack[STACK_SIZE];
    orc_raw_hit hit; memset(&hit, 0, sizeof(hit)); hit.dist = NO_HIT;
    uint32_t i = 0;
    const orc_node *top = &sc->blas[index_top];
    float dt = intersect_aabb(ray, node_min(top), node_max(top));
    hit.n_aabb += 1;
    if (dt < hit.dist) { stack[i].index = index_top; stack[i].dist = dt; ++i; }
    while (i > 0) {
        --i;
        stack_entry e = stack[i];
        if (e.dist >= hit.dist) continue;
        const orc_node *node = &sc->blas[e.index];
        if (node->end > 0) {
            for (uint32_t j = node->start * 3u; j < node->end * 3u; j += 3u) {
                uint32_t i0 = sc->indices[j], i1 = sc->indices[j + 1], i2 = sc->indices[j + 2];
                v3 t = intersect_triangle(ray, ld3(sc->vertices[i0].position), ld3(sc->vertices[i1].position),
                                          ld3(sc->vertices[i2].position));
                hit.n_tri += 1;
                if (t.x < hit.dist) {
                    hit.dist = t.x;
                    hit.bary[0] = 1.0f - t.y - t.z; hit.bary[1] = t.y; hit.bary[2] = t.z;
                    hit.i0 = i0; hit.i1 = i1; hit.i2 = i2;
                }
            }
        } else {
            uint32_t li = node->start, ri = li + 1;
            stack_entry left = {li, intersect_aabb(ray, node_min(&sc->blas[li]), node_max(&sc->blas[li]))};
            stack_entry right = {ri, intersect_aabb(ray, node_min(&sc->blas[ri]), node_max(&sc->blas[ri]))};
            hit.n_aabb += 2;
            stack_entry far = left, near = right;
            if (left.dist < right.dist) { far = right; near = left; }
            if (far.dist < hit.dist) stack[i++] = far;
            if (near.dist < hit.dist) stack[i++] = near;
        }
    }
    return hit;
}

/* local ray: (W2L * (o,1)).xyz and mat3(W2L) * d, each component ((c0*x + c1*y) + c2*z) [+ c3] */
static ray_t to_local(const orc_instance *in, const ray_t *r) {
    const float *m = in->world_to_local;
    ray_t l;
    l.o = mk3(((m[0] * r->o.x + m[4] * r->o.y) + m[8] * r->o.z) + m[12] * 1.0f,
              ((m[1] * r->o.x + m[5] * r->o.y) + m[9] * r->o.z) + m[13] * 1.0f,
              ((m[2] * r->o.x + m[6] * r->o.y) + m[10] * r->o.z) + m[14] * 1.0f);
    l.d = mk3((m[0] * r->d.x + m[4] * r->d.y) + m[8] * r->d.z,
              (m[1] * r->d.x + m[5] * r->d.y) + m[9] * r->d.z,
              (m[2] * r->d.x + m[6] * r->d.y) + m[10] * r->d.z);
    l.inv = div3(mk3(1.0f, 1.0f, 1.0f), l.d);
    return l;
}

/* intersect_TLAS, raytracing_sw.wgsl:176-250 */
static orc_raw_hit intersect_tlas(const orc_scene *sc, const ray_t *ray) {
    stack_entry stack[STACK_SIZE];
    orc_raw_hit hit; memset(&hit, 0, sizeof(hit)); hit.dist = NO_HIT;
    uint32_t i = 0;
    if (sc->n_tlas == 0) return hit;
    const orc_node *top = &sc->tlas[0];
    float dt = intersect_aabb(ray, node_min(top), node_max(top));
    hit.n_aabb += 1;
    if (dt < hit.dist) { stack[i].index = 0; stack[i].dist = dt; ++i; }
    while (i > 0) {
        --i;
        stack_entry e = stack[i];
        if (e.dist >= hit.dist) continue;
        const orc_node *node = &sc->tlas[e.index];
        if (node->end > 0) {
            for (uint32_t j = node->start; j < node->end; ++j) {
                const orc_instance *in = &sc->instances[j];
                ray_t lr = to_local(in, ray);
                orc_raw_hit hl = intersect_blas(sc, &lr, in->node);
                hit.n_aabb += hl.n_aabb;
                hit.n_tri += hl.n_tri;
                if (hl.dist < hit.dist) {
                    hit.dist = hl.dist; hit.instance = j;
                    hit.i0 = hl.i0; hit.i1 = hl.i1; hit.i2 = hl.i2;
                    memcpy(hit.bary, hl.bary, sizeof(hit.bary));
                }
            }
        } else {
            uint32_t li = node->start, ri = li + 1;
            stack_entry left = {li, intersect_aabb(ray, node_min(&sc->tlas[li]), node_max(&sc->tlas[li]))};
            stack_entry right = {ri, intersect_aabb(ray, node_min(&sc->tlas[ri]), node_max(&sc->tlas[ri]))};
            hit.n_aabb += 2;
            stack_entry far = left, near = right;
            if (left.dist < right.dist) { far = right; near = left; }
            if (far.dist < hit.dist) stack[i++] = far;
            if (near.dist < hit.dist) stack[i++] = near;
        }
    }
    return hit;
}

void orc_trace(const orc_scene *scene, const float o[3], const float d[3], orc_raw_hit *hit) {
    ray_t r; r.o = ld3(o); r.d = ld3(d); r.inv = div3(mk3(1.0f, 1.0f, 1.0f), r.d);
    *hit = intersect_tlas(scene, &r);
}

/* Brute force: every instance, every triangle of its BLAS (collected from the BLAS leaves),
 * closest strictly-smaller t, ties kept at the first instance/triangle encountered. */
static void collect_leaf_tris(const orc_scene *sc, uint32_t node, uint32_t *lo, uint32_t *hi) {
    const orc_node *n = &sc->blas[node];
    if (n->end > 0) { if (n->start < *lo) *lo = n->start; if (n->end > *hi) *hi = n->end; return; }
    collect_leaf_tris(sc, n->start, lo, hi);
    collect_leaf_tris(sc, n->start + 1, lo, hi);
}

void orc_trace_brute(const orc_scene *sc, const float o[3], const float d[3], orc_raw_hit *hit) {
    ray_t r; r.o = ld3(o); r.d = ld3(d); r.inv = div3(mk3(1.0f, 1.0f, 1.0f), r.d);
    memset(hit, 0, sizeof(*hit)); hit->dist = NO_HIT;
    for (uint32_t j = 0; j < sc->n_instances; ++j) {
        const orc_instance *in = &sc->instances[j];
        ray_t lr = to_local(in, &r);
        uint32_t lo = 0xFFFFFFFFu, hi = 0;
        collect_leaf_tris(sc, in->node, &lo, &hi);
        for (uint32_t t = lo; t < hi; ++t) {
            uint32_t i0 = sc->indices[3 * t], i1 = sc->indices[3 * t + 1], i2 = sc->indices[3 * t + 2];
            v3 h = intersect_triangle(&lr, ld3(sc->vertices[i0].position), ld3(sc->vertices[i1].position),
                                      ld3(sc->vertices[i2].position));
            if (h.x < hit->dist) {
                hit->dist = h.x; hit->instance = j; hit->i0 = i0; hit->i1 = i1; hit->i2 = i2;
                hit->bary[0] = 1.0f - h.y - h.z; hit->bary[1] = h.y; hit->bary[2] = h.z;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Camera ray, pathtracing.wgsl:27-41
 * ------------------------------------------------------------------------------------------ */
static ray_t generate_ray(const orc_camera *cam, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float jx, float jy) {
    const float *m = cam->clip_to_world;
    float dimx = (float)w, dimy = (float)h;
    float uvx = (2.0f * ((float)x + jx)) / dimx - 1.0f;
    float uvy = (2.0f * ((float)y + jy)) / dimy - 1.0f;
    v3 pos = mk3(m[12] / m[15], m[13] / m[15], m[14] / m[15]);
    float cx = -uvx, cy = -uvy, cz = -1.0f, cw = 1.0f;
    float wd[4];
    for (int r = 0; r < 4; ++r) wd[r] = ((m[0 + r] * cx + m[4 + r] * cy) + m[8 + r] * cz) + m[12 + r] * cw;
    ray_t ray;
    ray.o = pos;
    ray.d = sub3(pos, mk3(wd[0] / wd[3], wd[1] / wd[3], wd[2] / wd[3]));
    ray.inv = div3(mk3(1.0f, 1.0f, 1.0f), ray.d);
    return ray;
}

void orc_generate_ray(const orc_camera *cam, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                      const float jitter[2], float o[3], float d[3]) {
    ray_t r = generate_ray(cam, x, y, w, h, jitter[0], jitter[1]);
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; d[0] = r.d.x; d[1] = r.d.y; d[2] = r.d.z;
}

static void pixel_shift(uint32_t x, uint32_t y, float shift[4]) {
    uint32_t s[4] = {x, y, x, y}, hsh[4];
    orc_hash4u(s, hsh);
    orc_map4f(hsh, shift);
}

static inline const float *sobol_at(const float *table, uint32_t idx) { return table + 4u * idx; }

void orc_primary_hits(const orc_scene *scene, const orc_camera *cam, uint32_t width, uint32_t height,
                      const float *table, uint32_t stride, uint32_t sample, uint32_t *out) {
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            float shift[4];
            pixel_shift(x, y, shift);
            const float *row = sobol_at(table, sample * stride);
            float jx = fract1(row[0] + shift[0]), jy = fract1(row[1] + shift[1]);
            ray_t r = generate_ray(cam, x, y, width, height, jx, jy);
            orc_raw_hit h = intersect_tlas(scene, &r);
            uint32_t *o = out + 5u * (y * width + x);
            if (h.dist == NO_HIT) { o[0] = 0xFFFFFFFFu; o[1] = o[2] = o[3] = 0; }
            else { o[0] = h.instance; o[1] = h.i0; o[2] = h.i1; o[3] = h.i2; }
            memcpy(&o[4], &h.dist, 4);
        }
}

/* ------------------------------------------------------------------------------------------
 * Sobol sequence with Owen scrambling, restating Burley 2020 ("Practical Hash-based Owen
 * Scrambling", JCGT 9(4)) as used through sobol_burley::sample_4d(i, set, seed).
 * Direction numbers are Joe & Kuo's new-joe-kuo-6.21201 (primitive polynomials and initial
 * m_i supplied by the loader; the first 256 dimensions are enough for 64 dimension sets).
 * ------------------------------------------------------------------------------------------ */
static uint32_t g_dirs[256][32];
static int g_dirs_ready = 0;

/* poly[d]: primitive polynomial with leading and trailing bits (scipy's layout); m[d][k]: initial m_k */
int orc_sobol_init(const int64_t *poly, const int64_t *vinit, uint32_t n_dims, uint32_t vinit_stride) {
    if (n_dims > 256) n_dims = 256;
    for (uint32_t b = 0; b < 32; ++b) g_dirs[0][b] = 1u << (31 - b);
    for (uint32_t d = 1; d < n_dims; ++d) {
        int64_t p = poly[d];
        int s = 0;
        while ((p >> (s + 1)) != 0) ++s;            /* degree */
        uint32_t m[32];
        for (int k = 0; k < s && k < 32; ++k) m[k] = (uint32_t)vinit[d * vinit_stride + k];
        for (int k = s; k < 32; ++k) {
            uint32_t v = m[k - s] ^ (m[k - s] << s);
            for (int j = 1; j < s; ++j)
                if ((p >> (s - j)) & 1) v ^= m[k - j] << j;
            m[k] = v;
        }
        for (int b = 0; b < 32; ++b) g_dirs[d][b] = m[b] << (31 - b);
    }
    g_dirs_ready = 1;
    return 0;
}

uint32_t orc_sobol_unscrambled(uint32_t index, uint32_t dim) {
    uint32_t r = 0;
    for (int b = 0; index; ++b, index >>= 1)
        if (index & 1u) r ^= g_dirs[dim][b];
    return r;
}

static uint32_t reverse_bits32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}
/* Laine–Karras style permutation (Burley 2020, listing 3) */
static uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
static uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits32(lk_permute(reverse_bits32(x), seed));
}
/* hash_combine (Burley 2020, listing 4) */
static uint32_t hash_combine(uint32_t seed, uint32_t v) {
    return seed ^ (v + (seed << 6) + (seed >> 2));
}
/* A 32-bit integer hash to decorrelate seeds (Wellons' lowbias32). */
static uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

void orc_sobol_sample4d(uint32_t index, uint32_t dim_set, uint32_t seed, float out[4]) {
    uint32_t s = hash_u32(hash_combine(seed, dim_set));
    uint32_t shuffled = nested_uniform_scramble(index, s);
    for (uint32_t k = 0; k < 4; ++k) {
        uint32_t v = orc_sobol_unscrambled(shuffled, 4u * dim_set + k);
        v = nested_uniform_scramble(v, hash_u32(hash_combine(s, k + 1u)));
        out[k] = (float)(v >> 8) * (1.0f / 16777216.0f);
    }
}

/* pathtracer.rs:62-64: table[i * dims + set] = sample_4d(i, set, seed) */
void orc_sobol_table(uint32_t rows, uint32_t dims, uint32_t seed, float *out) {
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t d = 0; d < dims; ++d)
            orc_sobol_sample4d(i, d, seed, out + 4u * ((size_t)i * dims + d));
}

/* ------------------------------------------------------------------------------------------
 * Cube environment map: Vulkan face selection and bilinear filtering on level 0 with
 * seamless cube edges (texture.rs:126-141 binds a cube view with a linear sampler;
 * pathtracing.wgsl:168 samples it at LOD 0).
 * ------------------------------------------------------------------------------------------ */
/* face: major axis index, its sign, and the (axis, sign) of sc and tc (Vulkan face table) */
static const int FACE_MA[6] = {0, 0, 1, 1, 2, 2};
static const float FACE_MS[6] = {1, -1, 1, -1, 1, -1};
static const int FACE_SA[6] = {2, 2, 0, 0, 0, 0};
static const float FACE_SS[6] = {-1, 1, 1, 1, 1, -1};
static const int FACE_TA[6] = {1, 1, 2, 2, 1, 1};
static const float FACE_TS[6] = {-1, -1, 1, -1, -1, -1};

static int select_face(v3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return d.x >= 0.0f ? 0 : 1;
    if (ay >= az) return d.y >= 0.0f ? 2 : 3;
    return d.z >= 0.0f ? 4 : 5;
}

static const float *texel(const orc_env *env, int f, int i, int j) {
    return env->texels + 4u * (((size_t)f * env->face_size + (size_t)j) * env->face_size + (size_t)i);
}

/* Texel (i, j) of face f where exactly one coordinate is one step outside the face: the texel
 * of the adjacent face that shares that edge. */
static void wrap_edge(int f, int i, int j, int n, int *fo, int *io, int *jo) {
    float d[3];
    int out_s = (i < 0 || i >= n);
    float edge = out_s ? (i < 0 ? -1.0f : 1.0f) : (j < 0 ? -1.0f : 1.0f);
    int along = out_s ? j : i;
    d[FACE_MA[f]] = FACE_MS[f];
    if (out_s) { d[FACE_SA[f]] = edge * FACE_SS[f]; d[FACE_TA[f]] = 0.0f; }
    else { d[FACE_TA[f]] = edge * FACE_TS[f]; d[FACE_SA[f]] = 0.0f; }
    int oa = out_s ? FACE_SA[f] : FACE_TA[f];            /* adjacent face's major axis */
    int g = oa * 2 + (d[oa] > 0.0f ? 0 : 1);
    /* along-edge axis of f and its sign */
    int la = out_s ? FACE_TA[f] : FACE_SA[f];
    float ls = out_s ? FACE_TS[f] : FACE_SS[f];
    /* On g, the old major axis is at the edge, the along axis keeps its texel. */
    int gi, gj;
    if (FACE_SA[g] == FACE_MA[f]) {
        float sc = FACE_SS[g] * d[FACE_MA[f]];
        gi = sc > 0.0f ? n - 1 : 0;
        gj = (FACE_TS[g] * ls > 0.0f) ? along : n - 1 - along;
        (void)la;
    } else {
        float tc = FACE_TS[g] * d[FACE_MA[f]];
        gj = tc > 0.0f ? n - 1 : 0;
        gi = (FACE_SS[g] * ls > 0.0f) ? along : n - 1 - along;
    }
    *fo = g; *io = gi; *jo = gj;
}

static void fetch_texel(const orc_env *env, int f, int i, int j, float out[3]) {
    int n = (int)env->face_size;
    int oi = (i < 0 || i >= n), oj = (j < 0 || j >= n);
    if (!oi && !oj) { const float *t = texel(env, f, i, j); out[0] = t[0]; out[1] = t[1]; out[2] = t[2]; return; }
    if (oi && oj) {
        /* corner: average of the face's corner texel and its two edge neighbours */
        int ci = i < 0 ? 0 : n - 1, cj = j < 0 ? 0 : n - 1;
        int f1, i1, j1, f2, i2, j2;
        wrap_edge(f, i, cj, n, &f1, &i1, &j1);
        wrap_edge(f, ci, j, n, &f2, &i2, &j2);
        const float *a = texel(env, f, ci, cj), *b = texel(env, f1, i1, j1), *c = texel(env, f2, i2, j2);
        for (int k = 0; k < 3; ++k) out[k] = ((a[k] + b[k]) + c[k]) / 3.0f;
        return;
    }
    int g, gi, gj;
    wrap_edge(f, i, j, n, &g, &gi, &gj);
    const float *t = texel(env, g, gi, gj);
    out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}

/* The lat-long form (face_size 0): u = atan2(d.z, d.x)/2pi + 1/2, v = atan2(|d.xz|, d.y)/pi,
 * bilinear at LOD 0; texel columns wrap across the u seam and rows clamp at the poles. */
static void latlong_texel(const orc_env *env, int i, int j, float out[3]) {
    int w = (int)env->width, h = (int)env->height;
    if (i < 0) i = w - 1;
    else if (i >= w) i = 0;
    if (j < 0) j = 0;
    else if (j >= h) j = h - 1;
    const float *t = env->texels + ((size_t)j * (size_t)w + (size_t)i) * 4;
    out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}
static void latlong_sample(const orc_env *env, v3 d, float out[3]) {
    float r = sqrtf(d.x * d.x + d.z * d.z);
    float u = orc_atan2(d.z, d.x) * 0.15915494309189535f + 0.5f;
    float v = orc_atan2(r, d.y) * 0.31830988618379067f;
    float w = (float)env->width, h = (float)env->height;
    float x = u * w - 0.5f, y = v * h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    if (!(fx >= -1.0f)) fx = -1.0f;
    if (fx > w - 1.0f) fx = w - 1.0f;
    if (!(fy >= -1.0f)) fy = -1.0f;
    if (fy > h - 1.0f) fy = h - 1.0f;
    float a = x - fx, b = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    float t00[3], t10[3], t01[3], t11[3];
    latlong_texel(env, i0, j0, t00);
    latlong_texel(env, i0 + 1, j0, t10);
    latlong_texel(env, i0, j0 + 1, t01);
    latlong_texel(env, i0 + 1, j0 + 1, t11);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int k = 0; k < 3; ++k) out[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
}

void orc_env_sample(const orc_env *env, const float dir[3], float out[3]) {
    v3 d = ld3(dir);
    if (env->face_size == 0) {
        latlong_sample(env, d, out);
        return;
    }
    int f = select_face(d);
    float dc[3] = {d.x, d.y, d.z};
    float ma = fabsf(dc[FACE_MA[f]]);
    float sc = FACE_SS[f] * dc[FACE_SA[f]];
    float tc = FACE_TS[f] * dc[FACE_TA[f]];
    float s = 0.5f * (sc / ma + 1.0f);
    float t = 0.5f * (tc / ma + 1.0f);
    float n = (float)env->face_size;
    float u = s * n - 0.5f, v = t * n - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    int i0 = (int)fu, j0 = (int)fv;
    float t00[3], t10[3], t01[3], t11[3];
    fetch_texel(env, f, i0, j0, t00);
    fetch_texel(env, f, i0 + 1, j0, t10);
    fetch_texel(env, f, i0, j0 + 1, t01);
    fetch_texel(env, f, i0 + 1, j0 + 1, t11);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int k = 0; k < 3; ++k) out[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
}

/* ------------------------------------------------------------------------------------------
 * Integrator, pathtracing.wgsl:43-272
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    v3 position; float dist; v3 normal; v3 tangent; float tangent_w;
    v3 color; float roughness, metallic; int emissive;
} hit_info;

/* intersect_scene, raytracing_sw.wgsl:138-173 */
static hit_info intersect_scene(const orc_scene *sc, const ray_t *ray, orc_counters *cnt) {
    orc_raw_hit h = intersect_tlas(sc, ray);
    hit_info info; memset(&info, 0, sizeof(info));
    info.dist = h.dist;
    cnt->rays += 1; cnt->n_aabb += h.n_aabb; cnt->n_tri += h.n_tri;
    if (info.dist == NO_HIT) return info;
    const orc_instance *in = &sc->instances[h.instance];
    info.position = add3(ray->o, scl3(h.dist, ray->d));
    const orc_vertex *v0 = &sc->vertices[h.i0], *v1 = &sc->vertices[h.i1], *v2 = &sc->vertices[h.i2];
    float bx = h.bary[0], by = h.bary[1], bz = h.bary[2];
    v3 ln = add3(add3(scl3(bx, ld3(v0->normal)), scl3(by, ld3(v1->normal))), scl3(bz, ld3(v2->normal)));
    float lt[4];
    for (int k = 0; k < 4; ++k) lt[k] = (v0->tangent[k] * bx + v1->tangent[k] * by) + v2->tangent[k] * bz;
    const float *w2l = in->world_to_local, *l2w = in->local_to_world;
    /* transpose(mat3(W2L)) * n: component c = dot(W2L column c, n) */
    info.normal = mk3((w2l[0] * ln.x + w2l[1] * ln.y) + w2l[2] * ln.z,
                      (w2l[4] * ln.x + w2l[5] * ln.y) + w2l[6] * ln.z,
                      (w2l[8] * ln.x + w2l[9] * ln.y) + w2l[10] * ln.z);
    info.tangent = mk3((l2w[0] * lt[0] + l2w[4] * lt[1]) + l2w[8] * lt[2],
                       (l2w[1] * lt[0] + l2w[5] * lt[1]) + l2w[9] * lt[2],
                       (l2w[2] * lt[0] + l2w[6] * lt[1]) + l2w[10] * lt[2]);
    info.tangent_w = lt[3];
    info.color = ld3(in->color);
    info.roughness = in->roughness;
    info.metallic = in->metallic;
    info.emissive = in->emissive > 0.0f;
    return info;
}

/* sample_vndf_iso, pathtracing.wgsl:68-98 */
static v3 sample_vndf_iso(float r0, float r1, v3 wi, float alpha, v3 n) {
    if (alpha == 0.0f) return n;
    v3 wi_z = scl3(dot3(wi, n), n);
    v3 wi_xy = sub3(wi, wi_z);
    v3 wi_std = normalize3(sub3(wi_z, scl3(alpha, wi_xy)));
    float wi_std_z = dot3(wi_std, n);
    float phi = (2.0f * r0 - 1.0f) * PI_F;
    float z = (1.0f - r1) * (1.0f + wi_std_z) - wi_std_z;
    float sin_theta = sqrtf(fmin_(fmax_(1.0f - z * z, 0.0f), 1.0f));
    float x = sin_theta * orc_cos(phi);
    float y = sin_theta * orc_sin(phi);
    v3 c_std = mk3(x, y, z);
    v3 wr = add3(n, mk3(0.0f, 0.0f, 1.0f));
    float safe_wrz = fmax_(wr.z, 1e-6f);
    float dwc = dot3(wr, c_std);
    v3 c = sub3(divs3(scl3(dwc, wr), safe_wrz), c_std);
    v3 wm_std = add3(c, wi_std);
    v3 wm_std_z = scl3(dot3(n, wm_std), n);
    v3 wm_std_xy = sub3(wm_std_z, wm_std);
    return normalize3(add3(wm_std_z, scl3(alpha, wm_std_xy)));
}

/* sample_cosine_hemisphere, pathtracing.wgsl:100-105 */
static v3 sample_cosine_hemisphere(float r0, float r1) {
    float phi = TWO_PI_F * r0;
    float st = sqrtf(1.0f - r1), ct = sqrtf(r1);
    return mk3(orc_cos(phi) * st, orc_sin(phi) * st, ct);
}

/* F_SchlickApprox, pathtracing.wgsl:113-115 */
static v3 f_schlick(float hdotv, v3 r0) {
    float p = pow5f(1.0f - hdotv);
    return mk3(r0.x + (1.0f - r0.x) * p, r0.y + (1.0f - r0.y) * p, r0.z + (1.0f - r0.z) * p);
}

/* Lambda_TrowbridgeReitz, pathtracing.wgsl:119-125 */
static float lambda_tr(float ndotv, float alpha2) {
    float c2 = ndotv * ndotv;
    float s2 = 1.0f - c2;
    float t2 = s2 / c2;
    return (-1.0f + sqrtf(1.0f + alpha2 * t2)) / 2.0f;
}

/* pathtracing.wgsl:151-159: fract(table[entry] + shift); entries past the end of the table
 * (row 1024, or bounce sets beyond the table's last row) read as zero. */
static void sobol_fetch(const float *table, uint32_t entries, uint32_t entry, const float shift[4], float out[4]) {
    static const float zero[4] = {0, 0, 0, 0};
    const float *s = entry < entries ? sobol_at(table, entry) : zero;
    for (int k = 0; k < 4; ++k) out[k] = fract1(s[k] + shift[k]);
}

/* First bounce with Russian roulette (orc_set_roulette_first): 0 is the reference's rule
 * (pathtracing.wgsl:227-236); a larger value lets every path continue, throughput unchanged,
 * through the bounces before it (the library's nb_render_options). */
static uint32_t g_roulette_first = 0;
void orc_set_roulette_first(uint32_t bounce) { g_roulette_first = bounce; }

/* sample_rendering_eq, pathtracing.wgsl:161-241 */
/* Shading points of a path as the learned cache's training records see them (no counterpart in
 * the reference): up to max_rec non-emissive hits, each as 16 floats (position, outgoing
 * direction, normal, roughness, albedo, metallic, 0, 0) with the throughput that arrived there. */
/* Shading-point records of a path: the first `max` non-emissive shading points (training
 * records), or, with query set, the learned cache's query point alone: the first shading point
 * reached at bounce >= 1 after a lobe whose alpha is at least mirror_alpha, or at bounce
 * max_cache_bounce, where the path then stops (queried = 1) without shading it. */
typedef struct {
    float *raw; v3 *thr; uint32_t n, max;
    int query; float mirror_alpha; uint32_t max_cache_bounce; int queried;
} path_rec;

static v3 sample_rendering_eq_rec(const orc_scene *sc, const orc_globals *g, const float *table, uint32_t stride,
                                  uint32_t table_entries, const orc_env *env, const float shift[4], ray_t ray,
                                  orc_counters *cnt, path_rec *rec) {
    v3 thr = mk3(1.0f, 1.0f, 1.0f);
    float prev_alpha = 0.0f;   /* alpha of the previous bounce's lobe (1 for the diffuse lobe) */
    for (uint32_t bounce = 0; bounce <= g->bounces; ++bounce) {
        const int stop = rec && rec->query && bounce >= 1 &&
                         (prev_alpha >= rec->mirror_alpha || bounce >= rec->max_cache_bounce);
        hit_info hit = intersect_scene(sc, &ray, cnt);
        if (hit.dist == NO_HIT) {
            float d[3] = {ray.d.x, ray.d.y, ray.d.z}, e[3];
            orc_env_sample(env, d, e);
            return mul3(thr, ld3(e));
        }
        if (hit.emissive) return mul3(thr, hit.color);
        float alpha = hit.roughness * hit.roughness;
        float alpha2 = alpha * alpha;
        v3 n = normalize3(hit.normal);
        if (rec && rec->n < rec->max && (!rec->query || stop)) {
            v3 wo_ = normalize3(mk3(-ray.d.x, -ray.d.y, -ray.d.z));
            float *r = rec->raw + 16u * rec->n;
            r[0] = hit.position.x; r[1] = hit.position.y; r[2] = hit.position.z;
            r[3] = wo_.x; r[4] = wo_.y; r[5] = wo_.z;
            r[6] = n.x; r[7] = n.y; r[8] = n.z;
            r[9] = hit.roughness;
            r[10] = hit.color.x; r[11] = hit.color.y; r[12] = hit.color.z;
            r[13] = hit.metallic; r[14] = 0.0f; r[15] = 0.0f;
            rec->thr[rec->n] = thr;
            rec->n += 1;
        }
        if (stop) { rec->queried = 1; return mk3(0.0f, 0.0f, 0.0f); }
        float s0[4], s1[4];
        uint32_t e0 = g->sample * stride + 1u + bounce * 2u;
        sobol_fetch(table, table_entries, e0, shift, s0);
        sobol_fetch(table, table_entries, e0 + 1u, shift, s1);
        v3 wo = normalize3(mk3(-ray.d.x, -ray.d.y, -ray.d.z));
        float cos_o = dot3(wo, n);
        v3 wi;
        float metallic = hit.metallic;
        v3 albedo = hit.color;
        v3 f0 = mk3(0.04f * (1.0f - metallic) + albedo.x * metallic,
                    0.04f * (1.0f - metallic) + albedo.y * metallic,
                    0.04f * (1.0f - metallic) + albedo.z * metallic);
        float spec_w = luminance(f_schlick(dot3(wo, n), f0));
        float diff_w = (1.0f - metallic) * luminance(albedo);
        float p_spec = spec_w / (spec_w + diff_w);
        float p_diff = 1.0f - p_spec;
        if (s0[0] < p_spec) {
            v3 wm = sample_vndf_iso(s0[1], s0[2], wo, alpha, n);
            /* reflect(-wo, wm) = -wo - (2 * dot(wm, -wo)) * wm */
            v3 mwo = mk3(-wo.x, -wo.y, -wo.z);
            float k2 = 2.0f * dot3(wm, mwo);
            wi = sub3(mwo, scl3(k2, wm));
            float cos_d = dot3(wo, wm);
            float cos_i = dot3(wi, n);
            v3 F = f_schlick(cos_d, f0);
            float ll = lambda_tr(cos_i, alpha2);
            float lv = lambda_tr(cos_o, alpha2);
            float den = (1.0f + ll) + lv;
            v3 spec = mk3((F.x * (1.0f + lv)) / den, (F.y * (1.0f + lv)) / den, (F.z * (1.0f + lv)) / den);
            thr = mul3(thr, divs3(spec, p_spec));
            prev_alpha = alpha;
        } else {
            v3 t = hit.tangent;
            v3 nt = normalize3(sub3(t, scl3(dot3(t, n), n)));
            v3 b = cross3(n, nt);
            v3 c = sample_cosine_hemisphere(s1[1], s1[2]);
            wi = add3(add3(scl3(c.x, nt), scl3(c.y, b)), scl3(c.z, n));
            v3 wm = normalize3(add3(wi, wo));
            float cos_d = dot3(wi, wm);
            float cos_i = dot3(wi, n);
            float fd90 = 0.5f + (2.0f * alpha) * pow2f(cos_d);
            float resp = (1.0f + (fd90 - 1.0f) * pow5f(1.0f - cos_i)) * (1.0f + (fd90 - 1.0f) * pow5f(1.0f - cos_o));
            float om = 1.0f - metallic;
            v3 diff = mk3((om * albedo.x) * resp, (om * albedo.y) * resp, (om * albedo.z) * resp);
            thr = mul3(thr, divs3(diff, p_diff));
            prev_alpha = 1.0f;
        }
        if (bounce >= g_roulette_first) {   /* the reference: from bounce 0 */
            float p_cont = fmin_(1.0f - pow8f((float)bounce / (float)g->bounces), 1.0f);
            p_cont *= fmin_(luminance(thr) * g->contribution_factor, 1.0f);
            if (s0[2] < p_cont) thr = divs3(thr, p_cont);
            else return mk3(0.0f, 0.0f, 0.0f);
        }
        ray.o = hit.position; ray.d = wi; ray.inv = div3(mk3(1.0f, 1.0f, 1.0f), wi);
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

static v3 sample_rendering_eq(const orc_scene *sc, const orc_globals *g, const float *table, uint32_t stride,
                              uint32_t table_entries, const orc_env *env, const float shift[4], ray_t ray,
                              orc_counters *cnt) {
    return sample_rendering_eq_rec(sc, g, table, stride, table_entries, env, shift, ray, cnt, NULL);
}

/* One path of pixel (x, y) with its first three shading points as training records: target =
 * the path's radiance divided by the throughput that arrived at the point (0 where that is 0). */
void orc_train_path(const orc_scene *sc, const orc_camera *cam, const orc_globals *g, const float *table,
                    uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h,
                    uint32_t x, uint32_t y, float raw[48], float target[9], uint32_t *n_rec, float radiance[3]) {
    float shift[4];
    pixel_shift(x, y, shift);
    float jit[4];
    sobol_fetch(table, rows * stride, g->sample * stride, shift, jit);
    ray_t ray = generate_ray(cam, x, y, w, h, jit[0], jit[1]);
    v3 thr[3];
    path_rec rec = {raw, thr, 0u, 3u, 0, 0.0f, 0u, 0};
    orc_counters cnt = {0, 0, 0, 0};
    v3 out = sample_rendering_eq_rec(sc, g, table, stride, rows * stride, env, shift, ray, &cnt, &rec);
    for (uint32_t k = 0; k < rec.n; ++k) {
        target[3 * k + 0] = thr[k].x > 0.0f ? out.x / thr[k].x : 0.0f;
        target[3 * k + 1] = thr[k].y > 0.0f ? out.y / thr[k].y : 0.0f;
        target[3 * k + 2] = thr[k].z > 0.0f ? out.z / thr[k].z : 0.0f;
    }
    *n_rec = rec.n;
    radiance[0] = out.x; radiance[1] = out.y; radiance[2] = out.z;
}

/* The learned cache's query for pixel (x, y) of frame g->sample: the same path as
 * orc_sample_pixel up to its query point.  Returns 1 with the point's record (raw[16]) and the
 * throughput arriving there (thr[3]) when the path reaches it; direct[3] = the radiance met
 * before (sky, emitter), 0 when queried. */
uint32_t orc_cache_query(const orc_scene *sc, const orc_camera *cam, const orc_globals *g, const float *table,
                         uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h, uint32_t x,
                         uint32_t y, float mirror_alpha, uint32_t max_cache_bounce, float direct[3], float raw[16],
                         float thr_out[3]) {
    float shift[4];
    pixel_shift(x, y, shift);
    float jit[4];
    sobol_fetch(table, rows * stride, g->sample * stride, shift, jit);
    ray_t ray = generate_ray(cam, x, y, w, h, jit[0], jit[1]);
    v3 thr[1];
    path_rec rec = {raw, thr, 0u, 1u, 1, mirror_alpha, max_cache_bounce, 0};
    orc_counters cnt = {0, 0, 0, 0};
    v3 out = sample_rendering_eq_rec(sc, g, table, stride, rows * stride, env, shift, ray, &cnt, &rec);
    direct[0] = out.x; direct[1] = out.y; direct[2] = out.z;
    if (rec.queried) { thr_out[0] = thr[0].x; thr_out[1] = thr[0].y; thr_out[2] = thr[0].z; }
    return (uint32_t)rec.queried;
}

void orc_sample_pixel(const orc_scene *sc, const orc_camera *cam, const orc_globals *g, const float *table,
                      uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h,
                      uint32_t x, uint32_t y, float out[3], orc_counters *cnt) {
    float shift[4];
    pixel_shift(x, y, shift);
    float jit[4];
    sobol_fetch(table, rows * stride, g->sample * stride, shift, jit);
    ray_t ray = generate_ray(cam, x, y, w, h, jit[0], jit[1]);
    v3 s = sample_rendering_eq(sc, g, table, stride, rows * stride, env, shift, ray, cnt);
    out[0] = s.x; out[1] = s.y; out[2] = s.z;
    cnt->samples += 1;
}

/* main, pathtracing.wgsl:243-260 */
void orc_render(const orc_scene *sc, const orc_camera *cam, const orc_globals *g, const float *table,
                uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h,
                uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, float *rgba, orc_counters *counters,
                int threads) {
    uint64_t rays = 0, aabb = 0, tri = 0, samples = 0;
    (void)threads;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(+:rays, aabb, tri, samples)
#endif
    for (int yy = (int)y0; yy < (int)y1; ++yy) {
        orc_counters c = {0, 0, 0, 0};
        for (uint32_t x = x0; x < x1; ++x) {
            float *px = rgba + 4u * ((size_t)yy * w + x);
            float col[3] = {0, 0, 0};
            if (g->weight > 0.0f) { col[0] = px[0]; col[1] = px[1]; col[2] = px[2]; }
            float s[3];
            orc_sample_pixel(sc, cam, g, table, stride, rows, env, w, h, x, (uint32_t)yy, s, &c);
            /* mix(color, s, weight) = color * (1 - weight) + s * weight */
            for (int k = 0; k < 3; ++k) px[k] = col[k] * (1.0f - g->weight) + s[k] * g->weight;
            px[3] = 1.0f;
        }
        rays += c.rays; aabb += c.n_aabb; tri += c.n_tri; samples += c.samples;
    }
    if (counters) {
        counters->rays += rays; counters->n_aabb += aabb; counters->n_tri += tri; counters->samples += samples;
    }
}

void orc_render_mean_f64(const orc_scene *sc, const orc_camera *cam, const orc_globals *g0, const float *table,
                         uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h,
                         uint32_t first, uint32_t n, double *out64, int threads) {
    (void)threads;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for (int yy = 0; yy < (int)h; ++yy) {
        orc_counters c = {0, 0, 0, 0};
        orc_globals g = *g0;
        for (uint32_t x = 0; x < w; ++x) {
            double acc[3] = {0, 0, 0};
            for (uint32_t k = 0; k < n; ++k) {
                g.sample = first + k;
                float s[3];
                orc_sample_pixel(sc, cam, &g, table, stride, rows, env, w, h, x, (uint32_t)yy, s, &c);
                acc[0] += s[0]; acc[1] += s[1]; acc[2] += s[2];
            }
            double *o = out64 + 3u * ((size_t)yy * w + x);
            o[0] = acc[0] / n; o[1] = acc[1] / n; o[2] = acc[2] / n;
        }
    }
}

/* Mean of samples first..first+n-1 (float64 accumulation) at a list of pixels (x, y pairs). */
void orc_render_pixels_mean_f64(const orc_scene *sc, const orc_camera *cam, const orc_globals *g0, const float *table,
                                uint32_t stride, uint32_t rows, const orc_env *env, uint32_t w, uint32_t h,
                                const uint32_t *xy, uint32_t n_pixels, uint32_t first, uint32_t n, double *out64,
                                int threads) {
    (void)threads;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
    for (int k = 0; k < (int)n_pixels; ++k) {
        orc_counters c = {0, 0, 0, 0};
        orc_globals g = *g0;
        double acc[3] = {0, 0, 0};
        for (uint32_t i = 0; i < n; ++i) {
            g.sample = first + i;
            float s[3];
            orc_sample_pixel(sc, cam, &g, table, stride, rows, env, w, h, xy[2 * k], xy[2 * k + 1], s, &c);
            acc[0] += s[0]; acc[1] += s[1]; acc[2] += s[2];
        }
        out64[3 * k] = acc[0] / n; out64[3 * k + 1] = acc[1] / n; out64[3 * k + 2] = acc[2] / n;
    }
}
