/*
 * nb_oracle.h — CPU restatement of nBounce's path-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product never links, loads or calls it.
 *
 * Build: plain C11, -O2 -ffp-contract=off, no fast-math (IEEE binary32, no FMA
 * contraction), so every float operation happens in the order written here.
 *
 * What each function follows (reference file:line) and what pins it
 * (the test that checks it):
 *   orc_hash4u / orc_map4f    common.wgsl:7-9,27-41     inverse round trip, closed-form map values
 *   orc_sin / orc_cos         (own Cody-Waite + Cephes)  float64 numpy sin/cos, <= 2 ulp
 *   orc_atan / orc_atan2      (own, Cephes atanf; the lat-long environment form, no reference)
 *                                                        float64 numpy arctan2 over all quadrants,
 *                                                        axes and both reduction branches
 *   orc_mat4_inverse          glam Mat4::inverse         M*M^-1 = I, numpy.linalg.inv;
 *                                                        bit order vs glam 0.29: parity unpinned
 *   orc_camera                camera.rs:121-142          numpy restatement of look_at/perspective
 *                                                        closed forms; bits vs glam: parity unpinned
 *   orc_build_scene           bvh.rs:15-330, scene.rs:239-305
 *                                                        hand-computed tiny trees, invariants
 *                                                        (coverage, tight boxes, LIFO numbering,
 *                                                        cL+cR < cP), brute-force SAH
 *   orc_intersect_triangle    raytracing_sw.wgsl:51-78  closed-form hits / misses / culls
 *   orc_intersect_aabb        raytracing_sw.wgsl:81-89  closed-form slab cases
 *   orc_trace                 raytracing_sw.wgsl:138-319 brute-force closest hit over all
 *                                                        instanced triangles (identical except ties)
 *   orc_sobol_*               sobol_burley::sample_4d (pathtracer.rs:62-64), restated from
 *                             Burley 2020; unscrambled bits = scipy Sobol (Joe-Kuo);
 *                             Owen scrambling pinned by (0,m,2)-net stratification;
 *                             bits vs sobol_burley 0.5.0: parity unpinned
 *   orc_env_sample            texture.rs:126-141, pathtracing.wgsl:168 (Vulkan cube rules)
 *                                                        constant / linear fields, face centres;
 *                             lat-long form (no reference): constant field, linear-in-u rows,
 *                             one-hot texels at +x / +z / -z / +y, seam continuity, the baked
 *                             sky against the cube bake of the same sky
 *   orc_cache_query           (no reference: the learned cache's query point) the same path as
 *                             orc_sample_pixel (equal radiance bits where it ends before the
 *                             point); the record equals orc_train_path's k-th record, k = 1 for
 *                             rough materials, max_cache_bounce for smooth metal, 1 again with a
 *                             lower mirror_alpha (tests/test_oracle_cache_query.py)
 *   orc_set_roulette_first    (no reference: nb_render_options) forced roulette numbers give a
 *                             path of exactly first + 1 rays; 0 leaves the reference's rule
 *   orc_render                pathtracing.wgsl:27-272    white-furnace / closed-form single-bounce
 *                                                        expectations, RR unbiasedness; the
 *                                                        multi-bounce loop by hand-derived path
 *                                                        lengths in closed boxes: roulette length
 *                                                        term 1-(b/B)^8 (exact scan), per-bounce
 *                                                        Sobol entries i*17+1+2b / +2+2b (marker
 *                                                        tables, dependency probes), mean path
 *                                                        length = sum_b prod_k<b (1-(k/B)^8)
 *   orc_train_path            (no reference: the learned cache's training records) the same
 *                                                        path as orc_sample_pixel (equal radiance),
 *                                                        records = its shading points, target x
 *                                                        throughput = radiance
 */
#ifndef NB_ORACLE_H
#define NB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference byte layouts (bytemuck casts of the Rust structs). */
typedef struct { float min[3]; uint32_t start; float max[3]; uint32_t end; } orc_node;           /* bvh.rs:15-24, 32 B */
typedef struct { float position[3]; float u; float normal[3]; float v; float tangent[4]; } orc_vertex; /* scene.rs:13-21, 48 B */
typedef struct {
    float world_to_local[16]; float local_to_world[16];  /* column-major glam Mat4 */
    float color[4]; float roughness, metallic, emissive; uint32_t node;
} orc_instance;                                                                                   /* scene.rs:220-230, 160 B */
typedef struct {
    float local_to_world[16]; float color[4]; float roughness, metallic, emissive;
    uint32_t index_start, index_end;                                                              /* Range<u32> of indices */
} orc_primitive;                                                                                  /* scene.rs:59-67 */
typedef struct { float world_to_clip[16]; float clip_to_world[16]; } orc_camera;                 /* camera.rs:7-12, 128 B */
typedef struct { uint32_t sample; float weight; uint32_t bounces; float contribution_factor; } orc_globals; /* pathtracer.rs:26-33 */

typedef struct {
    const orc_node *blas; uint32_t n_blas;
    const orc_node *tlas; uint32_t n_tlas;
    const orc_instance *instances; uint32_t n_instances;
    const orc_vertex *vertices; uint32_t n_vertices;
    const uint32_t *indices; uint32_t n_indices;
} orc_scene;

typedef struct {
    const float *texels;   /* cube: 6 faces, each face_size x face_size RGBA32F, Vulkan layer order +X,-X,+Y,-Y,+Z,-Z;
                              lat-long (face_size 0): height rows of width RGBA32F, row 0 at +y */
    uint32_t face_size;
    uint32_t width, height;
} orc_env;

typedef struct {
    uint32_t instance, i0, i1, i2;
    float bary[3];
    float dist;
    uint32_t n_aabb, n_tri;
} orc_raw_hit;

typedef struct { uint64_t rays, n_aabb, n_tri, samples; } orc_counters;

/* ---- hashing ---- */
void orc_hash4u(const uint32_t s[4], uint32_t out[4]);
void orc_map4f(const uint32_t x[4], float out[4]);

/* ---- math ---- */
float orc_sin(float x);
float orc_cos(float x);
float orc_atan(float x);
float orc_atan2(float y, float x);
void orc_mat4_inverse(const float m[16], float out[16]);
void orc_mat4_mul(const float a[16], const float b[16], float out[16]);
void orc_camera_compute(const float eye[3], const float target[3], const float up[3],
                        float fov_y, float aspect, float z_near, orc_camera *out);

/* ---- BVH / scene build ---- */
/* Upper bounds: BLAS nodes <= sum over primitives of (2*tris-1); TLAS nodes <= 2*n_primitives-1. */
int orc_build_scene(const orc_vertex *vertices, uint32_t n_vertices,
                    uint32_t *indices, uint32_t n_indices,
                    const orc_primitive *prims, uint32_t n_prims,
                    orc_node *blas_out, uint32_t blas_cap, uint32_t *n_blas,
                    orc_node *tlas_out, uint32_t tlas_cap, uint32_t *n_tlas,
                    orc_instance *instances_out);
/* One BVHTree::append over triangles [tri_start, tri_end) of (vertices, indices); permutes indices in place. */
int orc_build_blas(const orc_vertex *vertices, uint32_t n_vertices,
                   uint32_t *indices, uint32_t n_indices,
                   uint32_t tri_start, uint32_t tri_end,
                   orc_node *nodes_out, uint32_t cap, uint32_t *n_nodes);

/* ---- intersection ---- */
void orc_intersect_triangle(const float o[3], const float d[3], const float v0[3], const float v1[3],
                            const float v2[3], float out_tuv[3]);
float orc_intersect_aabb(const float o[3], const float inv_d[3], const float bmin[3], const float bmax[3]);
void orc_trace(const orc_scene *scene, const float o[3], const float d[3], orc_raw_hit *hit);
void orc_trace_brute(const orc_scene *scene, const float o[3], const float d[3], orc_raw_hit *hit);
void orc_generate_ray(const orc_camera *cam, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
                      const float jitter[2], float o[3], float d[3]);
/* Primary hits of a whole image with jitter from the Sobol row `sample`: out = per pixel
 * {instance, i0, i1, i2, dist-bits} (5 x u32), instance = 0xFFFFFFFF on a miss. */
void orc_primary_hits(const orc_scene *scene, const orc_camera *cam, uint32_t width, uint32_t height,
                      const float *sobol_table, uint32_t sobol_stride, uint32_t sample,
                      uint32_t *out);

/* ---- Sobol-Burley ---- */
/* Joe-Kuo direction numbers: poly[d] with leading and trailing bits, vinit[d*stride + k] = m_{k+1}. */
int orc_sobol_init(const int64_t *poly, const int64_t *vinit, uint32_t n_dims, uint32_t vinit_stride);
uint32_t orc_sobol_unscrambled(uint32_t index, uint32_t dim);
void orc_sobol_sample4d(uint32_t index, uint32_t dim_set, uint32_t seed, float out[4]);
void orc_sobol_table(uint32_t rows, uint32_t dims, uint32_t seed, float *out);

/* ---- environment ---- */
void orc_env_sample(const orc_env *env, const float dir[3], float out[3]);

/* ---- options (the library's nb_render_options; not in the reference) ---- */
void orc_set_roulette_first(uint32_t bounce);

/* ---- integrator ---- */
/* Renders one sample per pixel into `rgba` (width*height*4, row-major) exactly as one dispatch of
 * pathtracing.wgsl:main, restricted to the pixel rectangle [x0,x1) x [y0,y1). `threads` > 1 uses
 * OpenMP over rows. When `f64_accum` is nonzero, `rgba64` (double) accumulates instead. */
void orc_render(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                const orc_env *env, uint32_t width, uint32_t height,
                uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                float *rgba, orc_counters *counters, int threads);
/* One path (one sample) for pixel (x,y): returns the radiance estimate. */
void orc_sample_pixel(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                      const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                      const orc_env *env, uint32_t width, uint32_t height,
                      uint32_t x, uint32_t y, float out[3], orc_counters *counters);
/* One path for pixel (x, y) with its first three non-emissive shading points as training records
 * of the learned cache (no counterpart in the reference): raw[16k..] = position, outgoing
 * direction, normal, roughness, albedo, metallic, 0, 0; target[3k..] = radiance / throughput
 * arriving there (0 where that throughput is 0). */
void orc_train_path(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                    const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                    const orc_env *env, uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                    float raw[48], float target[9], uint32_t *n_rec, float radiance[3]);
/* The learned cache's query for pixel (x, y) of frame g->sample (no counterpart in the
 * reference): the path of orc_sample_pixel up to the first shading point reached at bounce >= 1
 * after a lobe whose GGX alpha is at least mirror_alpha (the diffuse lobe counts as 1), or at
 * bounce max_cache_bounce.  Returns 1 when the path reaches that point, with its record in
 * raw[16] (as for orc_train_path) and the throughput arriving there in thr[3]; direct[3] is the
 * radiance met before it (sky or emitter; 0 when queried or ended by roulette or the limit). */
uint32_t orc_cache_query(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                         const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                         const orc_env *env, uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                         float mirror_alpha, uint32_t max_cache_bounce, float direct[3], float raw[16],
                         float thr[3]);
/* Same integrator but sums n_samples samples per pixel (rows first..first+n-1) in float64:
 * out64[p*3+c] = mean radiance. Used as the high-spp reference image. */
void orc_render_mean_f64(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                         const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                         const orc_env *env, uint32_t width, uint32_t height,
                         uint32_t first_sample, uint32_t n_samples, double *out64, int threads);

void orc_render_pixels_mean_f64(const orc_scene *scene, const orc_camera *cam, const orc_globals *g,
                                const float *sobol_table, uint32_t sobol_stride, uint32_t sobol_rows,
                                const orc_env *env, uint32_t width, uint32_t height,
                                const uint32_t *xy, uint32_t n_pixels, uint32_t first_sample, uint32_t n_samples,
                                double *out64, int threads);

#ifdef __cplusplus
}
#endif
#endif
