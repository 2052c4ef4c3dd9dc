/*
 * nbounce.h — C ABI of the MI355X (gfx950) path-tracing library.
 *
 * It replaces the renderer surface of the reference's `pathtracing` module.  Each entry
 * point names the reference interface it stands in for (paths relative to the reference
 * repository).  All structs below are byte-for-byte the reference's bytemuck layouts, so a
 * Rust caller passes `bytemuck::cast_slice` bytes straight through.
 *
 * Conventions
 *   - Every call returns nb_status; NB_OK == 0.  On failure nb_last_error() returns a
 *     thread-local message.  No call aborts the process.
 *   - Host arrays stay owned by the caller; the library copies what it needs during the call.
 *   - A context is externally synchronized (one host thread at a time).  Rendering calls are
 *     asynchronous on the context's HIP stream (and, for pipelined nb_render frames, its two
 *     side streams); nb_target_read() and nb_context_synchronize() wait for them.  A destroy
 *     function waits for the work that may still read what it frees: for scenes, environment
 *     maps, targets and nb_sobol_set()'s old table, all of the context's streams.
 *   - Destroy functions accept NULL.  Scenes, targets and environment maps keep their context
 *     alive: nb_context_destroy() takes effect when the last of them is destroyed.
 */
#ifndef NBOUNCE_H
#define NBOUNCE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NB_OK = 0,
    NB_ERR_INVALID_ARGUMENT = 1,
    NB_ERR_OUT_OF_MEMORY = 2,
    NB_ERR_HIP = 3,
    NB_ERR_BAD_LAYOUT = 4,
    NB_ERR_UNSUPPORTED = 5,
    NB_ERR_BUDGET = 6,
    NB_ERR_FORMAT = 7,      /* a scene or image file was rejected (MeshError, scene.rs:69-98) */
    NB_ERR_IO = 8           /* a file could not be read or written */
} nb_status;

/* ---- reference byte layouts ---------------------------------------------------------- */

/* BVHNode, src/pathtracing/bvh.rs:15-24 (32 B). Leaf <=> end > 0, range [start, end);
 * inner: end == 0, children start and start + 1. */
typedef struct { float min[3]; uint32_t start; float max[3]; uint32_t end; } nb_bvh_node;

/* Vertex, src/pathtracing/scene.rs:13-21 (48 B). */
typedef struct { float position[3]; float u; float normal[3]; float v; float tangent[4]; } nb_vertex;

/* Instance, src/pathtracing/scene.rs:220-230 (160 B); matrices column-major. */
typedef struct {
    float world_to_local[16];
    float local_to_world[16];
    float color[4];
    float roughness, metallic, emissive;
    uint32_t node;
} nb_instance;

/* Primitive, src/pathtracing/scene.rs:59-67 (index_range as [index_start, index_end)). */
typedef struct {
    float local_to_world[16];
    float color[4];
    float roughness, metallic, emissive;
    uint32_t index_start, index_end;
} nb_primitive;

/* CameraBuffer, src/common/camera.rs:7-12 (128 B). */
typedef struct { float world_to_clip[16]; float clip_to_world[16]; } nb_camera;

/* Globals push constants, src/pathtracing/pathtracer.rs:26-33 (16 B); sample is 1-based. */
typedef struct { uint32_t sample; float weight; uint32_t bounces; float contribution_factor; } nb_globals;

/* Per-ray counters (raytracing_sw.wgsl:99-101 n_aabb / n_tri, plus rays = intersect_scene
 * calls, instance entries, finished samples and rays that hit), summed since the last reset. */
typedef struct { uint64_t rays, n_aabb, n_tri, n_inst, samples, hits; } nb_counters;

typedef struct nb_context nb_context;
typedef struct nb_scene nb_scene;
typedef struct nb_target nb_target;
typedef struct nb_envmap nb_envmap;
typedef struct nb_cache nb_cache;

/* ---- errors / library ------------------------------------------------------------------ */
const char *nb_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
uint32_t nb_version(void);

/* ---- context: replaces WGPUContext (src/common/wgpu_context.rs:13-79) ------------------- */
nb_status nb_context_create(int device, nb_context **out);
nb_status nb_context_destroy(nb_context *ctx);
nb_status nb_context_synchronize(nb_context *ctx);

/* ---- host scene and glTF ingestion: Scene (src/pathtracing/scene.rs:100-105) and
 * Scene::parse_gltf (scene.rs:108-217) -----------------------------------------------------
 * A host scene holds the reference's three arrays (Vertex, u32 indices, Primitive).  Each parse
 * appends one glTF 2.0 document (.glb, or .gltf with data-URI or file buffers): every mesh
 * primitive's vertices and offset indices, then one Primitive per mesh primitive of every node
 * that has a mesh, with the node's local matrix (parents ignored, as in the reference).  A
 * primitive that is not a triangle list or lacks POSITION, NORMAL, TANGENT or indices fails
 * with NB_ERR_FORMAT and the reference's MeshError text; the scene is then left unchanged.
 * Accessor and view bounds are checked before anything is allocated, and a sparse accessor
 * without a bufferView may take at most 256 MiB, so a small file cannot make the reader claim
 * much more memory than it holds.
 * flags: NB_GLTF_SHARE_GEOMETRY appends the geometry of a mesh primitive once per distinct
 * accessor set, so primitives that reuse accessors share one index range (pair it with
 * NB_BUILD_SHARE_BLAS); the reference appends a copy per mesh primitive. */
#define NB_GLTF_SHARE_GEOMETRY 1u
typedef struct nb_host_scene nb_host_scene;
nb_status nb_host_scene_create(nb_host_scene **out);
nb_status nb_host_scene_destroy(nb_host_scene *scene);
nb_status nb_host_scene_parse_gltf(nb_host_scene *scene, const char *path, uint32_t flags);
/* The same from memory; relative buffer URIs resolve against base_dir (NULL: none allowed). */
nb_status nb_host_scene_parse_gltf_memory(nb_host_scene *scene, const void *data, size_t size,
                                          const char *base_dir, uint32_t flags);
nb_status nb_host_scene_counts(const nb_host_scene *scene, uint32_t *n_vertices, uint32_t *n_indices,
                               uint32_t *n_prims);
nb_status nb_host_scene_copy(const nb_host_scene *scene, nb_vertex *vertices, uint32_t *indices,
                             nb_primitive *prims);

/* ---- host-side scene build ------------------------------------------------------------- */
/* SceneBuffers::from_scene (src/pathtracing/scene.rs:280-307) without the upload: builds one
 * BLAS per primitive into one shared node array (bvh.rs:162-190), the instance bounds
 * (scene.rs:239-260), the TLAS (scene.rs:302) and writes the triangle permutation back into
 * `indices` (scene.rs:305).  Bit-exact with the reference builder.
 * flags: NB_BUILD_SHARE_BLAS reuses the first BLAS built over an index range for every later
 * primitive over the same range instead of rebuilding it (the reference rebuilds).
 * Capacities: see nb_build_scene_capacity. */
#define NB_BUILD_SHARE_BLAS 1u
nb_status nb_build_scene_capacity(uint32_t n_indices, const nb_primitive *prims, uint32_t n_prims,
                                  uint32_t flags, uint32_t *blas_cap, uint32_t *tlas_cap);
nb_status nb_build_scene(const nb_vertex *vertices, uint32_t n_vertices,
                         uint32_t *indices, uint32_t n_indices,
                         const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                         nb_bvh_node *blas, uint32_t blas_cap, uint32_t *n_blas,
                         nb_bvh_node *tlas, uint32_t tlas_cap, uint32_t *n_tlas,
                         nb_instance *instances);

/* ---- device scene: group-1 bindings blas, tlas, instances, vertices, indices ------------
 * (src/pathtracing/raytracing_sw.wgsl:33-37, uploaded at scene.rs:309-341).  The arrays are
 * validated (indices in range, children after parents, depth <= 32) and re-laid out for the
 * GPU; the caller's arrays are not kept.  The TLAS keeps its topology, but the boxes the kernels
 * traverse are refitted to each instance's transformed vertices (padded outward), so rays enter
 * fewer instances they miss; the closest hits are the same.  NB_TLAS_REFIT=0 in the environment
 * keeps the given boxes. */
nb_status nb_scene_create(nb_context *ctx,
                          const nb_bvh_node *blas, uint32_t n_blas,
                          const nb_bvh_node *tlas, uint32_t n_tlas,
                          const nb_instance *instances, uint32_t n_instances,
                          const nb_vertex *vertices, uint32_t n_vertices,
                          const uint32_t *indices, uint32_t n_indices,
                          nb_scene **out);
nb_status nb_scene_destroy(nb_scene *scene);
/* Device bytes held by the scene. */
nb_status nb_scene_device_bytes(const nb_scene *scene, uint64_t *bytes);

/* ---- Sobol-Burley table (pathtracer.rs:57-71, binding pathtracing.wgsl:13) -------------
 * nb_sobol_generate fills rows*dims float4 entries, entry [i*dims + set] = Owen-scrambled
 * Sobol point i of dimension set `set` (four dimensions), Burley 2020.
 * nb_sobol_set uploads a table (generated here or supplied by the caller, e.g. the exact
 * table a reference run used); `stride` is the number of float4 entries per sample row the
 * kernel steps by: 17 reproduces the reference (pathtracing.wgsl:2-4) whatever `bounces`. */
nb_status nb_sobol_generate(uint32_t rows, uint32_t dims, uint32_t seed, float *out);
nb_status nb_sobol_set(nb_context *ctx, const float *table, uint32_t rows, uint32_t dims, uint32_t stride);

/* ---- environment map (EnvMap::load, src/pathtracing/envmap.rs:12; cube view + linear
 * sampler, src/common/texture.rs:126-141) ----------------------------------------------
 * texels: 6 faces in Vulkan layer order +X,-X,+Y,-Y,+Z,-Z, each face_size^2 RGBA32F, row-major. */
nb_status nb_envmap_create_cube(nb_context *ctx, const float *texels, uint32_t face_size, nb_envmap **out);
nb_status nb_envmap_destroy(nb_envmap *env);
/* An equirectangular (lat-long) RGBA32F map, height rows of width texels, row 0 at the top.  No
 * counterpart in the reference, which samples only the cube (pathtracing.wgsl:14,168); this is
 * the native form of a 512x256-style sky.  Direction d (y up) looks up u = atan2(d.z, d.x)/2pi +
 * 1/2 (+x at u = 1/2, +z at 3/4, -z at 1/4), v = atan2(|d.xz|, d.y)/pi (+y at row 0), bilinear
 * at LOD 0, wrapping across the u seam and clamping at the poles.  Accepted wherever a cube is
 * (nb_render*, nb_server_add_scene, nb_render_cached). */
nb_status nb_envmap_create_latlong(nb_context *ctx, const float *texels, uint32_t width, uint32_t height,
                                   nb_envmap **out);

/* ---- render options (no counterpart in the reference) ------------------------------------
 * roulette_first_bounce: Russian roulette (pathtracing.wgsl:227-236) applies from this bounce
 * on; paths continue unconditionally, throughput unchanged, before it.  0 (the default) is the
 * reference's rule; 3 is "Russian roulette from bounce 3".  Applies to nb_render and
 * nb_render_frames on the context from the next call; served and cache-assisted renders keep
 * the reference's rule. */
typedef struct { uint32_t roulette_first_bounce; } nb_render_options;
nb_status nb_render_options_set(nb_context *ctx, const nb_render_options *options);
/* EnvMap::load (src/pathtracing/envmap.rs:12 -> Texture::create_cubemap, src/common/texture.rs:89-141):
 * a cube DDS file, decoded on the host to RGBA32F (mip 0 of each face; the shader samples LOD 0,
 * pathtracing.wgsl:168) and uploaded as nb_envmap_create_cube does.  Formats: BC6H_UF16 (the
 * reference's only one), BC6H_SF16, R16G16B16A16_FLOAT, R32G32B32A32_FLOAT (DX10 header, or the
 * legacy D3DFMT codes 113 / 116).  The six faces are read face-major with their mip chains
 * whatever the header's array size (the reference forces it to 6, texture.rs:95).  Unsupported
 * or truncated files fail with NB_ERR_FORMAT, unreadable ones with NB_ERR_IO. */
nb_status nb_envmap_load(nb_context *ctx, const char *path, nb_envmap **out);
/* The decoding alone, into 6 * face_size^2 RGBA32F texels (rgba == NULL: only *face_size). */
nb_status nb_dds_read_cube(const char *path, float *rgba, uint64_t cap_floats, uint32_t *face_size);
nb_status nb_dds_read_cube_memory(const void *data, size_t size, float *rgba, uint64_t cap_floats,
                                  uint32_t *face_size);
/* Writes a cube DDS (DX10 header, one mip level) from 6 faces of RGBA32F texels: as RGBA32F,
 * RGBA16F (rounded to nearest), or BC6H_UF16 (one-region blocks, for tests and tools).  No
 * counterpart in the reference, which only reads. */
#define NB_DDS_RGBA32F 0u
#define NB_DDS_RGBA16F 1u
#define NB_DDS_BC6H_UF16 2u
nb_status nb_dds_write_cube(const char *path, const float *rgba, uint32_t face_size, uint32_t format);
/* One BC6H block (16 bytes) decoded to 16 texels x RGB half-float bits, row-major; is_signed
 * selects BC6H_SF16.  Reserved modes decode to zero. */
nb_status nb_bc6h_decode_block(const uint8_t *block, uint32_t is_signed, uint16_t *rgb_half);

/* ---- camera (CameraController::calc_camera_data, src/common/camera.rs:121-142) ---------- */
nb_status nb_camera_look_at(const float eye[3], const float target[3], const float up[3],
                            float fov_y, float aspect, float z_near, nb_camera *out);

/* ---- render target (Pathtracer::create_output_texture, pathtracer.rs:192-202) ----------
 * RGBA32F, width and height multiples of 8 (pathtracer.rs:194). */
nb_status nb_target_create(nb_context *ctx, uint32_t width, uint32_t height, nb_target **out);
nb_status nb_target_destroy(nb_target *target);
nb_status nb_target_read(nb_target *target, float *rgba);         /* waits for pending work */
nb_status nb_target_write(nb_target *target, const float *rgba);
nb_status nb_target_clear(nb_target *target);

/* ---- display and image files ------------------------------------------------------------
 * nb_display stands for the RGBA16F surface (src/common/wgpu_context.rs:62) that BlitRenderer
 * draws the target onto (src/pathtracing/blit_renderer.rs, blit.wgsl:27-38): a fullscreen
 * triangle sampling the target bilinearly with clamp-to-edge addressing (texture.rs:64-74);
 * texel weights carry 8 fractional bits.  Target row 0 is the bottom of the view; display row 0
 * is the top of the surface.  nb_display_blit is asynchronous on the context stream;
 * nb_display_read waits and copies width * height * 4 half floats. */
typedef struct nb_display nb_display;
nb_status nb_display_create(nb_context *ctx, uint32_t width, uint32_t height, nb_display **out);
nb_status nb_display_destroy(nb_display *display);
nb_status nb_display_blit(nb_display *display, const nb_target *target);
nb_status nb_display_read(nb_display *display, uint16_t *rgba16f);

/* Files of an RGBA32F image in target row order (as nb_target_read returns it): Portable Float
 * Map (RGB, little-endian, rows bottom to top) and OpenEXR 2 (single-part scanline, no
 * compression, channels A B G R as FLOAT, or HALF with NB_EXR_HALF; scanline 0 at the top). */
#define NB_EXR_HALF 1u
nb_status nb_image_write_pfm(const char *path, const float *rgba, uint32_t width, uint32_t height);
nb_status nb_image_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height,
                             uint32_t flags);

/* ---- rendering (Pathtracer::dispatch, pathtracer.rs:225-239 -> pathtracing.wgsl:main) ----
 * nb_render: one sample per pixel of frame globals->sample with weight globals->weight,
 * color = mix(previous, sample, weight) (pathtracing.wgsl:247-260).  The frame's path kernel
 * runs on one of two library-owned side streams and writes the frame's samples; the mix runs on
 * the context stream after it, so work queued on the context stream after nb_render (and every
 * wait on it) follows the whole frame, while consecutive frames' paths overlap.  The target
 * keeps two sample buffers (32 bytes per pixel) from its first nb_render on.  NB_PIPELINE=0 in
 * the environment renders on the context stream alone (the same image bit for bit).
 * nb_render_frames: frames first->sample .. first->sample + n_frames - 1 in one launch, each
 * with weight 1/sample; the image equals n_frames consecutive nb_render calls bit for bit.
 * Targets of up to 2^19 pixels render one path per (frame, pixel) with the serving kernels;
 * larger ones keep a pixel in one lane for every frame. */
nb_status nb_render(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                    const nb_globals *globals, const nb_envmap *env);
nb_status nb_render_frames(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                           const nb_globals *first, uint32_t n_frames, const nb_envmap *env);

/* Per-render timing, off by default: with it on, nb_render, nb_render_frames and
 * nb_render_cached record HIP events around their launches on the context stream (the records
 * between launches cost a frame stream about 0.5 %).  No counterpart in the reference, which has
 * no timing API. */
nb_status nb_render_timing(nb_context *ctx, uint32_t on);
/* Device time of the last render call in milliseconds (its HIP events); NB_ERR_INVALID_ARGUMENT
 * when no render was timed since per-render timing was switched on. */
nb_status nb_last_render_ms(nb_context *ctx, float *ms);
/* HIP events around a region of work on the context stream: start records the first event,
 * stop records the second, waits for it and returns the elapsed milliseconds. */
nb_status nb_stream_timer_start(nb_context *ctx);
nb_status nb_stream_timer_stop(nb_context *ctx, float *ms);

/* ---- counters / debug ------------------------------------------------------------------- */
nb_status nb_counters_get(nb_context *ctx, nb_counters *out);
nb_status nb_counters_reset(nb_context *ctx);
/* Counting is on by default; off, the render kernels skip the counter atomics (the image is
 * the same).  The reference's n_aabb / n_tri are debug variables of the shader. */
nb_status nb_counters_enable(nb_context *ctx, uint32_t on);
/* The closest hit of every primary ray of frame `sample` (the commented-out debug view of
 * pathtracing.wgsl:262 over intersect_TLAS, raytracing_sw.wgsl:176): out[5*p..] =
 * {instance, i0, i1, i2, float bits of dist}, instance = 0xFFFFFFFF on a miss. */
nb_status nb_primary_hits(nb_context *ctx, const nb_scene *scene, const nb_camera *camera,
                          uint32_t width, uint32_t height, uint32_t sample, uint32_t *out);
/* nb_primary_hits without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (allocation-free interactive picking loops, code
 * that must not call hipMalloc).  The kernel writes into a fixed library-owned scratch array on
 * the device (40 MiB per GPU, held until the process ends), so width * height is at most
 * 2097152 pixels (larger sizes fail with NB_ERR_INVALID_ARGUMENT).  Every call on one GPU uses
 * that GPU's one scratch array: calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialized by the caller.  Same output as nb_primary_hits. */
nb_status nb_primary_hits_noalloc(nb_context *ctx, const nb_scene *scene, const nb_camera *camera,
                                  uint32_t width, uint32_t height, uint32_t sample, uint32_t *out);

/* ---- serving: continuous batching over resident scenes -------------------------------------
 * No counterpart in the reference (its frame loop, app.rs:177-219, renders one scene for one
 * window).  A server keeps several scenes resident, admits render requests while earlier ones
 * are on the GPU (as request slots and HBM budget free up), and renders every admitted request's
 * samples together in steps: one captured HIP graph per step (one persistent kernel with wave64
 * ray regeneration over every admitted request's paths of the step, then the ordered
 * accumulation).  Each
 * request's image equals what nb_render_frames produces for it alone (same camera, size,
 * samples 1..spp, bounces, contribution factor, the context's Sobol table).  A server is used
 * from one host thread at a time. */
typedef struct nb_server nb_server;

typedef struct {
    nb_camera camera;
    uint32_t scene;                 /* index from nb_server_add_scene */
    uint32_t width, height;         /* nonzero multiples of 8 */
    uint32_t spp;                   /* samples 1..spp */
    uint32_t bounces;
    float contribution_factor;
    double arrival_s;               /* server clock; negative = the time of submission */
    uint32_t mode;                  /* NB_REQUEST_PLAIN, or NB_REQUEST_CACHED (the scene needs a cache) */
    uint32_t train_paths;           /* cached: training paths traced per sample (0 = inference only) */
    float lr;                       /* cached: Adam learning rate of that training */
    uint32_t pad;
} nb_request;

#define NB_REQUEST_PLAIN 0u
#define NB_REQUEST_CACHED 1u

typedef struct { uint64_t id; double arrival_s, admit_s, done_s; } nb_completion;

typedef struct {
    uint64_t steps, rays, samples, n_aabb, n_tri, n_inst, hits;   /* plain requests' kernels */
    uint64_t bytes_in_use, bytes_budget;
    uint32_t active, queued, completed, pad;
    double gpu_busy_s;              /* summed device time of the steps */
    uint64_t cached_samples;        /* samples rendered for cached requests */
    double cached_busy_s;           /* their device time (part of gpu_busy_s) */
} nb_server_stats;

nb_status nb_server_create(nb_context *ctx, uint64_t hbm_budget_bytes, uint32_t max_paths_per_step,
                           uint32_t max_bounces, nb_server **out);
nb_status nb_server_destroy(nb_server *srv);
nb_status nb_server_add_scene(nb_server *srv, const nb_scene *scene, const nb_envmap *env, uint32_t *index);
/* Gives resident scene `scene` a learned radiance cache (created on the server's context, BF16 or
 * FP16): NB_REQUEST_CACHED requests for the scene render with nb_render_cached's method on it and
 * train it online.  The cache is shared by those requests and kept for the server's life. */
nb_status nb_server_set_cache(nb_server *srv, uint32_t scene, nb_cache *cache);
nb_status nb_server_submit(nb_server *srv, const nb_request *req, uint64_t *id);
/* Admits what fits, renders one step of the admitted requests and waits for it. */
nb_status nb_server_step(nb_server *srv, uint32_t *n_active, uint32_t *n_queued);
/* Completed requests since the last poll (up to cap). */
nb_status nb_server_poll(nb_server *srv, nb_completion *out, uint32_t cap, uint32_t *n);
nb_status nb_server_read(nb_server *srv, uint64_t id, float *rgba);
/* Frees a completed request's image (its HBM returns to the admission budget). */
nb_status nb_server_release(nb_server *srv, uint64_t id);
nb_status nb_server_now(nb_server *srv, double *seconds);
nb_status nb_server_stats_get(nb_server *srv, nb_server_stats *out);

/* ---- learned radiance cache ------------------------------------------------------------------
 * No counterpart in the reference (its README plans Neural Radiance Caching, README.md:12,23).
 * An MLP (96 encoded inputs -> 5 x 64 ReLU -> 3) maps a shading point to outgoing radiance; BF16
 * weights and activations on the matrix cores, FP32 accumulation and FP32 master parameters,
 * Adam on a relative L2 loss.  Raw records are 16 floats: position xyz, outgoing direction xyz,
 * normal xyz, roughness, albedo rgb, metallic, 0, 0.  Parameters are one flat fp32 vector:
 * W0 (64x96), b0, W1..W4 (64x64) with biases, W5 (3x64), b5; W is (out x in) row-major. */
uint32_t nb_cache_param_count(void);
/* BF16 operands (no loss scaling). */
nb_status nb_cache_create(nb_context *ctx, const float scene_min[3], const float scene_max[3], uint32_t seed,
                          uint32_t max_batch, nb_cache **out);
/* flags: NB_CACHE_FP16 selects FP16 operands, trained with a dynamic loss scale kept on the device
 * (starts at 1024; a step whose gradient is not finite is skipped and halves it; 2000 clean steps
 * double it, up to 65536). */
#define NB_CACHE_FP16 1u
nb_status nb_cache_create_ex(nb_context *ctx, const float scene_min[3], const float scene_max[3], uint32_t seed,
                             uint32_t max_batch, uint32_t flags, nb_cache **out);
/* The current loss scale (1 for a BF16 cache); waits for the cache's pending work. */
nb_status nb_cache_loss_scale(nb_cache *cache, float *scale);
nb_status nb_cache_destroy(nb_cache *cache);
nb_status nb_cache_get_params(nb_cache *cache, float *out);
nb_status nb_cache_set_params(nb_cache *cache, const float *in);
/* Host arrays in and out (used by tests and offline tools): the 16-bit input encoding (bf16 or
 * fp16 bits), the network output, the loss and its gradient (no update; divided by the loss
 * scale), and Adam steps over minibatches. */
nb_status nb_cache_encode(nb_cache *cache, const float *raw, uint32_t n, uint16_t *encoded);
nb_status nb_cache_infer(nb_cache *cache, const float *raw, uint32_t n, float *rgb);
nb_status nb_cache_gradient(nb_cache *cache, const float *raw, const float *target, uint32_t n, float *loss,
                            float *grad);
nb_status nb_cache_train(nb_cache *cache, const float *raw, const float *target, uint32_t n, uint32_t batch,
                         float lr, float *last_loss);
/* One cache-assisted frame: one sample per pixel whose path ends at its second shading point
 * (later after near-mirror bounces, at bounce 3 at the latest) with the radiance the cache
 * predicts there, accumulated into the target with globals->weight as nb_render does.  When
 * train_paths > 0, that many full paths of the plain integrator are traced first at a spread of
 * pixels; their shading points (up to three per path) with the path's suffix radiance become a
 * training set for one Adam step over all of them (at most the cache's max_batch records; the
 * environment variable NB_CACHE_STEP_ROWS=n splits them into steps of n) with learning rate lr.
 * *loss (optional) gets the last step's loss.  The cache must have been created on the same
 * context. */
nb_status nb_render_cached(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                           const nb_globals *globals, const nb_envmap *env, nb_cache *cache, uint32_t train_paths,
                           float lr, float *loss);
/* The training records nb_render_cached would train on for frame globals->sample of a width x
 * height view (for inspection and tests; the cache's parameters are not changed).  Training path
 * i of n traces pixel (i * max(floor(W*H / n), 1) + (sample * 2654435761 mod W*H)) mod W*H with Sobol
 * row 1 + (sample + 511) mod 1023; each of its first three non-emissive shading points gives a
 * record: 16 raw floats (as for nb_cache_infer) and the target radiance = the path's radiance /
 * the throughput that arrived there.  Records come in no particular order; *n is the number
 * traced (at most the cache's max_batch), of which min(*n, cap) are copied. */
nb_status nb_cache_training_records(nb_context *ctx, const nb_scene *scene, uint32_t width, uint32_t height,
                                    const nb_camera *camera, const nb_globals *globals, const nb_envmap *env,
                                    nb_cache *cache, uint32_t train_paths, float *raw, float *target, uint32_t cap,
                                    uint32_t *n);
/* The cache queries nb_render_cached makes for frame globals->sample of a width x height view,
 * before any network runs (for inspection and tests; nothing is rendered or trained).  Each
 * pixel's path runs to its query point: the first shading point reached at bounce >= 1 after a
 * lobe whose GGX alpha is at least 0.05 (the diffuse lobe counts as 1), or at bounce 3.
 * direct[4 * pixel] = the radiance met before it (sky or emitter; 0 when queried), alpha 1.  For
 * each pixel that reached its point, one record: 16 raw floats (as for nb_cache_infer) and
 * thr[4 * i] = (throughput arriving there, pixel index as float bits).  Records come in no
 * particular order; *n is the number of queries, of which min(*n, cap) are copied. */
nb_status nb_cache_query_records(nb_context *ctx, const nb_scene *scene, uint32_t width, uint32_t height,
                                 const nb_camera *camera, const nb_globals *globals, const nb_envmap *env,
                                 nb_cache *cache, float *direct, float *raw, float *thr, uint32_t cap, uint32_t *n);

#ifdef __cplusplus
}
#endif
#endif
