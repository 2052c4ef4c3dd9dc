// Internal declarations shared by the host files of libnbounce and the HIP kernels.
#pragma once

#include "nbounce.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nb {

// Records the thread-local error message and returns `st`.
nb_status set_error(nb_status st, const char *fmt, ...);

nb_status build_scene_capacity(uint32_t n_indices, const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                               uint32_t *blas_cap, uint32_t *tlas_cap);
nb_status build_scene(const nb_vertex *vertices, uint32_t n_vertices, uint32_t *indices, uint32_t n_indices,
                      const nb_primitive *prims, uint32_t n_prims, uint32_t flags,
                      nb_bvh_node *blas, uint32_t blas_cap, uint32_t *n_blas,
                      nb_bvh_node *tlas, uint32_t tlas_cap, uint32_t *n_tlas, nb_instance *instances);
void camera_look_at(const float *eye, const float *target, const float *up, float fov_y, float aspect,
                    float z_near, nb_camera *out);
nb_status sobol_generate(uint32_t rows, uint32_t dims, uint32_t seed, float *out);

// ---- device layout (HBM) --------------------------------------------------------------------
//
// Node records: one 64-B child pair per inner BVH node holding both children's boxes and
// references, so a traversal step is one record fetch (four dwordx4 loads).  Root records hold
// the root box as the left child and an all-NaN right box (the slab test misses NaN boxes), so
// entering a tree is the same step as any other.  BLAS and TLAS records share one array, each
// tree in depth-first preorder after its root record.  (Records that also carried both
// children's pairs, descending two levels per fetch, measured slower.)
//   q0 = (L.min.xyz, L.max.x)  q1 = (L.max.yz, R.min.xy)  q2 = (R.min.z, R.max.xyz)
//   q3 = (L.ref, R.ref, is_root, 0)  as bits
// A child reference is a record id (bit 31 clear) or a leaf: bit 31 set, primitive count in
// bits 26..30 (0 = more than 31: the count sits in the leaf-count side table at the leaf's
// start), first primitive in bits 0..25.
constexpr uint32_t kRecordFloat4 = 4;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafStartBits = 26;
constexpr uint32_t kLeafStartMask = (1u << kLeafStartBits) - 1u;
constexpr uint32_t kLeafMaxInline = 31u;
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;
constexpr uint32_t kMaxTreeDepth = 32;  // reference MAX_DEPTH (bvh.rs:153)
constexpr uint32_t kStackEntries = 2 * kMaxTreeDepth;
// The largest bounce limit a render accepts: the kernels keep it in 16 bits beside the first
// roulette bounce (PathConst::bounces).
constexpr uint32_t kMaxBounceLimit = 0xFFFFu;

// Device counters: kCounterSlots slots (one per XCD), kCounterStride u64 apart (4 KiB, so the
// slots' atomics do not share a cache line or memory channel), each of kCounterWords u64:
// rays, n_aabb, n_tri, n_inst, samples, hits, 0, 0.  Readers sum the slots.
constexpr uint32_t kCounterSlots = 8, kCounterWords = 8, kCounterStride = 512;
constexpr size_t kCounterBytes = kCounterSlots * kCounterStride * sizeof(unsigned long long);

// Per-XCD work counters of the persistent render kernel, kWorkStride words apart: device-scope
// atomics on one cache line (or one memory channel) serialise, so each counter gets its own
// 4 KiB block.
constexpr uint32_t kWorkStride = 1024;
constexpr size_t kWorkBytes = 8 * kWorkStride * sizeof(uint32_t);

// Triangles in BVH (permuted) order, 48 B: (v0.xyz, i0) (e1.xyz, i1) (e2.xyz, i2) with
// e1 = v1 - v0, e2 = v2 - v0 computed in binary32 exactly as the kernel would.
// Instance traversal record, 64 B: world_to_local rows 0..2 as (m[0][r], m[1][r], m[2][r], m[3][r])
// and (root record id, 0, 0, 0).
// Instance shading record, 96 B: (W2L col0.xyz, roughness) (W2L col1.xyz, metallic)
// (W2L col2.xyz, emissive) (L2W col0.xyz, color.r) (L2W col1.xyz, color.g) (L2W col2.xyz, color.b).
// Vertex shading record, 32 B: (normal.xyz, 0) (tangent.xyzw).
struct DevScene {
    const float4 *pairs;
    const float4 *tris;
    const float4 *inst;
    const float4 *ishade;
    const float4 *vshade;
    const uint32_t *tri_leaf_count;
    const uint32_t *inst_leaf_count;
    uint32_t tlas_root;  // record id, or kNoRecord for an empty TLAS
    uint32_t n_pairs;
};

struct DevEnv {
    // cube (n > 0): 6 faces of (n+2)^2 texels, one texel of seamless border around each face;
    // lat-long (n == 0): a header texel holding (width, height) as integers, then height+2 rows
    // of width+2 texels, the border columns wrapped from the opposite side and the border rows
    // repeating the edge rows.  (The size sits in memory, read only on a lookup, so that the
    // kernels do not hold it in registers.)
    const float4 *texels;
    uint32_t n;
};

struct RenderParams {
    DevScene scene;
    DevEnv env;
    float4 *target;
    uint32_t width, height;
    float c2w[16];
    const float4 *sobol;
    uint32_t sobol_entries, sobol_stride;
    uint32_t first_sample, n_frames;
    float first_weight;      // the caller's weight for the first frame (nb_render)
    uint32_t weights_from_sample;  // 1: every frame uses 1/sample (nb_render_frames)
    uint32_t bounces;
    float contribution_factor;
    unsigned long long *counters;  // rays, n_aabb, n_tri, n_inst, samples
    uint32_t tiles_x, n_tiles, tiles_per_xcd;
    uint32_t *hits_out;      // primary-hit debug output (5 words per pixel) or null
    uint32_t *work;          // 8 per-XCD work-unit counters (kWorkStride apart), zeroed per launch
    uint32_t regen_min_idle; // the persistent kernel refills idle lanes once at least this many are idle
    uint32_t inner_quorum;   // inner-node steps run alone while more lanes than this have one
    uint32_t tail_units;     // single work units are handed out once a band has at most this many
    uint32_t grab_units;     // work units per grab before that
    const uint32_t *unit_order;  // position -> work unit, a permutation within each XCD band, or null
    uint32_t *unit_cost;     // rays traced per work unit, added as pixels finish, or null
    uint32_t rr_first;       // first bounce with Russian roulette (the reference: 0)
    float4 *sample_out;      // or null; else (n_frames 1) each pixel's sample goes here, the target
                             // is neither read nor written
};

// ---- serving (serve.hip, server.cpp) ----------------------------------------------------------
struct ServeSceneDev {
    DevScene scene;
    DevEnv env;
};

// One admitted request, indexed by its slot.
struct ServeRequestDev {
    float c2w[16];
    float4 *target;          // RGBA32F output, zeroed at admission
    uint32_t width, height;
    uint32_t scene;          // scene slot
    uint32_t bounces;
    float contribution_factor;
    uint32_t rr_first;       // first bounce with Russian roulette (the reference: 0)
    uint32_t pad[2];
};

// The samples one request renders in one step: samples first_sample .. first_sample + count - 1
// of every pixel.  Its paths (one per sample and pixel) have ids [path_base, path_base + count *
// pixels), sample-major; its pixels occupy [pixel_base, pixel_base + pixels) of the step's
// accumulation range.  Its paths form 16-path work units (one sample of two rows of an 8x8 tile,
// as in the progressive renderer), count * pixels / 16 of them from unit_base in the step's list.
struct ServePass {
    uint32_t slot, first_sample, count, pixels;
    uint32_t path_base, pixel_base, unit_base, pad;
};

// The passes of one scene in a step: passes [pass_begin, pass_end), units [unit_begin, unit_end).
struct ServeGroup {
    uint32_t scene;
    uint32_t unit_begin, unit_end, pass_begin, pass_end;
    uint32_t pad[3];
};

// Per-step description, at a fixed device address so that one captured graph serves every step.
struct ServeStep {
    const ServeSceneDev *scenes;
    const ServeRequestDev *requests;
    const ServePass *passes;
    const ServeGroup *groups;
    uint32_t n_passes, n_groups, n_units, n_pixels;
    const float4 *sobol;
    uint32_t sobol_entries, sobol_stride;
    float4 *results;         // radiance per path id
    uint32_t *work;          // per group: units taken (kWorkStride apart), zeroed by the step's reset
    unsigned long long *counters;
    uint32_t regen_min_idle, inner_quorum, grab_units, pad1;
};

constexpr uint32_t kServeMaxBounces = 32;
constexpr uint32_t kServeMaxGroups = 64;   // one per resident scene at most

nb_status serve_launch_reset(const ServeStep *step, hipStream_t s);
// waves_per_simd: 5 for the server's steps, 4 for the sample-parallel nb_render_frames (the
// occupancy each measured best at).
nb_status serve_launch_step(const ServeStep *step, hipStream_t s, uint32_t waves_per_simd = 5);
nb_status serve_launch_accumulate(const ServeStep *step, uint32_t grid, hipStream_t s);

// Kernel launchers (trace.hip).
nb_status launch_render(const RenderParams &p, void *stream);
nb_status launch_primary_hits(const RenderParams &p, void *stream);
// target[i] = mix(target[i] (0 when weight is not positive), samples[i], weight) for n pixels: the
// step of the render kernels' accumulation, as a kernel of its own for frames whose paths were
// rendered into a sample buffer (RenderParams::sample_out).
nb_status launch_mix_samples(float4 *target, const float4 *samples, uint32_t n, float weight, hipStream_t s);

}  // namespace nb

// ---- opaque handles of the C ABI ---------------------------------------------------------------
struct nb_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_region_start = nullptr, ev_region_stop = nullptr;
    bool timed = false, region_started = false;
    bool render_timing = false;               // nb_render_timing: per-render events
    float4 *sobol = nullptr;
    uint32_t sobol_entries = 0, sobol_stride = 17;
    unsigned long long *counters = nullptr;   // kCounterSlots x kCounterWords
    bool counting = true;                     // nb_counters_enable
    uint32_t rr_first = 0;                    // nb_render_options_set: first bounce with roulette
    uint32_t *work = nullptr;       // tile counters of the persistent render kernel
    uint32_t *hits_scratch = nullptr;   // device address of nb_primary_hits_noalloc's scratch (not owned)
    // Sample-parallel nb_render_frames (small targets): the serving kernels' step description
    // (one request, one pass, one group) and per-path results, allocated at first use.
    void *sp_block = nullptr;
    float4 *sp_results = nullptr;
    uint32_t sp_max_paths = 0;
    // Pipelined nb_render (abi.cpp): each frame's path kernel runs on one of two side streams and
    // writes the frame's samples; the mix into the target runs on `stream`.  Created at first use.
    hipStream_t pipe_stream[2] = {nullptr, nullptr};
    uint32_t *pipe_work[2] = {nullptr, nullptr};   // work counters of each side stream's launches
    hipEvent_t pipe_gate = nullptr;   // `stream` as it was before the last mix was enqueued
    hipEvent_t pipe_done = nullptr;   // end of the last path kernel
    uint64_t pipe_seq = 0;            // pipelined frames so far (the side stream alternates)
    bool pipe_gate_valid = false;
    // Set when work on `stream` since the last mix reset what a path kernel adds to (the
    // counters): the next path kernel then waits for all of `stream`.
    bool pipe_fence = false;
    // Work-unit order sorts enqueued on `stream` so far (order.hip), and how many of them the
    // recorded pipe_gate follows: a pipelined path kernel may read an order only once its gate
    // follows the sort that wrote it.
    uint64_t order_sorts = 0;
    uint64_t pipe_gate_sorts = 0;
    // Scenes, targets and environments keep their context alive: nb_context_destroy only marks
    // it, and the last child's destroy releases it.
    uint32_t children = 0;
    bool closing = false;
};

struct nb_scene {
    nb_context *ctx = nullptr;
    void *block = nullptr;
    uint64_t bytes = 0;
    nb::DevScene dev{};
};

struct nb_target {
    nb_context *ctx = nullptr;
    float4 *pixels = nullptr;
    uint32_t width = 0, height = 0;
    // Cost-ordered work units of the persistent kernel (order.hip): rays traced per 16-pixel unit
    // over the frames since the costs were reset, and the order built from them.
    uint32_t n_units = 0;
    uint32_t *unit_cost = nullptr, *sort_vals = nullptr;
    // Two order buffers: a rebuild writes the one no launch in flight may read, and the target
    // switches to it (order_cur) once the launches that read orders can see the sort.
    uint32_t *unit_order[2] = {nullptr, nullptr};
    uint32_t order_cur = 0;
    bool order_pending = false;     // unit_order[order_cur ^ 1] holds a newer order
    uint64_t pending_sort = 0;      // the context's sort number that wrote it
    bool rebuild_wanted = false;    // a rebuild fell due while one was pending
    uint8_t *sort_keys = nullptr;   // 2 x n_units
    int *sort_offsets = nullptr;    // 9 band boundaries
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint32_t cost_frames = 0;       // frames accumulated in unit_cost (halved with the costs)
    uint32_t view_frames = 0;       // frames rendered since the camera last changed
    bool order_valid = false;
    const nb_scene *cost_scene = nullptr;   // the view the costs were gathered with
    nb_camera cost_camera{};
    uint32_t cost_bounces = 0;
    float cost_contribution = 0.0f;
    // Pipelined nb_render: the two frames' sample buffers (RGBA32F per pixel), at first use.
    float4 *samples[2] = {nullptr, nullptr};
};

namespace nb {
// Work-unit order of the persistent kernel (order.hip).  unit_order_alloc / unit_order_free pair
// with the target's lifetime.  unit_order_prepare, before a launch on `t`, resets the costs when
// the scene, bounce limit or contribution factor differs from the ones they were gathered with
// (halves them when only the camera differs), and points p at the order and the cost array (both
// null with NB_UNIT_ORDER=0).  unit_order_update, after the launch of n frames, rebuilds the order
// from the costs once the view has held for 2, 4, 8, ... frames.  The order changes only which
// wave renders a pixel when, never a pixel's samples.
// Both run on the context stream `s` and may overlap a pipelined path kernel of the same target:
// the costs are halved and read with atomics (a path kernel adds to them as its pixels finish),
// and a rebuild writes the order buffer no launch in flight reads.  `visible_sorts` is the number
// of the context's sorts the coming launch follows (all of them for a launch on `s`, the gate's
// count for a pipelined one); the rebuilt order is used from the first launch that follows its
// sort.
nb_status unit_order_alloc(nb_target *t);
void unit_order_free(nb_target *t);
nb_status unit_order_prepare(nb_target *t, const nb_scene *scene, const nb_camera *cam, const nb_globals *g,
                             RenderParams &p, hipStream_t s, uint64_t visible_sorts);
nb_status unit_order_update(nb_target *t, const RenderParams &p, uint32_t n_frames, hipStream_t s);
}  // namespace nb

struct nb_envmap {
    nb_context *ctx = nullptr;
    float4 *texels = nullptr;
    uint32_t n = 0;                     // cube face size; 0 for a lat-long map
    uint32_t width = 0, height = 0;     // lat-long size
};

inline nb::DevEnv dev_env(const nb_envmap *e) {
    return e ? nb::DevEnv{e->texels, e->n} : nb::DevEnv{nullptr, 0u};
}

// Device bytes of an environment map's padded texels.
inline uint64_t envmap_device_bytes(const nb_envmap *e) {
    return e->n ? 6ull * (e->n + 2) * (e->n + 2) * 16 : (static_cast<uint64_t>(e->width + 2) * (e->height + 2) + 1) * 16;
}


namespace nb {
// A scene, target, environment or server created on a context keeps it alive.
void context_add_child(nb_context *ctx);
void context_release_child(nb_context *ctx);
}  // namespace nb

// ---- learned radiance cache (mlp.hip, cache.cpp) ------------------------------------------------
namespace nb {
// Network: 96 encoded inputs -> 5 hidden layers of 64 (ReLU) -> 3 outputs.  Weights follow the
// (out x in) row-major convention; the last layer is padded to 16 output rows with zeros.
constexpr int kMlpIn = 96, kMlpHidden = 64, kMlpOut = 3, kMlpLayers = 6, kMlpOutPad = 16;
constexpr int kMlpRaw = 16;     // raw record: pos3, dir3, normal3, roughness, albedo3, metallic, 0, 0
constexpr int kMlpWeights = kMlpHidden * kMlpIn + 4 * kMlpHidden * kMlpHidden + kMlpOut * kMlpHidden;
constexpr int kMlpBiases = 5 * kMlpHidden + kMlpOut;
constexpr int kMlpParams = kMlpWeights + kMlpBiases;

struct MlpDev {
    const uint16_t *w[kMlpLayers];   // 16-bit (bf16, or fp16 when half), (out x in) row-major; layer 5 padded to 16 rows
    const float *b[kMlpLayers];      // fp32 biases (layer 5 padded to 16)
    float scene_min[3], scene_inv_extent[3];
    uint32_t half;                   // operands are fp16 (else bf16)
};

// One training step's rows (16-bit, row-major): encoded inputs, hidden activations, the output
// gradient (padded to 16 channels) and the backward pass's data gradients dz_L (of layer L's
// pre-activation); plus the targets, the loss sum and the fp16 loss scale (null for bf16).
struct MlpTrainRows {
    const float *target;          // n x 3
    uint16_t *x;                  // n x 96
    uint16_t *h[5];               // n x 64 each
    uint16_t *dy;                 // n x 16
    uint16_t *dz[5];              // n x 64 each
    float *loss;
    const float *loss_scale;
};

// The weight-gradient jobs of all layers: layer L's (out_p / 16) x (k_in / 16) tiles times the
// row chunks are jobs job_begin[L] .. job_begin[L + 1] - 1.
struct MlpGradJobs {
    const uint16_t *dz[kMlpLayers];       // layer L's output gradient rows (dy for the last layer)
    const uint16_t *h_prev[kMlpLayers];   // layer L's input rows (x for the first layer)
    float *dw[kMlpLayers], *db[kMlpLayers];
    uint32_t out_p[kMlpLayers], out_valid[kMlpLayers], k_in[kMlpLayers];
    uint32_t job_begin[kMlpLayers + 1];
    uint32_t rows_per_chunk;
};

// Where each layer's weights sit in the fp32 parameters and in the 16-bit copy, and the last
// layer's bias (copied to the padded bias).
struct MlpRefresh {
    uint32_t w_off[kMlpLayers], w_size[kMlpLayers], wb_off[kMlpLayers], b5_off;
};

// One frame of cache-assisted rendering plus the training paths traced beside it.
struct CacheFrame {
    DevScene scene;
    DevEnv env;
    const float4 *sobol;
    uint32_t sobol_entries, sobol_stride;
    float c2w[16];
    uint32_t width, height;
    uint32_t sample;
    float weight;
    uint32_t bounces;
    float contribution_factor;
    float mirror_alpha;          // lobes with GGX alpha below this do not end the path in the cache
    uint32_t max_cache_bounce;   // the path ends in the cache at this bounce at the latest
    float4 *target, *direct;
    float *q_raw, *q_thr;        // queries: raw records, (throughput rgb, pixel bits)
    uint32_t *q_count;
    uint32_t train_paths, train_offset, train_sample, t_capacity;
    float *t_raw, *t_target;
    uint32_t *t_count;
    unsigned long long *counters;
    uint32_t tiles_x, n_tiles, tiles_per_xcd;
};

// A cache-assisted frame in two parts: the path tracing up to the queries (independent of the
// network), then inference, combination and accumulation.
nb_status cache_launch_query(const CacheFrame &p, hipStream_t s);
nb_status cache_launch_finish(const CacheFrame &p, const MlpDev &m, float *rgb, hipStream_t s);
nb_status cache_render_frame(nb_context *ctx, const DevScene &scene, const DevEnv &env, float4 *target, uint32_t width,
                             uint32_t height, const float *c2w, const nb_globals &g, nb_cache *c, uint32_t train_paths,
                             float lr, float *loss);
nb_context *cache_context(const nb_cache *c);
// Device bytes a cache holds (parameters, Adam state, batch buffers, frame workspace), the
// pixels its frame workspace currently covers, and the workspace bytes a frame of n pixels needs.
uint64_t cache_device_bytes(const nb_cache *c);
uint64_t cache_workspace_pixels(const nb_cache *c);
uint64_t cache_workspace_bytes(uint64_t pixels);
nb_status cache_launch_train_paths(const CacheFrame &p, hipStream_t s);
nb_status mlp_launch_infer_count(const MlpDev &m, const float *raw, const uint32_t *n_dev, uint32_t n_max, float *rgb,
                                 hipStream_t s);
nb_status mlp_launch_encode(const MlpDev &m, const float *raw, uint32_t n, uint16_t *x, hipStream_t s);
nb_status mlp_launch_infer(const MlpDev &m, const float *raw, uint32_t n, float *rgb, hipStream_t s);
nb_status mlp_launch_to16(bool half, const float *src, uint16_t *dst, uint32_t n, hipStream_t s);
// One training step in four launches: forward with loss (stores tr's rows), backward through the
// data path, all weight gradients (added to j's dw / db), Adam with the 16-bit refresh.
nb_status mlp_launch_train_forward(const MlpDev &m, const float *raw, uint32_t n, const MlpTrainRows &tr, hipStream_t s);
nb_status mlp_launch_backward(const MlpDev &m, uint32_t n, const MlpTrainRows &tr, hipStream_t s);
nb_status mlp_launch_weight_grad(bool half, const MlpGradJobs &j, uint32_t n, hipStream_t s);
nb_status mlp_launch_adam_refresh(bool half, float *w, float *m, float *v, const float *g, uint32_t n, float lr, float b1,
                                  float b2, float eps, float c1, float c2, uint32_t *scale_state, const MlpRefresh &r,
                                  uint16_t *wbf, float *b5, hipStream_t s);
nb_status mlp_launch_fill(float *p, uint32_t n, float v, hipStream_t s);
}  // namespace nb
