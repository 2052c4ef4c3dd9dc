// Learned radiance cache: parameters, Adam state, training and inference buffers of the MLP in
// mlp.hip, and the nb_cache_* entry points.  The MLP maps a shading point (position, outgoing
// direction, normal, roughness, albedo, metallic) to outgoing radiance; it is trained online with
// Adam on a relative L2 loss against radiance estimates from the tracer's own path suffixes.
#include "nbounce.h"
#include "nb_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace nb;

namespace {

#define C_HIP(call)                                                                                    \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_error(e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "%s: %s", \
                             #call, hipGetErrorString(e_));                                            \
    } while (0)
#define C_TRY(expr)                    \
    do {                               \
        nb_status s_ = (expr);         \
        if (s_ != NB_OK) return s_;    \
    } while (0)

// Offsets of each layer's weights and biases in the flat fp32 parameter vector, and of each
// layer's weights in the bf16 copy (the last layer padded to 16 rows).
struct Layout {
    uint32_t w[kMlpLayers], b[kMlpLayers], wb[kMlpLayers], in[kMlpLayers], out[kMlpLayers];
    Layout() {
        uint32_t off = 0, offb = 0;
        for (int L = 0; L < kMlpLayers; ++L) {
            in[L] = L == 0 ? kMlpIn : kMlpHidden;
            out[L] = L == kMlpLayers - 1 ? kMlpOut : kMlpHidden;
            w[L] = off; off += in[L] * out[L];
            b[L] = off; off += out[L];
            wb[L] = offb; offb += in[L] * (L == kMlpLayers - 1 ? kMlpOutPad : out[L]);
        }
    }
};
const Layout kL;
constexpr uint32_t kBf16Weights = kMlpHidden * kMlpIn + 4 * kMlpHidden * kMlpHidden + kMlpOutPad * kMlpHidden;

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct nb_cache {
    nb_context *ctx = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                    // training paths beside a frame's query tracing
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint32_t *h_count = nullptr;                   // pinned: the training records of a frame
    uint32_t max_batch = 0;
    float scene_min[3], scene_inv[3];
    float *params = nullptr, *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    bool half = false;           // fp16 operands with a dynamic loss scale (else bf16)
    uint32_t *scale_state = nullptr;   // fp16: [0] loss scale bits, [1] clean steps, [2] overflow flag
    uint16_t *wbf = nullptr;     // 16-bit weights, last layer padded
    float *b5 = nullptr;         // last-layer bias padded to 16
    uint32_t adam_t = 0;
    // frame workspace (grows with the target)
    size_t ws_pixels = 0;
    float4 *direct = nullptr;
    float *q_raw = nullptr, *q_thr = nullptr, *q_rgb = nullptr;
    uint32_t *counts = nullptr;    // [0] queries, [1] training records
    // batch buffers
    float *raw = nullptr, *target = nullptr, *rgb = nullptr, *loss = nullptr;
    uint16_t *x = nullptr, *h[5] = {}, *dz[5] = {}, *dy = nullptr;

    MlpDev dev() const {
        MlpDev m;
        for (int L = 0; L < kMlpLayers; ++L) {
            m.w[L] = wbf + kL.wb[L];
            m.b[L] = L == kMlpLayers - 1 ? b5 : params + kL.b[L];
        }
        for (int d = 0; d < 3; ++d) { m.scene_min[d] = scene_min[d]; m.scene_inv_extent[d] = scene_inv[d]; }
        m.half = half ? 1u : 0u;
        return m;
    }
};

namespace nb {

// 16-bit weight copies and the padded last-layer bias from the fp32 parameters.
nb_status cache_refresh(nb_cache *c) {
    for (int L = 0; L < kMlpLayers; ++L)
        C_TRY(mlp_launch_to16(c->half, c->params + kL.w[L], c->wbf + kL.wb[L], kL.in[L] * kL.out[L], c->stream));
    C_HIP(hipMemcpyAsync(c->b5, c->params + kL.b[kMlpLayers - 1], kMlpOut * sizeof(float), hipMemcpyDeviceToDevice,
                         c->stream));
    return NB_OK;
}

// Rows per weight-gradient job (a multiple of 16): NB_WGRAD_ROWS overrides the default, read once.
// Over the cache bench's training steps (16,384 rows) the kernel took 18.7 / 16.5 / 19.8 / 33.7 ms
// in all with 128 / 256 / 512 / 1024 rows per job.
uint32_t wgrad_rows() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_WGRAD_ROWS");
        const long r = e ? std::strtol(e, nullptr, 10) : 256;
        return static_cast<uint32_t>(r < 16 ? 16 : (r > 4096 ? 4096 : r / 16 * 16));
    }();
    return v;
}

// Rows per Adam step of a cache-assisted frame's training, 0 = all of the frame's records in
// one step (the default); NB_CACHE_STEP_ROWS=n sets n (at least 256), read once.  On the
// learned-cache workload (1080p spheres orbit, 2^16 training paths, up to 196,608 records per
// frame) one step per frame fits better and costs less than steps of 16,384 rows: rel-MSE at
// 1 spp 0.885 against 0.913, equal-error frame-time ratio 0.238-0.242 against 0.244-0.255, and
// 0.349-0.356 against 0.409-0.463 with the training (98,304 rows: 0.894, 0.238-0.245, 0.357-0.362).
uint32_t train_step_rows() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_CACHE_STEP_ROWS");
        const long r = e ? std::strtol(e, nullptr, 10) : 0;
        return r <= 0 ? 0u : static_cast<uint32_t>(r < 256 ? 256 : (r > (1l << 24) ? (1l << 24) : r / 16 * 16));
    }();
    return v;
}

// Forward and backward over the batch in c->raw / c->target (n rows); gradients (times the loss
// scale of an fp16 cache) into c->grad, mean loss into c->loss.  No parameter update.
nb_status cache_forward_backward(nb_cache *c, const float *raw, const float *target, uint32_t n) {
    const MlpDev m = c->dev();
    hipStream_t s = c->stream;
    MlpTrainRows tr{};
    tr.target = target;
    tr.x = c->x;
    tr.dy = c->dy;
    for (int L = 0; L < 5; ++L) { tr.h[L] = c->h[L]; tr.dz[L] = c->dz[L]; }
    tr.loss = c->loss;
    tr.loss_scale = c->half ? reinterpret_cast<const float *>(c->scale_state) : nullptr;
    C_TRY(mlp_launch_fill(c->grad, kMlpParams + 1, 0.0f, s));   // the gradient and the loss after it
    C_TRY(mlp_launch_train_forward(m, raw, n, tr, s));
    C_TRY(mlp_launch_backward(m, n, tr, s));
    // layer L's weight gradient: its output gradient rows (dy for the last layer) against its
    // input rows (x for the first), in jobs of 256 rows so that a batch of a few thousand rows
    // still fills the GPU
    MlpGradJobs j{};
    j.rows_per_chunk = wgrad_rows();
    const uint32_t chunks = (n + j.rows_per_chunk - 1) / j.rows_per_chunk;
    uint32_t jobs = 0;
    for (int L = 0; L < kMlpLayers; ++L) {
        const bool last = L == kMlpLayers - 1;
        j.dz[L] = last ? c->dy : c->dz[L];
        j.h_prev[L] = L == 0 ? c->x : c->h[L - 1];
        j.dw[L] = c->grad + kL.w[L];
        j.db[L] = c->grad + kL.b[L];
        j.out_p[L] = last ? kMlpOutPad : kMlpHidden;
        j.out_valid[L] = kL.out[L];
        j.k_in[L] = kL.in[L];
        j.job_begin[L] = jobs;
        jobs += (j.out_p[L] / 16) * (j.k_in[L] / 16) * chunks;
    }
    j.job_begin[kMlpLayers] = jobs;
    return mlp_launch_weight_grad(c->half, j, n, s);
}

nb_status cache_adam_step(nb_cache *c, float lr) {
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;   // Adam with the usual momentum constants
    ++c->adam_t;
    const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, c->adam_t)));
    const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(0.99, c->adam_t)));
    MlpRefresh r{};
    for (int L = 0; L < kMlpLayers; ++L) {
        r.w_off[L] = kL.w[L];
        r.w_size[L] = kL.in[L] * kL.out[L];
        r.wb_off[L] = kL.wb[L];
    }
    r.b5_off = kL.b[kMlpLayers - 1];
    return mlp_launch_adam_refresh(c->half, c->params, c->adam_m, c->adam_v, c->grad, kMlpParams, lr, b1, b2, eps, c1, c2,
                                   c->half ? c->scale_state : nullptr, r, c->wbf, c->b5, c->stream);
}

MlpDev cache_dev(const nb_cache *c) { return c->dev(); }
hipStream_t cache_stream(const nb_cache *c) { return c->stream; }
float *cache_raw(nb_cache *c) { return c->raw; }
float *cache_target(nb_cache *c) { return c->target; }
uint32_t cache_max_batch(const nb_cache *c) { return c->max_batch; }

}  // namespace nb

extern "C" {

uint32_t nb_cache_param_count(void) { return kMlpParams; }

nb_status nb_cache_create(nb_context *ctx, const float scene_min[3], const float scene_max[3], uint32_t seed,
                          uint32_t max_batch, nb_cache **out) {
    return nb_cache_create_ex(ctx, scene_min, scene_max, seed, max_batch, 0u, out);
}

nb_status nb_cache_create_ex(nb_context *ctx, const float scene_min[3], const float scene_max[3], uint32_t seed,
                             uint32_t max_batch, uint32_t flags, nb_cache **out) {
    if (!ctx || !scene_min || !scene_max || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (flags & ~NB_CACHE_FP16) return set_error(NB_ERR_INVALID_ARGUMENT, "unknown cache flags %#x", flags);
    if (max_batch < 16 || max_batch > (1u << 24)) return set_error(NB_ERR_INVALID_ARGUMENT, "max_batch must be in 16..2^24");
    for (int d = 0; d < 3; ++d)
        if (!(scene_max[d] > scene_min[d])) return set_error(NB_ERR_INVALID_ARGUMENT, "empty scene box");
    C_HIP(hipSetDevice(ctx->device));
    nb_cache *c = new nb_cache();
    c->ctx = ctx;
    c->half = (flags & NB_CACHE_FP16) != 0;
    context_add_child(ctx);
    c->max_batch = (max_batch + 15) / 16 * 16;
    for (int d = 0; d < 3; ++d) { c->scene_min[d] = scene_min[d]; c->scene_inv[d] = 1.0f / (scene_max[d] - scene_min[d]); }
    const size_t B = c->max_batch;
    c->stream = ctx->stream;      // the cache's work is ordered with the context's renders
    hipError_t e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc(&c->h_count, 16, hipHostMallocDefault);
    auto alloc = [&](auto **p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), bytes); };
    alloc(&c->params, kMlpParams * 4); alloc(&c->adam_m, kMlpParams * 4); alloc(&c->adam_v, kMlpParams * 4);
    // the loss sum is the word after the gradient, so one fill clears both before a step
    alloc(&c->grad, (kMlpParams + 1) * 4); alloc(&c->wbf, kBf16Weights * 2); alloc(&c->b5, kMlpOutPad * 4);
    alloc(&c->raw, B * kMlpRaw * 4); alloc(&c->target, B * 3 * 4); alloc(&c->rgb, B * 3 * 4);
    if (e == hipSuccess) c->loss = c->grad + kMlpParams;
    alloc(&c->x, B * kMlpIn * 2); alloc(&c->dy, B * kMlpOutPad * 2);
    alloc(&c->counts, 16);
    alloc(&c->scale_state, 16);
    for (int L = 0; L < 5; ++L) { alloc(&c->h[L], B * kMlpHidden * 2); alloc(&c->dz[L], B * kMlpHidden * 2); }
    if (e == hipSuccess) e = hipMemset(c->wbf, 0, kBf16Weights * 2);
    if (e == hipSuccess) e = hipMemset(c->b5, 0, kMlpOutPad * 4);
    if (e == hipSuccess) e = hipMemset(c->adam_m, 0, kMlpParams * 4);
    if (e == hipSuccess) e = hipMemset(c->adam_v, 0, kMlpParams * 4);
    if (e == hipSuccess) {
        const float initial_scale = 1024.0f;   // fp16 loss scale at the start (unused for bf16)
        uint32_t st[4] = {0u, 0u, 0u, 0u};
        std::memcpy(&st[0], &initial_scale, 4);
        e = hipMemcpy(c->scale_state, st, 16, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        nb_cache_destroy(c);
        return set_error(NB_ERR_OUT_OF_MEMORY, "cache allocation: %s", hipGetErrorString(e));
    }
    // He-uniform weights for the ReLU layers, a 10x smaller last layer, zero biases.
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t st = 0x5EEDull ^ (static_cast<uint64_t>(seed) << 32);
    for (int L = 0; L < kMlpLayers; ++L) {
        const float bound = std::sqrt(6.0f / static_cast<float>(kL.in[L])) * (L == kMlpLayers - 1 ? 0.1f : 1.0f);
        for (uint32_t i = 0; i < kL.in[L] * kL.out[L]; ++i) {
            const double u = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[kL.w[L] + i] = static_cast<float>((2.0 * u - 1.0) * bound);
        }
    }
    nb_status s = nb_cache_set_params(c, p.data());
    if (s != NB_OK) { nb_cache_destroy(c); return s; }
    *out = c;
    return NB_OK;
}

nb_status nb_cache_destroy(nb_cache *c) {
    if (!c) return NB_OK;
    hipSetDevice(c->ctx->device);
    hipStreamSynchronize(c->stream);
    if (c->side) { hipStreamSynchronize(c->side); hipStreamDestroy(c->side); }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->h_count) hipHostFree(c->h_count);
    void *ps[] = {c->params, c->adam_m, c->adam_v, c->grad, c->wbf, c->b5, c->raw, c->target, c->rgb,
                  c->x, c->dy, c->direct, c->q_raw, c->q_thr, c->q_rgb, c->counts, c->scale_state};
    for (void *p : ps) hipFree(p);
    for (int L = 0; L < 5; ++L) { hipFree(c->h[L]); hipFree(c->dz[L]); }
    nb_context *ctx = c->ctx;
    delete c;
    context_release_child(ctx);
    return NB_OK;
}

nb_status nb_cache_get_params(nb_cache *c, float *out) {
    if (!c || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipStreamSynchronize(c->stream));
    C_HIP(hipMemcpy(out, c->params, kMlpParams * 4, hipMemcpyDeviceToHost));
    return NB_OK;
}

nb_status nb_cache_set_params(nb_cache *c, const float *in) {
    if (!c || !in) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipStreamSynchronize(c->stream));
    C_HIP(hipMemcpy(c->params, in, kMlpParams * 4, hipMemcpyHostToDevice));
    C_TRY(cache_refresh(c));
    C_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

nb_status nb_cache_encode(nb_cache *c, const float *raw, uint32_t n, uint16_t *out) {
    if (!c || !raw || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (n > c->max_batch) return set_error(NB_ERR_INVALID_ARGUMENT, "batch %u > max_batch %u", n, c->max_batch);
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipMemcpyAsync(c->raw, raw, static_cast<size_t>(n) * kMlpRaw * 4, hipMemcpyHostToDevice, c->stream));
    C_TRY(mlp_launch_encode(c->dev(), c->raw, n, c->x, c->stream));
    C_HIP(hipMemcpyAsync(out, c->x, static_cast<size_t>(n) * kMlpIn * 2, hipMemcpyDeviceToHost, c->stream));
    C_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

nb_status nb_cache_infer(nb_cache *c, const float *raw, uint32_t n, float *rgb) {
    if (!c || !raw || !rgb) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (n > c->max_batch) return set_error(NB_ERR_INVALID_ARGUMENT, "batch %u > max_batch %u", n, c->max_batch);
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipMemcpyAsync(c->raw, raw, static_cast<size_t>(n) * kMlpRaw * 4, hipMemcpyHostToDevice, c->stream));
    C_TRY(mlp_launch_infer(c->dev(), c->raw, n, c->rgb, c->stream));
    C_HIP(hipMemcpyAsync(rgb, c->rgb, static_cast<size_t>(n) * 12, hipMemcpyDeviceToHost, c->stream));
    C_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

nb_status nb_cache_gradient(nb_cache *c, const float *raw, const float *target, uint32_t n, float *loss, float *grad) {
    if (!c || !raw || !target || !loss || !grad) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0 || n > c->max_batch) return set_error(NB_ERR_INVALID_ARGUMENT, "batch must be in 1..%u", c->max_batch);
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipMemcpyAsync(c->raw, raw, static_cast<size_t>(n) * kMlpRaw * 4, hipMemcpyHostToDevice, c->stream));
    C_HIP(hipMemcpyAsync(c->target, target, static_cast<size_t>(n) * 12, hipMemcpyHostToDevice, c->stream));
    C_TRY(cache_forward_backward(c, c->raw, c->target, n));
    C_HIP(hipMemcpyAsync(loss, c->loss, 4, hipMemcpyDeviceToHost, c->stream));
    C_HIP(hipMemcpyAsync(grad, c->grad, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
    float scale = 1.0f;
    if (c->half) C_HIP(hipMemcpyAsync(&scale, c->scale_state, 4, hipMemcpyDeviceToHost, c->stream));
    C_HIP(hipStreamSynchronize(c->stream));
    if (scale != 1.0f)
        for (uint32_t i = 0; i < kMlpParams; ++i) grad[i] /= scale;
    return NB_OK;
}

nb_status nb_cache_loss_scale(nb_cache *c, float *scale) {
    if (!c || !scale) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *scale = 1.0f;
    if (!c->half) return NB_OK;
    C_HIP(hipSetDevice(c->ctx->device));
    C_HIP(hipStreamSynchronize(c->stream));
    C_HIP(hipMemcpy(scale, c->scale_state, 4, hipMemcpyDeviceToHost));
    return NB_OK;
}

nb_status nb_cache_train(nb_cache *c, const float *raw, const float *target, uint32_t n, uint32_t batch, float lr,
                         float *last_loss) {
    if (!c || !raw || !target) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (batch == 0 || batch > c->max_batch) return set_error(NB_ERR_INVALID_ARGUMENT, "batch must be in 1..%u", c->max_batch);
    C_HIP(hipSetDevice(c->ctx->device));
    for (uint32_t off = 0; off < n; off += batch) {
        const uint32_t k = n - off < batch ? n - off : batch;
        C_HIP(hipMemcpyAsync(c->raw, raw + static_cast<size_t>(off) * kMlpRaw, static_cast<size_t>(k) * kMlpRaw * 4,
                             hipMemcpyHostToDevice, c->stream));
        C_HIP(hipMemcpyAsync(c->target, target + static_cast<size_t>(off) * 3, static_cast<size_t>(k) * 12,
                             hipMemcpyHostToDevice, c->stream));
        C_TRY(cache_forward_backward(c, c->raw, c->target, k));
        C_TRY(cache_adam_step(c, lr));
    }
    if (last_loss) C_HIP(hipMemcpyAsync(last_loss, c->loss, 4, hipMemcpyDeviceToHost, c->stream));
    C_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

}  // extern "C"

namespace nb {

// Where a frame's training paths start: path i of n traces pixel (i * (W*H / n) + offset) mod W*H
// with Sobol row train_sample(sample), a row of the 1024-row table away from the frame's own
// (so the training paths do not repeat the frame's samples).
static uint32_t train_offset(uint32_t sample, size_t npix) {
    return static_cast<uint32_t>((static_cast<uint64_t>(sample) * 2654435761ull) % npix);
}
static uint32_t train_sample(uint32_t sample) { return 1u + (sample + 511u) % 1023u; }

// One cache-assisted frame into a float4 target of width x height pixels on the context's stream
// (nb_render_cached's work; also used by the server for cached requests).
// The per-pixel workspace of cache frames (direct radiance, query records, predictions) for at
// least npix pixels.
static nb_status ensure_workspace(nb_context *ctx, nb_cache *c, size_t npix) {
    if (npix <= c->ws_pixels) return NB_OK;
    C_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(c->direct); hipFree(c->q_raw); hipFree(c->q_thr); hipFree(c->q_rgb);
    c->direct = nullptr; c->q_raw = c->q_thr = c->q_rgb = nullptr; c->ws_pixels = 0;
    C_HIP(hipMalloc(&c->direct, npix * 16));
    C_HIP(hipMalloc(&c->q_raw, npix * kMlpRaw * 4));
    C_HIP(hipMalloc(&c->q_thr, npix * 16));
    C_HIP(hipMalloc(&c->q_rgb, npix * 12));
    c->ws_pixels = npix;
    return NB_OK;
}

nb_status cache_render_frame(nb_context *ctx, const DevScene &scene, const DevEnv &env, float4 *target, uint32_t width,
                             uint32_t height, const float *c2w, const nb_globals &g, nb_cache *c, uint32_t train_paths,
                             float lr, float *loss) {
    const size_t npix = static_cast<size_t>(width) * height;
    C_TRY(ensure_workspace(ctx, c, npix));
    CacheFrame f{};
    f.scene = scene;
    f.env = env;
    f.sobol = ctx->sobol;
    f.sobol_entries = ctx->sobol_entries;
    f.sobol_stride = ctx->sobol_stride;
    std::memcpy(f.c2w, c2w, sizeof(f.c2w));
    f.width = width;
    f.height = height;
    f.sample = g.sample;
    f.weight = g.weight;
    f.bounces = g.bounces;
    f.contribution_factor = g.contribution_factor;
    f.mirror_alpha = 0.05f;
    f.max_cache_bounce = 3;
    f.target = target;
    f.direct = c->direct;
    f.q_raw = c->q_raw;
    f.q_thr = c->q_thr;
    f.q_count = c->counts;
    f.train_paths = train_paths;
    f.train_offset = train_offset(g.sample, npix);
    f.train_sample = train_sample(g.sample);
    f.t_capacity = c->max_batch;
    f.t_raw = c->raw;
    f.t_target = c->target;
    f.t_count = c->counts + 1;
    f.counters = ctx->counting ? ctx->counters : nullptr;
    f.tiles_x = width / 8;
    f.n_tiles = (width / 8) * (height / 8);
    f.tiles_per_xcd = (f.n_tiles + 7) / 8;
    if (train_paths > 0) {
        // The training paths do not depend on the network and neither does the frame's tracing up
        // to its queries: the two run side by side (the side stream starts after the work already
        // queued on the context's stream), and the training steps and the inference that uses
        // their weights follow both.
        C_HIP(hipEventRecord(c->ev_fork, ctx->stream));
        C_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        C_TRY(cache_launch_train_paths(f, c->side));
        C_TRY(cache_launch_query(f, ctx->stream));
        C_HIP(hipMemcpyAsync(c->h_count, c->counts + 1, 4, hipMemcpyDeviceToHost, c->side));
        C_HIP(hipEventRecord(c->ev_join, c->side));
        C_HIP(hipStreamWaitEvent(ctx->stream, c->ev_join, 0));
        C_HIP(hipEventSynchronize(c->ev_join));
        uint32_t n = *c->h_count;
        if (n > c->max_batch) n = c->max_batch;
        const uint32_t batch = train_step_rows() ? train_step_rows() : c->max_batch;
        for (uint32_t off = 0; off + 256 <= n; off += batch) {
            const uint32_t k = n - off < batch ? n - off : batch;
            C_TRY(cache_forward_backward(c, c->raw + static_cast<size_t>(off) * kMlpRaw, c->target + 3 * static_cast<size_t>(off), k));
            C_TRY(cache_adam_step(c, lr));
        }
        if (loss) C_HIP(hipMemcpyAsync(loss, c->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        C_TRY(cache_launch_query(f, ctx->stream));
    }
    return cache_launch_finish(f, c->dev(), c->q_rgb, ctx->stream);
}

nb_context *cache_context(const nb_cache *c) { return c->ctx; }

uint64_t cache_workspace_bytes(uint64_t pixels) { return pixels * (16 + kMlpRaw * 4 + 16 + 12); }

uint64_t cache_workspace_pixels(const nb_cache *c) { return c->ws_pixels; }

uint64_t cache_device_bytes(const nb_cache *c) {
    const uint64_t B = c->max_batch;
    const uint64_t fixed = 3ull * kMlpParams * 4 + (kMlpParams + 1) * 4 + kBf16Weights * 2 + kMlpOutPad * 4 + 16 + 16;
    const uint64_t batch = B * kMlpRaw * 4 + B * 3 * 4 * 2 + B * kMlpIn * 2 + B * kMlpOutPad * 2 + 10ull * B * kMlpHidden * 2;
    return fixed + batch + cache_workspace_bytes(c->ws_pixels);
}

}  // namespace nb

extern "C" {

nb_status nb_cache_training_records(nb_context *ctx, const nb_scene *scene, uint32_t width, uint32_t height,
                                    const nb_camera *camera, const nb_globals *g, const nb_envmap *env, nb_cache *c,
                                    uint32_t train_paths, float *raw, float *target, uint32_t cap, uint32_t *n) {
    if (!ctx || !scene || !camera || !g || !env || !c || !n || (cap && (!raw || !target)))
        return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx || env->ctx != ctx || c->ctx != ctx)
        return set_error(NB_ERR_INVALID_ARGUMENT, "all objects must belong to the context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (g->sample == 0) return set_error(NB_ERR_INVALID_ARGUMENT, "samples are 1-based");
    if (!width || !height || width % 8 || height % 8 || train_paths == 0)
        return set_error(NB_ERR_INVALID_ARGUMENT, "need a nonzero multiple-of-8 size and training paths");
    if (g->bounces > kMaxBounceLimit)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bounces %u exceeds %u", g->bounces, kMaxBounceLimit);
    C_HIP(hipSetDevice(ctx->device));
    const size_t npix = static_cast<size_t>(width) * height;
    CacheFrame f{};
    f.scene = scene->dev;
    f.env = dev_env(env);
    f.sobol = ctx->sobol;
    f.sobol_entries = ctx->sobol_entries;
    f.sobol_stride = ctx->sobol_stride;
    std::memcpy(f.c2w, camera->clip_to_world, sizeof(f.c2w));
    f.width = width;
    f.height = height;
    f.sample = g->sample;
    f.bounces = g->bounces;
    f.contribution_factor = g->contribution_factor;
    f.train_paths = train_paths;
    f.train_offset = train_offset(g->sample, npix);
    f.train_sample = train_sample(g->sample);
    f.t_capacity = c->max_batch;
    f.t_raw = c->raw;
    f.t_target = c->target;
    f.t_count = c->counts + 1;
    C_TRY(cache_launch_train_paths(f, ctx->stream));
    uint32_t count = 0;
    C_HIP(hipMemcpyAsync(&count, c->counts + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    C_HIP(hipStreamSynchronize(ctx->stream));
    if (count > c->max_batch) count = c->max_batch;
    const uint32_t k = count < cap ? count : cap;
    if (k) {
        C_HIP(hipMemcpy(raw, c->raw, static_cast<size_t>(k) * kMlpRaw * 4, hipMemcpyDeviceToHost));
        C_HIP(hipMemcpy(target, c->target, static_cast<size_t>(k) * 12, hipMemcpyDeviceToHost));
    }
    *n = count;
    return NB_OK;
}

nb_status nb_cache_query_records(nb_context *ctx, const nb_scene *scene, uint32_t width, uint32_t height,
                                 const nb_camera *camera, const nb_globals *g, const nb_envmap *env, nb_cache *c,
                                 float *direct, float *raw, float *thr, uint32_t cap, uint32_t *n) {
    if (!ctx || !scene || !camera || !g || !env || !c || !n || !direct || (cap && (!raw || !thr)))
        return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx || env->ctx != ctx || c->ctx != ctx)
        return set_error(NB_ERR_INVALID_ARGUMENT, "all objects must belong to the context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (g->sample == 0) return set_error(NB_ERR_INVALID_ARGUMENT, "samples are 1-based");
    if (!width || !height || width % 8 || height % 8)
        return set_error(NB_ERR_INVALID_ARGUMENT, "need a nonzero multiple-of-8 size");
    if (g->bounces > kMaxBounceLimit)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bounces %u exceeds %u", g->bounces, kMaxBounceLimit);
    C_HIP(hipSetDevice(ctx->device));
    const size_t npix = static_cast<size_t>(width) * height;
    C_TRY(ensure_workspace(ctx, c, npix));
    // the query part of cache_render_frame's frame, as nb_render_cached launches it
    CacheFrame f{};
    f.scene = scene->dev;
    f.env = dev_env(env);
    f.sobol = ctx->sobol;
    f.sobol_entries = ctx->sobol_entries;
    f.sobol_stride = ctx->sobol_stride;
    std::memcpy(f.c2w, camera->clip_to_world, sizeof(f.c2w));
    f.width = width;
    f.height = height;
    f.sample = g->sample;
    f.bounces = g->bounces;
    f.contribution_factor = g->contribution_factor;
    f.mirror_alpha = 0.05f;
    f.max_cache_bounce = 3;
    f.direct = c->direct;
    f.q_raw = c->q_raw;
    f.q_thr = c->q_thr;
    f.q_count = c->counts;
    f.tiles_x = width / 8;
    f.n_tiles = (width / 8) * (height / 8);
    f.tiles_per_xcd = (f.n_tiles + 7) / 8;
    C_TRY(cache_launch_query(f, ctx->stream));
    uint32_t count = 0;
    C_HIP(hipMemcpyAsync(&count, c->counts, 4, hipMemcpyDeviceToHost, ctx->stream));
    C_HIP(hipStreamSynchronize(ctx->stream));
    C_HIP(hipMemcpy(direct, c->direct, npix * 16, hipMemcpyDeviceToHost));
    const uint32_t k = count < cap ? count : cap;
    if (k) {
        C_HIP(hipMemcpy(raw, c->q_raw, static_cast<size_t>(k) * kMlpRaw * 4, hipMemcpyDeviceToHost));
        C_HIP(hipMemcpy(thr, c->q_thr, static_cast<size_t>(k) * 16, hipMemcpyDeviceToHost));
    }
    *n = count;
    return NB_OK;
}

nb_status nb_render_cached(nb_context *ctx, const nb_scene *scene, nb_target *target, const nb_camera *camera,
                           const nb_globals *g, const nb_envmap *env, nb_cache *c, uint32_t train_paths, float lr,
                           float *loss) {
    if (!ctx || !scene || !target || !camera || !g || !env || !c) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != ctx || target->ctx != ctx || env->ctx != ctx || c->ctx != ctx)
        return set_error(NB_ERR_INVALID_ARGUMENT, "all objects must belong to the context");
    if (!ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set (nb_sobol_set)");
    if (g->sample == 0) return set_error(NB_ERR_INVALID_ARGUMENT, "samples are 1-based");
    if (g->bounces > kMaxBounceLimit)
        return set_error(NB_ERR_INVALID_ARGUMENT, "bounces %u exceeds %u", g->bounces, kMaxBounceLimit);
    C_HIP(hipSetDevice(ctx->device));
    if (ctx->render_timing) C_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    C_TRY(cache_render_frame(ctx, scene->dev, dev_env(env), target->pixels, target->width,
                             target->height, camera->clip_to_world, *g, c, train_paths, lr, loss));
    if (ctx->render_timing) {
        C_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        ctx->timed = true;
    }
    if (loss && train_paths > 0) C_HIP(hipStreamSynchronize(ctx->stream));
    return NB_OK;
}

}  // extern "C"
