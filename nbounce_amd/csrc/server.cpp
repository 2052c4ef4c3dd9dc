// Serving host side (nb_server_* in include/nbounce.h): resident scenes, request admission by
// slot and HBM budget, continuous batching of every admitted request's samples into steps, and
// one captured HIP graph that renders a step (reset, one persistent kernel and the ordered
// accumulation, serve.hip).
//
// Scheduling: admitted requests are visited round-robin in admission order; each round gives a
// request one more sample pass (all its pixels) while the step's path budget lasts, so small
// requests finish in a step or two even while a large one is being rendered, and a request that
// arrives while a step runs joins the next step.
#include "nbounce.h"
#include "nb_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <set>
#include <vector>

using namespace nb;

namespace {

#define SRV_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_error(e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "%s: %s", \
                             #call, hipGetErrorString(e_));                                            \
    } while (0)

#define SRV_TRY(expr)                  \
    do {                               \
        nb_status s_ = (expr);         \
        if (s_ != NB_OK) return s_;    \
    } while (0)

constexpr uint32_t kMaxSlots = 256;
constexpr uint32_t kMaxScenes = 64;

struct Request {
    uint64_t id = 0;
    nb_request req{};
    double arrival = 0, admit = -1, done = -1;
    uint32_t next_sample = 1;     // next sample to schedule
    int slot = -1;
    float4 *target = nullptr;
    uint64_t bytes = 0;
};

}  // namespace

struct nb_server {
    nb_context *ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // HBM admission: in_use counts the server's own step buffers, the resident scenes and
    // environments, the caches (parameters, Adam state, batch buffers and the frame workspace
    // the largest admitted cached request needs) and the admitted requests' images.
    uint64_t budget = 0, in_use = 0, resident = 0;
    std::set<const void *> counted;             // scenes, environments and caches already charged
    std::map<const nb_cache *, uint64_t> cache_ws_pixels;   // frame workspace charged per cache
    uint32_t max_paths = 0, max_bounces = 0, cu = 0;
    std::chrono::steady_clock::time_point t0;

    // device buffers
    uint32_t *work = nullptr;                   // per scene group of a step, kWorkStride apart
    float4 *results = nullptr;                  // radiance per path of a step
    unsigned long long *counters = nullptr;
    ServeStep *d_step = nullptr;
    ServePass *d_passes = nullptr;
    ServeGroup *d_groups = nullptr;
    ServeRequestDev *d_requests = nullptr;
    ServeSceneDev *d_scenes = nullptr;
    // pinned staging
    ServeStep *h_step = nullptr;
    ServePass *h_passes = nullptr;
    ServeGroup *h_groups = nullptr;
    ServeRequestDev *h_requests = nullptr;
    ServeSceneDev *h_scenes = nullptr;
    uint32_t n_scenes = 0;
    nb_cache *caches[kMaxScenes] = {};          // per resident scene, for cached requests
    uint64_t cached_samples = 0;
    double cached_busy = 0;

    hipGraphExec_t graph = nullptr;
    uint64_t next_id = 1;
    std::deque<Request *> queued;
    std::vector<Request *> active;              // admission order
    Request *slots[kMaxSlots] = {};
    std::map<uint64_t, Request *> finished;     // completed, image still held
    std::vector<nb_completion> unpolled;
    uint64_t steps = 0;
    double busy = 0;
    uint32_t completed = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// The serve kernel's scheduling parameters (the progressive renderer's meaning): lanes waiting
// before a wave shades and refills, the inner-node quorum and the units per grab.
// NB_SERVE_MIN_IDLE, NB_SERVE_INNER_QUORUM and NB_SERVE_GRAB_UNITS override them, read once.
// On the serving stream, waiting for 24 lanes instead of the renderer's 16 gave p50 2.35 vs
// 2.7 ms and p99 117 vs 120 ms; quorums of 24 and 32 and grabs of 2 and 8 gave nothing.
uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
    const char *e = std::getenv(name);
    if (!e) return dflt;
    const long v = std::strtol(e, nullptr, 10);
    return static_cast<uint32_t>(v < static_cast<long>(lo) ? lo : (v > static_cast<long>(hi) ? hi : v));
}
uint32_t serve_min_idle() { static const uint32_t v = env_u32("NB_SERVE_MIN_IDLE", 24, 1, 64); return v; }
uint32_t serve_inner_quorum() { static const uint32_t v = env_u32("NB_SERVE_INNER_QUORUM", 28, 0, 64); return v; }
uint32_t serve_grab_units() { static const uint32_t v = env_u32("NB_SERVE_GRAB_UNITS", 4, 1, 64); return v; }

nb_status capture_graph(nb_server *s) {
    SRV_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    nb_status st = serve_launch_reset(s->d_step, s->stream);
    if (st == NB_OK) st = serve_launch_step(s->d_step, s->stream);
    if (st == NB_OK) st = serve_launch_accumulate(s->d_step, s->cu * 8, s->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (st != NB_OK) { if (g) hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    return NB_OK;
}

void free_server(nb_server *s) {
    hipSetDevice(s->ctx->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->graph) hipGraphExecDestroy(s->graph);
    for (Request *r : s->queued) delete r;
    for (Request *r : s->active) { hipFree(r->target); delete r; }
    for (auto &kv : s->finished) { hipFree(kv.second->target); delete kv.second; }
    hipFree(s->work); hipFree(s->results); hipFree(s->counters);
    hipFree(s->d_step); hipFree(s->d_passes); hipFree(s->d_groups); hipFree(s->d_requests); hipFree(s->d_scenes);
    hipHostFree(s->h_step); hipHostFree(s->h_passes); hipHostFree(s->h_groups); hipHostFree(s->h_requests);
    hipHostFree(s->h_scenes);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->stream) hipStreamDestroy(s->stream);
    nb_context *ctx = s->ctx;
    delete s;
    context_release_child(ctx);
}

nb_status admit(nb_server *s) {
    while (!s->queued.empty()) {
        Request *r = s->queued.front();
        int slot = -1;
        for (uint32_t k = 0; k < kMaxSlots; ++k) if (!s->slots[k]) { slot = static_cast<int>(k); break; }
        // a cached request may grow its cache's frame workspace (kept for the server's life)
        uint64_t ws = 0, ws_px = 0;
        nb_cache *cache = r->req.mode == NB_REQUEST_CACHED ? s->caches[r->req.scene] : nullptr;
        if (cache) {
            ws_px = static_cast<uint64_t>(r->req.width) * r->req.height;
            const uint64_t have = s->cache_ws_pixels[cache];
            ws = ws_px > have ? cache_workspace_bytes(ws_px) - cache_workspace_bytes(have) : 0;
        }
        if (slot < 0 || s->in_use + r->bytes + ws > s->budget) break;   // FIFO: wait for slots / memory
        SRV_HIP(hipMalloc(&r->target, r->bytes));
        SRV_HIP(hipMemsetAsync(r->target, 0, r->bytes, s->stream));
        if (ws) {
            s->cache_ws_pixels[cache] = ws_px;
            s->resident += ws;
            s->in_use += ws;
        }
        s->in_use += r->bytes;
        r->slot = slot;
        r->admit = s->now();
        s->slots[slot] = r;
        s->active.push_back(r);
        s->queued.pop_front();
        ServeRequestDev &d = s->h_requests[slot];
        std::memcpy(d.c2w, r->req.camera.clip_to_world, sizeof(d.c2w));
        d.target = r->target;
        d.width = r->req.width;
        d.height = r->req.height;
        d.scene = r->req.scene;
        d.bounces = r->req.bounces;
        d.contribution_factor = r->req.contribution_factor;
        d.rr_first = 0;   // served requests use the reference's roulette
    }
    return NB_OK;
}

}  // namespace

extern "C" {

nb_status nb_server_create(nb_context *ctx, uint64_t budget, uint32_t max_paths, uint32_t max_bounces, nb_server **out) {
    if (!ctx || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (max_paths < 4096 || max_paths > (1u << 26)) return set_error(NB_ERR_INVALID_ARGUMENT, "max_paths must be in 4096..2^26");
    if (max_bounces > kServeMaxBounces) return set_error(NB_ERR_INVALID_ARGUMENT, "max_bounces must be <= %u", kServeMaxBounces);
    SRV_HIP(hipSetDevice(ctx->device));
    nb_server *s = new nb_server();
    s->ctx = ctx;
    context_add_child(ctx);
    s->budget = budget;
    s->max_paths = max_paths;
    s->max_bounces = max_bounces;
    s->t0 = std::chrono::steady_clock::now();
    int cu = 0;
    hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    s->cu = cu > 0 ? static_cast<uint32_t>(cu) : 64u;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e == hipSuccess) e = hipMalloc(&s->work, kServeMaxGroups * kWorkStride * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s->results, static_cast<size_t>(max_paths) * 16);
    if (e == hipSuccess) e = hipMalloc(&s->counters, kCounterBytes);
    if (e == hipSuccess) e = hipMemset(s->counters, 0, kCounterBytes);
    if (e == hipSuccess) e = hipMalloc(&s->d_step, sizeof(ServeStep));
    if (e == hipSuccess) e = hipMalloc(&s->d_passes, kMaxSlots * sizeof(ServePass));
    if (e == hipSuccess) e = hipMalloc(&s->d_groups, kServeMaxGroups * sizeof(ServeGroup));
    if (e == hipSuccess) e = hipMalloc(&s->d_requests, kMaxSlots * sizeof(ServeRequestDev));
    if (e == hipSuccess) e = hipMalloc(&s->d_scenes, kMaxScenes * sizeof(ServeSceneDev));
    if (e == hipSuccess) e = hipHostMalloc(&s->h_step, sizeof(ServeStep), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&s->h_passes, kMaxSlots * sizeof(ServePass), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&s->h_groups, kServeMaxGroups * sizeof(ServeGroup), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&s->h_requests, kMaxSlots * sizeof(ServeRequestDev), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&s->h_scenes, kMaxScenes * sizeof(ServeSceneDev), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_server(s);
        return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "server allocation: %s",
                         hipGetErrorString(e));
    }
    // the server's own device buffers count against the budget
    s->resident = kServeMaxGroups * kWorkStride * sizeof(uint32_t) + static_cast<uint64_t>(max_paths) * 16 +
                  kCounterBytes + sizeof(ServeStep) + kMaxSlots * (sizeof(ServePass) + sizeof(ServeRequestDev)) +
                  kServeMaxGroups * sizeof(ServeGroup) + kMaxScenes * sizeof(ServeSceneDev);
    s->in_use = s->resident;
    if (s->resident > budget) {
        const unsigned long long need = s->resident;
        free_server(s);
        return set_error(NB_ERR_BUDGET, "the server's step buffers (%llu B) exceed the HBM budget", need);
    }
    std::memset(s->h_requests, 0, kMaxSlots * sizeof(ServeRequestDev));
    std::memset(s->h_step, 0, sizeof(ServeStep));
    nb_status st = capture_graph(s);
    if (st != NB_OK) { free_server(s); return st; }
    *out = s;
    return NB_OK;
}

nb_status nb_server_destroy(nb_server *s) {
    if (!s) return NB_OK;
    free_server(s);
    return NB_OK;
}

nb_status nb_server_add_scene(nb_server *s, const nb_scene *scene, const nb_envmap *env, uint32_t *index) {
    if (!s || !scene || !env || !index) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->ctx != s->ctx || env->ctx != s->ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "scene and environment must belong to the server's context");
    if (s->n_scenes >= kMaxScenes) return set_error(NB_ERR_UNSUPPORTED, "at most %u resident scenes", kMaxScenes);
    // a scene or environment shared by several resident scenes is charged once
    uint64_t add = 0;
    if (!s->counted.count(scene)) add += scene->bytes;
    if (!s->counted.count(env)) add += envmap_device_bytes(env);
    if (s->in_use + add > s->budget)
        return set_error(NB_ERR_BUDGET, "scene %u needs %llu B, %llu B of the HBM budget are left", s->n_scenes,
                         static_cast<unsigned long long>(add), static_cast<unsigned long long>(s->budget - s->in_use));
    try {
        s->counted.insert(scene);
        s->counted.insert(env);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    s->resident += add;
    s->in_use += add;
    s->h_scenes[s->n_scenes] = ServeSceneDev{scene->dev, dev_env(env)};
    SRV_HIP(hipSetDevice(s->ctx->device));
    SRV_HIP(hipMemcpy(s->d_scenes, s->h_scenes, (s->n_scenes + 1) * sizeof(ServeSceneDev), hipMemcpyHostToDevice));
    *index = s->n_scenes++;
    return NB_OK;
}

nb_status nb_server_set_cache(nb_server *s, uint32_t scene, nb_cache *cache) {
    if (!s || !cache) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (scene >= s->n_scenes) return set_error(NB_ERR_INVALID_ARGUMENT, "scene %u not added", scene);
    if (cache_context(cache) != s->ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "the cache belongs to another context");
    if (s->caches[scene]) return set_error(NB_ERR_INVALID_ARGUMENT, "scene %u already has a cache", scene);
    const uint64_t add = s->counted.count(cache) ? 0 : cache_device_bytes(cache);
    if (s->in_use + add > s->budget)
        return set_error(NB_ERR_BUDGET, "the cache needs %llu B, %llu B of the HBM budget are left",
                         static_cast<unsigned long long>(add), static_cast<unsigned long long>(s->budget - s->in_use));
    try {
        s->counted.insert(cache);
        s->cache_ws_pixels[cache] = cache_workspace_pixels(cache);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    s->resident += add;
    s->in_use += add;
    s->caches[scene] = cache;
    return NB_OK;
}

nb_status nb_server_submit(nb_server *s, const nb_request *req, uint64_t *id) {
    if (!s || !req || !id) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (req->scene >= s->n_scenes) return set_error(NB_ERR_INVALID_ARGUMENT, "scene %u not added", req->scene);
    if (!req->width || !req->height || req->width % 8 || req->height % 8)
        return set_error(NB_ERR_INVALID_ARGUMENT, "size %ux%u: sides must be nonzero multiples of 8", req->width, req->height);
    const uint64_t pixels = static_cast<uint64_t>(req->width) * req->height;
    if (pixels > s->max_paths) return set_error(NB_ERR_INVALID_ARGUMENT, "%llu pixels exceed the step budget of %u paths",
                                                 static_cast<unsigned long long>(pixels), s->max_paths);
    if (req->spp == 0) return set_error(NB_ERR_INVALID_ARGUMENT, "spp must be >= 1");
    if (req->bounces > s->max_bounces) return set_error(NB_ERR_INVALID_ARGUMENT, "bounces %u > server maximum %u", req->bounces, s->max_bounces);
    if (req->mode != NB_REQUEST_PLAIN && req->mode != NB_REQUEST_CACHED)
        return set_error(NB_ERR_INVALID_ARGUMENT, "unknown request mode %u", req->mode);
    if (req->mode == NB_REQUEST_CACHED && !s->caches[req->scene])
        return set_error(NB_ERR_INVALID_ARGUMENT, "scene %u has no cache (nb_server_set_cache)", req->scene);
    const uint64_t bytes = pixels * 16;
    // a request that cannot fit beside the resident data even on an idle server is refused
    uint64_t ws = 0;
    if (req->mode == NB_REQUEST_CACHED) {
        const uint64_t have = s->cache_ws_pixels[s->caches[req->scene]];
        ws = pixels > have ? cache_workspace_bytes(pixels) - cache_workspace_bytes(have) : 0;
    }
    if (s->resident + bytes + ws > s->budget)
        return set_error(NB_ERR_BUDGET, "request needs %llu B, %llu B of the HBM budget are left beside the resident data",
                         static_cast<unsigned long long>(bytes + ws), static_cast<unsigned long long>(s->budget - s->resident));
    Request *r = new (std::nothrow) Request();
    if (!r) return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    r->id = s->next_id;
    r->req = *req;
    r->arrival = req->arrival_s >= 0 ? req->arrival_s : s->now();
    r->bytes = bytes;
    try {
        s->queued.push_back(r);
    } catch (const std::bad_alloc &) {
        delete r;
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    ++s->next_id;
    *id = r->id;
    return NB_OK;
}

static nb_status server_step(nb_server *s, uint32_t *n_active, uint32_t *n_queued);

// No exception crosses the C ABI: a host allocation failure inside a step is reported as
// NB_ERR_OUT_OF_MEMORY.
nb_status nb_server_step(nb_server *s, uint32_t *n_active, uint32_t *n_queued) {
    try {
        return server_step(s, n_active, n_queued);
    } catch (const std::bad_alloc &) {
        return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed in a server step");
    }
}

static nb_status server_step(nb_server *s, uint32_t *n_active, uint32_t *n_queued) {
    if (!s) return set_error(NB_ERR_INVALID_ARGUMENT, "null server");
    if (!s->ctx->sobol) return set_error(NB_ERR_INVALID_ARGUMENT, "no Sobol table set on the context");
    SRV_HIP(hipSetDevice(s->ctx->device));
    SRV_TRY(admit(s));
    // plan: round-robin sample passes within the path budget
    uint32_t n_passes = 0;
    uint64_t paths = 0;
    std::vector<uint32_t> take(s->active.size(), 0);
    for (bool progress = true; progress;) {
        progress = false;
        for (size_t k = 0; k < s->active.size(); ++k) {
            Request *r = s->active[k];
            const uint64_t px = static_cast<uint64_t>(r->req.width) * r->req.height;
            const uint32_t left = r->req.spp - (r->next_sample - 1) - take[k];
            if (left == 0 || paths + px > s->max_paths) continue;
            ++take[k];
            paths += px;
            progress = true;
        }
    }
    // plain passes grouped by scene (groups in order of first appearance, passes in admission
    // order within a group); 16-path units numbered across the step in that order
    uint32_t n_groups = 0;
    std::vector<int> group_of_scene(s->n_scenes, -1);
    std::vector<std::vector<size_t>> members;
    for (size_t k = 0; k < s->active.size(); ++k) {
        Request *r = s->active[k];
        if (!take[k] || r->req.mode == NB_REQUEST_CACHED) continue;
        int &gi = group_of_scene[r->req.scene];
        if (gi < 0) { gi = static_cast<int>(n_groups++); members.emplace_back(); }
        members[gi].push_back(k);
    }
    uint32_t units = 0, path_sum = 0, pixel_sum = 0;
    for (uint32_t gi = 0; gi < n_groups; ++gi) {
        ServeGroup &G = s->h_groups[gi];
        G = ServeGroup{};
        G.scene = s->active[members[gi][0]]->req.scene;
        G.unit_begin = units;
        G.pass_begin = n_passes;
        for (size_t k : members[gi]) {
            Request *r = s->active[k];
            ServePass &p = s->h_passes[n_passes++];
            p = ServePass{};
            p.slot = static_cast<uint32_t>(r->slot);
            p.first_sample = r->next_sample;
            p.count = take[k];
            p.pixels = r->req.width * r->req.height;
            p.path_base = path_sum;
            p.pixel_base = pixel_sum;
            p.unit_base = units;
            path_sum += p.count * p.pixels;   // within max_paths (the plan above)
            pixel_sum += p.pixels;
            units += p.count * (p.pixels / 16u);
        }
        G.unit_end = units;
        G.pass_end = n_passes;
    }
    bool any = false;
    if (n_passes > 0) {
        ServeStep &st = *s->h_step;
        st.scenes = s->d_scenes;
        st.requests = s->d_requests;
        st.passes = s->d_passes;
        st.groups = s->d_groups;
        st.n_passes = n_passes;
        st.n_groups = n_groups;
        st.n_units = units;
        st.n_pixels = pixel_sum;
        st.results = s->results;
        st.sobol = s->ctx->sobol;
        st.sobol_entries = s->ctx->sobol_entries;
        st.sobol_stride = s->ctx->sobol_stride;
        st.work = s->work;
        st.counters = s->counters;
        st.regen_min_idle = serve_min_idle();
        st.inner_quorum = serve_inner_quorum();
        st.grab_units = serve_grab_units();
        SRV_HIP(hipMemcpyAsync(s->d_requests, s->h_requests, kMaxSlots * sizeof(ServeRequestDev), hipMemcpyHostToDevice, s->stream));
        SRV_HIP(hipMemcpyAsync(s->d_passes, s->h_passes, n_passes * sizeof(ServePass), hipMemcpyHostToDevice, s->stream));
        SRV_HIP(hipMemcpyAsync(s->d_groups, s->h_groups, n_groups * sizeof(ServeGroup), hipMemcpyHostToDevice, s->stream));
        SRV_HIP(hipMemcpyAsync(s->d_step, s->h_step, sizeof(ServeStep), hipMemcpyHostToDevice, s->stream));
        SRV_HIP(hipEventRecord(s->ev0, s->stream));
        SRV_HIP(hipGraphLaunch(s->graph, s->stream));
        SRV_HIP(hipEventRecord(s->ev1, s->stream));
        SRV_HIP(hipEventSynchronize(s->ev1));
        float ms = 0;
        hipEventElapsedTime(&ms, s->ev0, s->ev1);
        s->busy += ms * 1e-3;
        any = true;
    }
    // cached requests: their sample passes, one cache-assisted frame per sample, in admission
    // order, on the context's stream (the caches' stream)
    bool cached = false;
    for (size_t k = 0; k < s->active.size(); ++k) {
        Request *r = s->active[k];
        if (!take[k] || r->req.mode != NB_REQUEST_CACHED) continue;
        nb_context *ctx = s->ctx;
        if (!cached) {
            SRV_HIP(hipStreamSynchronize(s->stream));   // admission zeroed the targets on the server stream
            SRV_HIP(hipEventRecord(ctx->ev_region_start, ctx->stream));
        }
        cached = true;
        const ServeSceneDev &sc = s->h_scenes[r->req.scene];
        for (uint32_t j = 0; j < take[k]; ++j) {
            nb_globals g;
            g.sample = r->next_sample + j;
            g.weight = 1.0f / static_cast<float>(g.sample);
            g.bounces = r->req.bounces;
            g.contribution_factor = r->req.contribution_factor;
            SRV_TRY(cache_render_frame(ctx, sc.scene, sc.env, r->target, r->req.width, r->req.height,
                                       r->req.camera.clip_to_world, g, s->caches[r->req.scene], r->req.train_paths,
                                       r->req.lr, nullptr));
        }
        s->cached_samples += static_cast<uint64_t>(take[k]) * r->req.width * r->req.height;
    }
    if (cached) {
        SRV_HIP(hipEventRecord(s->ctx->ev_region_stop, s->ctx->stream));
        SRV_HIP(hipEventSynchronize(s->ctx->ev_region_stop));
        float ms = 0;
        hipEventElapsedTime(&ms, s->ctx->ev_region_start, s->ctx->ev_region_stop);
        s->busy += ms * 1e-3;
        s->cached_busy += ms * 1e-3;
        any = true;
    }
    if (any) {
        ++s->steps;
        const double t = s->now();
        for (size_t k = 0; k < s->active.size(); ++k) s->active[k]->next_sample += take[k];
        std::vector<Request *> still;
        for (Request *r : s->active) {
            if (r->next_sample > r->req.spp) {
                r->done = t;
                s->slots[r->slot] = nullptr;
                s->finished[r->id] = r;
                s->unpolled.push_back(nb_completion{r->id, r->arrival, r->admit, r->done});
                ++s->completed;
            } else {
                still.push_back(r);
            }
        }
        s->active.swap(still);
    }
    if (n_active) *n_active = static_cast<uint32_t>(s->active.size());
    if (n_queued) *n_queued = static_cast<uint32_t>(s->queued.size());
    return NB_OK;
}

nb_status nb_server_poll(nb_server *s, nb_completion *out, uint32_t cap, uint32_t *n) {
    if (!s || !n || (cap && !out)) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t k = std::min<uint32_t>(cap, static_cast<uint32_t>(s->unpolled.size()));
    for (uint32_t i = 0; i < k; ++i) out[i] = s->unpolled[i];
    s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + k);
    *n = k;
    return NB_OK;
}

nb_status nb_server_read(nb_server *s, uint64_t id, float *rgba) {
    if (!s || !rgba) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return set_error(NB_ERR_INVALID_ARGUMENT, "request %llu is not completed or was released", static_cast<unsigned long long>(id));
    SRV_HIP(hipSetDevice(s->ctx->device));
    SRV_HIP(hipMemcpy(rgba, it->second->target, it->second->bytes, hipMemcpyDeviceToHost));
    return NB_OK;
}

nb_status nb_server_release(nb_server *s, uint64_t id) {
    if (!s) return set_error(NB_ERR_INVALID_ARGUMENT, "null server");
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return set_error(NB_ERR_INVALID_ARGUMENT, "request %llu is not completed or was released", static_cast<unsigned long long>(id));
    SRV_HIP(hipSetDevice(s->ctx->device));
    hipFree(it->second->target);
    s->in_use -= it->second->bytes;
    delete it->second;
    s->finished.erase(it);
    return NB_OK;
}

nb_status nb_server_now(nb_server *s, double *t) {
    if (!s || !t) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *t = s->now();
    return NB_OK;
}

nb_status nb_server_stats_get(nb_server *s, nb_server_stats *out) {
    if (!s || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    SRV_HIP(hipSetDevice(s->ctx->device));
    SRV_HIP(hipStreamSynchronize(s->stream));
    std::vector<unsigned long long> v(kCounterBytes / sizeof(unsigned long long));
    unsigned long long sum[kCounterWords] = {};
    SRV_HIP(hipMemcpy(v.data(), s->counters, kCounterBytes, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < kCounterSlots; ++k)
        for (uint32_t i = 0; i < kCounterWords; ++i) sum[i] += v[k * kCounterStride + i];
    std::memset(out, 0, sizeof(*out));
    out->steps = s->steps;
    out->rays = sum[0]; out->n_aabb = sum[1]; out->n_tri = sum[2]; out->n_inst = sum[3]; out->samples = sum[4];
    out->hits = sum[5];
    out->bytes_in_use = s->in_use;
    out->bytes_budget = s->budget;
    out->active = static_cast<uint32_t>(s->active.size());
    out->cached_samples = s->cached_samples;
    out->cached_busy_s = s->cached_busy;
    out->queued = static_cast<uint32_t>(s->queued.size());
    out->completed = s->completed;
    out->gpu_busy_s = s->busy;
    return NB_OK;
}

}  // extern "C"
