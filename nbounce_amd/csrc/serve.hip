// Serving kernels: one step renders the sample passes the server has planned for every admitted
// request, for requests that use different scenes, cameras, resolutions and sample counts, as
// reset -> serve_pt_kernel -> accumulate in one captured HIP graph; the step is described at a
// fixed device address, so the same graph is replayed for every step.
//
// serve_pt_kernel is the progressive renderer's persistent kernel (ray-level regeneration, wave64
// service and traverse phases; render_pt_kernel in trace.hip) over the step's paths (one per
// sample and pixel) in 16-path work units, grouped by scene.  A wave serves one scene at a time,
// so the scene's arrays stay wave-uniform and the traversal is the renderer's; waves start spread
// over the scenes in proportion to their units, take units from their scene's counter and move
// on to another scene with units left once their own is exhausted and their lanes have drained.
// A path's radiance goes to its slot of the results, and the accumulation mixes each pixel's
// samples into the request's image in sample order with weight 1/sample: every sample is a pure
// function of (pixel, Sobol row, scene), so a request's image is that of nb_render_frames of its
// samples alone, whatever shares the step.
#include "pt_device.h"

namespace nb {
namespace {
using namespace dev;

constexpr int kServeK = 16;
constexpr uint32_t kNoGroup = 0xFFFFFFFFu;

// The step's arrays are global memory, but their pointers are read from the step description, so
// the compiler cannot tell; reads through these pointers are global loads instead of flat ones.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gp(const T *p) {
    return (const __attribute__((address_space(1))) T *)(p);
}

__global__ void serve_reset_kernel(const ServeStep *step) {
    const ServeStep &S = *step;
    for (uint32_t g = threadIdx.x; g < S.n_groups; g += blockDim.x) S.work[g * kWorkStride] = 0u;
}

// The group holding step unit u (groups cover [0, n_units) in order).
__device__ __forceinline__ uint32_t group_of_unit(const ServeStep &S, uint32_t u) {
    uint32_t g = 0;
    while (g + 1u < S.n_groups && gp(S.groups + g)->unit_end <= u) ++g;
    return g;
}

// The pass holding step unit u of group G (its passes are ordered by unit_base).
__device__ __forceinline__ uint32_t pass_of_unit(const ServeStep &S, uint32_t g, uint32_t u) {
    uint32_t lo = gp(S.groups + g)->pass_begin, hi = gp(S.groups + g)->pass_end;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (gp(S.passes + mid)->unit_base <= u) lo = mid; else hi = mid;
    }
    return lo;
}

// A scene's array pointers as wave-uniform (scalar) values.  They are read from the step
// description through vector loads; only an explicit broadcast lets the traversal's loads take a
// scalar base plus a 32-bit offset instead of a 64-bit address per lane (124 -> 110 VGPRs).
template <typename T>
__device__ __forceinline__ const T *uniform_ptr(const T *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return reinterpret_cast<const T *>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ DevScene uniform_scene(const DevScene &d) {
    DevScene u = d;
    u.pairs = uniform_ptr(d.pairs);
    u.tris = uniform_ptr(d.tris);
    u.inst = uniform_ptr(d.inst);
    u.ishade = uniform_ptr(d.ishade);
    u.vshade = uniform_ptr(d.vshade);
    u.tri_leaf_count = uniform_ptr(d.tri_leaf_count);
    u.inst_leaf_count = uniform_ptr(d.inst_leaf_count);
    u.tlas_root = __builtin_amdgcn_readfirstlane(d.tlas_root);
    u.n_pairs = __builtin_amdgcn_readfirstlane(d.n_pairs);
    return u;
}

// Held to kWaves waves per SIMD.  The server runs 5 (96 VGPRs, 31 dwords spilled): over the
// serving stream (20 QPS, 200 mixed requests) the steps' kernel time is 1393-1395 ms against
// 1401 ms at 4 waves.  The sample-parallel nb_render_frames runs 4 (no spills): 256x256 x 16
// spp 2183-2186 Mrays/s against 2107 at 5 waves.  The renderer's kernel, the same traversal,
// loses 11 % at 5 waves.
template <int K, int kWaves>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves))) void serve_pt_kernel(const ServeStep *step) {
    __shared__ unsigned long long s_stack[K * 64];
    const ServeStep &S = *step;
    if (S.n_units == 0u || S.n_groups == 0u) return;
    const uint32_t lane = threadIdx.x;
    SpillArray spill;
    Stack<K> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};

    // the wave's group and scene (wave-uniform)
    uint32_t g = __builtin_amdgcn_readfirstlane(
        group_of_unit(S, static_cast<uint32_t>(static_cast<uint64_t>(blockIdx.x) * S.n_units / gridDim.x)));
    DevScene scene = uniform_scene(S.scenes[gp(S.groups + g)->scene].scene);
    PathConst pc{};
    pc.scene = scene;
    pc.env = S.scenes[gp(S.groups + g)->scene].env;
    pc.sobol = SobolView{S.sobol, S.sobol_entries, S.sobol_stride};

    // wave-uniform path stream: step positions [pos, end) (unit * 16 + path of the unit)
    uint32_t pos = 0, end = 0;
    bool more = true;   // the wave's group may still hold units
    uint32_t mode = kIdle;
    uint32_t x = 0, y = 0, bounce = 0, samples_done = 0;
    uint32_t path = 0;   // the lane's path id
    // every per-lane value starts defined: idle lanes take part in the wave's ballots
    Ray ray{};
    TraceState ts{};
    F3 thr = f3(1.0f, 1.0f, 1.0f);
    for (;;) {
        bool start = false;
        // ---- service: shade finished hits
        if (mode == kReady) {
            const Hit &h = ts.hit;
            ++c.rays;
            F3 out = f3(0.0f, 0.0f, 0.0f);
            StepResult r = kDone;
            if (h.t == kNoHit) {
                out = mul(thr, env_sample(pc.env, ray.d));
                r = kDone;
            } else {
                ++c.hits;
                Vertex dummy{};
                pc.shift = hash4f(x, y, x, y);
                r = shade<false>(pc, bounce, h, ray, thr, out, dummy);
                if (r == kContinue && bounce == (pc.bounces & 0xFFFFu)) { r = kDone; out = f3(0.0f, 0.0f, 0.0f); }
            }
            if (r == kContinue) {
                ++bounce;
                mode = kTracing;
                start = true;
            } else {
                *(__attribute__((address_space(1))) g_v4f *)(S.results + path) = g_v4f{out.x, out.y, out.z, 1.0f};
                ++samples_done;
                mode = kIdle;
            }
        }
        // ---- service: idle lanes take the next paths of the group's units, in lane order
        const unsigned long long idle = __ballot(mode == kIdle);
        if (idle != 0ull && more) {
            const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
            const uint32_t avail = end - pos;
            uint32_t nb = 0, ne = 0;   // units taken now, for the lanes the current ones do not cover
            if (avail < n_idle) {
                uint32_t b0 = 0, b1 = 0;
                if (lane == 0) {
                    const auto *G = gp(S.groups + g);
                    const uint32_t units = G->unit_end - G->unit_begin;
                    const uint32_t i = atomicAdd(S.work + g * kWorkStride, S.grab_units);
                    if (i < units) {
                        b0 = (G->unit_begin + i) * 16u;
                        b1 = (G->unit_begin + min(i + S.grab_units, units)) * 16u;
                    }
                }
                nb = __shfl(b0, 0, 64);
                ne = __shfl(b1, 0, 64);
                if (nb == ne) more = false;
            }
            if (mode == kIdle) {
                const uint32_t rank = static_cast<uint32_t>(__popcll(idle & ((1ull << lane) - 1ull)));
                uint32_t q = 0xFFFFFFFFu;
                if (rank < avail) q = pos + rank;
                else if (rank - avail < ne - nb) q = nb + (rank - avail);
                if (q != 0xFFFFFFFFu) {
                    // unit -> (pass, sample k of the pass, 16 pixels of the request's image)
                    const uint32_t unit = q >> 4;
                    const uint32_t pi = pass_of_unit(S, g, unit);
                    const auto *P = gp(S.passes + pi);
                    const auto *R = gp(S.requests + P->slot);
                    const uint32_t width = R->width, pixels = P->pixels;
                    const uint32_t units_per_sample = pixels / 16u, local = unit - P->unit_base;
                    const uint32_t k = local / units_per_sample;
                    const uint32_t pq = (local - k * units_per_sample) * 16u + (q & 15u);
                    const uint32_t tiles_x = width / 8u, tile = pq >> 6, kk = pq & 63u;
                    x = (tile % tiles_x) * 8u + (kk & 7u);
                    y = (tile / tiles_x) * 8u + (kk >> 3);
                    path = P->path_base + k * pixels + y * width + x;
                    pc.sample = P->first_sample + k;
                    pc.bounces = R->bounces | (R->rr_first << 16);
                    pc.contribution_factor = R->contribution_factor;
                    const float4 jit = sobol_fetch(pc.sobol, pc.sample * S.sobol_stride, hash4f(x, y, x, y));
                    ray = camera_ray(R->c2w, width, R->height, x, y, jit.x, jit.y);
                    thr = f3(1.0f, 1.0f, 1.0f);
                    bounce = 0;
                    mode = kTracing;
                    start = true;
                }
            }
            if (n_idle <= avail) {
                pos += n_idle;
            } else if (ne > nb) {
                pos = nb + min(n_idle - avail, ne - nb);
                end = ne;
            } else {
                pos = end;
            }
        }
        // ---- start the traversal of new rays
        if (start) trace_begin<K>(scene, ray, ts, st);
        if (__ballot(mode != kIdle) == 0ull) {
            // every lane idle and the group exhausted: move on to a group with units left, if any
            uint32_t next = kNoGroup;
            for (uint32_t k = 1; k <= S.n_groups && next == kNoGroup; ++k) {
                const uint32_t h = (g + k) % S.n_groups;
                const auto *G = gp(S.groups + h);
                const uint32_t taken = __hip_atomic_load(S.work + h * kWorkStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (taken < G->unit_end - G->unit_begin) next = h;
            }
            next = __builtin_amdgcn_readfirstlane(next);
            if (next == kNoGroup) break;
            g = next;
            scene = uniform_scene(S.scenes[gp(S.groups + g)->scene].scene);
            pc.scene = scene;
            pc.env = S.scenes[gp(S.groups + g)->scene].env;
            more = true;
            pos = end = 0;
            continue;
        }
        // ---- traverse
        const bool more_pixels = more;
        for (;;) {
            // inner-node visits and plain pops of every lane that has one, until no more than
            // inner_quorum lanes do: lanes that reached a leaf, a BLAS exit or the end wait, so
            // the leaf and instance code below runs for many lanes at once
            for (;;) {
                const bool inner = mode == kTracing && inner_pending<K>(ts, st);
                if (static_cast<uint32_t>(__popcll(__ballot(inner))) <= S.inner_quorum) break;
                if (inner) inner_step<K>(scene, ts, st, c);
            }
            if (mode == kTracing && trace_step<K>(scene, ray, ts, st, c)) mode = kReady;
            const uint32_t n_trace = static_cast<uint32_t>(__popcll(__ballot(mode == kTracing)));
            const uint32_t n_ready = static_cast<uint32_t>(__popcll(__ballot(mode == kReady)));
            const uint32_t n_wait = n_ready + (more_pixels ? 64u - n_trace - n_ready : 0u);
            if (n_trace == 0 || n_wait >= S.regen_min_idle || (n_trace < S.regen_min_idle && n_wait > 0)) break;
        }
    }
    flush_counters(S.counters, c, samples_done);
}

// color = mix(color, s, 1/sample) for the step's samples of each pixel, in sample order.
__global__ __launch_bounds__(256) void serve_accumulate_kernel(const ServeStep *step) {
    const ServeStep &S = *step;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < S.n_pixels; j += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = S.n_passes;   // the pass holding accumulation pixel j
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S.passes[mid].pixel_base <= j) lo = mid; else hi = mid;
        }
        const ServePass &ps = S.passes[lo];
        const ServeRequestDev &R = S.requests[ps.slot];
        const uint32_t pixel = j - ps.pixel_base;
        float4 *px = R.target + pixel;
        const float4 prev = *px;
        F3 col = f3(prev.x, prev.y, prev.z);
        for (uint32_t k = 0; k < ps.count; ++k) {
            const float w = 1.0f / static_cast<float>(ps.first_sample + k);
            const float4 s = S.results[ps.path_base + k * ps.pixels + pixel];
            // mix(color, s, weight) = color * (1 - weight) + s * weight
            col = f3(col.x * (1.0f - w) + s.x * w, col.y * (1.0f - w) + s.y * w, col.z * (1.0f - w) + s.z * w);
        }
        *px = make_float4(col.x, col.y, col.z, 1.0f);
    }
}

nb_status check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return NB_OK;
}

// Resident waves of serve_pt_kernel on the device (occupancy x CUs), a multiple of 8.
template <int kWaves>
uint32_t serve_grid() {
    static const uint32_t grid = [] {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_pt_kernel<kServeK, kWaves>, 64, 0) != hipSuccess)
            return 0u;
        const uint32_t g = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu < 1 ? 1 : per_cu) / 8u * 8u;
        return g < 8u ? 8u : g;
    }();
    return grid;
}

}  // namespace

nb_status serve_launch_reset(const ServeStep *step, hipStream_t s) {
    hipLaunchKernelGGL(serve_reset_kernel, dim3(1), dim3(64), 0, s, step);
    return check_launch("serve reset");
}

nb_status serve_launch_accumulate(const ServeStep *step, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_accumulate_kernel, dim3(grid), dim3(256), 0, s, step);
    return check_launch("serve accumulate");
}

nb_status serve_launch_step(const ServeStep *step, hipStream_t s, uint32_t waves_per_simd) {
    const uint32_t grid = waves_per_simd == 4u ? serve_grid<4>() : serve_grid<5>();
    if (grid == 0u) return set_error(NB_ERR_HIP, "serve kernel occupancy query failed");
    if (waves_per_simd == 4u) hipLaunchKernelGGL((serve_pt_kernel<kServeK, 4>), dim3(grid), dim3(64), 0, s, step);
    else hipLaunchKernelGGL((serve_pt_kernel<kServeK, 5>), dim3(grid), dim3(64), 0, s, step);
    return check_launch("serve step");
}

}  // namespace nb
