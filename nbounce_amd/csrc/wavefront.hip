// Wavefront form of the progressive renderer: the integrator of pathtracing.wgsl:161-272 split
// into a lean traversal kernel and a shading kernel that meet through ray queues in HBM, one
// bounce level at a time (Laine, Karras, Aila 2013, "Megakernels considered harmful").
//
// Per frame:  generate (camera rays of every pixel, tile order)
//             then for level 0..bounces:  trace(level) -> shade(level)
// * trace: persistent wave64 workers with a resumable traversal (trace_step), each owning a
//   contiguous range of the level's queue.  A lane that finishes its ray writes the hit; once
//   enough lanes of the wave are idle they take the next rays of the range together (ballot +
//   prefix count), so lanes are not held idle by the longest traversal of the wave.  No atomics:
//   device-scope atomics on one address serialise across the XCDs.
// * shade: one lane per queued ray: miss -> environment, emitter, or the bounce of shade()
//   (the megakernel's arithmetic); survivors are compacted into the next level's queue with a
//   wave ballot and one atomic per wave; finished paths apply mix(prev, s, weight) to their
//   pixel.  Every pixel has one path per frame, so the image equals the megakernel's bit for bit.
#include "pt_device.h"

#include <cstdlib>

namespace nb {
namespace {
using namespace dev;

constexpr int kWfK = 8;   // LDS stack entries per lane in the traversal kernel

// path id i -> pixel: ids run over 8x8 tiles in row-major tile order, row-major inside a tile
__device__ __forceinline__ void path_pixel(const WfParams &p, uint32_t i, uint32_t &x, uint32_t &y) {
    const uint32_t tile = i >> 6, k = i & 63u;
    x = (tile % p.tiles_x) * 8u + (k & 7u);
    y = (tile / p.tiles_x) * 8u + (k >> 3);
}

__global__ __launch_bounds__(256) void wf_generate_kernel(WfParams p) {
    const uint32_t n = p.width * p.height;
    if (blockIdx.x == 0) {
        for (uint32_t j = threadIdx.x; j < 2u * (kWfMaxLevels + 1u); j += blockDim.x) p.counts[j] = 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            p.counts[0] = n;
            if (p.counters) atomicAdd(p.counters + 4, static_cast<unsigned long long>(n));   // slot 0
        }
    }
    const SobolView sob{p.sobol, p.sobol_entries, p.sobol_stride};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t x, y;
        path_pixel(p, i, x, y);
        const float4 shift = hash4f(x, y, x, y);
        const float4 jit = sobol_fetch(sob, p.sample * p.sobol_stride, shift);
        const Ray r = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
        float4 *q = p.queue[0] + 2 * static_cast<size_t>(i);
        q[0] = make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(i));
        q[1] = make_float4(r.d.x, r.d.y, r.d.z, 0.0f);
        p.path[i] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
    }
}

template <int K, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void wf_trace_kernel(WfParams p, uint32_t level) {
    __shared__ uint32_t s_ref[K * 64];
    __shared__ float s_dist[K * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t n = p.counts[2 * level];
    const float4 *queue = p.queue[level & 1u];
    const uint32_t min_fetch = p.min_fetch;
    // this wave's contiguous range of the queue; the ranges of one XCD's waves are adjacent, so
    // coherent neighbouring rays share that XCD's L2 (workgroup b runs on XCD b % 8)
    const uint32_t G = gridDim.x;   // a multiple of 8
    const uint32_t range = (blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3);
    const uint32_t hi = static_cast<uint32_t>(static_cast<uint64_t>(range + 1u) * n / G);
    uint32_t cursor = static_cast<uint32_t>(static_cast<uint64_t>(range) * n / G);   // wave-uniform
    SpillArray spill;
    Stack<K> st;
    st.spill = spill;
    st.ref = (lds_u32 *)(s_ref + lane);
    st.dist = (lds_f32 *)(s_dist + lane);
    LaneCounters c{0, 0, 0, 0, 0};
    bool busy = false;
    uint32_t idx = 0;
    Ray world;
    TraceState ts;
    for (;;) {
        const bool more = cursor < hi;
        const unsigned long long idle = __ballot(!busy);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (more && n_idle >= min_fetch) {
            const uint32_t base = cursor;
            cursor = min(hi, cursor + n_idle);
            if (!busy) {
                const uint32_t i = base + static_cast<uint32_t>(__popcll(idle & ((1ull << lane) - 1ull)));
                if (i < hi) {
                    const float4 *q = queue + 2 * static_cast<size_t>(i);
                    const float4 q0 = q[0], q1 = q[1];
                    world.o = f3(q0.x, q0.y, q0.z);
                    world.d = f3(q1.x, q1.y, q1.z);
                    world.inv = rcp3(world.d);
                    trace_begin<K>(p.scene, world, ts, st);
                    busy = true;
                    idx = i;
                }
            }
        }
        if (__ballot(busy) == 0ull) {
            if (cursor >= hi) break;
            continue;   // every lane idle: the next iteration fetches (64 >= min_fetch)
        }
        // traverse until enough lanes are idle to refill (all of them once the range is drained)
        const uint32_t stop_idle = cursor < hi ? min_fetch : 64u;
        for (;;) {
            if (busy && trace_step<K>(p.scene, world, ts, st, c)) {
                busy = false;
                ++c.rays;
                if (ts.hit.t != kNoHit) ++c.hits;
                p.hit[idx] = make_uint4(__float_as_uint(ts.hit.t), __float_as_uint(ts.hit.u),
                                        __float_as_uint(ts.hit.v), ts.hit.tri);
                p.hit_inst[idx] = ts.hit.inst;
            }
            const unsigned long long b = __ballot(busy);
            if (b == 0ull || 64u - static_cast<uint32_t>(__popcll(b)) >= stop_idle) break;
        }
    }
    flush_counters(p.counters, c, 0);
}

__global__ __launch_bounds__(256) void wf_shade_kernel(WfParams p, uint32_t level) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t n = p.counts[2 * level];
    const float4 *in = p.queue[level & 1u];
    float4 *out = p.queue[(level + 1u) & 1u];
    uint32_t *out_count = p.counts + 2 * (level + 1u);
    PathConst pc;
    pc.scene = p.scene;
    pc.env = p.env;
    pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
    pc.sample = p.sample;
    pc.bounces = p.bounces;
    pc.contribution_factor = p.contribution_factor;
    for (uint32_t base = (blockIdx.x * 4u + wave) * 64u; base < n; base += gridDim.x * 256u) {
        const uint32_t i = base + lane;
        bool alive = false;
        uint32_t path = 0, bounce = 0;
        Ray ray;
        F3 thr = f3(0.0f, 0.0f, 0.0f);
        if (i < n) {
            const float4 *q = in + 2 * static_cast<size_t>(i);
            const float4 q0 = q[0], q1 = q[1];
            path = __float_as_uint(q0.w);
            ray.o = f3(q0.x, q0.y, q0.z);
            ray.d = f3(q1.x, q1.y, q1.z);
            ray.inv = rcp3(ray.d);
            const float4 ps = p.path[path];
            thr = f3(ps.x, ps.y, ps.z);
            bounce = __float_as_uint(ps.w);
            uint32_t x, y;
            path_pixel(p, path, x, y);
            const uint4 hb = p.hit[i];
            Hit h;
            h.t = __uint_as_float(hb.x);
            h.u = __uint_as_float(hb.y);
            h.v = __uint_as_float(hb.z);
            h.tri = hb.w;
            h.inst = p.hit_inst[i];
            F3 rad = f3(0.0f, 0.0f, 0.0f);
            StepResult r;
            if (h.t == kNoHit) {
                rad = mul(thr, env_sample(pc.env, ray.d));
                r = kDone;
            } else {
                pc.shift = hash4f(x, y, x, y);
                Vertex dummy;
                r = shade<false>(pc, bounce, h, ray, thr, rad, dummy);
                if (r == kContinue && bounce == p.bounces) { r = kDone; rad = f3(0.0f, 0.0f, 0.0f); }
            }
            if (r == kContinue) {
                alive = true;
            } else {
                // mix(color, s, weight) = color * (1 - weight) + s * weight
                float4 *px = p.target + static_cast<size_t>(y) * p.width + x;
                const float w = p.weight;
                F3 col = f3(0.0f, 0.0f, 0.0f);
                if (w > 0.0f) { const float4 prev = *px; col = f3(prev.x, prev.y, prev.z); }
                col = f3(col.x * (1.0f - w) + rad.x * w, col.y * (1.0f - w) + rad.y * w, col.z * (1.0f - w) + rad.z * w);
                *px = make_float4(col.x, col.y, col.z, 1.0f);
            }
        }
        // wave64 ballot compaction of the surviving paths into the next level's queue
        const unsigned long long mask = __ballot(alive);
        const uint32_t n_alive = static_cast<uint32_t>(__popcll(mask));
        uint32_t obase = 0;
        if (lane == 0 && n_alive) obase = atomicAdd(out_count, n_alive);
        obase = __shfl(obase, 0, 64);
        if (alive) {
            const uint32_t j = obase + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
            float4 *q = out + 2 * static_cast<size_t>(j);
            q[0] = make_float4(ray.o.x, ray.o.y, ray.o.z, __uint_as_float(path));
            q[1] = make_float4(ray.d.x, ray.d.y, ray.d.z, 0.0f);
            p.path[path] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(bounce + 1u));
        }
    }
}

nb_status launched(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return NB_OK;
}

uint32_t device_cus() {
    static const uint32_t cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0u;
        return static_cast<uint32_t>(n);
    }();
    return cus;
}

int trace_wpe_setting() {
    static const int v = [] {
        const char *e = std::getenv("NB_WF_WPE");
        return e ? std::atoi(e) : 4;
    }();
    return v;
}

uint32_t min_fetch_setting() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_WF_MIN_FETCH");
        const int n = e ? std::atoi(e) : 16;
        return static_cast<uint32_t>(n < 1 ? 1 : (n > 64 ? 64 : n));
    }();
    return v;
}

}  // namespace

// One frame (p.sample, p.weight) of the wavefront renderer on stream s.
nb_status wavefront_frame(const WfParams &p_in, uint32_t levels, hipStream_t s) {
    WfParams p = p_in;
    p.min_fetch = min_fetch_setting();
    const uint32_t cus = device_cus();
    if (cus == 0) return set_error(NB_ERR_HIP, "device attribute query failed");
    if (levels > kWfMaxLevels) return set_error(NB_ERR_INVALID_ARGUMENT, "at most %u bounce levels", kWfMaxLevels);
    const uint32_t n = p.width * p.height;
    const uint32_t gen_blocks = (n + 255u) / 256u < 4096u ? (n + 255u) / 256u : 4096u;
    hipLaunchKernelGGL(wf_generate_kernel, dim3(gen_blocks), dim3(256), 0, s, p);
    nb_status st = launched("wavefront generate");
    if (st != NB_OK) return st;
    for (uint32_t level = 0; level < levels; ++level) {
        switch (trace_wpe_setting()) {
        case 5: hipLaunchKernelGGL((wf_trace_kernel<kWfK, 5>), dim3(cus * 32u), dim3(64), 0, s, p, level); break;
        case 6: hipLaunchKernelGGL((wf_trace_kernel<kWfK, 6>), dim3(cus * 32u), dim3(64), 0, s, p, level); break;
        case 8: hipLaunchKernelGGL((wf_trace_kernel<kWfK, 8>), dim3(cus * 32u), dim3(64), 0, s, p, level); break;
        default: hipLaunchKernelGGL((wf_trace_kernel<kWfK, 1>), dim3(cus * 32u), dim3(64), 0, s, p, level); break;
        }
        if ((st = launched("wavefront trace")) != NB_OK) return st;
        hipLaunchKernelGGL(wf_shade_kernel, dim3(cus * 8u), dim3(256), 0, s, p, level);
        if ((st = launched("wavefront shade")) != NB_OK) return st;
    }
    return NB_OK;
}

}  // namespace nb
