// Serving kernels: a wavefront form of the integrator for batches of requests that use different
// scenes, cameras, resolutions and sample counts.  One step runs
//   reset -> generate -> bounce(0) -> ... -> bounce(max_bounces) -> accumulate
// as one captured HIP graph; every kernel reads the step from a fixed device address, so the
// same graph is replayed for every step.  The bounce kernels are persistent: a fixed grid of
// wave64 workers claims 64 paths at a time from the level's queue, advances each by one bounce
// (bounce_step, the same arithmetic as the progressive megakernel) and compacts the survivors
// into the next level's queue with a wave ballot and one atomic per wave.  Each sample is a pure
// function of (pixel, Sobol row, scene), and the accumulation applies a request's samples in
// sample order, so a request's image does not depend on what else shares its steps.
#include "pt_device.h"

namespace nb {
namespace {
using namespace dev;

constexpr int kServeK = 16;

// Path record: q0 = (o.xyz, path id)  q1 = (d.xyz, bounce)  q2 = (thr.xyz, pass index)
__device__ __forceinline__ uint32_t find_pass(const ServeStep &S, uint32_t path) {
    uint32_t lo = 0, hi = S.n_passes;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (S.passes[mid].path_base <= path) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t find_pass_pixel(const ServeStep &S, uint32_t pix) {
    uint32_t lo = 0, hi = S.n_passes;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (S.passes[mid].pixel_base <= pix) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void serve_reset_kernel(const ServeStep *step) {
    const ServeStep &S = *step;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * (kServeMaxBounces + 2); i += gridDim.x * blockDim.x)
        S.level[i] = 0u;
}

__global__ __launch_bounds__(256) void serve_generate_kernel(const ServeStep *step) {
    const ServeStep &S = *step;
    if (blockIdx.x == 0 && threadIdx.x == 0) S.level[0] = S.n_paths;
    const SobolView sob{S.sobol, S.sobol_entries, S.sobol_stride};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < S.n_paths; i += gridDim.x * blockDim.x) {
        const uint32_t pi = find_pass(S, i);
        const ServePass &ps = S.passes[pi];
        const ServeRequestDev &R = S.requests[ps.slot];
        const uint32_t local = i - ps.path_base;
        const uint32_t k = local / ps.pixels, pixel = local % ps.pixels;
        const uint32_t x = pixel % R.width, y = pixel / R.width;
        const uint32_t sample = ps.first_sample + k;
        const float4 shift = hash4f(x, y, x, y);
        float4 jit = sobol_fetch(sob, sample * S.sobol_stride, shift);
        Ray r = camera_ray(R.c2w, R.width, R.height, x, y, jit.x, jit.y);
        float4 *q = S.queue[0] + 4 * static_cast<size_t>(i);
        q[0] = make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(i));
        q[1] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(0u));
        q[2] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(pi));
    }
}

__global__ __launch_bounds__(64) void serve_bounce_kernel(const ServeStep *step, uint32_t level) {
    __shared__ unsigned long long s_stack[kServeK * 64];
    const ServeStep &S = *step;
    const uint32_t lane = threadIdx.x;
    const uint32_t n_in = S.level[2 * level];
    const float4 *in = S.queue[level & 1u];
    float4 *out = S.queue[(level + 1u) & 1u];
    SpillArray spill;
    Stack<kServeK> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};
    uint32_t samples_done = 0;
    const SobolView sob{S.sobol, S.sobol_entries, S.sobol_stride};
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&S.level[2 * level + 1], 64u);
        base = __shfl(base, 0, 64);
        if (base >= n_in) break;
        const uint32_t i = base + lane;
        bool alive = false;
        float4 q0, q1, q2;
        if (i < n_in) {
            const float4 *q = in + 4 * static_cast<size_t>(i);
            q0 = q[0]; q1 = q[1]; q2 = q[2];
            const uint32_t path = __float_as_uint(q0.w), bounce = __float_as_uint(q1.w), pi = __float_as_uint(q2.w);
            const ServePass &ps = S.passes[pi];
            const ServeRequestDev &R = S.requests[ps.slot];
            const ServeSceneDev &SC = S.scenes[R.scene];
            const uint32_t local = path - ps.path_base;
            const uint32_t pixel = local % ps.pixels;
            const uint32_t x = pixel % R.width, y = pixel / R.width;
            PathConst pc;
            pc.scene = SC.scene;
            pc.env = SC.env;
            pc.sobol = sob;
            pc.sample = ps.first_sample + local / ps.pixels;
            pc.bounces = R.bounces;
            pc.contribution_factor = R.contribution_factor;
            pc.shift = hash4f(x, y, x, y);
            Ray ray;
            ray.o = f3(q0.x, q0.y, q0.z);
            ray.d = f3(q1.x, q1.y, q1.z);
            ray.inv = rcp3(ray.d);
            F3 thr = f3(q2.x, q2.y, q2.z), rad = f3(0.0f, 0.0f, 0.0f);
            StepResult r = bounce_step<kServeK>(pc, bounce, ray, thr, rad, st, c);
            if (r == kContinue && bounce + 1u > R.bounces) { r = kDone; rad = f3(0.0f, 0.0f, 0.0f); }
            if (r == kDone) {
                S.results[path] = make_float4(rad.x, rad.y, rad.z, 1.0f);
                ++samples_done;
            } else {
                alive = true;
                q0 = make_float4(ray.o.x, ray.o.y, ray.o.z, q0.w);
                q1 = make_float4(ray.d.x, ray.d.y, ray.d.z, __uint_as_float(bounce + 1u));
                q2 = make_float4(thr.x, thr.y, thr.z, q2.w);
            }
        }
        // wave64 ballot compaction of the surviving paths into the next level's queue
        const unsigned long long mask = __ballot(alive);
        const uint32_t n_alive = static_cast<uint32_t>(__popcll(mask));
        uint32_t obase = 0;
        if (lane == 0 && n_alive) obase = atomicAdd(&S.level[2 * (level + 1u)], n_alive);
        obase = __shfl(obase, 0, 64);
        if (alive) {
            const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
            float4 *q = out + 4 * static_cast<size_t>(obase + rank);
            q[0] = q0; q[1] = q1; q[2] = q2;
        }
    }
    flush_counters(S.counters, c, samples_done);
}

// color = mix(color, s, 1/sample) for the step's samples of each pixel, in sample order.
__global__ __launch_bounds__(256) void serve_accumulate_kernel(const ServeStep *step) {
    const ServeStep &S = *step;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < S.n_pixels; j += gridDim.x * blockDim.x) {
        const ServePass &ps = S.passes[find_pass_pixel(S, j)];
        const ServeRequestDev &R = S.requests[ps.slot];
        const uint32_t pixel = j - ps.pixel_base;
        float4 *px = R.target + pixel;
        float4 prev = *px;
        F3 col = f3(prev.x, prev.y, prev.z);
        for (uint32_t k = 0; k < ps.count; ++k) {
            const uint32_t sample = ps.first_sample + k;
            const float w = 1.0f / static_cast<float>(sample);
            const float4 s = S.results[ps.path_base + k * ps.pixels + pixel];
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

}  // namespace

nb_status serve_launch_reset(const ServeStep *step, uint32_t, hipStream_t s) {
    hipLaunchKernelGGL(serve_reset_kernel, dim3(1), dim3(128), 0, s, step);
    return check_launch("serve reset");
}

nb_status serve_launch_generate(const ServeStep *step, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_generate_kernel, dim3(grid), dim3(256), 0, s, step);
    return check_launch("serve generate");
}

nb_status serve_launch_bounce(const ServeStep *step, uint32_t level, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_bounce_kernel, dim3(grid), dim3(64), 0, s, step, level);
    return check_launch("serve bounce");
}

nb_status serve_launch_accumulate(const ServeStep *step, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_accumulate_kernel, dim3(grid), dim3(256), 0, s, step);
    return check_launch("serve accumulate");
}

}  // namespace nb
