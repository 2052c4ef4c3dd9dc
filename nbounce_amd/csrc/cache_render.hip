// Cache-assisted path tracing: paths end at their second shading point (after the first bounce)
// or, after near-mirror bounces, at a later one, and take the radiance the MLP predicts there;
// training paths are full unbiased paths of the plain integrator whose suffix radiance becomes
// the MLP's training targets.
#include "pt_device.h"

namespace nb {
namespace {
using namespace dev;

// LDS stack entries per lane: the query kernel's paths end after two or three traces, and an
// 8-entry stack (4 KB per wave) with 80 VGPRs lets 6 waves share a SIMD (16 entries: 5 waves,
// 11 % slower); the training paths are full paths and keep 16.  Held to 72 VGPRs for 7 waves,
// the query kernel is 2.5 % faster alone, but the training-path kernel beside it on the side
// stream slows down and a training frame takes 5 % longer.
constexpr int kQueryK = 8;
constexpr int kTrainK = 16;

// Shading-point features of a non-emissive hit (the same expressions as the shading in
// bounce_step): position, outgoing direction, normal, roughness, albedo, metallic.
__device__ __forceinline__ bool hit_vertex(const DevScene &S, const Ray &ray, const Hit &h, float *raw, F3 &emit) {
    const float4 *sh = S.ishade + 6 * static_cast<size_t>(h.inst);
    float4 s0 = sh[0], s1 = sh[1], s2 = sh[2], s3 = sh[3], s4 = sh[4], s5 = sh[5];
    F3 albedo = f3(s3.w, s4.w, s5.w);
    if (s2.w > 0.0f) { emit = albedo; return false; }
    const float4 *tr = S.tris + 3 * static_cast<size_t>(h.tri);
    uint32_t i0 = __float_as_uint(tr[0].w), i1 = __float_as_uint(tr[1].w), i2 = __float_as_uint(tr[2].w);
    float4 n0 = S.vshade[2 * i0], n1 = S.vshade[2 * i1], n2 = S.vshade[2 * i2];
    float bx = 1.0f - h.u - h.v, by = h.u, bz = h.v;
    F3 pos = add(ray.o, scale(h.t, ray.d));
    F3 ln = add(add(scale(bx, xyz(n0)), scale(by, xyz(n1))), scale(bz, xyz(n2)));
    F3 nrm = f3((s0.x * ln.x + s0.y * ln.y) + s0.z * ln.z, (s1.x * ln.x + s1.y * ln.y) + s1.z * ln.z,
                (s2.x * ln.x + s2.y * ln.y) + s2.z * ln.z);
    F3 n = normalize(nrm);
    F3 wo = normalize(f3(-ray.d.x, -ray.d.y, -ray.d.z));
    raw[0] = pos.x; raw[1] = pos.y; raw[2] = pos.z;
    raw[3] = wo.x; raw[4] = wo.y; raw[5] = wo.z;
    raw[6] = n.x; raw[7] = n.y; raw[8] = n.z;
    raw[9] = s0.w;
    raw[10] = albedo.x; raw[11] = albedo.y; raw[12] = albedo.z;
    raw[13] = s1.w;
    raw[14] = 0.0f; raw[15] = 0.0f;
    return true;
}

__device__ __forceinline__ void store_raw(float *dst, const float *raw) {
    float4 *d = reinterpret_cast<float4 *>(dst);
    for (int q = 0; q < 4; ++q) d[q] = make_float4(raw[4 * q], raw[4 * q + 1], raw[4 * q + 2], raw[4 * q + 3]);
}

__device__ __forceinline__ uint32_t wave_prefix_alloc(bool want, uint32_t *counter, uint32_t &slot) {
    const unsigned long long mask = __ballot(want);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total = static_cast<uint32_t>(__popcll(mask));
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    slot = base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    return total;
}


// One sample per pixel: the path runs as in sample_rendering_eq until the first shading point
// after a bounce whose lobe was not a near mirror (GGX alpha >= mirror_alpha) or until bounce
// max_cache_bounce; there it stops and queues a cache query with its throughput.  Radiance met
// before (sky, emitters) goes to `direct`.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void cache_query_kernel(CacheFrame p) {
    __shared__ unsigned long long s_stack[kQueryK * 64];
    const uint32_t tile = (blockIdx.x & 7u) * p.tiles_per_xcd + (blockIdx.x >> 3);
    if (tile >= p.n_tiles) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % p.tiles_x) * 8u + (lane & 7u), y = (tile / p.tiles_x) * 8u + (lane >> 3);
    const uint32_t pixel = y * p.width + x;
    SpillArray spill;
    Stack<kQueryK> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};
    PathConst pc{};
    pc.scene = p.scene;
    pc.env = p.env;
    pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
    pc.sample = p.sample;
    pc.bounces = p.bounces;
    pc.contribution_factor = p.contribution_factor;
    pc.shift = hash4f(x, y, x, y);
    float4 jit = sobol_fetch(pc.sobol, p.sample * p.sobol_stride, pc.shift);
    Ray ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
    F3 thr = f3(1.0f, 1.0f, 1.0f), out = f3(0.0f, 0.0f, 0.0f);
    float prev_alpha = 0.0f;
    bool query = false;
    float raw[kMlpRaw];
    for (uint32_t bounce = 0; bounce <= p.bounces; ++bounce) {
        if (bounce >= 1 && (prev_alpha >= p.mirror_alpha || bounce >= p.max_cache_bounce)) {
            Hit h{};
            trace_counted<kQueryK>(pc, ray, h, st, c);
            if (h.t == kNoHit) { out = mul(thr, env_sample(pc.env, ray.d)); break; }
            F3 emit;
            if (!hit_vertex(pc.scene, ray, h, raw, emit)) { out = mul(thr, emit); break; }
            query = true;
            break;
        }
        Vertex v;
        if (bounce_step<kQueryK, true>(pc, bounce, ray, thr, out, st, c, &v) == kDone) break;
        prev_alpha = v.lobe_alpha;
    }
    p.direct[pixel] = make_float4(out.x, out.y, out.z, 1.0f);
    uint32_t slot;
    wave_prefix_alloc(query, p.q_count, slot);
    if (query) {
        store_raw(p.q_raw + static_cast<size_t>(slot) * kMlpRaw, raw);
        reinterpret_cast<float4 *>(p.q_thr)[slot] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(pixel));
    }
    flush_counters(p.counters, c, 1);
}

// Full unbiased paths at a spread of pixels; every shading point (up to three per path) becomes
// a training record whose target is the path's final radiance divided by the throughput that
// arrived at that point, i.e. the one-sample estimate of the point's outgoing radiance.
__global__ __launch_bounds__(64) void cache_train_paths_kernel(CacheFrame p) {
    __shared__ unsigned long long s_stack[kTrainK * 64];
    const uint32_t lane = threadIdx.x;
    SpillArray spill;
    Stack<kTrainK> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};
    const uint32_t npix = p.width * p.height;
    const uint32_t stride = npix / p.train_paths > 0 ? npix / p.train_paths : 1u;
    for (uint32_t base = blockIdx.x * 64u; base < p.train_paths; base += gridDim.x * 64u) {
        const uint32_t i = base + lane;
        const bool active = i < p.train_paths;
        constexpr int kRec = 3;
        float rec[kRec][kMlpRaw];
        F3 tin[kRec];
        int nrec = 0;
        F3 out = f3(0.0f, 0.0f, 0.0f);
        if (active) {
            const uint32_t pixel = (i * stride + p.train_offset) % npix;
            const uint32_t x = pixel % p.width, y = pixel / p.width;
            PathConst pc{};
            pc.scene = p.scene;
            pc.env = p.env;
            pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
            pc.sample = p.train_sample;
            pc.bounces = p.bounces;
            pc.contribution_factor = p.contribution_factor;
            pc.shift = hash4f(x, y, x, y);
            float4 jit = sobol_fetch(pc.sobol, p.train_sample * p.sobol_stride, pc.shift);
            Ray ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
            F3 thr = f3(1.0f, 1.0f, 1.0f);
            for (uint32_t bounce = 0; bounce <= p.bounces; ++bounce) {
                Hit h{};
                trace_counted<kTrainK>(pc, ray, h, st, c);
                if (h.t == kNoHit) { out = mul(thr, env_sample(pc.env, ray.d)); break; }
                F3 emit;
                float raw[kMlpRaw];
                const bool surface = hit_vertex(pc.scene, ray, h, raw, emit);
                if (surface && nrec < kRec) {
                    for (int q = 0; q < kMlpRaw; ++q) rec[nrec][q] = raw[q];
                    tin[nrec] = thr;
                    ++nrec;
                }
                Vertex v;
                if (shade<false>(pc, bounce, h, ray, thr, out, v) == kDone) break;
                if (bounce == p.bounces) out = f3(0.0f, 0.0f, 0.0f);
            }
        }
        for (int k = 0; k < kRec; ++k) {
            uint32_t slot;
            const bool want = active && k < nrec;
            wave_prefix_alloc(want, p.t_count, slot);
            if (want && slot < p.t_capacity) {
                store_raw(p.t_raw + static_cast<size_t>(slot) * kMlpRaw, rec[k]);
                const F3 t = tin[k];
                float *tg = p.t_target + 3 * static_cast<size_t>(slot);
                tg[0] = t.x > 0.0f ? out.x / t.x : 0.0f;
                tg[1] = t.y > 0.0f ? out.y / t.y : 0.0f;
                tg[2] = t.z > 0.0f ? out.z / t.z : 0.0f;
            }
        }
    }
    flush_counters(p.counters, c, 0);
}

// direct[pixel] += throughput * max(cache radiance, 0) for every query.
__global__ void cache_combine_kernel(const float *q_thr, const uint32_t *q_count, const float *rgb, float4 *direct) {
    const uint32_t n = *q_count;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const float4 t = reinterpret_cast<const float4 *>(q_thr)[q];
        const uint32_t pixel = __float_as_uint(t.w);
        const float *l = rgb + 3 * static_cast<size_t>(q);
        float4 d = direct[pixel];
        d.x = d.x + t.x * fmaxf(l[0], 0.0f);
        d.y = d.y + t.y * fmaxf(l[1], 0.0f);
        d.z = d.z + t.z * fmaxf(l[2], 0.0f);
        direct[pixel] = d;
    }
}

// target = mix(target, sample, weight) (pathtracing.wgsl:258) with the sample from `direct`.
__global__ void cache_accumulate_kernel(const float4 *direct, float4 *target, uint32_t n, float w) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 s = direct[i];
        float4 c = w > 0.0f ? target[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        target[i] = make_float4(c.x * (1.0f - w) + s.x * w, c.y * (1.0f - w) + s.y * w, c.z * (1.0f - w) + s.z * w, 1.0f);
    }
}

__global__ void zero_u32_kernel(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

nb_status ok_or(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return NB_OK;
}

}  // namespace

nb_status cache_launch_query(const CacheFrame &p, hipStream_t s) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p.q_count, 1u);
    hipLaunchKernelGGL(cache_query_kernel, dim3(8u * p.tiles_per_xcd), dim3(64), 0, s, p);
    return ok_or("cache query");
}

nb_status cache_launch_finish(const CacheFrame &p, const MlpDev &m, float *rgb, hipStream_t s) {
    nb_status st = mlp_launch_infer_count(m, p.q_raw, p.q_count, p.width * p.height, rgb, s);
    if (st != NB_OK) return st;
    hipLaunchKernelGGL(cache_combine_kernel, dim3(1024), dim3(256), 0, s, p.q_thr, p.q_count, rgb, p.direct);
    st = ok_or("cache combine");
    if (st != NB_OK) return st;
    hipLaunchKernelGGL(cache_accumulate_kernel, dim3(1024), dim3(256), 0, s, p.direct, p.target, p.width * p.height,
                       p.weight);
    return ok_or("cache accumulate");
}

nb_status cache_launch_train_paths(const CacheFrame &p, hipStream_t s) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p.t_count, 1u);
    const uint32_t blocks = (p.train_paths + 63) / 64;
    hipLaunchKernelGGL(cache_train_paths_kernel, dim3(blocks < 4096 ? blocks : 4096), dim3(64), 0, s, p);
    return ok_or("cache training paths");
}

}  // namespace nb
