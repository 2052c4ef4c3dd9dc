// Path-tracing kernels for gfx950 (wave64): the integrator megakernel of
// src/pathtracing/pathtracing.wgsl over the two-level BVH of src/pathtracing/raytracing_sw.wgsl.
//
// Numerics: compiled with -ffp-contract=off and correctly rounded divide / sqrt, every float
// expression is evaluated in binary32 in the reference's order (left to right, vector ops per
// component), so a sample's value depends only on (pixel, Sobol row, scene) and matches a CPU
// evaluation of the same expressions bit for bit.  sin / cos use a Cody–Waite reduction and the
// Cephes single-precision polynomials; pow with integer exponents is written as products.
//
// Traversal: one lane per pixel, an 8x8 pixel tile per wave.  Inner nodes are visited through
// 64-B child-pair records (both child boxes in one fetch).  Each lane keeps the top K entries of
// its traversal stack in LDS (lane-interleaved, so pushes and pops of a wave hit distinct
// banks) and spills older entries to private (scratch) memory.  TLAS and BLAS share the stack;
// entering an instance records the stack height, and the BLAS traversal of that instance ends
// when the stack falls back to it.  The current closest distance is handed into each BLAS,
// which prunes more than the reference's fresh per-instance search but selects the same hit.
#include "pt_device.h"

#include <cstdlib>

namespace nb {
namespace {
using namespace dev;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void flush_counters(const RenderParams &p, const LaneCounters &c, uint32_t samples) {
    uint32_t r = wave_sum(c.rays), a = wave_sum(c.aabb), t = wave_sum(c.tri), i = wave_sum(c.inst), s = wave_sum(samples);
    uint32_t h = wave_sum(c.hits);
    if (threadIdx.x == 0 && p.counters) {
        atomicAdd(p.counters + 0, static_cast<unsigned long long>(r));
        atomicAdd(p.counters + 1, static_cast<unsigned long long>(a));
        atomicAdd(p.counters + 2, static_cast<unsigned long long>(t));
        atomicAdd(p.counters + 3, static_cast<unsigned long long>(i));
        atomicAdd(p.counters + 4, static_cast<unsigned long long>(s));
        atomicAdd(p.counters + 5, static_cast<unsigned long long>(h));
    }
}

// XCD-aware tile order: workgroups are dealt round-robin over the 8 XCDs, so block b runs on
// XCD b % 8; give each XCD a contiguous run of tiles so neighbouring (coherent) tiles share an L2.
__device__ __forceinline__ uint32_t tile_of_block(const RenderParams &p) {
    return (blockIdx.x & 7u) * p.tiles_per_xcd + (blockIdx.x >> 3);
}

// main, pathtracing.wgsl:243-260, for frames first_sample .. first_sample + n_frames - 1.
template <int K>
__global__ __launch_bounds__(64) void render_kernel(RenderParams p) {
    __shared__ uint32_t s_ref[K * 64];
    __shared__ float s_dist[K * 64];
    const uint32_t tile = tile_of_block(p);
    if (tile >= p.n_tiles) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % p.tiles_x) * 8u + (lane & 7u), y = (tile / p.tiles_x) * 8u + (lane >> 3);
    Stack<K> st;
    st.ref = (lds_u32 *)(s_ref + lane);
    st.dist = (lds_f32 *)(s_dist + lane);
    LaneCounters c{0, 0, 0, 0, 0};
    float4 *px = p.target + static_cast<size_t>(y) * p.width + x;
    PathConst pc;
    pc.scene = p.scene;
    pc.env = p.env;
    pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
    pc.bounces = p.bounces;
    pc.contribution_factor = p.contribution_factor;
    pc.shift = hash4f(x, y, x, y);
    F3 col = f3(0.0f, 0.0f, 0.0f);
    for (uint32_t f = 0; f < p.n_frames; ++f) {
        const uint32_t sample = p.first_sample + f;
        const float w = (f == 0 && !p.weights_from_sample) ? p.first_weight : 1.0f / static_cast<float>(sample);
        if (f == 0 && w > 0.0f) { float4 prev = *px; col = f3(prev.x, prev.y, prev.z); }
        pc.sample = sample;
        float4 jit = sobol_fetch(pc.sobol, sample * p.sobol_stride, pc.shift);
        Ray ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
        F3 s = sample_path<K>(pc, ray, st, c);
        // mix(color, s, weight) = color * (1 - weight) + s * weight
        col = f3(col.x * (1.0f - w) + s.x * w, col.y * (1.0f - w) + s.y * w, col.z * (1.0f - w) + s.z * w);
    }
    *px = make_float4(col.x, col.y, col.z, 1.0f);
    flush_counters(p, c, p.n_frames);
}

template <int K>
__global__ __launch_bounds__(64) void primary_hits_kernel(RenderParams p) {
    __shared__ uint32_t s_ref[K * 64];
    __shared__ float s_dist[K * 64];
    const uint32_t tile = tile_of_block(p);
    if (tile >= p.n_tiles) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % p.tiles_x) * 8u + (lane & 7u), y = (tile / p.tiles_x) * 8u + (lane >> 3);
    Stack<K> st;
    st.ref = (lds_u32 *)(s_ref + lane);
    st.dist = (lds_f32 *)(s_dist + lane);
    LaneCounters c{0, 0, 0, 0, 0};
    const float4 shift = hash4f(x, y, x, y);
    float4 jit = sobol_fetch(SobolView{p.sobol, p.sobol_entries, p.sobol_stride}, p.first_sample * p.sobol_stride, shift);
    Ray ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
    Hit h;
    trace<K>(p.scene, ray, h, st, c);
    uint32_t *o = p.hits_out + 5 * (static_cast<size_t>(y) * p.width + x);
    if (h.t == kNoHit) {
        o[0] = 0xFFFFFFFFu; o[1] = 0; o[2] = 0; o[3] = 0;
    } else {
        const float4 *tr = p.scene.tris + 3 * static_cast<size_t>(h.tri);
        o[0] = h.inst;
        o[1] = __float_as_uint(tr[0].w); o[2] = __float_as_uint(tr[1].w); o[3] = __float_as_uint(tr[2].w);
    }
    o[4] = __float_as_uint(h.t);
}

}  // namespace

// Entries of each lane's traversal stack held in LDS (the rest spill to scratch): 16 by default,
// NB_LDS_STACK=4|8|16|32 selects another instantiation.
static int lds_entries() {
    static const int k = [] {
        const char *e = std::getenv("NB_LDS_STACK");
        int v = e ? std::atoi(e) : 16;
        return (v == 4 || v == 8 || v == 16 || v == 32) ? v : 16;
    }();
    return k;
}

template <template <int> class Kern>
static nb_status launch_with_k(const RenderParams &p, void *stream, const char *what) {
    const uint32_t grid = 8u * p.tiles_per_xcd;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (lds_entries()) {
    case 4: hipLaunchKernelGGL(Kern<4>::fn, dim3(grid), dim3(64), 0, s, p); break;
    case 8: hipLaunchKernelGGL(Kern<8>::fn, dim3(grid), dim3(64), 0, s, p); break;
    case 32: hipLaunchKernelGGL(Kern<32>::fn, dim3(grid), dim3(64), 0, s, p); break;
    default: hipLaunchKernelGGL(Kern<16>::fn, dim3(grid), dim3(64), 0, s, p); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return NB_OK;
}

namespace {
template <int K> struct RenderKernel { static constexpr auto fn = render_kernel<K>; };
template <int K> struct HitsKernel { static constexpr auto fn = primary_hits_kernel<K>; };
}  // namespace

nb_status launch_render(const RenderParams &p, void *stream) {
    return launch_with_k<RenderKernel>(p, stream, "render");
}

nb_status launch_primary_hits(const RenderParams &p, void *stream) {
    return launch_with_k<HitsKernel>(p, stream, "primary-hit");
}

}  // namespace nb
