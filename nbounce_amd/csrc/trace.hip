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
#include <cstring>

namespace nb {
namespace {
using namespace dev;


// XCD-aware tile order: workgroups are dealt round-robin over the 8 XCDs, so block b runs on
// XCD b % 8; give each XCD a contiguous run of tiles so neighbouring (coherent) tiles share an L2.
__device__ __forceinline__ uint32_t tile_of_block(const RenderParams &p) {
    return (blockIdx.x & 7u) * p.tiles_per_xcd + (blockIdx.x >> 3);
}

// main, pathtracing.wgsl:243-260, for frames first_sample .. first_sample + n_frames - 1.
template <int K>
__global__ __launch_bounds__(64) void render_kernel(RenderParams p) {
    __shared__ unsigned long long s_stack[K * 64];
    const uint32_t tile = tile_of_block(p);
    if (tile >= p.n_tiles) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % p.tiles_x) * 8u + (lane & 7u), y = (tile / p.tiles_x) * 8u + (lane >> 3);
    SpillArray spill;
    Stack<K> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};
    float4 *px = p.target + static_cast<size_t>(y) * p.width + x;
    PathConst pc{};
    pc.scene = p.scene;
    pc.env = p.env;
    pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
    pc.bounces = p.bounces | (p.rr_first << 16);
    pc.contribution_factor = p.contribution_factor;
    pc.shift = hash4f(x, y, x, y);
    F3 col = f3(0.0f, 0.0f, 0.0f), s = f3(0.0f, 0.0f, 0.0f);
    for (uint32_t f = 0; f < p.n_frames; ++f) {
        const uint32_t sample = p.first_sample + f;
        const float w = (f == 0 && !p.weights_from_sample) ? p.first_weight : 1.0f / static_cast<float>(sample);
        if (f == 0 && w > 0.0f && !p.sample_out) { float4 prev = *px; col = f3(prev.x, prev.y, prev.z); }
        pc.sample = sample;
        float4 jit = sobol_fetch(pc.sobol, sample * p.sobol_stride, pc.shift);
        Ray ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
        s = sample_path<K>(pc, ray, st, c);
        // mix(color, s, weight) = color * (1 - weight) + s * weight
        col = f3(col.x * (1.0f - w) + s.x * w, col.y * (1.0f - w) + s.y * w, col.z * (1.0f - w) + s.z * w);
    }
    if (p.sample_out) p.sample_out[static_cast<size_t>(y) * p.width + x] = make_float4(s.x, s.y, s.z, 1.0f);
    else *px = make_float4(col.x, col.y, col.z, 1.0f);
    flush_counters(p.counters, c, p.n_frames);
}

// Work units of the persistent kernel: 16 pixels (two rows of an 8x8 tile), counted per XCD
// band of tiles in p.work.  A wave takes the next units of its own XCD's band, then of the other
// bands (work stealing at the tail): p.grab_units units at a time while more than p.tail_units
// units remain in the band, one unit at a time after that.  Device-scope atomics are costly
// (about 30 ns each on one address), so grabs are coarse.  Sets the wave-uniform pixel range
// [begin, end) in tile order (pixel i is pixel i % 64 of tile i / 64); empty when all bands are
// exhausted.
__device__ __forceinline__ void grab_units(const RenderParams &p, uint32_t xcd, uint32_t &begin, uint32_t &end) {
    uint32_t b0 = 0, b1 = 0;
    if (threadIdx.x == 0) {
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t b = (xcd + k) & 7u;
            const uint32_t lo = b * p.tiles_per_xcd;
            if (lo >= p.n_tiles) continue;
            const uint32_t units = min(p.tiles_per_xcd, p.n_tiles - lo) * 4u;
            uint32_t *counter = p.work + b * kWorkStride;
            const uint32_t seen = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (seen >= units) continue;
            const uint32_t want = units - seen > p.tail_units ? p.grab_units : 1u;
            const uint32_t i = atomicAdd(counter, want);
            if (i < units) {
                b0 = (lo * 4u + i) * 16u;
                b1 = (lo * 4u + min(i + want, units)) * 16u;
                break;
            }
        }
    }
    begin = __shfl(b0, 0, 64);
    end = __shfl(b1, 0, 64);
}

// Persistent form of render_kernel with ray-level regeneration.  A fixed grid of wave64 workers
// claims 8x8 tiles from per-XCD counters.  The traversal of each lane is resumable
// (trace_step), and the wave alternates two phases:
//   traverse: every lane with a ray in flight takes one traversal step, repeatedly, until enough
//             lanes (min_service) have a finished hit or no work, or fewer lanes than that are
//             still traversing and some are waiting;
//   service:  lanes with a finished hit shade it (next ray, or path end and accumulation) and
//             idle lanes take the next pixels of the wave's tile stream (ballot + prefix count),
//             then start their traversal.
// So lanes are not held idle while the longest traversal or path of the wave finishes.  A lane
// keeps its pixel for all n_frames frames, in order: every sample and the mix chain are those of
// render_kernel, and the image does not depend on which lane or wave renders a pixel.
template <int K, bool kCount>
__global__ __launch_bounds__(64) void render_pt_kernel(RenderParams p) {
    __shared__ unsigned long long s_stack[K * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t xcd = blockIdx.x & 7u;
    const uint32_t min_service = p.regen_min_idle;
    SpillArray spill;
    Stack<K> st;
    st.init(s_stack, lane, spill);
    LaneCounters c{0, 0, 0, 0, 0};
    PathConst pc{};
    pc.scene = p.scene;
    pc.env = p.env;
    pc.sobol = SobolView{p.sobol, p.sobol_entries, p.sobol_stride};
    pc.bounces = p.bounces | (p.rr_first << 16);
    pc.contribution_factor = p.contribution_factor;

    // wave-uniform pixel stream: pixels [pos, end) of the wave's current units come next; the
    // first units are this wave's static share of its XCD's band
    uint32_t pos = 0, end = 0;
    bool more = true;   // the bands may still hold units
    {
        const uint32_t lo = xcd * p.tiles_per_xcd;
        if (lo < p.n_tiles) {
            const uint32_t units = min(p.tiles_per_xcd, p.n_tiles - lo) * 4u;
            const uint32_t u0 = (blockIdx.x >> 3) * p.grab_units;
            if (u0 < units) {
                pos = (lo * 4u + u0) * 16u;
                end = (lo * 4u + min(u0 + p.grab_units, units)) * 16u;
            }
        }
    }

    enum : uint32_t { kIdle = 0, kTracing = 1, kReady = 2 };
    uint32_t mode = kIdle;
    uint32_t x = 0, y = 0, f = 0, bounce = 0, frames_done = 0;
    uint32_t px_rays = 0;    // rays traced for the lane's pixel in this launch (its unit's cost)
    // every per-lane value starts defined: idle lanes take part in the wave's ballots, and an
    // undefined operand there would make the whole ballot undefined
    Ray ray{};               // the path's current ray (world space)
    TraceState ts{};
    F3 thr = f3(1.0f, 1.0f, 1.0f), col = f3(0.0f, 0.0f, 0.0f);
    float w = 0.0f;
    for (;;) {
        bool start = false;      // this lane begins a new ray in this service phase
        // ---- service: shade finished hits
        if (mode == kReady) {
            const Hit &h = ts.hit;
            ++c.rays;
            ++px_rays;
            F3 out = f3(0.0f, 0.0f, 0.0f);
            StepResult r = kDone;
            if (h.t == kNoHit) {
                out = mul(thr, env_sample(pc.env, ray.d));
                r = kDone;
            } else {
                ++c.hits;
                Vertex dummy{};
                pc.shift = hash4f(x, y, x, y);   // recomputed: not kept live across traversal
                r = shade<false>(pc, bounce, h, ray, thr, out, dummy);
                if (r == kContinue && bounce == p.bounces) { r = kDone; out = f3(0.0f, 0.0f, 0.0f); }
            }
            if (r == kContinue) {
                ++bounce;
                mode = kTracing;
                start = true;
            } else {
                // mix(color, s, weight) = color * (1 - weight) + s * weight
                col = f3(col.x * (1.0f - w) + out.x * w, col.y * (1.0f - w) + out.y * w, col.z * (1.0f - w) + out.z * w);
                bounce = 0;
                ++frames_done;
                if (++f == p.n_frames) {
                    const size_t i = static_cast<size_t>(y) * p.width + x;
                    if (p.sample_out) p.sample_out[i] = make_float4(out.x, out.y, out.z, 1.0f);
                    else p.target[i] = make_float4(col.x, col.y, col.z, 1.0f);
                    mode = kIdle;
                    if (p.unit_cost) {
                        const uint32_t tile = (y >> 3) * p.tiles_x + (x >> 3), k = ((y & 7u) << 3) | (x & 7u);
                        __hip_atomic_fetch_add(p.unit_cost + tile * 4u + (k >> 4), px_rays, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    }
                } else {
                    w = 1.0f / static_cast<float>(p.first_sample + f);
                    mode = kTracing;       // next frame of the same pixel: camera ray below
                    start = true;
                }
            }
        }
        // ---- service: idle lanes take the next pixels of the stream, in lane order
        const unsigned long long idle = __ballot(mode == kIdle);
        if (idle != 0ull && more) {
            const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
            const uint32_t avail = end - pos;
            uint32_t nb = 0, ne = 0;   // units taken now, for the lanes the current ones do not cover
            if (avail < n_idle) {
                grab_units(p, xcd, nb, ne);
                if (nb == ne) more = false;
            }
            if (mode == kIdle) {
                const uint32_t rank = static_cast<uint32_t>(__popcll(idle & ((1ull << lane) - 1ull)));
                uint32_t pix = 0xFFFFFFFFu;
                if (rank < avail) pix = pos + rank;
                else if (rank - avail < ne - nb) pix = nb + (rank - avail);
                if (pix != 0xFFFFFFFFu) {
                    // stream position -> work unit (cost order within the band, when given)
                    const uint32_t unit = p.unit_order ? p.unit_order[pix >> 4] : (pix >> 4);
                    const uint32_t q = unit * 16u + (pix & 15u);
                    const uint32_t tile = q >> 6, k = q & 63u;
                    x = (tile % p.tiles_x) * 8u + (k & 7u);
                    y = (tile / p.tiles_x) * 8u + (k >> 3);
                    mode = kTracing;
                    start = true;
                    f = 0;
                    bounce = 0;
                    px_rays = 0;
                    col = f3(0.0f, 0.0f, 0.0f);
                    // first frame: the caller's weight (nb_render) or 1/sample (nb_render_frames)
                    w = !p.weights_from_sample ? p.first_weight : 1.0f / static_cast<float>(p.first_sample);
                    if (w > 0.0f && !p.sample_out) {
                        const float4 prev = p.target[static_cast<size_t>(y) * p.width + x];
                        col = f3(prev.x, prev.y, prev.z);
                    }
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
        // ---- start the traversal of new rays (camera rays at bounce 0)
        if (start) {
            if (bounce == 0) {
                pc.sample = p.first_sample + f;
                const float4 jit = sobol_fetch(pc.sobol, pc.sample * p.sobol_stride, hash4f(x, y, x, y));
                ray = camera_ray(p.c2w, p.width, p.height, x, y, jit.x, jit.y);
                thr = f3(1.0f, 1.0f, 1.0f);
            }
            trace_begin<K>(p.scene, ray, ts, st);
        }
        // every lane idle after the refill: the stream is exhausted
        if (__ballot(mode != kIdle) == 0ull) break;
        // ---- traverse
        // (Written out here and in serve_pt_kernel: the same loop moved into a shared
        // force-inlined function compiles to code that renders wrong pixels.)
        const bool more_pixels = more;
        for (;;) {
            // inner-node visits and plain pops of every lane that has one, until no more than
            // inner_quorum lanes do: lanes that reached a leaf, a BLAS exit or the end wait, so
            // the leaf and instance code below runs for many lanes at once
            for (;;) {
                const bool inner = mode == kTracing && inner_pending<K>(ts, st);
                if (static_cast<uint32_t>(__popcll(__ballot(inner))) <= p.inner_quorum) break;
                if (inner) inner_step<K>(p.scene, ts, st, c);
            }
            if (mode == kTracing && trace_step<K>(p.scene, ray, ts, st, c)) mode = kReady;
            const uint32_t n_trace = static_cast<uint32_t>(__popcll(__ballot(mode == kTracing)));
            const uint32_t n_ready = static_cast<uint32_t>(__popcll(__ballot(mode == kReady)));
            const uint32_t n_wait = n_ready + (more_pixels ? 64u - n_trace - n_ready : 0u);
            if (n_trace == 0 || n_wait >= min_service || (n_trace < min_service && n_wait > 0)) break;
        }
    }
    if (kCount) flush_counters(p.counters, c, frames_done);   // else the counts are dead code
}

// One step of the accumulation (main, pathtracing.wgsl:243-260) for a frame rendered into a
// sample buffer: the same expression, in the same binary32 order, as the render kernels' mix.
__global__ __launch_bounds__(256) void mix_samples_kernel(float4 *target, const float4 *samples, uint32_t n, float w) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 s = samples[i];
    F3 col = f3(0.0f, 0.0f, 0.0f);
    if (w > 0.0f) {
        const float4 prev = target[i];
        col = f3(prev.x, prev.y, prev.z);
    }
    // mix(color, s, weight) = color * (1 - weight) + s * weight
    col = f3(col.x * (1.0f - w) + s.x * w, col.y * (1.0f - w) + s.y * w, col.z * (1.0f - w) + s.z * w);
    target[i] = make_float4(col.x, col.y, col.z, 1.0f);
}

__global__ void reset_work_kernel(uint32_t *work, uint32_t first) {
    if (threadIdx.x < 8) work[threadIdx.x * kWorkStride] = first;
}

template <int K>
__global__ __launch_bounds__(64) void primary_hits_kernel(RenderParams p) {
    __shared__ unsigned long long s_stack[K * 64];
    const uint32_t tile = tile_of_block(p);
    if (tile >= p.n_tiles) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % p.tiles_x) * 8u + (lane & 7u), y = (tile / p.tiles_x) * 8u + (lane >> 3);
    SpillArray spill;
    Stack<K> st;
    st.init(s_stack, lane, spill);
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

// Workgroups of the persistent kernel: as many one-wave workgroups as the CUs hold at once
// (the occupancy the compiled kernel allows), a multiple of 8 (XCDs), so every wave's first,
// statically assigned work units start at once.  NB_REGEN_WAVES_PER_CU overrides the per-CU count.
template <int K, bool kCount>
static uint32_t regen_grid() {
    static uint32_t grid = [] {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, render_pt_kernel<K, kCount>, 64, 0) != hipSuccess)
            return 0u;
        if (const char *e = std::getenv("NB_REGEN_WAVES_PER_CU")) per_cu = std::atoi(e);
        const uint32_t g = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu < 1 ? 1 : per_cu) / 8u * 8u;
        return g < 8u ? 8u : g;
    }();
    return grid;
}

// NB_INNER_QUORUM=n (0..64): the traversal phase runs inner-node steps alone while more than n
// lanes have one (64: never).  28 measured best on the instanced 1080p workload (24..32 within 1%).
// With the refitted TLAS, pipelined: 28 2232-2237 Mrays/s, 24 2209-2210, 32 2234-2241 (within
// the noise of 28), 36 2210-2212, 40 2141-2143.
static uint32_t inner_quorum() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_INNER_QUORUM");
        const int n = e ? std::atoi(e) : 28;
        return static_cast<uint32_t>(n < 0 ? 0 : (n > 64 ? 64 : n));
    }();
    return v;
}

// NB_TAIL_UNITS=n: single 16-pixel units are handed out once a band has at most n left.
static uint32_t tail_units() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_TAIL_UNITS");
        return e ? static_cast<uint32_t>(std::atoi(e)) : 0u;
    }();
    return v;
}

// NB_GRAB_UNITS=n: 16-pixel units per grab before the tail.  Defaults: 4 for a one-frame launch
// alone (finer grabs shorten its tail); 8 for a pipelined frame, whose tail the next frame's
// paths fill, so coarser grabs (fewer atomics) pay: instanced 1080p, pipelined, 1950 Mrays/s
// with 4, 1977 with 6, 1986-1988 with 8, 1968 with 12, 1944 with 16; 2 for a launch of several
// frames, where a unit carries every frame of its pixels (the atomics are amortised and the
// last units set a long tail): 64 frames 2202-2247 with 2 against 2169-2183 with 4 (1: 2225-
// 2232), 8 frames 2157-2168 against 2130-2171.
static uint32_t grab_units_setting(bool pipelined, uint32_t n_frames) {
    static const int v = [] {
        const char *e = std::getenv("NB_GRAB_UNITS");
        return e ? std::atoi(e) : 0;
    }();
    const int n = v > 0 ? v : (pipelined ? 8 : (n_frames > 1 ? 2 : 4));
    return static_cast<uint32_t>(n > 256 ? 256 : n);
}

// NB_REGEN_MIN_IDLE=n (1..64): a wave leaves its traversal phase to shade and refill once at
// least n lanes wait (finished hit, or idle while pixels remain).  24 measured best on the
// instanced 1080p workload with the refitted TLAS (pipelined 2227-2237 Mrays/s against
// 2206-2212 with 16, 2221 with 28, 2182 with 32; one launch after another 1765 against 1744;
// 8 frames per launch 2136 against 2097); 16 was best before the refit.
static uint32_t regen_min_idle() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_REGEN_MIN_IDLE");
        const int n = e ? std::atoi(e) : 24;
        return static_cast<uint32_t>(n < 1 ? 1 : (n > 64 ? 64 : n));
    }();
    return v;
}

template <int K>
static nb_status launch_regen(const RenderParams &p_in, hipStream_t s) {
    RenderParams p = p_in;
    p.regen_min_idle = regen_min_idle();
    p.inner_quorum = inner_quorum();
    p.tail_units = tail_units();
    p.grab_units = grab_units_setting(p.sample_out != nullptr, p.n_frames);
    const uint32_t grid = p.counters ? regen_grid<K, true>() : regen_grid<K, false>();
    if (grid == 0) return set_error(NB_ERR_HIP, "render: device attribute or occupancy query failed");
    // every wave's first grab_units units of its XCD's band are assigned statically (no atomic
    // at the start); the counters start after them
    const uint32_t first = (grid / 8u) * p.grab_units;
    hipLaunchKernelGGL(reset_work_kernel, dim3(1), dim3(64), 0, s, p.work, first);
    if (p.counters) hipLaunchKernelGGL((render_pt_kernel<K, true>), dim3(grid), dim3(64), 0, s, p);
    else hipLaunchKernelGGL((render_pt_kernel<K, false>), dim3(grid), dim3(64), 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "render launch: %s", hipGetErrorString(e));
    return NB_OK;
}

// NB_RENDER=tile selects the one-tile-per-wave kernel (no regeneration) for comparisons; the
// persistent kernel is the default.  Both render the same image.  Read at every launch, so a
// process can compare the two.
static bool use_tile_kernel() {
    const char *e = std::getenv("NB_RENDER");
    return e && std::strcmp(e, "tile") == 0;
}

nb_status launch_render(const RenderParams &p, void *stream) {
    if (use_tile_kernel() || !p.work) return launch_with_k<RenderKernel>(p, stream, "render");
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (lds_entries()) {
    case 4: return launch_regen<4>(p, s);
    case 8: return launch_regen<8>(p, s);
    case 32: return launch_regen<32>(p, s);
    default: return launch_regen<16>(p, s);
    }
}

nb_status launch_primary_hits(const RenderParams &p, void *stream) {
    return launch_with_k<HitsKernel>(p, stream, "primary-hit");
}

nb_status launch_mix_samples(float4 *target, const float4 *samples, uint32_t n, float weight, hipStream_t s) {
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(mix_samples_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, target, samples, n, weight);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "mix launch: %s", hipGetErrorString(e));
    return NB_OK;
}

}  // namespace nb
