// Cost-ordered work units of the persistent renderer (render_pt_kernel).
//
// The persistent kernel hands out 16-pixel work units of each XCD's band in order.  Once the
// units are all taken, the waves drain the paths still in flight and their lanes empty one by
// one; the last long paths set the end of the launch.  Handing out the expensive units first and
// the cheap ones (pixels whose paths end early, such as the sky) last lets the cheap units fill
// the lanes of waves whose long paths are still running.  A unit's cost is the number of rays its
// pixels traced in the earlier frames of the same view; the order is a permutation within each
// band (the band's L2 keeps its tiles), descending by cost class (the mean rays per unit and
// frame), stable within a class so that neighbouring units stay together.  A pixel's samples do not depend on the order: it changes
// only which wave renders a pixel when.
#include "nb_internal.h"

#include <hipcub/hipcub.hpp>

#include <cstdlib>
#include <cstring>

namespace nb {
namespace {

constexpr uint32_t kBands = 8;   // one per XCD, as in make_params / render_pt_kernel

// a[i] >>= shift (32: zero) for the value a[i] held when it was read: one atomic subtraction of
// the part that goes, so rays that a pipelined path kernel adds meanwhile are kept.
__global__ void shift_kernel(uint32_t *a, uint32_t n, uint32_t shift) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = __hip_atomic_load(a + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t drop = shift >= 32u ? v : v - (v >> shift);
    if (drop) __hip_atomic_fetch_sub(a + i, drop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sort key of unit i: its mean rays per frame >> shift, at most 255; value: the unit id.
__global__ void order_keys_kernel(const uint32_t *cost, uint32_t n, uint32_t frames, uint32_t shift,
                                  uint8_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = __hip_atomic_load(cost + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    keys[i] = static_cast<uint8_t>(min((c / frames) >> shift, 255u));
    vals[i] = i;
}

// NB_UNIT_ORDER=0 turns the ordering off (units in band order, no costs gathered).  Read at every
// launch, so a process can compare the two.
bool order_enabled() {
    const char *e = std::getenv("NB_UNIT_ORDER");
    return !(e && std::strcmp(e, "0") == 0);
}

// NB_UNIT_ORDER_SHIFT=s: cost classes are 2^s rays per unit and frame wide.  The default 0 (one
// ray per unit and frame, 1/16 ray per pixel) measured best on the instanced 1080p workload
// (1592-1602 Mrays/s; s=2: 1570, s=3: 1518, s=4: 1465, band order: 1427).
uint32_t order_shift() {
    static const uint32_t v = [] {
        const char *e = std::getenv("NB_UNIT_ORDER_SHIFT");
        const int s = e ? std::atoi(e) : 0;
        return static_cast<uint32_t>(s < 0 ? 0 : (s > 8 ? 8 : s));
    }();
    return v;
}

}  // namespace

#define NB_ORDER_HIP(call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return set_error(e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "%s: %s",    \
                             #call, hipGetErrorString(e_));                                             \
    } while (0)

nb_status unit_order_alloc(nb_target *t) {
    const uint32_t n_tiles = (t->width / 8u) * (t->height / 8u);
    const uint32_t per_band = (n_tiles + kBands - 1u) / kBands;
    const uint32_t n = n_tiles * 4u;
    int offsets[kBands + 1];
    for (uint32_t b = 0; b <= kBands; ++b) offsets[b] = static_cast<int>(std::min(b * per_band, n_tiles) * 4u);
    size_t tmp = 0;
    NB_ORDER_HIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        nullptr, tmp, static_cast<const uint8_t *>(nullptr), static_cast<uint8_t *>(nullptr),
        static_cast<const uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr), static_cast<int>(n),
        static_cast<int>(kBands), static_cast<const int *>(nullptr), static_cast<const int *>(nullptr) + 1, 0, 8,
        static_cast<hipStream_t>(nullptr)));
    t->n_units = n;
    t->sort_tmp_bytes = tmp;
    NB_ORDER_HIP(hipMalloc(&t->unit_cost, n * sizeof(uint32_t)));
    NB_ORDER_HIP(hipMalloc(&t->unit_order[0], n * sizeof(uint32_t)));
    NB_ORDER_HIP(hipMalloc(&t->unit_order[1], n * sizeof(uint32_t)));
    NB_ORDER_HIP(hipMalloc(&t->sort_vals, n * sizeof(uint32_t)));
    NB_ORDER_HIP(hipMalloc(&t->sort_keys, 2u * n));
    NB_ORDER_HIP(hipMalloc(&t->sort_offsets, sizeof(offsets)));
    NB_ORDER_HIP(hipMalloc(&t->sort_tmp, tmp ? tmp : 1u));
    NB_ORDER_HIP(hipMemcpy(t->sort_offsets, offsets, sizeof(offsets), hipMemcpyHostToDevice));
    NB_ORDER_HIP(hipMemset(t->unit_cost, 0, n * sizeof(uint32_t)));
    t->cost_frames = 0;
    t->view_frames = 0;
    t->order_valid = false;
    t->order_cur = 0;
    t->order_pending = false;
    t->rebuild_wanted = false;
    t->cost_scene = nullptr;
    return NB_OK;
}

void unit_order_free(nb_target *t) {
    hipFree(t->unit_cost);
    hipFree(t->unit_order[0]);
    hipFree(t->unit_order[1]);
    hipFree(t->sort_vals);
    hipFree(t->sort_keys);
    hipFree(t->sort_offsets);
    hipFree(t->sort_tmp);
    t->unit_cost = t->sort_vals = nullptr;
    t->unit_order[0] = t->unit_order[1] = nullptr;
    t->sort_keys = nullptr;
    t->sort_offsets = nullptr;
    t->sort_tmp = nullptr;
}

nb_status unit_order_prepare(nb_target *t, const nb_scene *scene, const nb_camera *cam, const nb_globals *g,
                             RenderParams &p, hipStream_t s, uint64_t visible_sorts) {
    p.unit_order = nullptr;
    p.unit_cost = nullptr;
    if (!order_enabled() || !t->unit_cost || t->n_units == 0) return NB_OK;
    if (t->order_pending && t->pending_sort <= visible_sorts) {
        // the coming launch follows the sort that wrote the other buffer: use it from now on
        t->order_cur ^= 1u;
        t->order_pending = false;
        t->order_valid = true;
    }
    const bool same_paths = scene == t->cost_scene && g->bounces == t->cost_bounces &&
                            g->contribution_factor == t->cost_contribution;
    if (!same_paths || std::memcmp(cam, &t->cost_camera, sizeof(nb_camera)) != 0) {
        // Another scene or path length limit: the costs would mislead, start over.  Another
        // camera: nearby views cost alike, so the costs are halved (the new view soon outweighs
        // the old ones) and the order is kept until the view has held for two frames.
        const uint32_t shift = same_paths ? 1u : 32u;
        hipLaunchKernelGGL(shift_kernel, dim3((t->n_units + 255u) / 256u), dim3(256), 0, s, t->unit_cost, t->n_units,
                           shift);
        NB_ORDER_HIP(hipGetLastError());
        t->cost_frames = same_paths ? t->cost_frames / 2u : 0u;
        t->view_frames = 0;
        if (!same_paths) {
            // an order (or one being built) from other paths would mislead
            t->order_valid = false;
            t->order_pending = false;
            t->rebuild_wanted = false;
        }
        t->cost_scene = scene;
        t->cost_camera = *cam;
        t->cost_bounces = g->bounces;
        t->cost_contribution = g->contribution_factor;
    }
    p.unit_cost = t->unit_cost;
    p.unit_order = t->order_valid ? t->unit_order[t->order_cur] : nullptr;
    return NB_OK;
}

nb_status unit_order_update(nb_target *t, const RenderParams &p, uint32_t n_frames, hipStream_t s) {
    if (!p.unit_cost) return NB_OK;
    const uint32_t before = t->view_frames;
    t->cost_frames += n_frames;
    t->view_frames += n_frames;
    // Rebuild once the view has held for 2, 4, 8, ... frames (and after its first frame when there
    // is no order yet), then every 64 frames: the ranking settles as costs add up, and a camera
    // that moves every frame does not pay for a sort per frame.
    const uint64_t lo = before, hi = t->view_frames;   // frames (lo, hi] were added
    const uint64_t first_pow2 = t->order_valid ? 2u : 1u;
    uint64_t pow2 = first_pow2;
    while (pow2 <= lo) pow2 <<= 1;                      // the first power of two after lo
    const bool rebuild = pow2 <= hi || (lo / 64u + 1u) * 64u <= hi || t->rebuild_wanted;
    if (!rebuild || t->cost_frames == 0) return NB_OK;
    if (t->order_pending) {
        // the other buffer still waits to be switched to: rebuild after that
        t->rebuild_wanted = true;
        return NB_OK;
    }
    // Into the buffer no launch in flight reads: launches that began before the switch to the
    // current buffer have all finished (each is followed on `s` by its mix or is itself on `s`,
    // and this call's own launch is already enqueued), and launches from now on read the
    // current buffer until they follow this sort.
    const uint32_t n = t->n_units;
    uint32_t *out = t->unit_order[t->order_cur ^ 1u];
    hipLaunchKernelGGL(order_keys_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, t->unit_cost, n, t->cost_frames,
                       order_shift(), t->sort_keys, t->sort_vals);
    NB_ORDER_HIP(hipGetLastError());
    size_t tmp = t->sort_tmp_bytes;
    NB_ORDER_HIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        t->sort_tmp, tmp, t->sort_keys, t->sort_keys + n, t->sort_vals, out, static_cast<int>(n),
        static_cast<int>(kBands), t->sort_offsets, t->sort_offsets + 1, 0, 8, s));
    t->order_pending = true;
    t->pending_sort = ++t->ctx->order_sorts;
    t->rebuild_wanted = false;
    return NB_OK;
}

}  // namespace nb
