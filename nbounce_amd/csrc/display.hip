// Display blit (BlitRenderer, src/pathtracing/blit_renderer.rs; blit.wgsl:27-38): the RGBA32F
// accumulation target drawn over the whole surface by one fullscreen triangle whose texcoord is
// ndc * 0.5 + 0.5, sampled bilinearly with clamp-to-edge addressing (texture.rs:64-74), written
// to an RGBA16F surface (wgpu_context.rs:62).  Target row 0 is the bottom of the view, so the
// surface's top row samples the target's last row.
//
// Sampling: for surface pixel (px, py) of a W x H surface over a Wt x Ht target the texel-space
// coordinate x = (px + 0.5) / W * Wt - 0.5 is taken to 8 fractional bits, rounded to nearest,
// in exact integer arithmetic, as a texture unit's fixed-point weights; rows likewise with
// v = 1 - (py + 0.5) / H.  At W == Wt the fraction is exactly 0 and the blit is a row flip.
#include "nb_internal.h"

#include <hip/hip_fp16.h>

#include <new>

struct nb_display {
    nb_context *ctx = nullptr;
    ushort4 *pixels = nullptr;   // RGBA16F, row 0 at the top of the surface
    uint32_t width = 0, height = 0;
};

namespace nb {
namespace {

// The sample coordinate ((p + 0.5) / n_dst * n_src - 0.5) in 1/256 texel units, rounded half
// up: floor(((2p + 1) * n_src * 256 + n_dst) / (2 n_dst)) - 128.
__device__ __forceinline__ int64_t fixed_coord(uint32_t p, uint32_t n_dst, uint32_t n_src) {
    const uint64_t num = (2ull * p + 1ull) * n_src * 256ull + n_dst;
    return static_cast<int64_t>(num / (2ull * n_dst)) - 128;
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float w) {
    const float k = 1.0f - w;
    return make_float4(a.x * k + b.x * w, a.y * k + b.y * w, a.z * k + b.z * w, a.w * k + b.w * w);
}

__device__ __forceinline__ uint16_t half_bits(float f) { return __half_as_ushort(__float2half_rn(f)); }

__global__ __launch_bounds__(256) void blit_kernel(const float4 *__restrict__ src, uint32_t sw, uint32_t sh,
                                                   ushort4 *__restrict__ dst, uint32_t dw, uint32_t dh) {
    const uint32_t px = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t py = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (px >= dw || py >= dh) return;
    const int64_t fx = fixed_coord(px, dw, sw);
    const int64_t fy = fixed_coord(dh - 1u - py, dh, sh);   // v = 1 - (py + 0.5) / H
    const int64_t x0 = fx >> 8, y0 = fy >> 8;                 // floor (arithmetic shift)
    const float wx = static_cast<float>(fx & 255) * (1.0f / 256.0f);
    const float wy = static_cast<float>(fy & 255) * (1.0f / 256.0f);
    auto cx = [&](int64_t x) { return static_cast<uint32_t>(x < 0 ? 0 : (x >= sw ? sw - 1 : x)); };
    auto cy = [&](int64_t y) { return static_cast<uint32_t>(y < 0 ? 0 : (y >= sh ? sh - 1 : y)); };
    const uint32_t xa = cx(x0), xb = cx(x0 + 1), ya = cy(y0), yb = cy(y0 + 1);
    const float4 t00 = src[static_cast<size_t>(ya) * sw + xa], t10 = src[static_cast<size_t>(ya) * sw + xb];
    const float4 t01 = src[static_cast<size_t>(yb) * sw + xa], t11 = src[static_cast<size_t>(yb) * sw + xb];
    const float4 c = lerp4(lerp4(t00, t10, wx), lerp4(t01, t11, wx), wy);
    dst[static_cast<size_t>(py) * dw + px] = make_ushort4(half_bits(c.x), half_bits(c.y), half_bits(c.z), half_bits(c.w));
}

nb_status hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return NB_OK;
    return set_error(e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace
}  // namespace nb

using namespace nb;

extern "C" {

nb_status nb_display_create(nb_context *ctx, uint32_t width, uint32_t height, nb_display **out) {
    if (!ctx || !out) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0 || width > 32768 || height > 32768)
        return set_error(NB_ERR_INVALID_ARGUMENT, "display %ux%u: sides must be in [1, 32768]", width, height);
    nb_status st = hip_status(hipSetDevice(ctx->device), "hipSetDevice");
    if (st != NB_OK) return st;
    nb_display *d = new (std::nothrow) nb_display();
    if (!d) return set_error(NB_ERR_OUT_OF_MEMORY, "out of host memory");
    d->ctx = ctx;
    d->width = width;
    d->height = height;
    st = hip_status(hipMalloc(&d->pixels, static_cast<size_t>(width) * height * sizeof(ushort4)), "display allocation");
    if (st != NB_OK) {
        delete d;
        return st;
    }
    context_add_child(ctx);
    *out = d;
    return NB_OK;
}

nb_status nb_display_destroy(nb_display *d) {
    if (!d) return NB_OK;
    nb_context *ctx = d->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipFree(d->pixels);
    delete d;
    context_release_child(ctx);
    return NB_OK;
}

nb_status nb_display_blit(nb_display *d, const nb_target *target) {
    if (!d || !target) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    if (target->ctx != d->ctx) return set_error(NB_ERR_INVALID_ARGUMENT, "target and display belong to different contexts");
    nb_status st = hip_status(hipSetDevice(d->ctx->device), "hipSetDevice");
    if (st != NB_OK) return st;
    dim3 grid((d->width + 15u) / 16u, (d->height + 15u) / 16u);
    hipLaunchKernelGGL(blit_kernel, grid, dim3(256), 0, d->ctx->stream, target->pixels, target->width, target->height,
                       d->pixels, d->width, d->height);
    return hip_status(hipGetLastError(), "blit_kernel launch");
}

nb_status nb_display_read(nb_display *d, uint16_t *rgba16f) {
    if (!d || !rgba16f) return set_error(NB_ERR_INVALID_ARGUMENT, "null argument");
    nb_status st = hip_status(hipSetDevice(d->ctx->device), "hipSetDevice");
    if (st == NB_OK) st = hip_status(hipStreamSynchronize(d->ctx->stream), "hipStreamSynchronize");
    if (st == NB_OK)
        st = hip_status(hipMemcpy(rgba16f, d->pixels, static_cast<size_t>(d->width) * d->height * sizeof(ushort4),
                                  hipMemcpyDeviceToHost),
                        "display read");
    return st;
}

}  // extern "C"
