// Radiance-cache MLP kernels for gfx950: 16-bit operands on the matrix cores, FP32 accumulation.
// Every kernel is instantiated for BF16 (v_mfma_f32_16x16x16_bf16) and for FP16
// (v_mfma_f32_16x16x16_f16); the FP16 form trains with a dynamic loss scale kept on the device
// (the output gradient is multiplied by it before rounding to 16 bits; the FP32 weight gradients
// are divided by it before Adam; a non-finite gradient skips the step and halves the scale,
// 2000 clean steps double it).
//
// Layout convention: every dense layer is computed transposed, Z^T = W H^T, with the weights
// W (out x in, row-major) as the MFMA A operand and the activations as the B operand.  For
// 16x16x16 MFMA, lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16]:
// the D tile of output features 16m..16m+15 for 16 batch rows is, lane for lane, exactly the B
// tile of input features 16m..16m+15 of the next layer, so the fused inference kernel keeps a
// row's activations in registers from the encoding to the output with no shuffles.
#include "pt_device.h"

#include <hip/hip_fp16.h>

namespace nb {
namespace {
using namespace dev;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// binary32 -> bfloat16, round to nearest even (gfx950's v_cvt_pk_bf16_f32; for finite inputs
// the same bits as adding 0x7FFF plus the kept lowest bit and truncating)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// binary32 -> the 16-bit operand format (round to nearest even) and back
template <bool kHalf>
__device__ __forceinline__ uint16_t to16(float f) {
    if constexpr (kHalf) return __half_as_ushort(__float2half_rn(f));
    else return f2bf(f);
}
template <bool kHalf>
__device__ __forceinline__ float from16(uint16_t h) {
    if constexpr (kHalf) return __half2float(__ushort_as_half(h));
    else return bf2f(h);
}

template <bool kHalf>
__device__ __forceinline__ f4v mfma(s4 a, s4 b, f4v c) {
    if constexpr (kHalf) return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(h4, a), __builtin_bit_cast(h4, b), c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ s4 load4(const uint16_t *p) {   // 4 consecutive bf16 (8-byte aligned)
    uint2 v = *reinterpret_cast<const uint2 *>(p);
    s4 r = {};
    r[0] = static_cast<short>(v.x & 0xFFFFu); r[1] = static_cast<short>(v.x >> 16);
    r[2] = static_cast<short>(v.y & 0xFFFFu); r[3] = static_cast<short>(v.y >> 16);
    return r;
}
__device__ __forceinline__ s4 load4_strided(const uint16_t *p, uint32_t stride) {
    s4 r = {};
    for (int i = 0; i < 4; ++i) r[i] = static_cast<short>(p[i * stride]);
    return r;
}

// Input encoding of one raw record r (pos3, dir3, normal3, roughness, albedo3, metallic):
//   0..71   sin, cos of 2^f pi p_d, f = 0..11, p = position normalised to the scene box
//   72..83  sin, cos of 2^f pi w_d, f = 0..1, w = direction
//   84..86  normal, 87 roughness, 88..90 albedo, 91 metallic, 92..95 zero
__device__ __forceinline__ float feature(const float *r, int f, const MlpDev &m) {
    if (f < 84) {
        float x, scale;
        int q;
        if (f < 72) {
            const int d = f / 24;
            q = f % 24;
            x = (r[d] - m.scene_min[d]) * m.scene_inv_extent[d];
        } else {
            const int g = f - 72, d = g / 4;
            q = g % 4;
            x = r[3 + d];
        }
        scale = kPi * static_cast<float>(1u << (q >> 1));
        float s, c;
        sin_cos(x * scale, s, c);
        return (q & 1) ? c : s;
    }
    if (f < 87) return r[6 + (f - 84)];
    if (f == 87) return r[9];
    if (f < 91) return r[10 + (f - 88)];
    if (f == 91) return r[13];
    return 0.0f;
}

// sin_cos with the quadrant applied by selects and sign flips instead of a switch, the same bits:
// the feature index, and with it the quadrant, differs between the lanes of a wave, and a switch
// compiles to divergent branches.
__device__ __forceinline__ void sin_cos_select(float x, float &s, float &c) {
    int q;
    const float r = reduce_quadrant(x, q);
    const float ps = poly_sin(r), pc = poly_cos(r);
    const uint32_t neg_s = (static_cast<uint32_t>(q) & 2u) << 30, neg_c = (static_cast<uint32_t>(q + 1) & 2u) << 30;
    s = __uint_as_float(__float_as_uint((q & 1) ? pc : ps) ^ neg_s);
    c = __uint_as_float(__float_as_uint((q & 1) ? ps : pc) ^ neg_c);
}

// Features f0 .. f0 + 3 (f0 a multiple of 4) of the raw record at `row` (global memory), the same
// values as feature(): the sin/cos features come in pairs of one argument (frequencies k and
// k + 1), so one sin_cos serves two features.  The record's values are read by index from memory
// (an index into a register array that differs between lanes would compile to select chains).
__device__ __forceinline__ void feature4(const float *row, int f0, const MlpDev &m, float out[4]) {
    if (f0 < 84) {
        float x;
        int k;
        if (f0 < 72) {
            const int d = f0 / 24;
            k = (f0 % 24) >> 1;
            const float lo = d == 0 ? m.scene_min[0] : (d == 1 ? m.scene_min[1] : m.scene_min[2]);
            const float inv = d == 0 ? m.scene_inv_extent[0] : (d == 1 ? m.scene_inv_extent[1] : m.scene_inv_extent[2]);
            x = (row[d] - lo) * inv;
        } else {
            const int g = f0 - 72, d = g / 4;
            k = (g % 4) >> 1;
            x = row[3 + d];
        }
        sin_cos_select(x * (kPi * static_cast<float>(1u << k)), out[0], out[1]);
        sin_cos_select(x * (kPi * static_cast<float>(1u << (k + 1))), out[2], out[3]);
    } else if (f0 < 92) {
        const int b = f0 == 84 ? 6 : 10;   // normal and roughness, or albedo and metallic
        for (int i = 0; i < 4; ++i) out[i] = row[b + i];
    } else {
        for (int i = 0; i < 4; ++i) out[i] = 0.0f;
    }
}

// Threads per workgroup of the forward kernel: 8 waves share one LDS copy of the weights (51 KB),
// so two workgroups and 4 waves per SIMD fit a CU (with 128 VGPRs).
constexpr uint32_t kFwdThreads = 512;
constexpr int kPad = 4;                 // LDS row padding (bf16) against bank conflicts
constexpr int kS0 = kMlpIn + kPad;      // 100
constexpr int kSH = kMlpHidden + kPad;  // 68

// Copies a (rows x cols) row-major 16-bit matrix from global memory into LDS rows of `pitch`
// values, 4 values (8 bytes) per access: cols, pitch and the source offset are multiples of 4,
// and the LDS array is 8-byte aligned.
__device__ __forceinline__ void stage_rows(uint16_t *dst, const uint16_t *src, int rows, int cols, int pitch) {
    const int q = cols / 4;
    for (int i = threadIdx.x; i < rows * q; i += blockDim.x) {
        const int r = i / q, c = (i % q) * 4;
        *reinterpret_cast<uint2 *>(dst + r * pitch + c) = *reinterpret_cast<const uint2 *>(src + r * cols + c);
    }
}

// 4 consecutive 16-bit values (8-byte aligned) to global memory
__device__ __forceinline__ void store4(uint16_t *p, s4 v) {
    *reinterpret_cast<uint2 *>(p) = make_uint2((static_cast<uint32_t>(static_cast<uint16_t>(v[0]))) |
                                                   (static_cast<uint32_t>(static_cast<uint16_t>(v[1])) << 16),
                                               (static_cast<uint32_t>(static_cast<uint16_t>(v[2]))) |
                                                   (static_cast<uint32_t>(static_cast<uint16_t>(v[3])) << 16));
}

// Fused forward pass: weights staged once per workgroup in LDS (47 KB), each wave64 takes 16
// rows at a time through all six layers in registers.  kTrain (one training step's forward
// pass) also stores the encoded inputs and the five hidden activations as 16-bit rows for the
// backward pass, and ends each row with the relative L2 loss: the loss sum (times 1/(3n)) and
// the output gradient 2 (y - t) / (y^2 + 0.01) / (3n), times the loss scale and rounded to
// 16 bits, padded to 16 channels with zeros.
template <bool kHalf, bool kTrain>
__global__ __launch_bounds__(kFwdThreads) __attribute__((amdgpu_waves_per_eu(4))) void mlp_forward_kernel(MlpDev m, const float *raw, uint32_t n_max, const uint32_t *n_dev,
                                                          float *rgb, MlpTrainRows tr) {
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    __shared__ __attribute__((aligned(16))) uint16_t sw0[kMlpHidden * kS0];
    __shared__ __attribute__((aligned(16))) uint16_t swh[4][kMlpHidden * kSH];
    __shared__ __attribute__((aligned(16))) uint16_t sw5[kMlpOutPad * kSH];
    __shared__ float sb[5 * kMlpHidden + kMlpOutPad];
    stage_rows(sw0, m.w[0], kMlpHidden, kMlpIn, kS0);
    for (int L = 0; L < 4; ++L) stage_rows(swh[L], m.w[L + 1], kMlpHidden, kMlpHidden, kSH);
    stage_rows(sw5, m.w[5], kMlpOutPad, kMlpHidden, kSH);
    for (int i = threadIdx.x; i < 5 * kMlpHidden + kMlpOutPad; i += blockDim.x)
        sb[i] = i < 5 * kMlpHidden ? m.b[i / kMlpHidden][i % kMlpHidden] : m.b[5][i - 5 * kMlpHidden];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const uint32_t per_block = kFwdThreads / 64u, waves = gridDim.x * per_block;
    float loss_part = 0.0f;
    for (uint32_t base = (blockIdx.x * per_block + wave) * 16u; base < n; base += waves * 16u) {
        const uint32_t row = base + col;
        // rows past n compute the features of row n - 1; their outputs are not stored
        const float *rowp = raw + static_cast<size_t>(row < n ? row : n - 1u) * kMlpRaw;
        s4 x[6];
        // unrolled, so that x is indexed by constants (a loop index into a register array
        // compiles to select chains)
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            float v[4];
            feature4(rowp, 16 * t + 4 * kq, m, v);
#pragma unroll
            for (int i = 0; i < 4; ++i) x[t][i] = static_cast<short>(to16<kHalf>(v[i]));
            if constexpr (kTrain) {
                if (row < n) store4(tr.x + static_cast<size_t>(row) * kMlpIn + 16 * t + 4 * kq, x[t]);
            }
        }
        s4 h[4];
        for (int mt = 0; mt < 4; ++mt) {
            f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int t = 0; t < 6; ++t) acc = mfma<kHalf>(load4(&sw0[(16 * mt + col) * kS0 + 16 * t + 4 * kq]), x[t], acc);
            for (int i = 0; i < 4; ++i) h[mt][i] = static_cast<short>(to16<kHalf>(fmaxf(acc[i] + sb[16 * mt + 4 * kq + i], 0.0f)));
            if constexpr (kTrain) {
                if (row < n) store4(tr.h[0] + static_cast<size_t>(row) * kMlpHidden + 16 * mt + 4 * kq, h[mt]);
            }
        }
        // not unrolled: unrolled, the compiler keeps all four layers' weight fragments in
        // registers (256 VGPRs, 2 waves per SIMD) instead of reading them from LDS
#pragma unroll 1
        for (int L = 0; L < 4; ++L) {
            s4 g[4];
            for (int mt = 0; mt < 4; ++mt) {
                f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int t = 0; t < 4; ++t) acc = mfma<kHalf>(load4(&swh[L][(16 * mt + col) * kSH + 16 * t + 4 * kq]), h[t], acc);
                const float *bl = sb + (L + 1) * kMlpHidden;
                for (int i = 0; i < 4; ++i) g[mt][i] = static_cast<short>(to16<kHalf>(fmaxf(acc[i] + bl[16 * mt + 4 * kq + i], 0.0f)));
                if constexpr (kTrain) {
                    if (row < n) store4(tr.h[L + 1] + static_cast<size_t>(row) * kMlpHidden + 16 * mt + 4 * kq, g[mt]);
                }
            }
            for (int mt = 0; mt < 4; ++mt) h[mt] = g[mt];
        }
        f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int t = 0; t < 4; ++t) acc = mfma<kHalf>(load4(&sw5[col * kSH + 16 * t + 4 * kq]), h[t], acc);
        const float *b5 = sb + 5 * kMlpHidden;
        if constexpr (kTrain) {
            // (y - t)^2 / (y^2 + 0.01) with y in the denominator held constant, and its gradient
            // 2 (y - t) / den / (3n) times the loss scale
            s4 dyv = {0, 0, 0, 0};
            if (kq == 0 && row < n) {
                const float scale = 1.0f / (3.0f * static_cast<float>(n));
                const float ls = tr.loss_scale ? *tr.loss_scale : 1.0f;
                for (int c = 0; c < kMlpOut; ++c) {
                    const float yv = acc[c] + b5[c], tv = tr.target[3 * static_cast<size_t>(row) + c];
                    const float den = yv * yv + 0.01f;
                    const float d = yv - tv;
                    loss_part += d * d / den;
                    dyv[c] = static_cast<short>(to16<kHalf>(2.0f * d / den * scale * ls));
                }
            }
            if (row < n) store4(tr.dy + static_cast<size_t>(row) * kMlpOutPad + 4 * kq, dyv);
        } else {
            if (kq == 0 && row < n)
                for (int c = 0; c < kMlpOut; ++c) rgb[3 * static_cast<size_t>(row) + c] = acc[c] + b5[c];
        }
    }
    if constexpr (kTrain) {
        // one atomic per workgroup: the waves' sums meet in LDS
        __shared__ float s_loss[kFwdThreads / 64];
        for (int o = 32; o > 0; o >>= 1) loss_part += __shfl_xor(loss_part, o, 64);
        if (lane == 0) s_loss[wave] = loss_part;
        __syncthreads();
        if (threadIdx.x == 0) {
            float sum = 0.0f;
            for (uint32_t k = 0; k < kFwdThreads / 64u; ++k) sum += s_loss[k];
            if (sum != 0.0f) atomicAdd(tr.loss, sum * (1.0f / (3.0f * static_cast<float>(n))));
        }
    }
}

// Backward pass through the data path of one training step: dH_{L-1}^T = W_L^T dZ_L^T, masked by
// ReLU'(h_{L-1}) and rounded to 16 bits, for L = 5 .. 1, each wave taking 16 rows at a time with
// the gradients in registers (the same lane layout as the forward pass) and the weights of layers
// 1-5 staged in LDS.  Stores dz_0 .. dz_4 (n x 64, 16-bit) for the weight gradients, starting
// from the forward pass's output gradient dy; each dz value is the fp32 sum over k tiles in order,
// masked and rounded once.
template <bool kHalf>
__global__ __launch_bounds__(256) void mlp_backward_kernel(MlpDev m, uint32_t n, MlpTrainRows tr) {
    __shared__ __attribute__((aligned(16))) uint16_t swh[4][kMlpHidden * kSH];
    __shared__ __attribute__((aligned(16))) uint16_t sw5[kMlpOutPad * kSH];
    for (int L = 0; L < 4; ++L) stage_rows(swh[L], m.w[L + 1], kMlpHidden, kMlpHidden, kSH);
    stage_rows(sw5, m.w[5], kMlpOutPad, kMlpHidden, kSH);
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const uint32_t waves = gridDim.x * 4u;
    for (uint32_t base = (blockIdx.x * 4u + wave) * 16u; base < n; base += waves * 16u) {
        const uint32_t row = base + col;
        const bool valid = row < n;
        const size_t hrow = static_cast<size_t>(valid ? row : 0u) * kMlpHidden;
        const s4 zero = {0, 0, 0, 0};
        // layer 5: one k tile over the 16 padded outputs
        const s4 g5 = valid ? load4(tr.dy + static_cast<size_t>(row) * kMlpOutPad + 4 * kq) : zero;
        s4 dz[4];
        for (int mt = 0; mt < 4; ++mt) {
            s4 a = {};
            for (int i = 0; i < 4; ++i) a[i] = static_cast<short>(sw5[(4 * kq + i) * kSH + 16 * mt + col]);
            const f4v acc = mfma<kHalf>(a, g5, f4v{0.0f, 0.0f, 0.0f, 0.0f});
            const s4 hp = valid ? load4(tr.h[4] + hrow + 16 * mt + 4 * kq) : zero;
            for (int i = 0; i < 4; ++i)
                dz[mt][i] = static_cast<short>(to16<kHalf>(from16<kHalf>(static_cast<uint16_t>(hp[i])) > 0.0f ? acc[i] : 0.0f));
            if (valid) store4(tr.dz[4] + hrow + 16 * mt + 4 * kq, dz[mt]);
        }
#pragma unroll 1
        for (int L = 4; L >= 1; --L) {
            s4 nz[4];
            for (int mt = 0; mt < 4; ++mt) {
                f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int t = 0; t < 4; ++t) {
                    s4 a = {};
                    for (int i = 0; i < 4; ++i) a[i] = static_cast<short>(swh[L - 1][(16 * t + 4 * kq + i) * kSH + 16 * mt + col]);
                    acc = mfma<kHalf>(a, dz[t], acc);
                }
                const s4 hp = valid ? load4(tr.h[L - 1] + hrow + 16 * mt + 4 * kq) : zero;
                for (int i = 0; i < 4; ++i)
                    nz[mt][i] = static_cast<short>(to16<kHalf>(from16<kHalf>(static_cast<uint16_t>(hp[i])) > 0.0f ? acc[i] : 0.0f));
                if (valid) store4(tr.dz[L - 1] + hrow + 16 * mt + 4 * kq, nz[mt]);
            }
            for (int mt = 0; mt < 4; ++mt) dz[mt] = nz[mt];
        }
    }
}

template <bool kHalf>
__global__ void mlp_encode_kernel(MlpDev m, const float *raw, uint32_t n, uint16_t *x) {
    const size_t total = static_cast<size_t>(n) * kMlpIn;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t row = i / kMlpIn;
        const int f = static_cast<int>(i % kMlpIn);
        float r[kMlpRaw];
        for (int q = 0; q < kMlpRaw; ++q) r[q] = raw[row * kMlpRaw + q];
        x[i] = to16<kHalf>(feature(r, f, m));
    }
}

// dW += dZ^T H_prev over rows, db += column sums of dZ: one (16-output x 16-input) tile per wave
// over a chunk of rows, then fp32 atomics.  Only outputs < out_valid are written.  Job `job` of a
// layer: tile job % tiles over the rows of chunk job / tiles.
template <bool kHalf>
__device__ __forceinline__ void weight_grad_job(const uint16_t *dz, uint32_t out_p, uint32_t out_valid, const uint16_t *h_prev,
                                                uint32_t k_in, uint32_t n, uint32_t rows_per_chunk, float *dw, float *db,
                                                uint32_t job) {
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, kq = lane >> 4;
    const uint32_t n_tiles = k_in / 16;
    const uint32_t tiles = (out_p / 16) * n_tiles;
    const uint32_t tile = job % tiles, chunk = job / tiles;
    const uint32_t mt = tile / n_tiles, nt = tile % n_tiles;
    const uint32_t r0 = chunk * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
    f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (uint32_t rb = r0; rb < r1; rb += 16) {
        s4 a = {}, bop = {};
        for (int i = 0; i < 4; ++i) {
            const uint32_t row = rb + 4 * kq + i;
            const bool ok = row < r1;
            a[i] = ok ? static_cast<short>(dz[static_cast<size_t>(row) * out_p + 16 * mt + col]) : short(0);
            bop[i] = ok ? static_cast<short>(h_prev[static_cast<size_t>(row) * k_in + 16 * nt + col]) : short(0);
            if (nt == 0 && ok) bsum += from16<kHalf>(static_cast<uint16_t>(a[i]));
        }
        acc = mfma<kHalf>(a, bop, acc);
    }
    for (int i = 0; i < 4; ++i) {
        const uint32_t o = 16 * mt + 4 * kq + i;
        if (o < out_valid) atomicAdd(&dw[static_cast<size_t>(o) * k_in + 16 * nt + col], acc[i]);
    }
    if (nt == 0) {
        // lanes col, col+16, col+32, col+48 hold partial sums of output 16mt+col
        bsum += __shfl_xor(bsum, 16, 64);
        bsum += __shfl_xor(bsum, 32, 64);
        const uint32_t o = 16 * mt + col;
        if (kq == 0 && o < out_valid) atomicAdd(&db[o], bsum);
    }
}

// The weight and bias gradients of all six layers in one launch: jobs job_begin[L] ..
// job_begin[L + 1] - 1 are layer L's (a wave's job, and so its layer, is wave-uniform).
template <bool kHalf>
__global__ __launch_bounds__(256) void mlp_weight_grad_kernel(MlpGradJobs J, uint32_t n) {
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t job = blockIdx.x * 4u + wave; job < J.job_begin[kMlpLayers]; job += gridDim.x * 4u) {
        int L = 0;
        while (job >= J.job_begin[L + 1]) ++L;
        weight_grad_job<kHalf>(J.dz[L], J.out_p[L], J.out_valid[L], J.h_prev[L], J.k_in[L], n, J.rows_per_chunk, J.dw[L],
                               J.db[L], job - J.job_begin[L]);
    }
}

// One Adam step on the gradient g / loss scale (skipped when the gradient check found a non-finite
// value; scale == nullptr means no scaling), then the parameter's 16-bit weight copy (or the padded
// last-layer bias) from the new value, as to16_kernel writes it.
template <bool kHalf>
__global__ void adam_refresh_kernel(float *w, float *m, float *v, const float *g, uint32_t n, float lr, float b1, float b2,
                                    float eps, float c1, float c2, const float *scale, const uint32_t *overflow,
                                    MlpRefresh R, uint16_t *wbf, float *b5) {
    if (overflow && *overflow) return;
    const float inv = scale ? 1.0f / *scale : 1.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i] * inv;
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        for (int L = 0; L < kMlpLayers; ++L)
            if (i - R.w_off[L] < R.w_size[L]) wbf[R.wb_off[L] + (i - R.w_off[L])] = to16<kHalf>(wi);
        if (i - R.b5_off < static_cast<uint32_t>(kMlpOut)) b5[i - R.b5_off] = wi;
    }
}

template <bool kHalf>
__global__ void to16_kernel(const float *src, uint16_t *dst, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = to16<kHalf>(src[i]);
}

// Loss-scale state: [0] scale (float bits), [1] clean steps since the last change, [2] overflow
// flag of the current step.  Sets the flag when any gradient value is not finite.
__global__ void grad_check_kernel(const float *g, uint32_t n, uint32_t *state) {
    bool bad = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) bad |= !isfinite(g[i]);
    if (__ballot(bad) != 0ull && (threadIdx.x & 63u) == 0) atomicOr(state + 2, 1u);
}

// After the step: halve the scale on overflow, double it after 2000 clean steps; clear the flag.
__global__ void scale_update_kernel(uint32_t *state) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = __uint_as_float(state[0]);
    if (state[2]) {
        s = fmaxf(s * 0.5f, 1.0f);
        state[1] = 0u;
    } else if (++state[1] >= 2000u) {
        s = fminf(s * 2.0f, 65536.0f);
        state[1] = 0u;
    }
    state[0] = __float_as_uint(s);
    state[2] = 0u;
}

__global__ void fill_kernel(float *p, uint32_t n, float v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

nb_status launched(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NB_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return NB_OK;
}

uint32_t blocks_for(size_t items, uint32_t per_block, uint32_t cap = 4096) {
    size_t b = (items + per_block - 1) / per_block;
    return static_cast<uint32_t>(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

#define NB_MLP_LAUNCH(kernel, grid, block, half, ...)                                              \
    do {                                                                                          \
        if (half) hipLaunchKernelGGL(kernel<true>, dim3(grid), dim3(block), 0, s, __VA_ARGS__);  \
        else hipLaunchKernelGGL(kernel<false>, dim3(grid), dim3(block), 0, s, __VA_ARGS__);      \
    } while (0)
#define NB_MLP_LAUNCH2(kernel, flag, grid, block, half, ...)                                                  \
    do {                                                                                                    \
        if (half) hipLaunchKernelGGL((kernel<true, flag>), dim3(grid), dim3(block), 0, s, __VA_ARGS__);   \
        else hipLaunchKernelGGL((kernel<false, flag>), dim3(grid), dim3(block), 0, s, __VA_ARGS__);       \
    } while (0)

nb_status mlp_launch_encode(const MlpDev &m, const float *raw, uint32_t n, uint16_t *x, hipStream_t s) {
    NB_MLP_LAUNCH(mlp_encode_kernel, blocks_for(static_cast<size_t>(n) * kMlpIn, 256), 256, m.half, m, raw, n, x);
    return launched("mlp encode");
}

nb_status mlp_launch_infer(const MlpDev &m, const float *raw, uint32_t n, float *rgb, hipStream_t s) {
    NB_MLP_LAUNCH2(mlp_forward_kernel, false, blocks_for(n, kFwdThreads / 4, 512), kFwdThreads, m.half, m, raw, n,
                   static_cast<const uint32_t *>(nullptr), rgb, MlpTrainRows{});
    return launched("mlp infer");
}

nb_status mlp_launch_infer_count(const MlpDev &m, const float *raw, const uint32_t *n_dev, uint32_t n_max, float *rgb,
                                 hipStream_t s) {
    NB_MLP_LAUNCH2(mlp_forward_kernel, false, blocks_for(n_max, kFwdThreads / 4, 512), kFwdThreads, m.half, m, raw, n_max, n_dev, rgb,
                   MlpTrainRows{});
    return launched("mlp infer");
}

nb_status mlp_launch_train_forward(const MlpDev &m, const float *raw, uint32_t n, const MlpTrainRows &tr, hipStream_t s) {
    NB_MLP_LAUNCH2(mlp_forward_kernel, true, blocks_for(n, kFwdThreads / 4, 1024), kFwdThreads, m.half, m, raw, n,
                   static_cast<const uint32_t *>(nullptr), static_cast<float *>(nullptr), tr);
    return launched("mlp training forward");
}

nb_status mlp_launch_backward(const MlpDev &m, uint32_t n, const MlpTrainRows &tr, hipStream_t s) {
    NB_MLP_LAUNCH(mlp_backward_kernel, blocks_for(n, 64, 1024), 256, m.half, m, n, tr);
    return launched("mlp backward");
}

nb_status mlp_launch_weight_grad(bool half, const MlpGradJobs &j, uint32_t n, hipStream_t s) {
    NB_MLP_LAUNCH(mlp_weight_grad_kernel, blocks_for(j.job_begin[kMlpLayers], 4), 256, half, j, n);
    return launched("mlp weight gradients");
}

nb_status mlp_launch_adam_refresh(bool half, float *w, float *m, float *v, const float *g, uint32_t n, float lr, float b1,
                                  float b2, float eps, float c1, float c2, uint32_t *scale_state, const MlpRefresh &r,
                                  uint16_t *wbf, float *b5, hipStream_t s) {
    const float *scale = scale_state ? reinterpret_cast<const float *>(scale_state) : nullptr;
    const uint32_t *overflow = scale_state ? scale_state + 2 : nullptr;
    if (scale_state) {
        hipLaunchKernelGGL(grad_check_kernel, dim3(blocks_for(n, 256, 256)), dim3(256), 0, s, g, n, scale_state);
        nb_status st = launched("gradient check");
        if (st != NB_OK) return st;
    }
    NB_MLP_LAUNCH(adam_refresh_kernel, blocks_for(n, 256), 256, half, w, m, v, g, n, lr, b1, b2, eps, c1, c2, scale,
                  overflow, r, wbf, b5);
    nb_status st = launched("adam");
    if (st != NB_OK || !scale_state) return st;
    hipLaunchKernelGGL(scale_update_kernel, dim3(1), dim3(64), 0, s, scale_state);
    return launched("loss-scale update");
}

nb_status mlp_launch_to16(bool half, const float *src, uint16_t *dst, uint32_t n, hipStream_t s) {
    NB_MLP_LAUNCH(to16_kernel, blocks_for(n, 256), 256, half, src, dst, n);
    return launched("to 16-bit");
}

nb_status mlp_launch_fill(float *p, uint32_t n, float v, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
    return launched("fill");
}

}  // namespace nb
