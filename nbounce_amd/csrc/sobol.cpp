// Owen-scrambled Sobol table for the sampler binding (pathtracing.wgsl:13), generated as
// src/pathtracing/pathtracer.rs:57-64 does with sobol_burley::sample_4d(i, set, seed):
// entry [i * dims + set] holds dimensions 4*set .. 4*set+3 of point i.
//
// Method: Burley 2020, "Practical Hash-based Owen Scrambling" (JCGT 9(4)): the point index is
// shuffled by a nested uniform scramble (a Laine–Karras permutation on the bit-reversed index),
// each Sobol coordinate gets its own nested uniform scramble, and the top 24 bits become the
// float.  Direction numbers: Joe & Kuo new-joe-kuo-6.21201.
#include "nbounce.h"
#include "nb_internal.h"

#include <cstdint>

namespace nb {
namespace {

#include "joe_kuo_256.inc"

constexpr int kDims = 256;

struct DirectionNumbers {
    uint32_t v[kDims][32];
    DirectionNumbers() {
        for (int b = 0; b < 32; ++b) v[0][b] = 1u << (31 - b);
        for (int d = 1; d < kDims; ++d) {
            const uint32_t s = kJoeKuo[d][0], a = kJoeKuo[d][1];
            uint32_t m[32];
            for (uint32_t k = 0; k < s; ++k) m[k] = kJoeKuo[d][2 + k];
            // m_k = 2 a_1 m_{k-1} ^ 4 a_2 m_{k-2} ^ ... ^ 2^s m_{k-s} ^ m_{k-s}
            for (uint32_t k = s; k < 32; ++k) {
                uint32_t x = m[k - s] ^ (m[k - s] << s);
                for (uint32_t j = 1; j < s; ++j)
                    if ((a >> (s - 1 - j)) & 1u) x ^= m[k - j] << j;
                m[k] = x;
            }
            for (int b = 0; b < 32; ++b) v[d][b] = m[b] << (31 - b);
        }
    }
};

const DirectionNumbers &dirs() {
    static const DirectionNumbers d;
    return d;
}

inline uint32_t sobol_bits(uint32_t index, int dim) {
    const uint32_t *v = dirs().v[dim];
    uint32_t x = 0;
    for (int b = 0; index != 0; ++b, index >>= 1) x ^= (index & 1u) ? v[b] : 0u;
    return x;
}

inline uint32_t bit_reverse(uint32_t x) {
    x = (x << 16) | (x >> 16);
    x = ((x & 0x00FF00FFu) << 8) | ((x >> 8) & 0x00FF00FFu);
    x = ((x & 0x0F0F0F0Fu) << 4) | ((x >> 4) & 0x0F0F0F0Fu);
    x = ((x & 0x33333333u) << 2) | ((x >> 2) & 0x33333333u);
    x = ((x & 0x55555555u) << 1) | ((x >> 1) & 0x55555555u);
    return x;
}

// Laine–Karras permutation constants from Burley 2020, listing 3.
inline uint32_t owen(uint32_t x, uint32_t seed) {
    uint32_t y = bit_reverse(x) + seed;
    y ^= y * 0x6c50b47cu;
    y ^= y * 0xb82f1e52u;
    y ^= y * 0xc7afe638u;
    y ^= y * 0x8d22f6e6u;
    return bit_reverse(y);
}

inline uint32_t combine(uint32_t seed, uint32_t v) { return seed ^ (v + (seed << 6) + (seed >> 2)); }

inline uint32_t mix32(uint32_t x) {  // lowbias32 (C. Wellons)
    x = (x ^ (x >> 16)) * 0x7feb352du;
    x = (x ^ (x >> 15)) * 0x846ca68bu;
    return x ^ (x >> 16);
}

}  // namespace

nb_status sobol_generate(uint32_t rows, uint32_t dims, uint32_t seed, float *out) {
    if (dims == 0 || 4u * dims > static_cast<uint32_t>(kDims))
        return set_error(NB_ERR_INVALID_ARGUMENT, "dimension sets must be in 1..%d", kDims / 4);
    for (uint32_t set = 0; set < dims; ++set) {
        const uint32_t s = mix32(combine(seed, set));
        for (uint32_t i = 0; i < rows; ++i) {
            const uint32_t idx = owen(i, s);
            float *o = out + 4 * (static_cast<size_t>(i) * dims + set);
            for (uint32_t k = 0; k < 4; ++k) {
                uint32_t x = owen(sobol_bits(idx, static_cast<int>(4 * set + k)), mix32(combine(s, k + 1)));
                o[k] = static_cast<float>(x >> 8) * (1.0f / 16777216.0f);
            }
        }
    }
    return NB_OK;
}

}  // namespace nb
