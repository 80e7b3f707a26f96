// ChaCha20 keystream blocks for the query and mask expansions, one block per lane.
//
// Matches the rand_chacha 0.3.1 stream the reference draws from (Cargo.lock:390):
// key = the 32-byte seed (little-endian words), block counter in words 12-13,
// stream id in words 14-15, counter starting at 0. The reference reads u64 draws
// low word first from consecutive keystream words (rand_core BlockRng), so draw D
// is words 2D, 2D+1, i.e. block D/8, words 2(D%8) and 2(D%8)+1.
// Call sites: src/math/poly.rs:657-685 (query a), src/inspiring/inspiring2.rs:866-887
// (w_mask rows).
#pragma once
#include <stdint.h>
#include "modq.hpp"

namespace inspire {

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define INSPIRE_QR(a, b, c, d)                                      \
    a += b; d ^= a; d = rotl32(d, 16); c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8); c += d; b ^= c; b = rotl32(b, 7);

__host__ __device__ __forceinline__ void chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                                                        uint32_t out[16]) {
    uint32_t x0 = 0x61707865u, x1 = 0x3320646eu, x2 = 0x79622d32u, x3 = 0x6b206574u;
    uint32_t x4 = key[0], x5 = key[1], x6 = key[2], x7 = key[3];
    uint32_t x8 = key[4], x9 = key[5], x10 = key[6], x11 = key[7];
    uint32_t x12 = (uint32_t)counter, x13 = (uint32_t)(counter >> 32);
    uint32_t x14 = (uint32_t)stream, x15 = (uint32_t)(stream >> 32);
    const uint32_t s12 = x12, s13 = x13, s14 = x14, s15 = x15;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        INSPIRE_QR(x0, x4, x8, x12) INSPIRE_QR(x1, x5, x9, x13)
        INSPIRE_QR(x2, x6, x10, x14) INSPIRE_QR(x3, x7, x11, x15)
        INSPIRE_QR(x0, x5, x10, x15) INSPIRE_QR(x1, x6, x11, x12)
        INSPIRE_QR(x2, x7, x8, x13) INSPIRE_QR(x3, x4, x9, x14)
    }
    out[0] = x0 + 0x61707865u; out[1] = x1 + 0x3320646eu; out[2] = x2 + 0x79622d32u; out[3] = x3 + 0x6b206574u;
    out[4] = x4 + key[0]; out[5] = x5 + key[1]; out[6] = x6 + key[2]; out[7] = x7 + key[3];
    out[8] = x8 + key[4]; out[9] = x9 + key[5]; out[10] = x10 + key[6]; out[11] = x11 + key[7];
    out[12] = x12 + s12; out[13] = x13 + s13; out[14] = x14 + s14; out[15] = x15 + s15;
}
#undef INSPIRE_QR

__host__ __device__ __forceinline__ void seed_key(const uint8_t *seed, uint32_t key[8]) {
    for (int i = 0; i < 8; i++)
        key[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                 (uint32_t)seed[4 * i + 3] << 24;
}

// rand 0.8 gen_range(0..q) acceptance for one u64 draw v: returns true and the
// value when the low product word is inside the zone (q << lz(q)) - 1.
__host__ __device__ __forceinline__ bool uniform_q_draw(uint64_t v, uint64_t &out) {
    uint64_t hi, lo = mul_wide(v, kQ, hi);
    const uint64_t zone = (kQ << 4) - 1;  // lz(q) = 4
    out = hi;
    return lo <= zone;
}

}  // namespace inspire
