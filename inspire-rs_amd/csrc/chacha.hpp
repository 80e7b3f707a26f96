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

// Acceptance bound of rand 0.8's gen_range(0..q) on the low product word: (q << lz(q)) - 1.
constexpr uint64_t kZoneQ = (kQ << 4) - 1;  // lz(q) = 4

// gen_range(0..q) for one u64 draw v against an acceptance bound on the low word of v*q:
// returns true and the value when the draw is accepted.
__host__ __device__ __forceinline__ bool uniform_q_draw_in(uint64_t v, uint64_t zone, uint64_t &out) {
    uint64_t hi, lo = mul_wide(v, kQ, hi);
    out = hi;
    return lo <= zone;
}
// rand 0.8 gen_range(0..q) acceptance for one u64 draw v (the reference's zone).
__host__ __device__ __forceinline__ bool uniform_q_draw(uint64_t v, uint64_t &out) {
    return uniform_q_draw_in(v, kZoneQ, out);
}

// Serial gen_range(0..q) expansion with rejection (poly.rs:657-685 through rand's
// sample_single_inclusive): u64 draws of stream 0 of ChaCha20(key) in order, two keystream
// words each, low word first (rand_core's BlockRng: every draw takes two words and a block
// holds 16, so no draw straddles a block), each accepted value handed to put(index, value)
// until n are out. The kernels' parallel expansion takes one block per lane and runs this only
// when some lane saw a rejected draw; zone is kZoneQ there (the host tests pass narrower
// zones to force rejections). The block's words are read at unrolled constant indices, so
// they stay in registers.
template <class Put>
__host__ __device__ inline void replay_gen_range_q(const uint32_t key[8], int n, uint64_t zone, Put &&put) {
    uint32_t w[16];
    int got = 0;
    for (uint64_t blk = 0; got < n; blk++) {
        chacha20_block(key, blk, 0, w);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            uint64_t a;
            if (got < n && uniform_q_draw_in((uint64_t)w[2 * u] | (uint64_t)w[2 * u + 1] << 32, zone, a))
                put(got++, a);
        }
    }
}

// u % m for a u64 draw and odd m < 2^32, by Barrett with recip = floor((2^64 - 1) / m): the
// estimate floor(u * recip / 2^64) is the quotient or one below it.
__host__ __device__ __forceinline__ uint32_t reduce64(uint64_t u, uint32_t m, uint64_t recip) {
    uint64_t hi;
    (void)mul_wide(u, recip, hi);
    const uint64_t r = u - hi * m;  // [0, 2m)
    return (uint32_t)(r >= m ? r - m : r);
}

// Serial expand_uniform (special_modulus/mod.rs:223-237) over `limbs` consecutive moduli of
// one ChaCha20 stream: `count` residues per modulus, a draw at or above zone[l] =
// (2^64 - 1) / m * m rejected, each accepted draw reduced mod m[l] and handed to
// put(l, index, value). Used where a lane of the parallel expansion saw a rejected draw.
template <class Put>
__host__ __device__ inline void replay_expand_uniform(const uint32_t key[8], uint64_t stream, int limbs, int count,
                                                      const uint32_t *m, const uint64_t *zone,
                                                      const uint64_t *recip, Put &&put) {
    uint32_t w[16];
    int l = 0, got = 0;
    for (uint64_t blk = 0; l < limbs; blk++) {
        chacha20_block(key, blk, stream, w);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t v = (uint64_t)w[2 * u] | (uint64_t)w[2 * u + 1] << 32;
            if (l < limbs && v < zone[l]) {
                put(l, got, reduce64(v, m[l], recip[l]));
                if (++got == count) {
                    l++;
                    got = 0;
                }
            }
        }
    }
}

}  // namespace inspire
