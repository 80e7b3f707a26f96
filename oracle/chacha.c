/*
 * Oracle (test infrastructure only): ChaCha20Rng, seed_from_u64 and the rand-0.8
 * uniform draws the reference relies on.
 *
 * The reference pins rand_chacha 0.3.1, rand_core 0.6.4 and rand 0.8.6
 * (Cargo.lock:369-420); their sources are not vendored under the reference, so
 * this restates their published algorithms:
 *   - ChaCha20 (RFC 7539 rounds) with a 64-bit block counter in words 12-13 and a
 *     64-bit stream id in words 14-15; the block RNG buffers 4 blocks (64 words)
 *     and next_u64 reads two words low-first, stitching across a refill.
 *   - seed_from_u64: PCG32 (mul 6364136223846793005, inc 11634580027462260723,
 *     XSH-RR output) filling the 32-byte seed 4 bytes at a time.
 *   - gen_range for integers: widening multiply by the range, accept when the
 *     low word is <= (range << lz(range)) - 1.
 *   - gen_range(0.0..1.0): 52 random bits under exponent 0, minus 1.0.
 * Call sites: src/math/poly.rs:657-685, src/inspiring/inspiring2.rs:866-887,
 * src/math/gaussian.rs:140-147,210-217. Pinned by the Gaussian stream hashes
 * (src/math/gaussian.rs:334-351).
 */
#include "oracle.h"
#include <string.h>

#define ROTL(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)            \
    a += b; d ^= a; d = ROTL(d, 16); \
    c += d; b ^= c; b = ROTL(b, 12); \
    a += b; d ^= a; d = ROTL(d, 8);  \
    c += d; b ^= c; b = ROTL(b, 7);

void orc_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, uint32_t out[16]) {
    uint32_t s[16], x[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream; s[15] = (uint32_t)(stream >> 32);
    memcpy(x, s, sizeof x);
    for (int r = 0; r < 10; r++) {
        QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13])
        QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12])
        QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

void orc_chacha_from_seed(orc_chacha *r, const uint8_t seed[32]) {
    for (int i = 0; i < 8; i++)
        r->key[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 |
                    (uint32_t)seed[4 * i + 2] << 16 | (uint32_t)seed[4 * i + 3] << 24;
    r->counter = 0;
    r->stream = 0;
    r->index = 64;
}

static uint32_t pcg32(uint64_t *state) {
    const uint64_t MUL = 6364136223846793005ULL, INC = 11634580027462260723ULL;
    *state = *state * MUL + INC;
    uint64_t st = *state;
    uint32_t xorshifted = (uint32_t)(((st >> 18) ^ st) >> 27);
    uint32_t rot = (uint32_t)(st >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

void orc_chacha_seed_from_u64(orc_chacha *r, uint64_t state) {
    uint8_t seed[32];
    for (int i = 0; i < 8; i++) {
        uint32_t x = pcg32(&state);
        seed[4 * i] = (uint8_t)x; seed[4 * i + 1] = (uint8_t)(x >> 8);
        seed[4 * i + 2] = (uint8_t)(x >> 16); seed[4 * i + 3] = (uint8_t)(x >> 24);
    }
    orc_chacha_from_seed(r, seed);
}

static void refill(orc_chacha *r) {
    for (int b = 0; b < 4; b++) orc_chacha_block(r->key, r->counter + (uint64_t)b, r->stream, r->buf + 16 * b);
    r->counter += 4;
}

void orc_chacha_set_stream(orc_chacha *r, uint64_t stream) {
    r->stream = stream;
    if (r->index != 64) {
        /* regenerate the current buffer under the new stream at the same word position */
        int idx = r->index;
        r->counter -= 4;
        refill(r);
        r->index = idx;
    }
}

uint32_t orc_chacha_next_u32(orc_chacha *r) {
    if (r->index >= 64) { refill(r); r->index = 0; }
    return r->buf[r->index++];
}

uint64_t orc_chacha_next_u64(orc_chacha *r) {
    int idx = r->index;
    if (idx < 63) {
        r->index += 2;
        return (uint64_t)r->buf[idx] | (uint64_t)r->buf[idx + 1] << 32;
    } else if (idx >= 64) {
        refill(r);
        r->index = 2;
        return (uint64_t)r->buf[0] | (uint64_t)r->buf[1] << 32;
    } else {
        uint64_t x = r->buf[63];
        refill(r);
        r->index = 1;
        return ((uint64_t)r->buf[0] << 32) | x;
    }
}

void orc_chacha_fill_bytes(orc_chacha *r, uint8_t *out, size_t len) {
    /* fill_via_u32_chunks: whole words consumed, the tail word's unused bytes dropped */
    size_t done = 0;
    while (done < len) {
        if (r->index >= 64) { refill(r); r->index = 0; }
        while (r->index < 64 && done < len) {
            uint32_t w = r->buf[r->index++];
            for (int k = 0; k < 4 && done < len; k++) out[done++] = (uint8_t)(w >> (8 * k));
        }
    }
}

static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

/* sample_single_inclusive's loop for a nonzero range and an acceptance bound `zone` on the
 * low word of the widening product */
static uint64_t gen_range_zone(orc_chacha *r, uint64_t low, uint64_t range, uint64_t zone) {
    for (;;) {
        uint64_t v = orc_chacha_next_u64(r);
        unsigned __int128 m = (unsigned __int128)v * range;
        uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
        if (lo <= zone) return low + hi;
    }
}

uint64_t orc_gen_range_u64(orc_chacha *r, uint64_t low, uint64_t high_excl) {
    uint64_t high = high_excl - 1;
    uint64_t range = high - low + 1;
    if (range == 0) return orc_chacha_next_u64(r);
    return gen_range_zone(r, low, range, (range << clz64(range)) - 1);
}

int64_t orc_gen_range_i64_incl(orc_chacha *r, int64_t low, int64_t high) {
    uint64_t range = (uint64_t)high - (uint64_t)low + 1;
    if (range == 0) return (int64_t)orc_chacha_next_u64(r);
    uint64_t zone = (range << clz64(range)) - 1;
    for (;;) {
        uint64_t v = orc_chacha_next_u64(r);
        unsigned __int128 m = (unsigned __int128)v * range;
        uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
        if (lo <= zone) return (int64_t)((uint64_t)low + hi);
    }
}

double orc_gen_f64_unit(orc_chacha *r) {
    uint64_t bits = (orc_chacha_next_u64(r) >> 12) | 0x3ff0000000000000ULL;
    double v;
    memcpy(&v, &bits, 8);
    return (v - 1.0) * 1.0 + 0.0;
}

void orc_expand_seed(const uint8_t seed[32], size_t n, uint64_t q, uint64_t *a) {
    orc_chacha r;
    orc_chacha_from_seed(&r, seed);
    for (size_t i = 0; i < n; i++) a[i] = orc_gen_range_u64(&r, 0, q);
}

/* Test hook: orc_expand_seed with the acceptance bound given instead of (q << lz(q)) - 1, so
 * the tests can force rejections and check the kernels' serial replay against this loop. */
void orc_expand_seed_zone(const uint8_t seed[32], size_t n, uint64_t q, uint64_t zone, uint64_t *a) {
    orc_chacha r;
    orc_chacha_from_seed(&r, seed);
    for (size_t i = 0; i < n; i++) a[i] = gen_range_zone(&r, 0, q, zone);
}
