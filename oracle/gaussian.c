/*
 * Oracle (test infrastructure only): the seeded discrete Gaussian sampler,
 * restating src/math/gaussian.rs:31-33 (tailcut = ceil(6 sigma)),
 * :100-114 (acceptance thresholds exp(-x^2 / 2 sigma^2) as f64 bit patterns),
 * :140-147 (with_seed = ChaCha20Rng::seed_from_u64) and :210-220
 * (x = gen_range(-B..=B), u = gen_range(0.0..1.0), accept when bits(p) > bits(u)).
 * Pinned by the five stream hashes and the centred-vector hash of
 * src/math/gaussian.rs:334-351.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

uint64_t fnv1a(const uint8_t *bytes, size_t len) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < len; i++) {
        h ^= bytes[i];
        h *= 0x00000100000001b3ULL;
    }
    return h;
}

static void gauss_build(orc_gauss *g, double sigma) {
    g->sigma = sigma;
    g->tailcut = (int64_t)ceil(sigma * 6.0);
    size_t cnt = (size_t)(2 * g->tailcut + 1);
    g->accept_bits = (uint64_t *)malloc(cnt * sizeof(uint64_t));
    double s2 = 2.0 * sigma * sigma;
    for (int64_t x = -g->tailcut; x <= g->tailcut; x++) {
        double xsq = (double)(x * x);
        double prob = exp(-xsq / s2);
        uint64_t bits;
        if (isnan(prob)) prob = 0.0;
        memcpy(&bits, &prob, 8);
        g->accept_bits[x + g->tailcut] = bits;
    }
}

void orc_gauss_with_seed(orc_gauss *g, double sigma, uint64_t seed) {
    gauss_build(g, sigma);
    orc_chacha_seed_from_u64(&g->rng, seed);
}

void orc_gauss_from_seed(orc_gauss *g, double sigma, const uint8_t seed[32]) {
    gauss_build(g, sigma);
    orc_chacha_from_seed(&g->rng, seed);
}

void orc_gauss_free(orc_gauss *g) { free(g->accept_bits); g->accept_bits = NULL; }

int64_t orc_gauss_sample(orc_gauss *g) {
    for (;;) {
        int64_t x = orc_gen_range_i64_incl(&g->rng, -g->tailcut, g->tailcut);
        uint64_t thr = g->accept_bits[x + g->tailcut];
        double u = orc_gen_f64_unit(&g->rng);
        uint64_t ub;
        memcpy(&ub, &u, 8);
        if (thr > ub) return x;
    }
}

uint64_t orc_gauss_stream_hash(double sigma, uint64_t seed, size_t n) {
    orc_gauss g;
    orc_gauss_with_seed(&g, sigma, seed);
    uint8_t *bytes = (uint8_t *)malloc(n * 8);
    for (size_t i = 0; i < n; i++) {
        int64_t s = orc_gauss_sample(&g);
        memcpy(bytes + 8 * i, &s, 8);
    }
    uint64_t h = fnv1a(bytes, n * 8);
    free(bytes);
    orc_gauss_free(&g);
    return h;
}

uint64_t orc_gauss_centered_hash_seed7(size_t n, uint64_t q) {
    uint8_t seed[32];
    memset(seed, 7, 32);
    orc_gauss g;
    orc_gauss_from_seed(&g, 3.2, seed);
    uint8_t *bytes = (uint8_t *)malloc(n * 8);
    for (size_t i = 0; i < n; i++) {
        int64_t s = orc_gauss_sample(&g);
        uint64_t v = s < 0 ? q + (uint64_t)s : (uint64_t)s;
        memcpy(bytes + 8 * i, &v, 8);
    }
    uint64_t h = fnv1a(bytes, n * 8);
    free(bytes);
    orc_gauss_free(&g);
    return h;
}
