/*
 * Oracle (test infrastructure only): the negacyclic NTT of src/math/ntt.rs.
 *
 *   compute_q_inv_neg / compute_r_squared   ntt.rs:698-711
 *   find_primitive_root (least QNR g, psi = g^((q-1)/2n))  ntt.rs:743-752
 *   compute_twiddle_factors (Montgomery, bit-reversed ladder) ntt.rs:797-834
 *   forward: lift to Montgomery, Cooley-Tukey             ntt.rs:201-247, 578-607
 *   inverse: Gentleman-Sande, times n^-1, strip R          ntt.rs:279-349, 610-640
 *   Montgomery product (== Solinas REDC at DEFAULT_Q)      ntt.rs:689-695, solinas_redc.rs:13-22
 * Pinned by brute-force negacyclic convolution and round trips (tests/test_oracle_math.py).
 */
#include "oracle.h"
#include <stdlib.h>

typedef unsigned __int128 u128;

static uint64_t csubq(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }
static uint64_t submod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

static uint64_t q_inv_neg_of(uint64_t q) {
    uint64_t y = 1;
    for (int i = 1; i < 64; i++) y |= (y * q) & (1ULL << i);
    return (uint64_t)0 - y;
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1) r = (uint64_t)((u128)r * b % m);
        b = (uint64_t)((u128)b * b % m);
        e >>= 1;
    }
    return r;
}

static uint64_t redc(u128 ab, uint64_t q, uint64_t qinv) {
    uint64_t m = (uint64_t)ab * qinv;
    u128 t = (ab + (u128)m * q) >> 64;
    /* ab + m*q can exceed 2^128 only if ab >= 2^128 - 2^124; inputs here are < q^2 < 2^120 */
    return csubq((uint64_t)t, q);
}

uint64_t orc_mont_mul(const orc_ntt *c, uint64_t a, uint64_t b) {
    return redc((u128)a * b, c->q, c->q_inv_neg);
}

uint64_t orc_to_mont(const orc_ntt *c, uint64_t a) { return redc((u128)a * c->r2, c->q, c->q_inv_neg); }

static void twiddles(size_t n, uint64_t psi_mont, const orc_ntt *c, uint64_t *f) {
    f[0] = 0;
    f[1] = orc_to_mont(c, 1);
    for (size_t m = 1; m < n; m++) {
        if ((m & (m - 1)) == 0) {
            size_t e = n / (2 * m);
            uint64_t pw = orc_to_mont(c, 1);
            for (size_t k = 0; k < e; k++) pw = orc_mont_mul(c, pw, psi_mont);
            f[m] = pw;
        } else {
            size_t prev = m & (m - 1), step = m & (~m + 1);
            f[m] = orc_mont_mul(c, f[prev], f[step]);
        }
    }
}

int orc_ntt_init(orc_ntt *c, size_t n, uint64_t q) {
    c->n = n;
    c->q = q;
    c->q_inv_neg = q_inv_neg_of(q);
    u128 rmod = ((u128)1 << 64) % q;
    c->r2 = (uint64_t)((rmod * rmod) % q);
    uint64_t two_n = 2 * (uint64_t)n, ex = (q - 1) / two_n;
    c->psi = 0;
    for (uint64_t g = 2; g < 65536; g++) {
        uint64_t cand = powmod(g, ex, q);
        if (powmod(cand, two_n, q) == 1 && powmod(cand, two_n / 2, q) != 1) { c->psi = cand; break; }
    }
    if (!c->psi) return -1;
    c->psi_inv = powmod(c->psi, q - 2, q);
    c->psi_pow = (uint64_t *)malloc(n * sizeof(uint64_t));
    c->psi_inv_pow = (uint64_t *)malloc(n * sizeof(uint64_t));
    twiddles(n, orc_to_mont(c, c->psi), c, c->psi_pow);
    twiddles(n, orc_to_mont(c, c->psi_inv), c, c->psi_inv_pow);
    c->n_inv_mont = orc_to_mont(c, powmod((uint64_t)n, q - 2, q));
    return 0;
}

void orc_ntt_free(orc_ntt *c) {
    free(c->psi_pow);
    free(c->psi_inv_pow);
    c->psi_pow = c->psi_inv_pow = NULL;
}

void orc_ntt_forward(const orc_ntt *c, uint64_t *a) {
    size_t n = c->n, t = n;
    uint64_t q = c->q;
    for (size_t j = 0; j < n; j++) a[j] = orc_to_mont(c, a[j]);
    for (size_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (size_t i = 0; i < m; i++) {
            size_t j1 = 2 * i * t;
            uint64_t w = c->psi_pow[m + i];
            for (size_t j = j1; j < j1 + t; j++) {
                uint64_t u = a[j], v = orc_mont_mul(c, a[j + t], w);
                a[j] = csubq(u + v, q);
                a[j + t] = submod(u, v, q);
            }
        }
    }
}

void orc_ntt_inverse(const orc_ntt *c, uint64_t *a) {
    size_t n = c->n, t = 1;
    uint64_t q = c->q;
    for (size_t m = n; m > 1;) {
        m >>= 1;
        for (size_t i = 0; i < m; i++) {
            size_t j2 = i * 2 * t;
            uint64_t w = c->psi_inv_pow[m + i];
            for (size_t j = j2; j < j2 + t; j++) {
                uint64_t u = a[j], v = a[j + t];
                a[j] = csubq(u + v, q);
                a[j + t] = orc_mont_mul(c, submod(u, v, q), w);
            }
        }
        t <<= 1;
    }
    for (size_t j = 0; j < n; j++) a[j] = orc_mont_mul(c, orc_mont_mul(c, a[j], c->n_inv_mont), 1);
}

void orc_negacyclic_mul_naive(size_t n, uint64_t q, const uint64_t *a, const uint64_t *b, uint64_t *out) {
    for (size_t k = 0; k < n; k++) out[k] = 0;
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
            uint64_t p = (uint64_t)((u128)a[i] * b[j] % q);
            size_t k = i + j;
            if (k < n) out[k] = csubq(out[k] + p, q);
            else out[k - n] = submod(out[k - n], p, q);
        }
}
