/*
 * Oracle (test infrastructure only): the special-modulus parameter set
 * (src/special_modulus), restated in C.
 *
 *   sm_prime_*   Prime32, arith.rs:13-102 (Montgomery R = 2^32, values in [0, q))
 *   sm_ntt_*     Ntt32, arith.rs:109-324 (psi = first g whose (q-1)/2n-th power has
 *                order 2n; forward CT over psi^bitrev(k), inverse GS with the last
 *                stage carrying n^-1; automorphism tables by search)
 *   sm_ring_*    Ring, mod.rs:256-363 (g = 2d/gamma + 1, key mask on stream
 *                "key mask" with expand_uniform mod.rs:216-237, the centred lifts
 *                and mod_down)
 *   sm_table     Server::table + step1_row_impl, server.rs:211-250, 330-383
 *   sm_seat      Server::seat, server.rs:177-203
 *   sm_respond   respond_impl, server.rs:385-499 (the respond path's arithmetic is in
 *                sm_arith.h, shared with the AVX-512 build in special_avx512.c)
 *   sm_client_*  Client::from_seed / drawn / build / query / decode,
 *                client.rs:137-330 (stream "key seed" for the key-mask seed)
 *   sm_pack      the tight codec at a prime's bit width, wire.rs:10-21
 *
 * Pinning: the whole chain (client key, query, table, seat, respond and the wire
 * codec) is pinned by the served-set digests of tests/special_modulus_kat.rs:349-380;
 * the pieces by Montgomery == plain products, O(n^2) negacyclic products, the
 * automorphism invariant, i128 brute force of the lifts and mod_down
 * (mod.rs:436-476) and decode of every record.
 */
#include "oracle.h"
#include "sm_arith.h"

#include <stdlib.h>
#include <string.h>

/* ---------------- Prime32 ---------------- */

static int sm_is_prime(uint64_t m) {
    if (m < 2) return 0;
    for (uint64_t d = 2; d * d <= m; d++)
        if (m % d == 0) return 0;
    return 1;
}

int sm_prime_init(sm_prime *p, uint64_t q) {
    if (q >= (1ULL << 31) || !sm_is_prime(q)) return -1;
    uint32_t q32 = (uint32_t)q, inv = 1;
    for (int i = 0; i < 5; i++) inv *= 2u - q32 * inv; /* Newton: q * inv == 1 mod 2^32 */
    p->q = q32;
    p->neg_qinv = 0u - inv;
    p->r2 = (uint32_t)(((unsigned __int128)1 << 64) % q);
    return 0;
}

uint32_t sm_csub(const sm_prime *p, uint32_t x) { return smi_csub(p, x); }
uint32_t sm_add(const sm_prime *p, uint32_t a, uint32_t b) { return smi_add(p, a, b); }
uint32_t sm_sub(const sm_prime *p, uint32_t a, uint32_t b) { return smi_sub(p, a, b); }
uint32_t sm_mont(const sm_prime *p, uint32_t a, uint32_t b) { return smi_mont(p, a, b); }

uint32_t sm_to_mont(const sm_prime *p, uint32_t a) { return sm_mont(p, a, p->r2); }

uint32_t sm_small(const sm_prime *p, int64_t x) {
    int64_t r = x % (int64_t)p->q;
    return (uint32_t)(r < 0 ? r + p->q : r);
}

int64_t sm_centred(const sm_prime *p, uint32_t x) { return smi_centred(p, x); }

uint32_t sm_pow(const sm_prime *p, uint64_t base, uint64_t e) {
    uint64_t acc = 1, b = base % p->q;
    while (e) {
        if (e & 1) acc = acc * b % p->q;
        b = b * b % p->q;
        e >>= 1;
    }
    return (uint32_t)acc;
}

/* ---------------- Ntt32 ---------------- */

static size_t sm_bitrev(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

int sm_ntt_init(sm_ntt *c, size_t n, uint64_t q) {
    memset(c, 0, sizeof *c);
    if (sm_prime_init(&c->p, q)) return -1;
    if (n < 2 || (n & (n - 1)) || (q - 1) % (2 * n)) return -1;
    const sm_prime *p = &c->p;
    uint32_t psi = 0;
    for (uint64_t g = 2; g < q; g++) {
        uint32_t cand = sm_pow(p, g, (q - 1) / (2 * n));
        if (sm_pow(p, cand, n) == p->q - 1) { psi = cand; break; }
    }
    if (!psi) return -1;
    uint32_t psi_inv = sm_pow(p, psi, 2 * n - 1);
    int bits = 0;
    while ((1ULL << bits) < n) bits++;
    c->n = n;
    c->fwd = (uint32_t *)malloc(n * 4);
    c->inv = (uint32_t *)malloc(n * 4);
    uint32_t *pw = (uint32_t *)malloc(n * 4), *pi = (uint32_t *)malloc(n * 4);
    uint64_t a = 1, b = 1;
    for (size_t k = 0; k < n; k++) {
        pw[k] = (uint32_t)a;
        pi[k] = (uint32_t)b;
        a = a * psi % q;
        b = b * psi_inv % q;
    }
    for (size_t k = 0; k < n; k++) {
        c->fwd[k] = sm_to_mont(p, pw[sm_bitrev(k, bits)]);
        c->inv[k] = sm_to_mont(p, pi[sm_bitrev(k, bits)]);
    }
    free(pw);
    free(pi);
    c->psi = psi;
    c->n_inv = sm_to_mont(p, sm_pow(p, n, q - 2));
    c->last = sm_mont(p, c->inv[1], c->n_inv);
    return 0;
}

void sm_ntt_free(sm_ntt *c) {
    free(c->fwd);
    free(c->inv);
    c->fwd = c->inv = NULL;
}

void sm_ntt_forward(const sm_ntt *c, uint32_t *a) { smi_ntt_forward(c, a); }
void sm_ntt_inverse(const sm_ntt *c, uint32_t *a) { smi_ntt_inverse(c, a); }

typedef struct { uint32_t v; uint16_t slot; } sm_vs;

static int sm_vs_cmp(const void *a, const void *b) {
    uint32_t x = ((const sm_vs *)a)->v, y = ((const sm_vs *)b)->v;
    return x < y ? -1 : x > y;
}

/* Slot k of tau_t(f) is slot out[k] of f (arith.rs:276-323). */
int sm_automorphism_tables(const sm_ntt *c, const size_t *exps, size_t count, uint16_t *out) {
    size_t n = c->n;
    if (n > (1u << 16)) return -1;
    uint32_t *x = (uint32_t *)calloc(n, 4);
    sm_vs *idx = (sm_vs *)malloc(n * sizeof(sm_vs));
    x[1] = 1;
    sm_ntt_forward(c, x);
    for (size_t k = 0; k < n; k++) { idx[k].v = x[k]; idx[k].slot = (uint16_t)k; }
    qsort(idx, n, sizeof(sm_vs), sm_vs_cmp);
    int rc = 0;
    for (size_t k = 1; k < n; k++)
        if (idx[k].v == idx[k - 1].v) rc = -1;
    for (size_t e = 0; e < count && !rc; e++) {
        size_t t = exps[e] % (2 * n);
        memset(x, 0, n * 4);
        if (t < n) x[t] = 1; else x[t - n] = c->p.q - 1;
        sm_ntt_forward(c, x);
        for (size_t k = 0; k < n && !rc; k++) {
            sm_vs key = {x[k], 0};
            sm_vs *hit = (sm_vs *)bsearch(&key, idx, n, sizeof(sm_vs), sm_vs_cmp);
            if (!hit) rc = -1; else out[e * n + k] = hit->slot;
        }
    }
    free(x);
    free(idx);
    return rc;
}

/* ---------------- Ring ---------------- */

#define SM_KEY_MASK_STREAM 0x6b6579206d61736bULL /* big-endian "key mask", mod.rs:58 */
#define SM_KEY_SEED_STREAM 0x6b65792073656564ULL /* big-endian "key seed", client.rs:73 */

int sm_ring_init(sm_ring *r, size_t d, size_t gamma, uint64_t q, uint64_t p, int with_tables) {
    memset(r, 0, sizeof *r);
    if (sm_ntt_init(&r->lq, d, q) || sm_ntt_init(&r->lp, d, p)) return -1;
    r->d = d;
    r->gamma = gamma;
    r->g = 2 * d / gamma + 1;
    r->delta = (uint32_t)(q / 16);
    r->p_mod_q = (uint32_t)(p % q);
    r->neg_off = (uint32_t)((q - p % q) % q);
    r->p_over_q = (uint32_t)(p / q);
    r->p_inv = sm_to_mont(&r->lq.p, sm_pow(&r->lq.p, p % q, q - 2));
    if (with_tables) {
        size_t *exps = (size_t *)malloc(gamma * sizeof(size_t));
        size_t v = 1;
        for (size_t i = 0; i < gamma; i++) { exps[i] = v; v = v * r->g % (2 * d); }
        r->tables = (uint16_t *)malloc(gamma * d * 2);
        uint16_t *other = (uint16_t *)malloc(gamma * d * 2);
        int rc = sm_automorphism_tables(&r->lq, exps, gamma, r->tables) ||
                 sm_automorphism_tables(&r->lp, exps, gamma, other) ||
                 memcmp(r->tables, other, gamma * d * 2) != 0;
        free(exps);
        free(other);
        if (rc) return -1;
    }
    return 0;
}

void sm_ring_free(sm_ring *r) {
    sm_ntt_free(&r->lq);
    sm_ntt_free(&r->lp);
    free(r->tables);
    r->tables = NULL;
}

/* expand_uniform (mod.rs:223-237) with the rejection bound as an argument */
static void sm_expand_uniform_in(orc_chacha *rng, uint64_t m, uint64_t zone, uint32_t *out, size_t count) {
    for (size_t i = 0; i < count; i++) {
        uint64_t u;
        do u = orc_chacha_next_u64(rng); while (u >= zone);
        out[i] = (uint32_t)(u % m);
    }
}

/* expand_uniform (mod.rs:223-237): reject the short last interval of u64. */
static void sm_expand_uniform(orc_chacha *rng, uint64_t m, uint32_t *out, size_t count) {
    sm_expand_uniform_in(rng, m, (UINT64_MAX / m) * m, out, count);
}

/* Test hook: expand_uniform over `limbs` consecutive moduli of one stream (as Ring::key_mask
 * draws its q limb then its p limb), each with the rejection bound given, so the tests can
 * force rejections and check the kernels' serial replay against this loop. */
void sm_expand_uniform_zone(const uint8_t seed[32], uint64_t stream, int limbs, size_t count, const uint32_t *m,
                            const uint64_t *zone, uint32_t *out) {
    orc_chacha rng;
    orc_chacha_from_seed(&rng, seed);
    if (stream) orc_chacha_set_stream(&rng, stream);
    for (int l = 0; l < limbs; l++) sm_expand_uniform_in(&rng, m[l], zone[l], out + (size_t)l * count, count);
}

/* A query's mask a from its seed, coefficient form (server.rs:408). */
void sm_expand_query_mask(const sm_ring *r, const uint8_t seed[32], uint32_t *a) {
    orc_chacha rng;
    orc_chacha_from_seed(&rng, seed);
    sm_expand_uniform(&rng, r->lq.p.q, a, r->d);
}

/* Ring::key_mask (mod.rs:313-330): both limbs, transform form, Montgomery operands. */
void sm_key_mask(const sm_ring *r, const uint8_t seed[32], uint32_t *w_q, uint32_t *w_p) {
    orc_chacha rng;
    orc_chacha_from_seed(&rng, seed);
    orc_chacha_set_stream(&rng, SM_KEY_MASK_STREAM);
    sm_expand_uniform(&rng, r->lq.p.q, w_q, r->d);
    sm_expand_uniform(&rng, r->lp.p.q, w_p, r->d);
    sm_ntt_forward(&r->lq, w_q);
    sm_ntt_forward(&r->lp, w_p);
    for (size_t k = 0; k < r->d; k++) {
        w_q[k] = sm_to_mont(&r->lq.p, w_q[k]);
        w_p[k] = sm_to_mont(&r->lp.p, w_p[k]);
    }
}

uint32_t sm_p_to_q_centred(const sm_ring *r, uint32_t x) { return smi_p_to_q_centred(r, x); }
uint32_t sm_q_to_p_centred(const sm_ring *r, uint32_t c) { return smi_q_to_p_centred(r, c); }
uint32_t sm_mod_down(const sm_ring *r, uint32_t xq, uint32_t xp) { return smi_mod_down(r, xq, xp); }

/* ---------------- Server ---------------- */

/* Signed nibbles of a cell, low nibble first (mod.rs:366-371). */
static void sm_nibbles(const uint8_t *cell, size_t bytes, int8_t *out) {
    for (size_t i = 0; i < bytes; i++) {
        out[2 * i] = (int8_t)((int8_t)(uint8_t)(cell[i] << 4) >> 4);
        out[2 * i + 1] = (int8_t)((int8_t)cell[i] >> 4);
    }
}

/* step1_row_impl (server.rs:330-383). columns: gamma x d signed nibbles. */
static void sm_step1_row(const sm_ring *r, size_t i, const int8_t *columns, uint32_t *out) {
    size_t d = r->d, gamma = r->gamma, two_d = 2 * d;
    size_t e = 1, base = r->g % two_d, k = (d - i) % d;
    while (k) {
        if (k & 1) e = e * base % two_d;
        base = base * base % two_d;
        k >>= 1;
    }
    int32_t *acc = (int32_t *)calloc(d, 4);
    for (size_t j = 0; j < gamma; j++) {
        const int8_t *col = columns + j * d;
        size_t kk = j * e % two_d;
        int flip = kk >= d;
        size_t sh = flip ? kk - d : kk;
        /* coefficient m of h_j lands at m + sh, negated once for X^d = -1 and once more by flip */
        for (size_t m = 0; m < d; m++) {
            int32_t v = col[m];
            size_t at = m + sh;
            int neg = flip;
            if (at >= d) { at -= d; neg = !neg; }
            acc[at] += neg ? -v : v;
        }
    }
    const sm_ntt *lq = &r->lq;
    uint32_t *poly = (uint32_t *)malloc(d * 4);
    for (size_t m = 0; m < d; m++) poly[m] = sm_small(&lq->p, acc[m]);
    sm_ntt_forward(lq, poly);
    uint32_t inv_gamma = sm_to_mont(&lq->p, sm_to_mont(&lq->p, sm_pow(&lq->p, gamma, lq->p.q - 2)));
    const uint16_t *tab = r->tables + i * d;
    for (size_t m = 0; m < d; m++) out[m] = sm_mont(&lq->p, poly[tab[m]], inv_gamma);
    free(acc);
    free(poly);
}

/* Server::table (server.rs:211-250). cells: rows x cell_bytes. step1: gamma x d,
 * values: d x gamma (rows past the cells zero). Returns -1 for a ragged or
 * oversized shard, -2 - m for cell m with bytes past its record. */
int sm_table(const sm_ring *r, size_t cell_bytes, size_t record_bytes, const uint8_t *cells,
             size_t len, uint32_t *step1, int8_t *values) {
    size_t d = r->d, gamma = r->gamma;
    if (len % cell_bytes || len / cell_bytes > d) return -1;
    size_t rows = len / cell_bytes;
    memset(values, 0, d * gamma);
    for (size_t m = 0; m < rows; m++) {
        const uint8_t *cell = cells + m * cell_bytes;
        for (size_t b = record_bytes; b < cell_bytes; b++)
            if (cell[b]) return -2 - (int)m;
        sm_nibbles(cell, cell_bytes, values + m * gamma);
    }
    int8_t *columns = (int8_t *)calloc(gamma * d, 1);
    for (size_t m = 0; m < rows; m++)
        for (size_t j = 0; j < gamma; j++) columns[j * d + m] = values[m * gamma + j];
#pragma omp parallel for schedule(dynamic)
    for (size_t i = 0; i < gamma; i++) sm_step1_row(r, i, columns, step1 + i * d);
    free(columns);
    return 0;
}

/* Server::seat (server.rs:177-203). Outputs transform form, Montgomery operands. */
int sm_seat(const sm_ring *r, const uint32_t *key_q, const uint32_t *key_p, const uint8_t mask_seed[32],
            uint32_t *sq, uint32_t *sp, uint32_t *w_q, uint32_t *w_p) {
    size_t d = r->d;
    for (size_t k = 0; k < d; k++)
        if (key_q[k] >= r->lq.p.q || key_p[k] >= r->lp.p.q) return -1;
    memcpy(sq, key_q, d * 4);
    memcpy(sp, key_p, d * 4);
    sm_ntt_forward(&r->lq, sq);
    sm_ntt_forward(&r->lp, sp);
    for (size_t k = 0; k < d; k++) {
        sq[k] = sm_to_mont(&r->lq.p, sq[k]);
        sp[k] = sm_to_mont(&r->lp.p, sp[k]);
    }
    sm_key_mask(r, mask_seed, w_q, w_p);
    return 0;
}

/* respond_impl (server.rs:385-499), portable build. nb body coefficients are returned. */
int sm_respond(const sm_ring *r, const uint32_t *step1, const int8_t *values, size_t rows,
               const uint32_t *sq, const uint32_t *sp, const uint32_t *w_q, const uint32_t *w_p,
               const uint8_t seed[32], const uint32_t *qbody, uint32_t *mask, uint32_t *body, size_t nb) {
    return smi_respond(r, step1, values, rows, sq, sp, w_q, w_p, seed, qbody, mask, body, nb);
}

/* Kernel::Avx512's availability test (server.rs:50-58): the respond_avx512 form runs
 * only where the CPU has all four feature sets. */
int sm_avx512_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl");
}

/* ---------------- Client ---------------- */

/* f(X^t) for a small signed polynomial (mod.rs:240-253). */
static void sm_tau_signed(const int64_t *v, size_t d, size_t t, int64_t *out) {
    memset(out, 0, d * 8);
    for (size_t i = 0; i < d; i++) {
        size_t k = t * i % (2 * d);
        if (k < d) out[k] = v[i]; else out[k - d] = -v[i];
    }
}

struct sm_client {
    const sm_ring *ring;
    uint32_t *s_q;      /* transform of the secret at q, Montgomery operands */
    uint32_t *key_q, *key_p;
    uint8_t mask_seed[32];
    orc_gauss noise;
    orc_chacha seeds;
};

/* Client::from_seed -> drawn -> build (client.rs:137-244). */
sm_client *sm_client_from_seed(const sm_ring *r, double sigma, const uint8_t seed[32]) {
    size_t d = r->d;
    sm_client *c = (sm_client *)calloc(1, sizeof *c);
    c->ring = r;
    orc_gauss_from_seed(&c->noise, sigma, seed);
    orc_chacha_from_seed(&c->seeds, seed);
    orc_chacha_set_stream(&c->seeds, 1);
    int64_t *secret = (int64_t *)malloc(d * 8), *err = (int64_t *)malloc(d * 8), *ts = (int64_t *)malloc(d * 8);
    for (size_t k = 0; k < d; k++) secret[k] = orc_gauss_sample(&c->noise);
    for (size_t k = 0; k < d; k++) err[k] = orc_gauss_sample(&c->noise);
    {
        /* key_mask_seed (client.rs:78-85): a copy of the seed source on stream "key seed",
         * word position 0 */
        orc_chacha km = c->seeds;
        km.stream = SM_KEY_SEED_STREAM;
        km.counter = 0;
        km.index = 64;
        orc_chacha_fill_bytes(&km, c->mask_seed, 32);
    }
    const sm_ntt *lq = &r->lq, *lp = &r->lp;
    uint32_t *w_q = (uint32_t *)malloc(d * 4), *w_p = (uint32_t *)malloc(d * 4);
    sm_key_mask(r, c->mask_seed, w_q, w_p);
    c->s_q = (uint32_t *)malloc(d * 4);
    uint32_t *s_p = (uint32_t *)malloc(d * 4);
    for (size_t k = 0; k < d; k++) { c->s_q[k] = sm_small(&lq->p, secret[k]); s_p[k] = sm_small(&lp->p, secret[k]); }
    sm_ntt_forward(lq, c->s_q);
    sm_ntt_forward(lp, s_p);
    uint32_t *sw_q = (uint32_t *)malloc(d * 4), *sw_p = (uint32_t *)malloc(d * 4);
    for (size_t k = 0; k < d; k++) {
        sw_q[k] = sm_mont(&lq->p, c->s_q[k], w_q[k]);
        sw_p[k] = sm_mont(&lp->p, s_p[k], w_p[k]);
    }
    sm_ntt_inverse(lq, sw_q);
    sm_ntt_inverse(lp, sw_p);
    sm_tau_signed(secret, d, r->g, ts);
    uint32_t p_mod_q = sm_to_mont(&lq->p, r->p_mod_q);
    c->key_q = (uint32_t *)malloc(d * 4);
    c->key_p = (uint32_t *)malloc(d * 4);
    for (size_t k = 0; k < d; k++) {
        uint32_t lead = sm_mont(&lq->p, sm_small(&lq->p, ts[k]), p_mod_q);
        c->key_q[k] = sm_add(&lq->p, sm_sub(&lq->p, lead, sw_q[k]), sm_small(&lq->p, err[k]));
        c->key_p[k] = sm_sub(&lp->p, sm_small(&lp->p, err[k]), sw_p[k]);
    }
    for (size_t k = 0; k < d; k++) c->s_q[k] = sm_to_mont(&lq->p, c->s_q[k]);
    free(secret); free(err); free(ts); free(w_q); free(w_p); free(s_p); free(sw_q); free(sw_p);
    return c;
}

void sm_client_free(sm_client *c) {
    if (!c) return;
    free(c->s_q); free(c->key_q); free(c->key_p);
    orc_gauss_free(&c->noise);
    free(c);
}

void sm_client_key(const sm_client *c, uint8_t mask_seed[32], uint32_t *key_q, uint32_t *key_p) {
    memcpy(mask_seed, c->mask_seed, 32);
    memcpy(key_q, c->key_q, c->ring->d * 4);
    memcpy(key_p, c->key_p, c->ring->d * 4);
}

/* Client::query -> query_with (client.rs:252-310): b = -a*s + e + delta*X^-row. */
int sm_client_query(sm_client *c, size_t row, uint8_t seed[32], uint32_t *body) {
    const sm_ring *r = c->ring;
    size_t d = r->d;
    orc_chacha_fill_bytes(&c->seeds, seed, 32);
    int64_t *err = (int64_t *)malloc(d * 8);
    for (size_t k = 0; k < d; k++) err[k] = orc_gauss_sample(&c->noise);
    if (row >= d) { free(err); return -1; }
    const sm_ntt *lq = &r->lq;
    uint32_t *a = (uint32_t *)malloc(d * 4);
    sm_expand_query_mask(r, seed, a);
    sm_ntt_forward(lq, a);
    for (size_t k = 0; k < d; k++) a[k] = sm_mont(&lq->p, a[k], c->s_q[k]);
    sm_ntt_inverse(lq, a);
    size_t target = (d - row) & (d - 1);
    uint32_t message = row == 0 ? r->delta : lq->p.q - r->delta;
    for (size_t k = 0; k < d; k++) {
        uint32_t m = k == target ? message : 0;
        body[k] = sm_add(&lq->p, sm_sub(&lq->p, sm_small(&lq->p, err[k]), a[k]), m);
    }
    free(err);
    free(a);
    return 0;
}

/* Client::decode (client.rs:312-340): the record a response carries. */
void sm_client_decode(const sm_client *c, const uint32_t *mask, const uint32_t *body, size_t nb, uint8_t *record) {
    const sm_ring *r = c->ring;
    size_t d = r->d;
    const sm_ntt *lq = &r->lq;
    uint32_t *ph = (uint32_t *)malloc(d * 4);
    memcpy(ph, mask, d * 4);
    sm_ntt_forward(lq, ph);
    for (size_t k = 0; k < d; k++) ph[k] = sm_mont(&lq->p, ph[k], c->s_q[k]);
    sm_ntt_inverse(lq, ph);
    uint64_t delta = r->delta, half = delta / 2;
    for (size_t k = 0; k + 1 < nb; k += 2) {
        uint8_t nib[2];
        for (int h = 0; h < 2; h++) {
            uint64_t x = sm_add(&lq->p, ph[k + h], body[k + h]);
            nib[h] = (uint8_t)(((x + half) / delta) & 15);
        }
        record[k / 2] = (uint8_t)(nib[0] | (nib[1] << 4));
    }
    free(ph);
}

/* ---------------- wire ---------------- */

static size_t sm_bits(uint64_t m) {
    size_t b = 0;
    while ((1ULL << b) < m) b++; /* bit length of m - 1 */
    return b;
}

size_t sm_packed_bytes(size_t count, uint64_t m) { return (count * sm_bits(m) + 7) / 8; }

/* pack_coefficients_tight at one modulus: bit b of value i at bit i*bits+b, LSB first. */
size_t sm_pack(const uint32_t *values, size_t count, uint64_t m, uint8_t *out) {
    size_t bits = sm_bits(m), len = sm_packed_bytes(count, m);
    memset(out, 0, len);
    for (size_t i = 0; i < count; i++)
        for (size_t b = 0; b < bits; b++)
            if ((values[i] >> b) & 1) {
                size_t at = i * bits + b;
                out[at / 8] |= (uint8_t)(1u << (at % 8));
            }
    return len;
}

/* Cells of the reference's KAT inputs (tests/special_modulus_kat.rs:37-46): each cell's
 * record bytes from one ChaCha20Rng::seed_from_u64 stream, the rest zero. */
void sm_kat_cells(uint64_t seed, size_t rows, size_t cell_bytes, size_t record_bytes, int worst, uint8_t *out) {
    orc_chacha rng;
    orc_chacha_seed_from_u64(&rng, seed);
    memset(out, 0, rows * cell_bytes);
    for (size_t m = 0; m < rows; m++) {
        if (worst) memset(out + m * cell_bytes, 0x88, record_bytes);
        else orc_chacha_fill_bytes(&rng, out + m * cell_bytes, record_bytes);
    }
}
