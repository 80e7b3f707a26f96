/*
 * Oracle (test infrastructure only): the client side of the reference's packing-key
 * known-answer cells, restated so that the recorded response digests can be replayed.
 * Nothing here is part of the product; it generates inputs (keys, queries, database
 * bytes) in the reference's exact draw order.
 *
 *   RlweSecretKey::generate                 src/rlwe/enc.rs:48-51
 *   Poly::sample_gaussian_moduli            src/math/poly.rs:607-631
 *   ModQ::from_signed                       src/math/modular.rs:234-241
 *   apply_automorphism                      src/rlwe/galois.rs:12-45
 *   generate_ks_matrix draws (error only;   src/ks/setup.rs:108-145, 182-192
 *     its `a` is OS entropy)
 *   setup_with_rng draw order               src/pir/setup.rs:408-494
 *   ClientPackingKeys::generate_once        src/inspiring/inspiring2.rs:794-831
 *   expand_mask (stream 0, next_u64 % q)    src/inspiring/inspiring2.rs:839-887
 *   generate_ksk_body                       src/inspiring/inspiring2.rs:891-929
 *   GadgetVector::power                     src/rgsw/types.rs:117-126
 *   ClientSession::seeded_row / query       src/pir/session.rs:399-433
 *   SeededRgswCiphertext::encrypt_with_rng  src/rgsw/types.rs:345-386 (orc_seeded_row)
 *   inverse_monomial                        src/pir/encode_db.rs:43-70
 *   ShardConfig::try_index_to_shard         src/params.rs:1367-1380
 *   encode_database / extract_column_value  src/pir/encode_db.rs:76-159
 *   the KAT cell (database bytes, seeds,    tests/packing_key_fixture/mod.rs:64-147
 *     probe query, next_query)
 *   ClientPackingKeys bincode               src/inspiring/inspiring2.rs:684-697, src/math/poly.rs:286-310
 *   kept_mask_rows (approximate gadget:     src/inspiring/inspiring2.rs:839-856
 *     stream gamma, lowest digit dropped)
 *
 * and the client's decode of a packed response (PIR.Extract):
 *
 *   RlweCiphertext::decrypt                 src/rlwe/enc.rs:91-107
 *   PlaintextRounding::round                src/rlwe/enc.rs:14-44
 *   mod_switch_secret_key                   src/pir/mod_switch.rs:221-252
 *   reconstruct_entry                       src/pir/encode_db.rs:161-176
 *
 * Pinned by tests/golden/client_packing_key_kat.txt (key bytes and the sampler position
 * after them), by the SESSIONKEY and RESP digests of
 * tests/golden/packing_seat_response_kat.txt, and (the decode) by the reference's
 * post-switch error known answers at four target moduli (src/pir/mod_switch.rs:1070-1114).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

struct orc_cell {
    size_t n, entry, gamma, records, dropped;
    uint64_t q, delta;
    uint8_t *db;
    uint8_t w_seed[32];
    uint64_t *s, *y_body;
    orc_gauss g;
    orc_chacha rng;
    orc_ntt ntt;
};

static uint64_t from_signed(int64_t v, uint64_t q) {
    uint64_t mag = v < 0 ? (uint64_t)(-v) : (uint64_t)v;
    uint64_t rem = mag % q;
    return v < 0 ? (rem ? q - rem : 0) : rem;
}

void orc_sample_poly(orc_gauss *g, size_t n, uint64_t q, uint64_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = from_signed(orc_gauss_sample(g), q);
}

void orc_apply_automorphism(const uint64_t *in, size_t n, size_t t, uint64_t q, uint64_t *out) {
    memset(out, 0, n * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) {
        size_t j = (size_t)(((u128)t * i) % (2 * n));
        if (j < n) out[j] = (out[j] + in[i]) % q;
        else out[j - n] = (out[j - n] + q - in[i]) % q;
    }
}

/* a * b mod (X^n + 1, q) through the oracle's NTT (standard in, standard out) */
static void negacyclic_mul(const orc_ntt *c, const uint64_t *a, const uint64_t *b, uint64_t *out) {
    size_t n = c->n;
    uint64_t *x = (uint64_t *)malloc(n * 8), *y = (uint64_t *)malloc(n * 8);
    memcpy(x, a, n * 8);
    memcpy(y, b, n * 8);
    orc_ntt_forward(c, x);
    orc_ntt_forward(c, y);
    for (size_t i = 0; i < n; i++) x[i] = orc_mont_mul(c, x[i], y[i]);
    orc_ntt_inverse(c, x);
    memcpy(out, x, n * 8);
    free(x);
    free(y);
}

/* ClientPackingKeys::generate_once (base 2^20, 3 digits): the mask expands 3 rows from ChaCha20(w_seed)
 * on stream 0 (exact gadget) or stream gamma (approximate gadget, which drops the lowest digit and
 * keeps rows 1 and 2). Kept row r is y_body[r] = tau_g(s) * 2^(20 (dropped + r)) - s * w_mask[r] + e_r,
 * the error rows drawn in row order. */
int orc_client_packing_keys_gadget(size_t n, size_t gamma, size_t dropped, const uint64_t *s,
                                   const uint8_t w_seed[32], orc_gauss *g, uint64_t *y_body) {
    if (gamma == 0 || (gamma & (gamma - 1)) || gamma > n / 2 || dropped > 1) return -1;
    const uint64_t q = ORC_Q;
    orc_ntt ntt;
    if (orc_ntt_init(&ntt, n, q)) return -1;
    size_t gen = (2 * n / gamma + 1) % (2 * n);
    uint64_t *tau = (uint64_t *)malloc(n * 8), *w = (uint64_t *)malloc(n * 8), *sw = (uint64_t *)malloc(n * 8);
    orc_apply_automorphism(s, n, gen, q, tau);
    orc_chacha r;
    orc_chacha_from_seed(&r, w_seed);
    orc_chacha_set_stream(&r, dropped ? (uint64_t)gamma : 0);
    uint64_t power = 1;
    for (size_t digit = 0; digit < 3; digit++) {
        for (size_t i = 0; i < n; i++) w[i] = orc_chacha_next_u64(&r) % q;
        if (digit >= dropped) {
            negacyclic_mul(&ntt, s, w, sw);
            uint64_t *y = y_body + (digit - dropped) * n;
            for (size_t i = 0; i < n; i++) {
                uint64_t e = from_signed(orc_gauss_sample(g), q);
                uint64_t t = (uint64_t)(((u128)tau[i] * power) % q);
                y[i] = (uint64_t)(((u128)t + (q - sw[i]) + e) % q);
            }
        }
        power = (uint64_t)(((u128)power << 20) % q);
    }
    free(tau);
    free(w);
    free(sw);
    orc_ntt_free(&ntt);
    return 0;
}

int orc_client_packing_keys(size_t n, size_t gamma, const uint64_t *s, const uint8_t w_seed[32], orc_gauss *g,
                            uint64_t *y_body) {
    return orc_client_packing_keys_gadget(n, gamma, 0, s, w_seed, g, y_body);
}

static size_t put_u64(uint8_t *o, size_t at, uint64_t v) {
    if (o) for (int k = 0; k < 8; k++) o[at + k] = (uint8_t)(v >> (8 * k));
    return at + 8;
}

/* bincode of ClientPackingKeys { y_body, z_body: [], full_key: false, num_to_pack }: each Poly is
 * the tight wire form (u64 byte count, coefficients LSB-first at 60 bits, moduli, q, dim,
 * crt_q0_inv_mod_q1 = 0, is_ntt = false). out may be NULL to get the length. */
size_t orc_packing_keys_bytes(size_t n, const uint64_t *y_body, size_t rows, size_t gamma, uint8_t *out) {
    const uint64_t q = ORC_Q;
    const unsigned bits = 64 - (unsigned)__builtin_clzll(q - 1);
    size_t at = put_u64(out, 0, rows);
    for (size_t r = 0; r < rows; r++) {
        size_t nbytes = (n * bits + 7) / 8;
        at = put_u64(out, at, nbytes);
        if (out) {
            memset(out + at, 0, nbytes);
            for (size_t i = 0; i < n; i++) {
                uint64_t v = y_body[r * n + i];
                size_t bit = i * bits;
                for (unsigned b = 0; b < bits; b++, bit++)
                    if ((v >> b) & 1) out[at + bit / 8] |= (uint8_t)(1u << (bit % 8));
            }
        }
        at += nbytes;
        at = put_u64(out, at, 1);
        at = put_u64(out, at, q);
        at = put_u64(out, at, q);
        at = put_u64(out, at, n);
        at = put_u64(out, at, 0);
        if (out) out[at] = 0;
        at += 1;
    }
    at = put_u64(out, at, 0);   /* z_body */
    if (out) out[at] = 0;       /* full_key */
    at += 1;
    at = put_u64(out, at, gamma);
    return at;
}

/* client_packing_key_kat cell (tests/client_packing_key_kat.rs:20-50): sampler from
 * seed_from_u64(seed), secret, keys under w_seed = [gamma as u8; 32], then the next sample. */
int orc_client_key_kat(size_t n, size_t gamma, uint64_t seed, uint64_t *y_body, int64_t *next) {
    orc_gauss g;
    orc_gauss_with_seed(&g, 6.4, seed);
    uint64_t *s = (uint64_t *)malloc(n * 8);
    orc_sample_poly(&g, n, ORC_Q, s);
    uint8_t w_seed[32];
    memset(w_seed, (int)(uint8_t)gamma, 32);
    int rc = orc_client_packing_keys(n, gamma, s, w_seed, &g, y_body);
    *next = orc_gauss_sample(&g);
    free(s);
    orc_gauss_free(&g);
    return rc;
}

static size_t log2_sz(size_t n) {
    size_t l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

orc_cell *orc_cell_new(size_t n, size_t entry, uint64_t seed) { return orc_cell_new_gadget(n, entry, seed, 0); }

orc_cell *orc_cell_new_gadget(size_t n, size_t entry, uint64_t seed, size_t dropped) {
    if (n < 2 || (n & (n - 1)) || entry == 0 || dropped > 1) return NULL;
    orc_cell *c = (orc_cell *)calloc(1, sizeof(orc_cell));
    c->n = n;
    c->dropped = dropped;
    c->entry = entry;
    c->gamma = (entry + 1) / 2;
    c->records = 3 * n;
    c->q = ORC_Q;
    c->delta = ORC_Q / 65537;
    size_t len = c->records * entry;
    c->db = (uint8_t *)malloc(len);
    for (size_t i = 0; i < len; i++) c->db[i] = (uint8_t)(((uint64_t)i * 0x9E3779B9ULL + seed) >> 7);
    if (orc_ntt_init(&c->ntt, n, c->q)) { orc_cell_free(c); return NULL; }
    orc_gauss_with_seed(&c->g, 6.4, seed);
    orc_chacha_seed_from_u64(&c->rng, seed ^ 0x5eed);
    c->s = (uint64_t *)malloc(n * 8);
    c->y_body = (uint64_t *)malloc(3 * n * 8);
    orc_sample_poly(&c->g, n, c->q, c->s);
    /* the tree-packing galois keys: log2(n) matrices of 3 rows, one error polynomial each */
    uint64_t *scratch = (uint64_t *)malloc(n * 8);
    for (size_t i = 0; i < log2_sz(n) * 3; i++) orc_sample_poly(&c->g, n, c->q, scratch);
    uint8_t v_seed[32];
    orc_chacha_fill_bytes(&c->rng, c->w_seed, 32);
    orc_chacha_fill_bytes(&c->rng, v_seed, 32);
    if (orc_client_packing_keys_gadget(n, c->gamma, dropped, c->s, c->w_seed, &c->g, c->y_body)) {
        free(scratch);
        orc_cell_free(c);
        return NULL;
    }
    /* the probe query's error (its row seed comes from OS entropy and does not matter) */
    orc_sample_poly(&c->g, n, c->q, scratch);
    free(scratch);
    return c;
}

void orc_cell_free(orc_cell *c) {
    if (!c) return;
    free(c->db);
    free(c->s);
    free(c->y_body);
    orc_gauss_free(&c->g);
    orc_ntt_free(&c->ntt);
    free(c);
}

void orc_cell_info(const orc_cell *c, uint8_t w_seed[32], uint64_t *y_body, uint8_t *db, uint64_t *secret) {
    if (w_seed) memcpy(w_seed, c->w_seed, 32);
    if (y_body) memcpy(y_body, c->y_body, (3 - c->dropped) * c->n * 8);
    if (db) memcpy(db, c->db, c->records * c->entry);
    if (secret) memcpy(secret, c->s, c->n * 8);
}

/* shard columns as encode_database lays them out: gamma x n, column j of record m is
 * entry[2j] | entry[2j+1] << 8 (a lone last byte stands alone) */
void orc_cell_columns(const orc_cell *c, uint32_t shard, uint64_t *columns) {
    size_t n = c->n, e = c->entry;
    for (size_t j = 0; j < c->gamma; j++)
        for (size_t m = 0; m < n; m++) {
            const uint8_t *rec = c->db + ((size_t)shard * n + m) * e;
            size_t off = 2 * j;
            uint64_t v = 0;
            if (off + 1 < e) v = rec[off] | ((uint64_t)rec[off + 1] << 8);
            else if (off < e) v = rec[off];
            columns[j * n + m] = v;
        }
}

/* SeededRgswCiphertext::encrypt_with_rng at a one-row gadget: the row seed from rng,
 * a = expand(seed), e from the sampler, b = -a*s + e + delta * X^(-local) */
void orc_seeded_row(size_t n, const uint64_t *s, size_t local, orc_chacha *rng, orc_gauss *g, uint8_t seed[32],
                    uint64_t *b) {
    const uint64_t q = ORC_Q, delta = ORC_Q / 65537;
    orc_ntt ntt;
    if (orc_ntt_init(&ntt, n, q)) return;
    orc_chacha_fill_bytes(rng, seed, 32);
    uint64_t *a = (uint64_t *)malloc(n * 8), *as = (uint64_t *)malloc(n * 8);
    orc_expand_seed(seed, n, q, a);
    negacyclic_mul(&ntt, a, s, as);
    for (size_t i = 0; i < n; i++) {
        uint64_t e = from_signed(orc_gauss_sample(g), q);
        b[i] = (uint64_t)(((u128)(q - as[i]) % q + e) % q);
    }
    /* X^(-local): 1 at 0 when local = 0, else -1 at n - local */
    local %= n;
    if (local == 0) b[0] = (uint64_t)(((u128)b[0] + delta) % q);
    else b[n - local] = (uint64_t)(((u128)b[n - local] + (q - delta)) % q);
    free(a);
    free(as);
    orc_ntt_free(&ntt);
}

/* next_query: the cell's stream at record index */
void orc_cell_query(orc_cell *c, uint64_t index, uint32_t *shard, uint8_t seed[32], uint64_t *b) {
    *shard = (uint32_t)(index / c->n);
    orc_seeded_row(c->n, c->s, (size_t)(index % c->n), &c->rng, &c->g, seed, b);
}

/* heap handles so that a test can replay a draw order of its own */
orc_gauss *orc_gauss_new(double sigma, uint64_t seed) {
    orc_gauss *g = (orc_gauss *)calloc(1, sizeof(orc_gauss));
    orc_gauss_with_seed(g, sigma, seed);
    return g;
}
void orc_gauss_delete(orc_gauss *g) {
    if (!g) return;
    orc_gauss_free(g);
    free(g);
}
orc_chacha *orc_chacha_new_u64(uint64_t seed) {
    orc_chacha *r = (orc_chacha *)calloc(1, sizeof(orc_chacha));
    orc_chacha_seed_from_u64(r, seed);
    return r;
}
void orc_chacha_delete(orc_chacha *r) { free(r); }

size_t orc_cell_gamma(const orc_cell *c) { return c->gamma; }

/* ---------------- PIR.Extract: the client's decode of a packed response ---------------- */

/* Coefficients 0..count-1 of the phase a*s + b mod (X^n + 1, q), where b holds count
 * coefficients and is zero above them (a packed response keeps only its b prefix). s is in
 * [0, q). With s centred below 2^32 in magnitude the negacyclic sums run exactly in i128;
 * otherwise each product is reduced on its own. */
void orc_phase(size_t n, uint64_t q, const uint64_t *s, const uint64_t *a, const uint64_t *b, size_t count,
               uint64_t *out) {
    int small = 1;
    int64_t *sc = (int64_t *)malloc(n * sizeof(int64_t));
    for (size_t i = 0; i < n; i++) {
        uint64_t v = s[i] % q;
        int64_t c = v > q / 2 ? -(int64_t)(q - v) : (int64_t)v;
        if (c >= (1ll << 32) || c <= -(1ll << 32)) small = 0;
        sc[i] = c;
    }
    for (size_t i = 0; i < count && i < n; i++) {
        uint64_t r;
        if (small) {
            __int128 acc = 0;
            for (size_t j = 0; j <= i; j++) acc += (__int128)a[j] * sc[i - j];
            for (size_t j = i + 1; j < n; j++) acc -= (__int128)a[j] * sc[n + i - j];
            __int128 m = acc % (__int128)q;
            if (m < 0) m += q;
            r = (uint64_t)m;
        } else {
            uint64_t pos = 0, neg = 0;
            for (size_t j = 0; j <= i; j++) pos = (uint64_t)(((u128)pos + (u128)a[j] * (s[i - j] % q)) % q);
            for (size_t j = i + 1; j < n; j++) neg = (uint64_t)(((u128)neg + (u128)a[j] * (s[n + i - j] % q)) % q);
            r = (pos + q - neg) % q;
        }
        out[i] = (uint64_t)(((u128)r + b[i] % q) % q);
    }
    free(sc);
}

/* PlaintextRounding::round: floor((v + floor(delta / 2)) / delta) mod p. The reference's
 * reciprocal estimate plus one correction is this exact quotient. */
uint64_t orc_round_plaintext(uint64_t v, uint64_t delta, uint64_t p) {
    u128 num = (u128)v + delta / 2;
    return (uint64_t)(num / delta) % p;
}

/* RlweCiphertext::decrypt restricted to coefficients 0..count-1. */
void orc_decrypt(size_t n, uint64_t q, const uint64_t *s, const uint64_t *a, const uint64_t *b, size_t count,
                 uint64_t delta, uint64_t p, uint64_t *out) {
    orc_phase(n, q, s, a, b, count, out);
    for (size_t i = 0; i < count && i < n; i++) out[i] = orc_round_plaintext(out[i], delta, p);
}

/* mod_switch_secret_key: each coefficient's centred lift mod q_old (negative above
 * floor(q_old / 2)), reduced mod q_new. */
void orc_switch_secret(size_t n, const uint64_t *s, uint64_t q_old, uint64_t q_new, uint64_t *out) {
    uint64_t half = q_old / 2;
    for (size_t i = 0; i < n; i++) {
        uint64_t c = s[i];
        int negative = c > half;
        uint64_t mag = negative ? q_old - c : c;
        uint64_t red = mag % q_new;
        out[i] = negative ? (red ? q_new - red : 0) : red;
    }
}

/* reconstruct_entry: column j fills bytes 2j (low) and 2j+1 (high) while they fit. */
void orc_reconstruct_entry(const uint64_t *values, size_t count, size_t entry, uint8_t *out) {
    memset(out, 0, entry);
    for (size_t j = 0; j < count; j++) {
        size_t off = 2 * j;
        if (off < entry) out[off] = (uint8_t)(values[j] & 0xff);
        if (off + 1 < entry) out[off + 1] = (uint8_t)((values[j] >> 8) & 0xff);
    }
}
