/*
 * Oracle (test infrastructure only): InspiRING TwoPacking respond, restating
 *   PackParams::build                      src/inspiring/inspiring2.rs:209-286
 *   derive_automorph_tables                src/inspiring/inspiring2.rs:330-379
 *   kept_mask_rows / expand_mask           src/inspiring/inspiring2.rs:839-887
 *   OfflinePackingKeys::generate           src/inspiring/inspiring2.rs:537-598
 *   rotation_sums                          src/inspiring/inspiring2.rs:964-1004
 *   fold_rotations_backward                src/inspiring/inspiring2.rs:1014-1085
 *   LastDigitNtt                           src/inspiring/inspiring2.rs:1094-1161
 *   rounded_digit_rows                     src/inspiring/inspiring2.rs:1184-1225
 *   gadget_decompose(_default_q)           src/rgsw/external_product.rs:30-99
 *   packing_online_gather                  src/inspiring/inspiring2.rs:1393-1485
 *   ShardPackingStore::build (step-1 sums) src/inspiring/shard_store.rs:86-156
 *   check_columns_and_fold_b / fold_stripe src/inspiring/shard_store.rs:197-265, 378-427
 *   CheckedColumns::packing_offline        src/inspiring/shard_store.rs:282-342
 *   pack_shard (both branches)             src/pir/respond.rs:770-863
 *   plaintext_multiply_columns, sample_extract_coeff0
 *                                          src/pir/respond.rs:585-599, src/rlwe/enc.rs:134-174
 *   mod_switch_poly                        src/pir/mod_switch.rs:148-175
 *   ServerResponse bincode (tight)         src/pir/respond.rs:111-176, src/math/poly.rs:166-225,286-310
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static uint64_t csubq(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }
static uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) { return csubq(a + b, q); }
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

typedef struct { uint64_t v; uint32_t e; } vexp;
static int cmp_vexp(const void *x, const void *y) {
    const vexp *a = (const vexp *)x, *b = (const vexp *)y;
    return a->v < b->v ? -1 : a->v > b->v ? 1 : 0;
}

/* slot exponent map shared by every table: exps[j] = e with w_j = w_0^e, and
 * slot_of_exp[e/2] = j */
static int slot_exponents(const orc_pack *p, uint32_t *exps, uint32_t *slot_of_exp) {
    size_t n = p->n, two_n = 2 * n;
    uint64_t q = p->q;
    vexp *pw = (vexp *)malloc(n * sizeof(vexp));
    for (size_t e = 1, k = 0; e < two_n; e += 2, k++) {
        uint64_t v = e < n ? p->mono_ntt[e * n + 0] % q : (q - p->mono_ntt[(e - n) * n + 0] % q) % q;
        pw[k].v = v;
        pw[k].e = (uint32_t)e;
    }
    qsort(pw, n, sizeof(vexp), cmp_vexp);
    for (size_t k = 1; k < n; k++)
        if (pw[k].v == pw[k - 1].v) { free(pw); return -1; }
    for (size_t j = 0; j < n; j++) slot_of_exp[j] = UINT32_MAX;
    for (size_t j = 0; j < n; j++) {
        uint64_t root = p->mono_ntt[1 * n + j] % q;
        size_t lo = 0, hi = n;
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (pw[mid].v < root) lo = mid + 1; else hi = mid;
        }
        if (lo >= n || pw[lo].v != root) { free(pw); return -1; }
        uint32_t e = pw[lo].e;
        if (slot_of_exp[e / 2] != UINT32_MAX) { free(pw); return -1; }
        slot_of_exp[e / 2] = (uint32_t)j;
        exps[j] = e;
    }
    free(pw);
    return 0;
}

int orc_automorph_table(const orc_pack *p, size_t t, uint32_t *out) {
    size_t n = p->n, two_n = 2 * n;
    uint32_t *exps = (uint32_t *)malloc(n * 4), *soe = (uint32_t *)malloc(n * 4);
    int rc = slot_exponents(p, exps, soe);
    if (rc == 0)
        for (size_t j = 0; j < n; j++) out[j] = soe[((t * exps[j]) % two_n) / 2];
    free(exps);
    free(soe);
    return rc;
}

int orc_pack_init(orc_pack *p, size_t n, size_t gamma, size_t gadget_len, size_t dropped) {
    memset(p, 0, sizeof *p);
    p->n = n;
    p->gamma = gamma;
    p->q = ORC_Q;
    p->gadget_base = 1ULL << 20;
    p->gadget_len = gadget_len;
    p->dropped = dropped;
    if (orc_ntt_init(&p->ntt, n, p->q)) return -1;
    size_t two_n = 2 * n;
    p->generator = two_n / gamma + 1;
    p->gen_pows = (size_t *)malloc(n * sizeof(size_t));
    size_t v = 1;
    for (size_t i = 0; i < n; i++) { p->gen_pows[i] = v; v = (v * p->generator) % two_n; }
    p->mod_inv_gamma = powmod(gamma, p->q - 2, p->q);
    p->mono_ntt = (uint64_t *)calloc(n * n, 8);
    p->neg_mono_ntt = (uint64_t *)calloc(n * n, 8);
    for (size_t j = 0; j < n; j++) {
        p->mono_ntt[j * n + j] = 1;
        orc_ntt_forward(&p->ntt, p->mono_ntt + j * n);
        p->neg_mono_ntt[j * n + j] = p->q - 1;
        orc_ntt_forward(&p->ntt, p->neg_mono_ntt + j * n);
    }
    uint32_t *exps = (uint32_t *)malloc(n * 4), *soe = (uint32_t *)malloc(n * 4);
    if (slot_exponents(p, exps, soe)) { free(exps); free(soe); return -2; }
    p->tables = (uint32_t *)malloc(gamma * n * 4);
    for (size_t i = 0; i < gamma; i++) {
        size_t t = p->gen_pows[i];
        for (size_t j = 0; j < n; j++) p->tables[i * n + j] = soe[((t * exps[j]) % two_n) / 2];
    }
    free(exps);
    free(soe);
    return 0;
}

void orc_pack_free(orc_pack *p) {
    free(p->gen_pows); free(p->mono_ntt); free(p->neg_mono_ntt); free(p->tables);
    orc_ntt_free(&p->ntt);
}

static size_t key_rows(const orc_pack *p) { return p->gadget_len - p->dropped; }

void orc_expand_w_mask(const orc_pack *p, const uint8_t w_seed[32], uint64_t *w_mask) {
    orc_chacha r;
    orc_chacha_from_seed(&r, w_seed);
    orc_chacha_set_stream(&r, p->dropped == 0 ? 0 : (uint64_t)p->gamma);
    size_t n = p->n;
    for (size_t row = 0; row < p->gadget_len; row++) {
        for (size_t j = 0; j < n; j++) {
            uint64_t c = orc_chacha_next_u64(&r) % p->q;
            if (row >= p->dropped) w_mask[(row - p->dropped) * n + j] = c;
        }
    }
}

/* constant poly c in NTT form: every slot holds c*R mod q (mono_ntt of X^0 times c) */
static void const_ntt(const orc_pack *p, uint64_t c, uint64_t *out) {
    memset(out, 0, p->n * 8);
    out[0] = c;
    orc_ntt_forward(&p->ntt, out);
}

int orc_pack_set_simd(orc_pack *p, int on) {
    p->simd = on && p->q == ((1ULL << 60) - (1ULL << 14) + 1) && p->n % 8 == 0 && orc_ifma_available();
    return p->simd;
}

/* Poly::mul_acc_ntt_domain (poly.rs:999-1040): the IFMA kernel when switched on */
static void mul_acc(const orc_pack *p, uint64_t *acc, const uint64_t *a, const uint64_t *b) {
    if (p->simd) {
        orc_ifma_mont_mul_acc(acc, a, b, p->n, p->q, p->ntt.q_inv_neg);
        return;
    }
    for (size_t j = 0; j < p->n; j++) acc[j] = addmod(acc[j], orc_mont_mul(&p->ntt, a[j], b[j]), p->q);
}

void orc_rotation_sums(const orc_pack *p, const uint64_t *inputs_ntt, size_t count, uint64_t *out) {
    size_t n = p->n, two_n = 2 * n;
    uint64_t *inv = (uint64_t *)malloc(n * 8), *r = (uint64_t *)malloc(n * 8);
    const_ntt(p, p->mod_inv_gamma, inv);
    for (size_t i = 0; i < p->gamma; i++) {
        memset(r, 0, n * 8); /* NTT(0) = 0 */
        for (size_t j = 0; j < count; j++) {
            size_t exp_index = (n - i) % n;
            size_t index = (j * p->gen_pows[exp_index]) % two_n;
            const uint64_t *mono = index < n ? p->mono_ntt + index * n : p->neg_mono_ntt + (index - n) * n;
            mul_acc(p, r, mono, inputs_ntt + j * n);
        }
        for (size_t k = 0; k < n; k++) r[k] = orc_mont_mul(&p->ntt, r[k], inv[k]);
        const uint32_t *tab = p->tables + i * n;
        for (size_t k = 0; k < n; k++) out[i * n + k] = r[tab[k]];
    }
    free(inv);
    free(r);
}

void orc_store_build(const orc_pack *p, const uint64_t *columns, uint64_t *sums_ntt) {
    size_t n = p->n;
    uint64_t *cn = (uint64_t *)malloc(p->gamma * n * 8);
    memcpy(cn, columns, p->gamma * n * 8);
    for (size_t j = 0; j < p->gamma; j++) orc_ntt_forward(&p->ntt, cn + j * n);
    orc_rotation_sums(p, cn, p->gamma, sums_ntt);
    free(cn);
}

void orc_fold_b(const orc_pack *p, const uint64_t *columns, const uint64_t *b, uint64_t *b_poly) {
    size_t n = p->n;
    uint64_t q = p->q;
    uint64_t *brev = (uint64_t *)malloc(n * 8);
    for (size_t k = 0; k < n; k++) brev[k] = k == 0 ? b[0] % q : (q - b[n - k] % q) % q;
    memset(b_poly, 0, n * 8);
    for (size_t j = 0; j < p->gamma; j++) {
        u128 low = 0, high = 0;
        for (size_t k = 0; k < n; k++) {
            u128 prod = (u128)columns[j * n + k] * brev[k];
            low += prod & (u128)UINT64_MAX;
            high += prod >> 64;
        }
        u128 qq = q;
        u128 r64 = (((u128)1) << 64) % qq;
        b_poly[j] = (uint64_t)(((high % qq) * r64 + low % qq) % qq);
    }
    free(brev);
}

static void decompose(const orc_pack *p, const uint64_t *c, uint64_t *digits /* rows x n */) {
    size_t n = p->n, rows = key_rows(p);
    uint64_t mask = p->gadget_base - 1;
    if (p->dropped == 0) {
        for (size_t j = 0; j < n; j++) {
            uint64_t v = c[j];
            for (size_t k = 0; k < p->gadget_len; k++) { digits[k * n + j] = v % p->gadget_base; v /= p->gadget_base; }
        }
    } else {
        unsigned dropped_bits = 20 * (unsigned)p->dropped;
        u128 half = (u128)1 << (dropped_bits - 1);
        for (size_t j = 0; j < n; j++) {
            uint64_t r = (uint64_t)(((u128)c[j] + half) >> dropped_bits);
            for (size_t k = 0; k + 1 < rows; k++) { digits[k * n + j] = r & mask; r >>= 20; }
            digits[(rows - 1) * n + j] = r;
        }
    }
}

/* Test hook: the digit split above over n coefficients (key_rows x n digits out). */
void orc_decompose_digits(const orc_pack *p, const uint64_t *c, uint64_t *digits) { decompose(p, c, digits); }

/* fold_rotations_backward fused with packing_online_gather: r_all (gamma x n, NTT) is
 * consumed; a_hat and sum_b (coefficient form) are produced. */
static void recursion_and_online(const orc_pack *p, const uint64_t *w_mask_ntt, const uint64_t *y_ntt,
                                 uint64_t *r_all, uint64_t *a_hat, uint64_t *sum_b) {
    size_t n = p->n, rows = key_rows(p);
    uint64_t q = p->q;
    int last_digit = p->gamma > p->gadget_len + 1 && p->dropped == 0;
    uint64_t *of_sum = NULL, *of_digit = NULL;
    if (last_digit) {
        /* powers of the base mod q, base^len >= q holds for (2^20, 3) */
        uint64_t pw[8];
        uint64_t x = 1;
        for (size_t k = 0; k < p->gadget_len; k++) { pw[k] = x; x = (uint64_t)((u128)x * (p->gadget_base % q) % q); }
        uint64_t inverse = powmod(pw[p->gadget_len - 1], q - 2, q);
        of_sum = (uint64_t *)malloc(n * 8);
        const_ntt(p, inverse, of_sum);
        of_digit = (uint64_t *)malloc((p->gadget_len - 1) * n * 8);
        for (size_t k = 0; k + 1 < p->gadget_len; k++)
            const_ntt(p, (q - (uint64_t)((u128)inverse * pw[k] % q)) % q, of_digit + k * n);
    }
    uint64_t *c = (uint64_t *)malloc(n * 8), *dg = (uint64_t *)malloc(rows * n * 8);
    uint64_t *acc = (uint64_t *)calloc(n, 8);
    for (size_t ii = p->gamma - 1; ii-- > 0;) {
        size_t i = ii;
        memcpy(c, r_all + (i + 1) * n, n * 8);
        orc_ntt_inverse(&p->ntt, c);
        decompose(p, c, dg);
        size_t transformed = last_digit ? rows - 1 : rows;
        for (size_t k = 0; k < transformed; k++) orc_ntt_forward(&p->ntt, dg + k * n);
        if (last_digit) {
            uint64_t *last = dg + (rows - 1) * n;
            memset(last, 0, n * 8);
            mul_acc(p, last, r_all + (i + 1) * n, of_sum);
            for (size_t k = 0; k + 1 < rows; k++) mul_acc(p, last, dg + k * n, of_digit + k * n);
        }
        const uint32_t *tab = p->tables + i * n;
        for (size_t k = 0; k < rows; k++) {
            const uint64_t *w = w_mask_ntt + k * n, *y = y_ntt + k * n, *t = dg + k * n;
            uint64_t *r = r_all + i * n;
            if (p->simd) {
                /* R_i += T_k w_all[i][k] through mul_acc_ntt_domain (inspiring2.rs:1061); the
                 * online gather stays a scalar loop (inspiring2.rs:1465-1478) */
                orc_ifma_mont_mul_acc_gather(r, t, w, tab, n, q, p->ntt.q_inv_neg);
                for (size_t j = 0; j < n; j++) acc[j] = addmod(acc[j], orc_mont_mul(&p->ntt, y[tab[j]], t[j]), q);
                continue;
            }
            for (size_t j = 0; j < n; j++) {
                /* w_all_ntt[i][k][j] = w_mask_ntt[k][tab_i[j]] (apply_automorphism_ntt) */
                r[j] = addmod(r[j], orc_mont_mul(&p->ntt, t[j], w[tab[j]]), q);
                acc[j] = addmod(acc[j], orc_mont_mul(&p->ntt, y[tab[j]], t[j]), q);
            }
        }
    }
    memcpy(a_hat, r_all, n * 8);
    orc_ntt_inverse(&p->ntt, a_hat);
    memcpy(sum_b, acc, n * 8);
    orc_ntt_inverse(&p->ntt, sum_b);
    free(c); free(dg); free(acc); free(of_sum); free(of_digit);
}

int orc_respond(const orc_pack *p, const uint8_t w_seed[32], const uint64_t *a, const uint64_t *b,
                const uint64_t *columns, const uint64_t *y_body, const uint64_t *sums_ntt,
                uint64_t *a_hat, uint64_t *b_out) {
    size_t n = p->n, g = p->gamma, rows = key_rows(p);
    uint64_t q = p->q;
    uint64_t *w_mask = (uint64_t *)malloc(rows * n * 8), *y_ntt = (uint64_t *)malloc(rows * n * 8);
    orc_expand_w_mask(p, w_seed, w_mask);
    for (size_t k = 0; k < rows; k++) orc_ntt_forward(&p->ntt, w_mask + k * n);
    memcpy(y_ntt, y_body, rows * n * 8);
    for (size_t k = 0; k < rows; k++) orc_ntt_forward(&p->ntt, y_ntt + k * n);

    uint64_t *r_all = (uint64_t *)malloc(g * n * 8), *b_poly = (uint64_t *)calloc(n, 8);
    if (sums_ntt) {
        /* store path: b fold, then R_i[j] = a_ntt[tab_i[j]] * S_i[j] */
        orc_fold_b(p, columns, b, b_poly);
        uint64_t *a_ntt = (uint64_t *)malloc(n * 8);
        memcpy(a_ntt, a, n * 8);
        orc_ntt_forward(&p->ntt, a_ntt);
        for (size_t i = 0; i < g; i++) {
            const uint32_t *tab = p->tables + i * n;
            for (size_t j = 0; j < n; j++) r_all[i * n + j] = orc_mont_mul(&p->ntt, a_ntt[tab[j]], sums_ntt[i * n + j]);
        }
        free(a_ntt);
    } else {
        /* no store: per-column plaintext products, coefficient 0 of b*h_j, packing_offline */
        uint64_t *an = (uint64_t *)malloc(n * 8), *bn = (uint64_t *)malloc(n * 8), *h = (uint64_t *)malloc(n * 8);
        uint64_t *acols = (uint64_t *)malloc(g * n * 8), *t = (uint64_t *)malloc(n * 8);
        memcpy(an, a, n * 8); orc_ntt_forward(&p->ntt, an);
        memcpy(bn, b, n * 8); orc_ntt_forward(&p->ntt, bn);
        for (size_t j = 0; j < g; j++) {
            memcpy(h, columns + j * n, n * 8);
            orc_ntt_forward(&p->ntt, h);
            for (size_t k = 0; k < n; k++) t[k] = orc_mont_mul(&p->ntt, an[k], h[k]);
            orc_ntt_inverse(&p->ntt, t);                        /* a*h_j, coefficient form */
            memcpy(acols + j * n, t, n * 8);
            orc_ntt_forward(&p->ntt, acols + j * n);             /* packing_offline re-transforms it */
            for (size_t k = 0; k < n; k++) t[k] = orc_mont_mul(&p->ntt, bn[k], h[k]);
            orc_ntt_inverse(&p->ntt, t);
            if (j < n) b_poly[j] = t[0];
        }
        orc_rotation_sums(p, acols, g, r_all);
        free(an); free(bn); free(h); free(acols); free(t);
    }
    uint64_t *sum_b = (uint64_t *)malloc(n * 8);
    recursion_and_online(p, w_mask, y_ntt, r_all, a_hat, sum_b);
    for (size_t j = 0; j < n; j++) b_out[j] = addmod(b_poly[j], sum_b[j], q);
    free(w_mask); free(y_ntt); free(r_all); free(b_poly); free(sum_b);
    return 0;
}

void orc_mod_switch(const uint64_t *in, size_t count, uint64_t q_old, uint64_t q_new, uint64_t *out) {
    u128 half = (u128)q_old / 2;
    for (size_t i = 0; i < count; i++) {
        u128 scaled = (u128)in[i] * q_new + half;
        out[i] = (uint64_t)((scaled / q_old) % q_new);
    }
}

static size_t put_u64(uint8_t *o, size_t at, uint64_t v) {
    if (o) for (int k = 0; k < 8; k++) o[at + k] = (uint8_t)(v >> (8 * k));
    return at + 8;
}
static size_t put_u32(uint8_t *o, size_t at, uint32_t v) {
    if (o) for (int k = 0; k < 4; k++) o[at + k] = (uint8_t)(v >> (8 * k));
    return at + 4;
}
static size_t tight(uint8_t *o, size_t at, const uint64_t *c, size_t count, unsigned bits) {
    size_t nbytes = (count * bits + 7) / 8;
    at = put_u64(o, at, nbytes);
    if (o) {
        memset(o + at, 0, nbytes);
        for (size_t i = 0; i < count; i++) {
            size_t bit = i * bits;
            for (unsigned b = 0; b < bits; b++, bit++)
                if ((c[i] >> b) & 1) o[at + bit / 8] |= (uint8_t)(1u << (bit % 8));
        }
    }
    return at + nbytes;
}

size_t orc_encode_response(size_t n, uint64_t modulus, const uint64_t *a_hat, const uint64_t *b,
                           size_t retained, uint8_t *out) {
    unsigned bits = 64 - (unsigned)__builtin_clzll(modulus - 1);
    size_t at = 0;
    at = put_u32(out, at, 1);                    /* TightResponseCiphertextWireRef::Packed */
    at = tight(out, at, a_hat, n, bits);         /* a: TightPolyWire.coeffs */
    at = put_u64(out, at, 1);                    /* moduli: vec len 1 */
    at = put_u64(out, at, modulus);
    at = put_u64(out, at, modulus);              /* q */
    at = put_u64(out, at, n);                    /* dim */
    at = put_u64(out, at, 0);                    /* crt_q0_inv_mod_q1 */
    if (out) out[at] = 0;                        /* is_ntt = false */
    at += 1;
    at = tight(out, at, b, retained, bits);      /* b_prefix */
    at = put_u32(out, at, (uint32_t)retained);   /* retained */
    at = put_u64(out, at, 0);                    /* column_ciphertexts: empty vec */
    if (out) out[at] = 1;                        /* Some( */
    at += 1;
    at = put_u32(out, at, 0);                    /* PackingMode::Inspiring ) */
    return at;
}
