/*
 * Oracle (test infrastructure only): the special-modulus respond path (src/special_modulus)
 * as static inline functions. special.c includes them for the portable build, and
 * special_avx512.c includes the same text under -mavx512f -mavx512bw -mavx512dq -mavx512vl:
 * the reference's Kernel::Avx512 is likewise respond_impl itself compiled under
 * #[target_feature(enable = "avx512f,avx512bw,avx512dq,avx512vl")] (server.rs:300-330),
 * with no hand-written vector code.
 *
 *   smi_csub/add/sub/mont  Prime32, arith.rs:13-102 (Montgomery R = 2^32, values in [0, q))
 *   smi_ntt_forward        Ntt32::forward, arith.rs:188-205 (CT over psi^bitrev(k))
 *   smi_ntt_inverse        Ntt32::inverse, arith.rs:247-272 (GS, last stage carrying n^-1)
 *   smi_*_centred          the centred lifts, mod.rs:333-356
 *   smi_mod_down           mod.rs:358-363
 *   smi_respond            respond_impl, server.rs:385-499
 */
#ifndef INSPIRE_ORACLE_SM_ARITH_H
#define INSPIRE_ORACLE_SM_ARITH_H

#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static inline uint32_t smi_csub(const sm_prime *p, uint32_t x) { return x >= p->q ? x - p->q : x; }
static inline uint32_t smi_add(const sm_prime *p, uint32_t a, uint32_t b) { return smi_csub(p, a + b); }
static inline uint32_t smi_sub(const sm_prime *p, uint32_t a, uint32_t b) { return smi_csub(p, a + p->q - b); }

static inline uint32_t smi_mont(const sm_prime *p, uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * p->neg_qinv;
    return smi_csub(p, (uint32_t)((t + (uint64_t)m * p->q) >> 32));
}

static inline int64_t smi_centred(const sm_prime *p, uint32_t x) {
    return x > p->q / 2 ? (int64_t)x - p->q : (int64_t)x;
}

static inline void smi_ntt_forward(const sm_ntt *c, uint32_t *a) {
    const sm_prime *p = &c->p;
    size_t n = c->n, t = n;
    for (size_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (size_t blk = 0; blk < m; blk++) {
            uint32_t s = c->fwd[m + blk];
            uint32_t *x = a + 2 * blk * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                uint32_t u = x[j], v = smi_mont(p, y[j], s);
                x[j] = smi_add(p, u, v);
                y[j] = smi_sub(p, u, v);
            }
        }
    }
}

static inline void smi_ntt_inverse(const sm_ntt *c, uint32_t *a) {
    const sm_prime *p = &c->p;
    size_t n = c->n, t = 1;
    for (size_t m = n; m > 2; m >>= 1) {
        size_t h = m >> 1;
        for (size_t blk = 0; blk < h; blk++) {
            uint32_t s = c->inv[h + blk];
            uint32_t *x = a + 2 * blk * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                uint32_t u = x[j], v = y[j];
                x[j] = smi_add(p, u, v);
                y[j] = smi_mont(p, u + p->q - v, s);
            }
        }
        t <<= 1;
    }
    for (size_t j = 0; j < t; j++) {
        uint32_t u = a[j], v = a[j + t];
        a[j] = smi_mont(p, u + v, c->n_inv);
        a[j + t] = smi_mont(p, u + p->q - v, c->last);
    }
}

static inline uint32_t smi_p_to_q_centred(const sm_ring *r, uint32_t x) {
    const sm_prime *pq = &r->lq.p;
    uint32_t base = x;
    for (uint32_t k = 0; k < r->p_over_q; k++) base = smi_csub(pq, base);
    return x > r->lp.p.q / 2 ? smi_add(pq, base, r->neg_off) : base;
}

static inline uint32_t smi_q_to_p_centred(const sm_ring *r, uint32_t c) {
    return c > r->lq.p.q / 2 ? c + (r->lp.p.q - r->lq.p.q) : c;
}

static inline uint32_t smi_mod_down(const sm_ring *r, uint32_t xq, uint32_t xp) {
    const sm_prime *pq = &r->lq.p;
    return smi_mont(pq, smi_sub(pq, xq, smi_p_to_q_centred(r, xp)), r->p_inv);
}

/* respond_impl (server.rs:385-499). nb body coefficients are returned. */
static inline int smi_respond(const sm_ring *r, const uint32_t *step1, const int8_t *values, size_t rows,
                              const uint32_t *sq, const uint32_t *sp, const uint32_t *w_q, const uint32_t *w_p,
                              const uint8_t seed[32], const uint32_t *qbody, uint32_t *mask, uint32_t *body,
                              size_t nb) {
    size_t d = r->d, gamma = r->gamma;
    const sm_ntt *lq = &r->lq, *lp = &r->lp;
    const sm_prime *pq = &lq->p, *pp = &lp->p;
    for (size_t k = 0; k < d; k++)
        if (qbody[k] >= pq->q) return -1;

    int64_t *fold = (int64_t *)calloc(nb, 8);
    for (size_t m = 0; m < rows; m++) {
        int64_t c = m == 0 ? smi_centred(pq, qbody[0]) : -smi_centred(pq, qbody[d - m]);
        const int8_t *row = values + m * gamma;
        for (size_t k = 0; k < nb; k++) fold[k] += c * row[k];
    }

    uint32_t *a = (uint32_t *)malloc(d * 4), *cur = (uint32_t *)malloc(d * 4),
             *next = (uint32_t *)malloc(d * 4), *tp = (uint32_t *)malloc(d * 4),
             *cp = (uint32_t *)malloc(d * 4), *rq = (uint32_t *)malloc(d * 4),
             *acc_q = (uint32_t *)calloc(d, 4), *acc_p = (uint32_t *)calloc(d, 4);
    sm_expand_query_mask(r, seed, a);
    smi_ntt_forward(lq, a);

    {
        const uint16_t *tab = r->tables + (gamma - 1) * d;
        const uint32_t *s = step1 + (gamma - 1) * d;
        for (size_t j = 0; j < d; j++) cur[j] = smi_mont(pq, a[tab[j]], s[j]);
    }
    for (size_t ii = gamma - 1; ii-- > 0;) {
        const uint16_t *tab = r->tables + ii * d;
        memcpy(tp, cur, d * 4);
        smi_ntt_inverse(lq, tp);
        for (size_t j = 0; j < d; j++) tp[j] = smi_q_to_p_centred(r, tp[j]);
        smi_ntt_forward(lp, tp);
        for (size_t j = 0; j < d; j++) cp[j] = smi_mont(pp, tp[j], w_p[tab[j]]);
        smi_ntt_inverse(lp, cp);
        for (size_t j = 0; j < d; j++) rq[j] = smi_p_to_q_centred(r, cp[j]);
        smi_ntt_forward(lq, rq);
        const uint32_t *s = step1 + ii * d;
        for (size_t j = 0; j < d; j++) {
            size_t src = tab[j];
            uint32_t cq = smi_mont(pq, cur[j], w_q[src]);
            next[j] = smi_add(pq, smi_mont(pq, a[src], s[j]), smi_mont(pq, smi_sub(pq, cq, rq[j]), r->p_inv));
            acc_q[j] = smi_add(pq, acc_q[j], smi_mont(pq, cur[j], sq[src]));
            acc_p[j] = smi_add(pp, acc_p[j], smi_mont(pp, tp[j], sp[src]));
        }
        uint32_t *t = cur; cur = next; next = t;
    }
    memcpy(mask, cur, d * 4);
    smi_ntt_inverse(lq, mask);
    smi_ntt_inverse(lq, acc_q);
    smi_ntt_inverse(lp, acc_p);
    for (size_t k = 0; k < nb; k++) {
        int64_t f = fold[k] % (int64_t)pq->q;
        if (f < 0) f += pq->q;
        body[k] = smi_add(pq, smi_mod_down(r, acc_q[k], acc_p[k]), (uint32_t)f);
    }
    free(fold); free(a); free(cur); free(next); free(tp); free(cp); free(rq); free(acc_q); free(acc_p);
    return 0;
}

#endif
