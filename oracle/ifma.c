/*
 * Oracle (test infrastructure only): the reference's optional AVX-512 IFMA52 path for the
 * packing recursion's multiply-accumulates, used for the CPU baseline's "with the SIMD
 * feature" leg. Under the simd-packing-offline feature Poly::mul_acc_ntt_domain
 * (src/math/poly.rs:999-1040) runs solinas_redc::avx512_ifma::solinas_mont_mul_acc_x8
 * (src/math/solinas_redc.rs:216-246) instead of its scalar loop; the NTTs and the online
 * gather stay scalar in the reference either way.
 *
 * Restated from the split-52 method the reference documents (src/math/ifma52.rs:1-26): a
 * 60-bit operand is a1 2^52 + a0 with a1 < 2^8, the partial products come from
 * vpmadd52luq / vpmadd52huq, and the Montgomery reduction at q = 2^60 - 2^14 + 1 forms
 * m q = m 2^60 - m (2^14 - 1) with shifts (solinas_redc.rs:11-22). Equal to the scalar
 * Montgomery product for every input below q (tests/test_oracle_ifma.py).
 *
 * The functions carry their own target attribute, so the rest of the oracle is compiled for
 * the baseline ISA and the host picks this path at run time only when it has the features.
 */
#include <immintrin.h>
#include <stddef.h>
#include <stdint.h>

#include "oracle.h"

#define IFMA_TARGET __attribute__((target("avx512f,avx512dq,avx512ifma")))

int orc_ifma_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") &&
           __builtin_cpu_supports("avx512ifma");
}

/* a b 2^-64 mod q per lane, canonical, for a, b < q = 2^60 - 2^14 + 1 */
static inline IFMA_TARGET __m512i mont_x8(__m512i a, __m512i b, __m512i qinv, __m512i q) {
    const __m512i m52 = _mm512_set1_epi64((long long)((1ULL << 52) - 1));
    const __m512i zero = _mm512_setzero_si512();
    const __m512i a0 = _mm512_and_si512(a, m52), a1 = _mm512_srli_epi64(a, 52);
    const __m512i b0 = _mm512_and_si512(b, m52), b1 = _mm512_srli_epi64(b, 52);
    /* a b = L0 + S1 2^52 + S2 2^104 with L0 = lo52(a0 b0) < 2^52,
     * S1 = hi52(a0 b0) + lo52(a1 b0) + lo52(a0 b1) < 3 2^52 and
     * S2 = hi52(a1 b0) + hi52(a0 b1) + a1 b1 < 2^17 */
    const __m512i L0 = _mm512_madd52lo_epu64(zero, a0, b0);
    __m512i S1 = _mm512_madd52hi_epu64(zero, a0, b0);
    S1 = _mm512_madd52lo_epu64(S1, a1, b0);
    S1 = _mm512_madd52lo_epu64(S1, a0, b1);
    __m512i S2 = _mm512_madd52hi_epu64(zero, a1, b0);
    S2 = _mm512_madd52hi_epu64(S2, a0, b1);
    S2 = _mm512_madd52lo_epu64(S2, a1, b1);
    /* as (hi, lo) words: S1's low 12 bits fill bits 52..63 of lo, nothing carries */
    const __m512i lo = _mm512_or_si512(L0, _mm512_slli_epi64(S1, 52));
    const __m512i hi = _mm512_add_epi64(_mm512_srli_epi64(S1, 12), _mm512_slli_epi64(S2, 40));
    /* m = lo (-q^-1) mod 2^64; the high word of m q = m 2^60 - P, P = m (2^14 - 1) */
    const __m512i m = _mm512_mullo_epi64(lo, qinv);
    const __m512i m14 = _mm512_slli_epi64(m, 14);
    const __m512i p_lo = _mm512_sub_epi64(m14, m);
    const __mmask8 p_borrow = _mm512_cmplt_epu64_mask(m14, m);
    const __m512i p_hi = _mm512_mask_sub_epi64(_mm512_srli_epi64(m, 50), p_borrow, _mm512_srli_epi64(m, 50),
                                               _mm512_set1_epi64(1));
    const __m512i m60 = _mm512_slli_epi64(m, 60);
    const __mmask8 x_borrow = _mm512_cmplt_epu64_mask(m60, p_lo);
    __m512i x_hi = _mm512_sub_epi64(_mm512_srli_epi64(m, 4), p_hi);
    x_hi = _mm512_mask_sub_epi64(x_hi, x_borrow, x_hi, _mm512_set1_epi64(1));
    /* (a b + m q) / 2^64: the low words sum to 0 mod 2^64, carrying one exactly when lo != 0 */
    __m512i t = _mm512_add_epi64(hi, x_hi);
    const __mmask8 nz = _mm512_test_epi64_mask(lo, lo);
    t = _mm512_mask_add_epi64(t, nz, t, _mm512_set1_epi64(1));
    /* t < 2q */
    return _mm512_mask_sub_epi64(t, _mm512_cmpge_epu64_mask(t, q), t, q);
}

static inline IFMA_TARGET __m512i addq_x8(__m512i acc, __m512i x, __m512i q) {
    const __m512i s = _mm512_add_epi64(acc, x);
    return _mm512_mask_sub_epi64(s, _mm512_cmpge_epu64_mask(s, q), s, q);
}

/* acc[j] = acc[j] + a[j] b[j] 2^-64 mod q, n a multiple of 8, all values below q */
IFMA_TARGET void orc_ifma_mont_mul_acc(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t n,
                                       uint64_t q, uint64_t q_inv_neg) {
    const __m512i qv = _mm512_set1_epi64((long long)q), qinv = _mm512_set1_epi64((long long)q_inv_neg);
    for (size_t j = 0; j < n; j += 8) {
        const __m512i p = mont_x8(_mm512_loadu_si512(a + j), _mm512_loadu_si512(b + j), qinv, qv);
        _mm512_storeu_si512(acc + j, addq_x8(_mm512_loadu_si512(acc + j), p, qv));
    }
}

/* acc[j] = acc[j] + a[j] w[tab[j]] 2^-64 mod q: the rotated operand read through the
 * automorphism table, which the reference materialises as w_all_ntt[i][k] */
IFMA_TARGET void orc_ifma_mont_mul_acc_gather(uint64_t *acc, const uint64_t *a, const uint64_t *w,
                                              const uint32_t *tab, size_t n, uint64_t q, uint64_t q_inv_neg) {
    const __m512i qv = _mm512_set1_epi64((long long)q), qinv = _mm512_set1_epi64((long long)q_inv_neg);
    for (size_t j = 0; j < n; j += 8) {
        const __m256i ix = _mm256_loadu_si256((const __m256i *)(tab + j));
        const __m512i wv = _mm512_i32gather_epi64(ix, (const void *)w, 8);
        const __m512i p = mont_x8(_mm512_loadu_si512(a + j), wv, qinv, qv);
        _mm512_storeu_si512(acc + j, addq_x8(_mm512_loadu_si512(acc + j), p, qv));
    }
}
