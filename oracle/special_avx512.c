/*
 * Oracle (test infrastructure only): the special-modulus respond compiled for
 * AVX-512F/BW/DQ/VL, the reference's Kernel::Avx512 (respond_avx512, server.rs:319-330:
 * respond_impl under #[target_feature], vectorised by the compiler). The Makefile builds
 * this file alone with -mavx512f -mavx512bw -mavx512dq -mavx512vl; callers check
 * sm_avx512_available() first.
 */
#include "sm_arith.h"

int sm_respond_avx512(const sm_ring *r, const uint32_t *step1, const int8_t *values, size_t rows,
                      const uint32_t *sq, const uint32_t *sp, const uint32_t *w_q, const uint32_t *w_p,
                      const uint8_t seed[32], const uint32_t *qbody, uint32_t *mask, uint32_t *body, size_t nb) {
    return smi_respond(r, step1, values, rows, sq, sp, w_q, w_p, seed, qbody, mask, body, nb);
}
