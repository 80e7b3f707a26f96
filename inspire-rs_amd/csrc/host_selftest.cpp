// Host build of the device arithmetic headers, exported for the CPU test suite
// (tests/test_device_arith_host.py checks them against big-integer arithmetic).
#include <stddef.h>
#include <stdint.h>
#include "chacha.hpp"
#include "modq.hpp"

using namespace inspire;

extern "C" {
uint64_t st_mulmod(uint64_t a, uint64_t b) { return mulmod(a, b); }
uint64_t st_reduce120(uint64_t hi, uint64_t lo) { return reduce120(hi, lo); }
void st_mul_wide(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) { *lo = mul_wide(a, b, *hi); }
uint64_t st_acc_reduce(const uint64_t *a, const uint64_t *b, int count, uint64_t init) {
    Acc128 acc;
    acc.add(init);
    for (int i = 0; i < count; i++) acc.mac(a[i], b[i]);
    return acc.reduce();
}
uint64_t st_mod_switch(uint64_t c, uint64_t qn) { return mod_switch_coeff(c, qn); }
uint64_t st_mod_p61(uint64_t hi, uint64_t lo) { return mod_p61(hi, lo); }
uint32_t st_automorph_slot(uint32_t t, uint32_t j, int logn) { return automorph_slot(t, j, logn); }
void st_chacha_block(const uint8_t *seed, uint64_t counter, uint64_t stream, uint32_t *out) {
    uint32_t key[8];
    seed_key(seed, key);
    chacha20_block(key, counter, stream, out);
}
uint64_t st_shoup_companion(uint64_t w) { return shoup_companion(w); }
uint64_t st_shoup(uint64_t a, uint64_t w) {
    uint64_t wp = shoup_companion(w), hi;
    (void)mul_wide(a, wp, hi);
    return a * w - hi * kQ;
}
uint64_t st_shoup_lazy(uint64_t a, uint64_t w) { return shoup_lazy(a, w, shoup_companion(w)); }
int st_uniform_q_draw(uint64_t v, uint64_t *out) { return uniform_q_draw(v, *out) ? 1 : 0; }
// the packing chain's digit split, both gadgets
void st_gadget_digits(uint64_t c, int exact, uint64_t *d0, uint64_t *d1) {
    if (exact)
        gadget_digits<true>(c, *d0, *d1);
    else
        gadget_digits<false>(c, *d0, *d1);
}
// the kernels' serial rejection replays, at any acceptance bound
void st_replay_gen_range_q(const uint8_t *seed, int n, uint64_t zone, uint64_t *out) {
    uint32_t key[8];
    seed_key(seed, key);
    replay_gen_range_q(key, n, zone, [&](int k, uint64_t a) { out[k] = a; });
}
void st_replay_expand_uniform(const uint8_t *seed, uint64_t stream, int limbs, int count, const uint32_t *m,
                              const uint64_t *zone, uint32_t *out) {
    uint32_t key[8];
    seed_key(seed, key);
    uint64_t recip[2];
    for (int l = 0; l < limbs && l < 2; l++) recip[l] = ~0ULL / m[l];  // as the server builds it
    replay_expand_uniform(key, stream, limbs, count, m, zone, recip,
                          [&](int l, int k, uint32_t v) { out[(size_t)l * count + k] = v; });
}
uint32_t st_reduce64(uint64_t u, uint32_t m) {
    return reduce64(u, m, ~0ULL / m);
}
}
