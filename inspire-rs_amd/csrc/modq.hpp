// Exact arithmetic mod q = 2^60 - 2^14 + 1 on 32-bit lanes (gfx950).
//
// The reference multiplies in Montgomery form with R = 2^64 and a Solinas REDC
// (src/math/solinas_redc.rs:13-22, src/math/ntt.rs:689-695). Every value that
// leaves the respond path is a canonical residue, so any exact product gives the
// same bytes. This build keeps residues in standard form and reduces the 120-bit
// product directly with the special form of q: 2^60 = 2^14 - 1 (mod q). The wide
// product is four 32x32->64 multiply-adds; the reduction is shifts and adds.
//
// Everything here is __host__ __device__ so the CPU test suite can exercise the
// exact same code against u128 arithmetic.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace inspire {

constexpr uint64_t kQ = 1152921504606830593ULL;  // 2^60 - 2^14 + 1
constexpr uint64_t kTwoQ = 2 * kQ;
constexpr uint64_t kM60 = (1ULL << 60) - 1;
constexpr uint64_t kM46 = (1ULL << 46) - 1;
constexpr uint64_t kTagPrime = (1ULL << 61) - 1;  // shard_store.rs:28

__host__ __device__ __forceinline__ uint64_t mul32w(uint32_t a, uint32_t b) { return (uint64_t)a * b; }

// 128-bit product of a, b < 2^60: returns the low word, hi receives the high word.
__host__ __device__ __forceinline__ uint64_t mul_wide(uint64_t a, uint64_t b, uint64_t &hi) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = mul32w(a0, b0);
    uint64_t p01 = mul32w(a0, b1) + (p00 >> 32);      // < 2^60 + 2^32
    uint64_t p10 = mul32w(a1, b0) + (uint32_t)p01;    // < 2^60 + 2^32
    hi = mul32w(a1, b1) + (p01 >> 32) + (p10 >> 32);
    return (p10 << 32) | (uint32_t)p00;
}

// hi*2^64 + lo (< 2^120) mod q, canonical.
__host__ __device__ __forceinline__ uint64_t reduce120(uint64_t hi, uint64_t lo) {
    uint64_t A = (hi << 4) | (lo >> 60);  // P = A*2^60 + B, A < 2^60
    uint64_t B = lo & kM60;
    uint64_t Ah = hi >> 42;               // A >> 46
    uint64_t Al = A & kM46;
    // A*2^60 = A*2^14 - A = Ah*2^60 + Al*2^14 - A = Ah*(2^14-1) + Al*2^14 - A
    uint64_t U = B + (Al << 14) + (Ah << 14);  // < 2^61 + 2^28
    uint64_t V = A + Ah;                       // < 2^60 + 2^14
    uint64_t r = U + kTwoQ - V;                // (0, 2^62 + 2^28)
    uint64_t C = r >> 60, D = r & kM60;
    uint64_t s = D + (C << 14) - C;            // < 2^60 + 2^16 < 2q
    return s >= kQ ? s - kQ : s;
}

__host__ __device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b) {
    uint64_t hi, lo = mul_wide(a, b, hi);
    return reduce120(hi, lo);
}

__host__ __device__ __forceinline__ uint64_t addmod(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s >= kQ ? s - kQ : s;
}

__host__ __device__ __forceinline__ uint64_t submod(uint64_t a, uint64_t b) {
    return a >= b ? a - b : a + kQ - b;
}

// Sum of up to 16 wide products accumulated as (hi, lo) before one reduction.
struct Acc128 {
    uint64_t lo, hi;
    __host__ __device__ __forceinline__ Acc128() : lo(0), hi(0) {}
    __host__ __device__ __forceinline__ void mac(uint64_t a, uint64_t b) {
        uint64_t h, l = mul_wide(a, b, h);
        uint64_t s = lo + l;
        hi += h + (s < l ? 1 : 0);
        lo = s;
    }
    __host__ __device__ __forceinline__ void add(uint64_t a) {
        uint64_t s = lo + a;
        hi += (s < a ? 1 : 0);
        lo = s;
    }
    // valid while hi < 2^60 (at most ~16 products of residues)
    __host__ __device__ __forceinline__ uint64_t reduce() const {
        // fold hi once so the value fits reduce120's 2^120 bound: 2^64 = 2^18 - 2^4 (mod q)
        if (hi >> 56) {
            uint64_t h2 = hi >> 56, hl = hi & ((1ULL << 56) - 1);
            // h2 * 2^120 = h2 * 2^(120-60*2) * (2^14-1)^2 = h2 * (2^14-1)^2 (mod q)
            uint64_t extra = h2 * ((1ULL << 14) - 1) * ((1ULL << 14) - 1);  // < 2^36
            uint64_t s = lo + extra;
            uint64_t h = hl + (s < extra ? 1 : 0);
            return reduce120(h, s);
        }
        return reduce120(hi, lo);
    }
};

__host__ __device__ __forceinline__ uint64_t mod_p61(uint64_t hi, uint64_t lo) {
    // hi * 2^64 + lo with hi < 2^58; 2^61 = 1 (mod 2^61 - 1)
    uint64_t v = (lo & kTagPrime) + (lo >> 61) + (hi << 3);
    v = (v & kTagPrime) + (v >> 61);
    return v >= kTagPrime ? v - kTagPrime : v;
}

// Switch moduli this path implements: every q' below 2^46, which covers the reference's
// 36- and 45-bit targets (MOD_SWITCH_TARGET_36BIT/45BIT, src/pir/mod_switch.rs:33,47).
constexpr uint64_t kSwitchModulusLimit = 1ULL << 46;

// floor((c * q' + floor(q/2)) / q) mod q' for c < q, q' < 2^46. With X = c q' + floor(q/2) < 2^106,
// e = floor(X / 2^60) is the quotient or one below it: X/q - X/2^60 = X (2^14 - 1) / (q 2^60)
// < (q' + 1/2)(2^14 - 1) / 2^60 < 1. The quotient is at most q', so one subtraction reduces it.
__host__ __device__ __forceinline__ uint64_t mod_switch_coeff(uint64_t c, uint64_t qn) {
    uint64_t hi, lo = mul_wide(c, qn, hi);
    const uint64_t half = kQ / 2;
    uint64_t l2 = lo + half;
    hi += (l2 < half);
    lo = l2;
    uint64_t e = (hi << 4) | (lo >> 60);  // floor(X / 2^60): the quotient or one below it
    uint64_t ph, pl = mul_wide(e + 1, kQ, ph);
    bool ge = (hi > ph) || (hi == ph && lo >= pl);
    uint64_t quot = ge ? e + 1 : e;
    return quot >= qn ? quot - qn : quot;
}

// Shoup companion floor(w * 2^64 / q) of a residue w < q. Estimate from
// 2^64/q = 16 (1 + d/2^60 + ...), d = 2^14 - 1, then correct against the exact remainder.
__host__ __device__ __forceinline__ uint64_t shoup_companion(uint64_t w) {
    uint64_t eh, el = mul_wide(w, (1ULL << 14) - 1, eh);
    uint64_t e = (w << 4) + ((eh << 8) | (el >> 56));
    for (int it = 0; it < 4; it++) {
        // r = w*2^64 - e*q as a signed 128-bit value (r_hi, r_lo)
        uint64_t ph, pl = mul_wide(e, kQ, ph);
        uint64_t r_lo = 0 - pl;
        uint64_t r_hi = w - ph - (pl != 0 ? 1 : 0);
        if ((int64_t)r_hi < 0) { e--; continue; }
        if (r_hi == 0 && r_lo < kQ) break;
        e++;
    }
    return e;
}

// Shoup product a*w mod q with the companion w' = floor(w * 2^64 / q), for any a < 2^64.
// The quotient estimate keeps a1*w'1 and the high halves of the two middle partial
// products, leaving out a0*w'0 and the middle carries: it is at most 2 short of
// floor(a*w'/2^64) (the dropped part is below 3 * 2^32 in units of 2^32), and that
// floor is itself within 1 of floor(a*w/q), so the result is in [0, 4q). Three 32-bit
// multiplies for the quotient instead of four plus their carry moves.
__host__ __device__ __forceinline__ uint64_t shoup_lazy(uint64_t a, uint64_t w, uint64_t wp) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t p0 = (uint32_t)wp, p1 = (uint32_t)(wp >> 32);
    const uint64_t est = mul32w(a1, p1) + (mul32w(a1, p0) >> 32) + (mul32w(a0, p1) >> 32);
    return a * w - est * kQ;
}

// Base-2^20 digits of a canonical coefficient c, two rows.
// Exact gadget (gadget_decompose, rgsw/external_product.rs:65-99): the rows of digits 0 and 1,
// d0 = c mod 2^20, d1 = floor(c / 2^20) mod 2^20; the third digit is not split out (the
// packing chain derives its transform from the others).
// Approximate gadget (gadget_decompose_rounded / rounded_digit_rows, inspiring2.rs:1167-1225,
// one dropped digit): r = floor((c + 2^19) / 2^20), the round-half-up quotient, then
// d0 = r mod 2^20 and d1 = floor(r / 2^20), the top row keeping the rounding carry (it
// equals 2^20 for c >= 2^60 - 2^19). So c = 2^20 (d0 + 2^20 d1) + e with e in [-2^19, 2^19).
template <bool EXACT>
__host__ __device__ __forceinline__ void gadget_digits(uint64_t c, uint64_t &d0, uint64_t &d1) {
    if (EXACT) {
        d0 = c & 0xfffff;
        d1 = (c >> 20) & 0xfffff;
    } else {
        const uint64_t r = (c + (1ULL << 19)) >> 20;
        d0 = r & 0xfffff;
        d1 = r >> 20;
    }
}

__host__ __device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_bitreverse32(x) >> (32 - bits);
#else
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
#endif
}

// Slot permutation of tau_t in NTT form: NTT(tau_t f)[j] = NTT(f)[slot].
// Slot j of the bit-reversed negacyclic NTT evaluates at psi^(2*brv(j)+1)
// (the closed form of derive_automorph_tables, inspiring2.rs:330-379).
__host__ __device__ __forceinline__ uint32_t automorph_slot(uint32_t t, uint32_t j, int logn) {
    uint32_t two_n_mask = (2u << logn) - 1;
    uint32_t e = 2 * bitrev(j, logn) + 1;
    return bitrev(((t * e) & two_n_mask) >> 1, logn);
}

}  // namespace inspire
