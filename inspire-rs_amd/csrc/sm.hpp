// Device-visible objects of the special-modulus set (src/special_modulus): a 29-bit
// ciphertext prime q and a 31-bit special prime p on 32-bit words, one step-1 table
// per shard, one seated key per client. Shared by sm.hip (kernels) and sm_capi.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace inspire {

// A constant w < m and its Shoup companion floor(w * 2^32 / m).
struct Pair32 {
    uint32_t w, wp;
};

// Largest ciphertext prime whose lazy residues fit a word: the recursion's widest lazy sum mod
// q, t2 + 4q - v, is below 6q. Larger primes (the reference accepts any below 2^31) run the
// kernels with canonical residues mod q.
constexpr uint64_t kSmLazyQMax = 0xffffffffULL / 6;

struct SmCtxDev {
    int logn, n, gamma, nb, cell_bytes;
    uint32_t q, p;
    uint32_t half_q;      // q / 2: the q -> p lift of c > q/2 is c + (p - q)
    uint32_t half_p;      // p / 2: the p -> q lift of r > p/2 is r - p
    uint32_t p_minus_q;
    uint32_t q_barrett;   // floor(2^32 / q)
    uint32_t neg_off;     // (q - p mod q) mod q
    uint32_t g;           // 2n / gamma + 1, generator of the rotation subgroup
    int wide_q;           // q > kSmLazyQMax: the kernels keep residues mod q canonical
    uint64_t q_recip, p_recip;  // floor(2^64 / m), reducing uniform u64 draws
    uint64_t zone_q, zone_p;    // (2^64 - 1) / m * m: draws at or above are rejected
    uint32_t qninv, pninv;  // -m^-1 mod 2^32 (Montgomery products, R = 2^32)
    Pair32 pinv_n;        // (n p)^-1 mod q: the recursion's division by p, with n^-1 folded in
    Pair32 pinv;          // p^-1 mod q: the body's final division by p
    Pair32 table_scale;   // (n gamma)^-1 R mod q: the step-1 table as Montgomery operands
    Pair32 seat_wq;       // n R mod q: the key mask's q limb
    Pair32 seat_kq;       // R mod q: the key's q limb
    Pair32 seat_p;        // n^-1 R mod p: the key mask's and the key's p limbs
    // psi^(+-bitrev(k)) with companions, unscaled (no n^-1 in the inverse)
    const Pair32 *twf_q, *twi_q, *twf_p, *twi_p;
};

// One query's resident inputs. Workgroup b answers jobs[b], whose query sits at position qi of
// the batch (seeds, bodies, outputs); the host lays the jobs out by sm_order_jobs.
struct SmJob {
    const uint32_t *step1;  // gamma x n: (n gamma)^-1 R NTT_q(P_i), canonical
    const uint8_t *cells;   // n x cell_bytes, rows past the shard's cells zero
    const uint32_t *seat;   // 4 x n Montgomery operands: n w_q, key_q (mod q), w_p / n, key_p / n (mod p)
    int qi;
};

struct SmBatchDev {
    int count;
    const SmJob *jobs;
    const uint8_t *seeds;    // count x 32
    const uint32_t *bodies;  // count x n
    uint32_t *mask_out;      // count x n (also where each query parks NTT(a) during the chain)
    uint32_t *body_out;      // count x nb, unused when wire_out is set
    // when non-null, each query's Response::to_bytes at wire_out + qi * wire_stride (stride a
    // multiple of 4 holding the length rounded up to 4); mask_out then is scratch only
    uint8_t *wire_out = nullptr;
    size_t wire_stride = 0;
};

// Response::to_bytes (wire.rs:125-131): mask then body, each packed LSB-first at the bit
// width of q - 1 (pack_coefficients_tight, poly.rs:166-225)
struct SmWire {
    uint32_t w, mask_bytes, body_bytes, total;
};
__host__ __device__ inline SmWire sm_wire(uint32_t n, uint32_t nb, uint32_t q) {
    SmWire L;
    uint32_t w = 0;
    while (w < 32 && ((q - 1) >> w) != 0) w++;
    L.w = w;
    L.mask_bytes = (n * w + 7) / 8;
    L.body_bytes = (nb * w + 7) / 8;
    L.total = L.mask_bytes + L.body_bytes;
    return L;
}

// Step-1 table of one shard from its cells (device), using cols (gamma x n bytes) as scratch.
hipError_t sm_launch_table(const SmCtxDev &c, const uint8_t *cells, int8_t *cols, uint32_t *step1, hipStream_t s);
// Seat of one key: mask seed (32 bytes, device), both key limbs in coefficient form (device).
hipError_t sm_launch_seat(const SmCtxDev &c, const uint8_t *mask_seed, const uint32_t *key_q, const uint32_t *key_p,
                          uint32_t *seat, hipStream_t s);
// One launch answering every query of the batch.
hipError_t sm_launch_respond(const SmCtxDev &c, const SmBatchDev &b, hipStream_t s);

}  // namespace inspire
