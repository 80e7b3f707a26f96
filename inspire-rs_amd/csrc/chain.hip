// The per-query packing chain on gfx950, one workgroup of n/4 threads per query.
//
// Restates, fused into one kernel:
//   CheckedColumns::packing_offline  (src/inspiring/shard_store.rs:282-342)
//       step 1: R_i = NTT(a)[tab_i] * S_i
//   fold_rotations_backward          (src/inspiring/inspiring2.rs:1014-1085)
//       for i = gamma-2 .. 0: T = G^-1(iNTT(R_{i+1})), R_i += sum_k NTT(T_k) * w_all[i][k]
//       (the last digit's transform read off the others, LastDigitNtt :1094-1161)
//   packing_online_gather            (src/inspiring/inspiring2.rs:1393-1485)
//       sum += sum_k y_k[tab_i] * NTT(T_k), fused into the same step
//   a_hat = iNTT(R_0), b = b_poly + iNTT(sum), then mod_switch_poly (mod_switch.rs:148-175)
//
// Arithmetic: every product on the chain has one operand fixed ahead of the query
// (twiddles, the store's S_i, the mask rows w, the key rows y, the last-digit
// constants), so all of them are Shoup products a*w - floor(a*w'/2^64)*q with the
// precomputed companion w' = floor(w*2^64/q), with a three-multiply quotient estimate.
// Residues stay lazy, in [0, 8q) through the inverse transforms and the state, in
// [0, 16q) through the forward transforms and [0, 4q) in the online sum, and are made
// canonical only where the reference's bytes depend on it (the digits and the outputs). The inverse transform carries no n^-1: S_i, w and y hold it already.
//
// Layout: thread t owns slots 4t..4t+3. The first inverse pass (bits 0-1) and the
// last forward pass (bits 1-0) work on exactly those slots, so the recursion state
// R and the online sum stay in registers; the passes over higher bits exchange
// through LDS. The pass over the top bit of the inverse transform, the digit split
// and the first pass of both forward transforms share one register round.
#include <hip/hip_runtime.h>

#include "chacha.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "modq.hpp"
#include "ntt.hpp"

namespace inspire {

// Every product here is shoup_lazy (modq.hpp): result in [0, 4q) for any a < 2^64.
__device__ __forceinline__ uint64_t shoup(uint64_t a, uint64_t w, uint64_t wp) { return shoup_lazy(a, w, wp); }
__device__ __forceinline__ uint64_t shoup(uint64_t a, const Pair &p) { return shoup(a, p.w, p.wp); }
// store rows and key rows, read through the global address space (glob, device.hpp)
__device__ __forceinline__ uint64_t shoup(uint64_t a, const __attribute__((address_space(1))) Pair &p) {
    return shoup(a, p.w, p.wp);
}

constexpr uint64_t kFourQ = 4 * kQ;
constexpr uint64_t kEightQ = 8 * kQ;  // < 2^63; 16q = 2^64 - 2^18 + 16 < 2^64 bounds every sum
__device__ __forceinline__ uint64_t csub8q(uint64_t x) { return x >= kEightQ ? x - kEightQ : x; }
__device__ __forceinline__ uint64_t csub4q(uint64_t x) { return x >= kFourQ ? x - kFourQ : x; }
__device__ __forceinline__ uint64_t csub2q(uint64_t x) { return x >= kTwoQ ? x - kTwoQ : x; }
__device__ __forceinline__ uint64_t csubq(uint64_t x) { return x >= kQ ? x - kQ : x; }
// [0, 4q) -> canonical
__device__ __forceinline__ uint64_t canon4(uint64_t x) { return csubq(csub2q(x)); }
// [0, 8q) -> canonical
__device__ __forceinline__ uint64_t canon8(uint64_t x) { return csubq(csub2q(csub4q(x))); }

// Lazy ranges. Shoup products land in [0, 4q) for any 64-bit operand, so only the operand
// that is added or subtracted needs a bound, and a reduction is spent only where a later sum
// could reach 16q:
//   inverse (Gentleman-Sande) values are in [0, 8q);
//   forward (Cooley-Tukey) values are in [0, 16q), and a radix-4 pass reduces its two inputs
//   that enter the first layer's sums to [0, 8q).
//
// Gentleman-Sande butterfly, x, y in [0, 8q): x <- (x + y) reduced to [0, 8q), y <- (x - y) w.
__device__ __forceinline__ void gs(uint64_t &x, uint64_t &y, const Pair &w) {
    uint64_t s = csub8q(x + y);   // x + y < 16q
    uint64_t d = x + kEightQ - y;  // (0, 16q)
    x = s;
    y = shoup(d, w);
}
// The same for x, y in [0, 4q) (two Shoup outputs): the sum stays below 8q unreduced.
__device__ __forceinline__ void gs_small(uint64_t &x, uint64_t &y, const Pair &w) {
    uint64_t s = x + y;           // [0, 8q)
    uint64_t d = x + kFourQ - y;  // (0, 8q)
    x = s;
    y = shoup(d, w);
}
// Cooley-Tukey butterfly (Harvey's form without the reduction of x): x in [0, 12q), y any
// 64-bit value; outputs x + t and x + 4q - t with t = y w in [0, 4q), both below 16q.
__device__ __forceinline__ void ct(uint64_t &x, uint64_t &y, const Pair &w) {
    uint64_t t = shoup(y, w);  // [0, 4q)
    uint64_t u = x;
    x = u + t;
    y = u + kFourQ - t;
}

// Radix-4 inverse pass over bits (b, b+1) on a group's registers v[k] = x[base + (k << b)],
// inputs in [0, 8q). The first layer's differences are Shoup outputs in [0, 4q), so the
// second layer's (v1, v3) butterfly needs no reduction. Outputs: v0, v1 in [0, 8q), v2, v3
// in [0, 4q).
__device__ __forceinline__ void gs4r(uint64_t v[4], const Pair &w0, const Pair &w1, const Pair &w2) {
    gs(v[0], v[1], w0);
    gs(v[2], v[3], w1);
    gs(v[0], v[2], w2);
    gs_small(v[1], v[3], w2);
}
template <int LOGN>
__device__ __forceinline__ void gs4(uint64_t v[4], int b, int hi, const Pair *__restrict__ tw) {
    constexpr int N = 1 << LOGN;
    const int m0 = N >> (b + 1), m1 = N >> (b + 2);
    const Pair w0 = tw[m0 + 2 * hi], w1 = tw[m0 + 2 * hi + 1], w2 = tw[m1 + hi];
    gs4r(v, w0, w1, w2);
}
// Radix-4 forward pass over bits (b+1, b), inputs in [0, 16q) (SMALL: v0, v1 already below
// 8q). v0 and v1 are the x operands of the first layer: reduced to [0, 8q), the first layer
// leaves every value below 12q and the second below 16q.
template <int LOGN, bool SMALL = false>
__device__ __forceinline__ void ct4(uint64_t v[4], int b, int hi, const Pair *__restrict__ tw) {
    constexpr int N = 1 << LOGN;
    const int m0 = N >> (b + 2);
    const Pair w0 = tw[m0 + hi], w1 = tw[2 * m0 + 2 * hi], w2 = tw[2 * m0 + 2 * hi + 1];
    if constexpr (!SMALL) {
        v[0] = csub8q(v[0]);
        v[1] = csub8q(v[1]);
    }
    ct(v[0], v[2], w0);
    ct(v[1], v[3], w0);
    ct(v[0], v[1], w1);
    ct(v[2], v[3], w2);
}

// Ordering between two radix-4 passes. Lane t owns slots 4t..4t+3, so each 64-lane wave
// owns 256 consecutive slots in every pass over bits below 8 (the groups of the pass at
// base bit b <= 6 are 4^(b/2) consecutive lanes); exchanges between two such passes need
// only the wave's own LDS order, and a pass touching bit 8 or above takes the barrier.
__device__ __forceinline__ void step_sync(bool wave_local) {
    if (wave_local) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

__device__ __forceinline__ void group_of(int g, int b, int &base, int &hi) {
    const int lo = g & ((1 << b) - 1);
    hi = g >> b;
    base = (hi << (b + 2)) | lo;
}

template <int LOGN>
__device__ __forceinline__ void lds_inv_pass(uint64_t *x, int b, const Pair *tw, int t) {
    int base, hi;
    group_of(t, b, base, hi);
    uint64_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = x[lds_ix(base + (k << b))];
    gs4<LOGN>(v, b, hi, tw);
#pragma unroll
    for (int k = 0; k < 4; k++) x[lds_ix(base + (k << b))] = v[k];
}

template <int LOGN>
__device__ __forceinline__ void lds_fwd_pass2(uint64_t *x, uint64_t *y, int b, const Pair *tw, int t) {
    int base, hi;
    group_of(t, b, base, hi);
    uint64_t u[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        u[k] = x[lds_ix(base + (k << b))];
        v[k] = y[lds_ix(base + (k << b))];
    }
    ct4<LOGN>(u, b, hi, tw);
    ct4<LOGN>(v, b, hi, tw);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[lds_ix(base + (k << b))] = u[k];
        y[lds_ix(base + (k << b))] = v[k];
    }
}

// Inverse transform of the register state s (own slots) into LDS x over bits [0, STOP), with
// a barrier at the end. Passes at base bit BL and up read their twiddles from the LDS copy
// tw_l, lower ones from tw (L2). warm receives one word loaded from warm_src, issued right
// behind the last L2 twiddle loads of the transform: a cache warm-up whose wait the caller
// places late (vector loads complete in order, so no vector load may come between).
template <int LOGN, int STOP, int BL>
__device__ __forceinline__ void inverse_to(uint64_t s[4], uint64_t *x, const Pair *tw, const Pair *tw_l,
                                           Glob<uint32_t> warm_src, uint32_t &warm, int t) {
    constexpr int N = 1 << LOGN;
    const Pair w0 = tw[N / 2 + 2 * t], w1 = tw[N / 2 + 2 * t + 1], w2 = tw[N / 4 + t];
    if constexpr (BL <= 2) warm = *warm_src;
    uint64_t v[4] = {s[0], s[1], s[2], s[3]};
    // the radix-4 inverse pass over bits 0-1 (gs4 at b = 0)
    gs4r(v, w0, w1, w2);
    const int base = 4 * t;
#pragma unroll
    for (int k = 0; k < 4; k++) x[lds_ix(base + k)] = v[k];
    step_sync(STOP > 2);  // wave-local when the pass over bits 2-3 follows
    int b = 2;
    if constexpr (BL > 2) {
        // the pass over bits 2-3 reads L2 twiddles: its twiddle loads, then the warm-up
        int gb, hi;
        group_of(t, 2, gb, hi);
        const int m0 = N >> 3, m1 = N >> 4;
        const Pair u0 = tw[m0 + 2 * hi], u1 = tw[m0 + 2 * hi + 1], u2 = tw[m1 + hi];
        warm = *warm_src;
        uint64_t y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) y[k] = x[lds_ix(gb + (k << 2))];
        gs4r(y, u0, u1, u2);
#pragma unroll
        for (int k = 0; k < 4; k++) x[lds_ix(gb + (k << 2))] = y[k];
        step_sync(6 <= 8 && 4 < STOP);
        b = 4;
    }
    for (; b < STOP; b += 2) {
        lds_inv_pass<LOGN>(x, b, tw_l, t);
        // the next pass is over bits b+2, b+3, or the top of the transform
        step_sync(b + 4 <= 8 && b + 2 < STOP);
    }
}

// The top of the inverse transform, the digit split and the top of both forward
// transforms for this thread's share, then the remaining forward passes down to
// bits 3-2. On return C and D hold the digit transforms after all passes but the
// last (bits 1-0), barrier included.
template <int LOGN, bool EXACT, int BL>
__device__ __forceinline__ void middle(uint64_t *C, uint64_t *D, const Pair *twi, const Pair *twf, const Pair *twf_g,
                                       int t) {
    constexpr int N = 1 << LOGN;
    if constexpr (LOGN & 1) {
        // single top bit: pairs (e, e + N/2), two per thread
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int e = t + r * (N / 4);
            uint64_t x = C[lds_ix(e)], y = C[lds_ix(e + N / 2)];
            gs(x, y, twi[1]);  // x in [0, 8q), y a Shoup output in [0, 4q)
            uint64_t a0, a1, b0, b1;
            gadget_digits<EXACT>(canon8(x), a0, a1);
            gadget_digits<EXACT>(canon4(y), b0, b1);
            ct(a0, b0, twf[1]);
            ct(a1, b1, twf[1]);
            C[lds_ix(e)] = a0;
            C[lds_ix(e + N / 2)] = b0;
            D[lds_ix(e)] = a1;
            D[lds_ix(e + N / 2)] = b1;
        }
        __syncthreads();
        for (int b = LOGN - 3; b >= 2; b -= 2) {
            if (b >= BL)
                lds_fwd_pass2<LOGN>(C, D, b, twf, t);
            else
                lds_fwd_pass2<LOGN>(C, D, b, twf_g, t);
            step_sync(b + 2 <= 8);  // the next pass is over bits b-2, b-1 (or the own slots)
        }
    } else {
        // top two bits: one radix-4 group per thread
        int base, hi;
        const int b = LOGN - 2;
        group_of(t, b, base, hi);
        uint64_t v[4], u[4], w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = C[lds_ix(base + (k << b))];
        gs4<LOGN>(v, b, hi, twi);
        // gs4 leaves v0, v1 in [0, 8q) and v2, v3 in [0, 4q)
        gadget_digits<EXACT>(canon8(v[0]), u[0], w[0]);
        gadget_digits<EXACT>(canon8(v[1]), u[1], w[1]);
        gadget_digits<EXACT>(canon4(v[2]), u[2], w[2]);
        gadget_digits<EXACT>(canon4(v[3]), u[3], w[3]);
        ct4<LOGN, true>(u, b, hi, twf);  // digits: below 2^21
        ct4<LOGN, true>(w, b, hi, twf);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            C[lds_ix(base + (k << b))] = u[k];
            D[lds_ix(base + (k << b))] = w[k];
        }
        __syncthreads();
        for (int bb = LOGN - 4; bb >= 2; bb -= 2) {
            if (bb >= BL)
                lds_fwd_pass2<LOGN>(C, D, bb, twf, t);
            else
                lds_fwd_pass2<LOGN>(C, D, bb, twf_g, t);
            step_sync(bb + 2 <= 8);
        }
    }
}

// Byte p of `count` coefficients packed LSB-first at w bits each (pack_coefficients_tight,
// poly.rs:166-225), read from an LDS array; zero past the last coefficient.
__device__ inline uint32_t tight_byte(const uint64_t *L, uint32_t count, uint32_t w, uint32_t p) {
    const uint32_t bit = 8 * p, i = bit / w, s = bit - i * w;
    if (i >= count) return 0;
    uint64_t v = L[lds_ix(i)] >> s;
    if (w - s < 8 && i + 1 < count) v |= L[lds_ix(i + 1)] << (w - s);
    return (uint32_t)(v & 0xff);
}

// Byte o of the response (layout in WireLayout); a from C, the b prefix from D.
__device__ inline uint32_t wire_byte(const WireLayout &L, const uint64_t *C, const uint64_t *D, uint32_t o) {
    auto le = [](uint64_t v, uint32_t k) -> uint32_t { return (uint32_t)(v >> (8 * k)) & 0xff; };
    if (o < 4) return le(1, o);
    o -= 4;
    if (o < 8) return le(L.a_bytes, o);
    o -= 8;
    if (o < L.a_bytes) return tight_byte(C, L.n, L.w, o);
    o -= L.a_bytes;
    if (o < 40) {
        const uint32_t f = o / 8;
        const uint64_t v = f == 0 ? 1 : f <= 2 ? L.modulus : f == 3 ? L.n : 0;
        return le(v, o % 8);
    }
    o -= 40;
    if (o < 1) return 0;
    o -= 1;
    if (o < 8) return le(L.b_bytes, o);
    o -= 8;
    if (o < L.b_bytes) return tight_byte(D, L.gamma, L.w, o);
    o -= L.b_bytes;
    if (o < 4) return le(L.gamma, o);
    o -= 4;
    if (o < 8) return 0;
    o -= 8;
    if (o < 1) return 1;
    return 0;  // PackingMode::Inspiring (u32 0) and the padding of the last word
}

// Registers are budgeted for 6 waves per SIMD (80 per lane), which lets three workgroups share
// a CU. The same build is also the fastest with one or two queries per CU (measured: 13%
// less time per query alone on a CU than a 128-register build): nothing is hoisted out of
// the recursion that the step can recompute.
template <int LOGN, int ROWS, bool EXACT>
__global__ __launch_bounds__((1 << LOGN) / 4) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_chain(CtxDev c, BatchDev bt) {
    constexpr int N = 1 << LOGN;
    constexpr int T = N / 4;
    // Three workgroups per CU must fit 160 KiB of LDS (53 KiB each at n = 2048): NTT(a)
    // unpadded (it is only gathered from), LDS twiddles for the passes at bit 4 and up only,
    // and the prologue's two flags in the first words of C, which nothing else touches before
    // the recursion.
    constexpr int BL = 4;                    // lowest pass base bit with LDS twiddles
    constexpr int TWN = (1 << LOGN) >> BL;   // ladder entries below n >> BL
    __shared__ uint64_t A[N];                      // NTT(a), unpadded
    __shared__ uint64_t C[LdsSize<LOGN>::value];   // transform buffer / digit 0
    __shared__ uint64_t D[LdsSize<LOGN>::value];   // digit 1
    int &bad_column = reinterpret_cast<int *>(C)[0];
    int &rejected = reinterpret_cast<int *>(C)[1];
    auto aix = [](int e) { return e; };

    // a grouped launch takes its queries in shard order, a contiguous slice of it per XCD
    const int qi = batch_grouped(c, bt) ? (int)bt.order[xcd_slot(blockIdx.x, bt.count)] : (int)blockIdx.x;
    const int tid = threadIdx.x, gamma = c.gamma;
    // the ids of device-buffer batches are not checked on the host: a query naming no shard
    // with a store is refused here, before anything reads the shard (the whole workgroup)
    // One thread reads the entry and the workgroup takes its copy from LDS, so every wave
    // decides alike even if the table changes under a running launch.
    __shared__ ShardDev SH;
    if (tid == 0) SH = shard_entry(c.shards, c.num_shards, bt.shard_ids[qi]);
    __syncthreads();
    const ShardDev sh{uniform(SH.records), uniform(SH.store), uniform(SH.tag_key), uniform(SH.tags)};
    if (!shard_ready(sh)) {
        if (tid == 0) bt.status[qi] = -1;
        return;
    }
    const Pair *twf = c.twf, *twi = c.twi;
    const Glob<Pair> Y = glob(bt.yprimes ? bt.yprimes[qi] : bt.yprime);
    // Twiddles of the passes at base bit BL and up (ladder entries below n >> BL: 16 KiB at
    // n = 2048 for BL = 2, 4 KiB for BL = 4) live in LDS for the whole kernel; the lower
    // passes read theirs from L2.
    __shared__ Pair TWL[2][TWN];
    for (int m = tid; m < TWN; m += T) {
        TWL[0][m] = twf[m];
        TWL[1][m] = twi[m];
    }
    const Pair *const twf_l = TWL[0], *const twi_l = TWL[1];

    // --- finish the scan: b_poly and the column tags, before any packing work
    if (tid == 0) { bad_column = 0x7fffffff; rejected = 0; }
    __syncthreads();
    // b_poly[j] is parked in the first word of the query's own scan partials for column j,
    // which only this thread reads, until the epilogue (keeps it out of LDS)
    uint64_t *const bpoly = bt.scan_part + (size_t)qi * bt.ks * gamma * 4;
    const int slices = scan_slices(c, bt);
    for (int j = tid; j < gamma; j += T) {
        uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int s = 0; s < slices; s++) {
            const uint64_t *p = bt.scan_part + (((size_t)qi * bt.ks + s) * gamma + j) * 4;
            s0 += p[0]; s1 += p[1]; s2 += p[2]; s3 += p[3];
        }
        uint64_t lo = s0 + (s1 << 32), hi = (s1 >> 32) + (lo < s0);
        bpoly[4 * j] = reduce120(hi, lo);
        uint64_t tlo = s2 + (s3 << 32), thi = (s3 >> 32) + (tlo < s2);
        if (mod_p61(thi, tlo) != sh.tags[j]) atomicMin(&bad_column, j);
    }
    __syncthreads();
    if (bad_column != 0x7fffffff) {
        if (tid == 0) bt.status[qi] = bad_column + 1;
        return;
    }

    // --- a: the query's expanded row, or gen_range(0..q) draws from ChaCha20(seed)
    if (bt.a_in) {
        for (int j = tid; j < N; j += T) A[aix(j)] = bt.a_in[(size_t)qi * N + j];
    } else {
        uint32_t key[8];
        seed_key(bt.seeds + (size_t)qi * 32, key);
        for (int blk = tid; blk < N / 8; blk += T) {
            uint32_t w[16];
            chacha20_block(key, (uint64_t)blk, 0, w);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                uint64_t v = (uint64_t)w[2 * u] | (uint64_t)w[2 * u + 1] << 32, a;
                if (!uniform_q_draw(v, a)) rejected = 1;
                A[aix(blk * 8 + u)] = a;
            }
        }
        __syncthreads();
        if (rejected) {
            // a draw outside the acceptance zone (probability ~2^-46 each): replay the
            // stream serially with rejection, as the reference's RNG does
            if (tid == 0) replay_gen_range_q(key, N, kZoneQ, [&](int k, uint64_t a) { A[aix(k)] = a; });
        }
    }
    __syncthreads();
    // NTT(a) with the radix-4 forward passes from the top bit down, in place in A
    {
        if constexpr (LOGN & 1) {
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int e = tid + r * (N / 4);
                uint64_t x = A[aix(e)], y = A[aix(e + N / 2)];
                ct(x, y, twf_l[1]);
                A[aix(e)] = x;
                A[aix(e + N / 2)] = y;
            }
            __syncthreads();
        }
        for (int b = (LOGN & 1) ? LOGN - 3 : LOGN - 2; b >= 0; b -= 2) {
            int base, hi;
            group_of(tid, b, base, hi);
            uint64_t v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = A[aix(base + (k << b))];
            if (b >= BL)
                ct4<LOGN>(v, b, hi, twf_l);
            else
                ct4<LOGN>(v, b, hi, twf);
#pragma unroll
            for (int k = 0; k < 4; k++) A[aix(base + (k << b))] = v[k];
            __syncthreads();
        }
    }

    // --- step 1 for the last rotation, on own slots
    uint64_t X[4], acc[4];
    {
        const uint32_t t = c.gpow[gamma - 1];
        const Glob<Pair> S = glob(sh.store + (size_t)(gamma - 1) * N);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = 4 * tid + k;
            X[k] = shoup(A[aix(automorph_slot(t, j, LOGN))], S[j]);
            acc[k] = 0;
        }
    }

    // --- backward recursion with the online gather
    for (int i = gamma - 2; i >= 0; i--) {
        // the thread index is re-read through an empty asm each step: its derived addresses
        // are then recomputed in the step instead of being hoisted out of the loop into
        // registers, which the 80-register budget cannot hold (no spills, against 45 spilled
        // registers when hoisted)
        int tl = tid;
        asm volatile("" : "+v"(tl));
        const Glob<Pair> S = glob(sh.store + (size_t)i * N);
        // this step's store row is read at the end of the step; touching its lines now (one
        // word per thread, unconditionally, so the wait counts stay exact) lets the HBM fetch
        // overlap the transforms
        uint32_t warm = 0;
        inverse_to<LOGN, (LOGN & 1) ? LOGN - 1 : LOGN - 2, BL>(X, C, twi, twi_l,
                         reinterpret_cast<Glob<uint32_t>>(S + 4 * tl), warm, tl);
        middle<LOGN, EXACT, BL>(C, D, twi_l, twf_l, twf, tl);
        // last forward pass (bits 1-0) on own slots, written back to those slots: the pointwise
        // work reads each slot's two digit transforms from LDS when it needs them, which keeps
        // 16 registers free for its row loads (no other lane touches these slots until the next
        // step's first pass)
        const int base = 4 * tl;
        {
            uint64_t t0[4], t1[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                t0[k] = C[lds_ix(base + k)];
                t1[k] = D[lds_ix(base + k)];
            }
            ct4<LOGN>(t0, 0, tl, twf);
            ct4<LOGN>(t1, 0, tl, twf);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                C[lds_ix(base + k)] = t0[k];
                D[lds_ix(base + k)] = t1[k];
            }
        }
        asm volatile("" ::"v"(warm));  // the warm-up load completes here, long after its issue
        const uint32_t t = c.gpow[i];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = base + k;
            const uint32_t src = automorph_slot(t, j, LOGN);
            // exact gadget: the third digit's transform is kx*R + k0*T0 + k1*T1, folded into the
            // rows (W_2 = kx*w_2, W_0 = w_0 + k0*w_2, W_1 = w_1 + k1*w_2; same for y), so its
            // operand here is R itself
            const uint64_t d[3] = {C[lds_ix(j)], D[lds_ix(j)], X[k]};
            uint64_t r = shoup(A[aix(src)], S[j]);
            uint64_t o = acc[k];
#pragma unroll
            for (int rr = 0; rr < ROWS; rr++) {
                r += shoup(d[rr], c.wprime[(size_t)rr * N + src]);
                o += shoup(d[rr], Y[(size_t)rr * N + src]);
            }
            // r < (1 + ROWS) 4q <= 16q < 2^64: the state back to [0, 8q); o < acc + ROWS 4q
            // <= 16q: the online sum back to [0, 4q), so the next step's o stays below 16q
            X[k] = csub8q(r);
            acc[k] = csub4q(csub8q(o));
        }
    }

    // --- a_hat = iNTT(R_0) into C, iNTT(sum) into D
    {
        uint64_t v[4] = {X[0], X[1], X[2], X[3]}, u[4] = {acc[0], acc[1], acc[2], acc[3]};
        gs4<LOGN>(v, 0, tid, twi);
        gs4<LOGN>(u, 0, tid, twi);
        const int base = 4 * tid;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            C[lds_ix(base + k)] = v[k];
            D[lds_ix(base + k)] = u[k];
        }
        __syncthreads();
        for (int b = 2; b < LOGN - 1; b += 2) {
            if (b >= BL) {
                lds_inv_pass<LOGN>(C, b, twi_l, tid);
                lds_inv_pass<LOGN>(D, b, twi_l, tid);
            } else {
                lds_inv_pass<LOGN>(C, b, twi, tid);
                lds_inv_pass<LOGN>(D, b, twi, tid);
            }
            __syncthreads();
        }
        if constexpr (LOGN & 1) {
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int e = tid + r * (N / 4);
                uint64_t x = C[lds_ix(e)], y = C[lds_ix(e + N / 2)];
                gs(x, y, twi_l[1]);  // x in [0, 8q), y in [0, 4q)
                C[lds_ix(e)] = canon8(x);
                C[lds_ix(e + N / 2)] = canon4(y);
                x = D[lds_ix(e)], y = D[lds_ix(e + N / 2)];
                gs(x, y, twi_l[1]);
                D[lds_ix(e)] = canon8(x);
                D[lds_ix(e + N / 2)] = canon4(y);
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int e = tid + r * T;
                C[lds_ix(e)] = canon8(C[lds_ix(e)]);  // inverse outputs: [0, 8q)
                D[lds_ix(e)] = canon8(D[lds_ix(e)]);
            }
        }
        __syncthreads();
    }
    const uint64_t qn = bt.switch_modulus;
    if (!bt.wire_out) {
        for (int j = tid; j < N; j += T) {
            uint64_t a = C[lds_ix(j)];
            bt.a_out[(size_t)qi * N + j] = qn ? mod_switch_coeff(a, qn) : a;
        }
        for (int j = tid; j < gamma; j += T) {
            uint64_t bv = addmod(bpoly[4 * j], D[lds_ix(j)]);
            bt.b_out[(size_t)qi * gamma + j] = qn ? mod_switch_coeff(bv, qn) : bv;
        }
    } else {
        // final coefficients back in place (each thread its own slots), then the response
        // bytes assembled from LDS, one 4-byte word per lane and pass
        for (int j = tid; j < N; j += T) {
            uint64_t a = C[lds_ix(j)];
            a = qn ? mod_switch_coeff(a, qn) : a;
            C[lds_ix(j)] = a;
            if (bt.a_out) bt.a_out[(size_t)qi * N + j] = a;
        }
        for (int j = tid; j < gamma; j += T) {
            uint64_t bv = addmod(bpoly[4 * j], D[lds_ix(j)]);
            bv = qn ? mod_switch_coeff(bv, qn) : bv;
            D[lds_ix(j)] = bv;
            if (bt.b_out) bt.b_out[(size_t)qi * gamma + j] = bv;
        }
        __syncthreads();
        const WireLayout L = wire_layout(N, gamma, qn ? qn : kQ);
        uint32_t *out = reinterpret_cast<uint32_t *>(bt.wire_out + (size_t)qi * bt.wire_stride);
        const uint32_t words = (L.total + 3) / 4;
        for (uint32_t wi = tid; wi < words; wi += T) {
            uint32_t word = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) word |= wire_byte(L, C, D, 4 * wi + k) << (8 * k);
            out[wi] = word;
        }
    }
    if (tid == 0) bt.status[qi] = 0;
}

// At most queries_per_cu workgroups per CU: a workgroup reserves more than
// 1/(queries_per_cu + 1) of the CU's 160 KiB LDS (dynamic padding, never touched), so the
// dispatcher spreads a launch over all CUs instead of stacking queries on some and
// idling others.
constexpr size_t kLdsPerCu = 160 * 1024;

template <int LOGN>
static hipError_t launch_chain_t(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    constexpr int T = (1 << LOGN) / 4;
    auto kern = c.rows == 3 ? k_chain<LOGN, 3, true> : k_chain<LOGN, 2, false>;
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(kern));
    if (e != hipSuccess) return e;
    size_t dyn = 0;
    if (c.queries_per_cu <= 2) {
        const size_t share = kLdsPerCu / (size_t)(c.queries_per_cu + 1) + 256;
        if (attr.sharedSizeBytes < share) dyn = share - attr.sharedSizeBytes;
    }
    hipLaunchKernelGGL(kern, dim3(b.count), dim3(T), dyn, s, c, b);
    return hipGetLastError();
}

hipError_t launch_chain_only(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (b.count == 0) return hipSuccess;
    switch (c.logn) {
        case 8: return launch_chain_t<8>(c, b, s);
        case 9: return launch_chain_t<9>(c, b, s);
        case 10: return launch_chain_t<10>(c, b, s);
        case 11: return launch_chain_t<11>(c, b, s);
        case 12: return launch_chain_t<12>(c, b, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
