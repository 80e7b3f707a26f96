// gfx950 kernels of the special-modulus set: step-1 table build, key seat, and the
// respond chain, one workgroup of n/8 threads per query.
//
// Restates src/special_modulus/server.rs (table :211-250 with step1_row_impl :330-383,
// seat :177-203, respond_impl :385-499) and the key mask of mod.rs:313-330.
//
// The reference recursion, for i = gamma-2 .. 0 with R = tau_{g^{gamma-1}}(a) * S_{gamma-1}:
//   tp = NTT_p(lift(iNTT_q(R)));  cp = tp * tau_i(w_p);  r = NTT_q(lift(iNTT_p(cp)))
//   R' = tau_i(a) * S_i + (R * tau_i(w_q) - r) / p
//   acc_q += R * tau_i(key_q);  acc_p += tp * tau_i(key_p)
// with tau_i the slot permutation of X -> X^(g^i). The two lifts are centred and odd,
// so transform, lift and transform commute with every tau. This kernel runs the same
// recursion in the frame V_i = tau_i^-1(R_i): the state is rotated by tau_g once per
// step (Y = tau_g(V)), and then every factor is un-rotated:
//   V_i = a * S'_i + (Y * w_q - r(tp(Y) * w_p)) / p,  S'_i = tau_i^-1(S_i) = gamma^-1 NTT(P_i)
// The accumulators follow by Horner, B <- tau_g(B) + Y * key each step, which leaves
// acc = B_0 in the reference's own frame. So the query's NTT(a) and the seat's four
// rows stay in registers on the thread's own slots, the table streams S'_i coalesced,
// and the only gather per step is one LDS permutation of (V, B_q, B_p).
//
// n^-1 is folded into the constants (the state is carried as V/n, the table holds
// (n gamma)^-1 NTT(P_i), the seat holds n w_q, w_p/n and key_p/n, the division by p is
// by n p), so every inverse transform here is unscaled and every lift still sees the
// exact coefficients. Products by the twiddles and the kernel-wide constants use Shoup
// companions, a*w - floor(a*w'/2^32)*m; products by the per-query rows (the table,
// the seat) are Montgomery products a*b*2^-32 with the rows stored as b = x*2^32, so
// they need no companion words in HBM or registers. Both land in [0, 2m).
// Residues mod q stay lazy in [0, 4q) (6q < 2^32); residues mod p (p < 2^31) stay canonical,
// except the forward transforms' outputs, in [0, 2p). A ciphertext prime above 2^32 / 6 (the
// reference accepts any below 2^31) runs the kernels' WQ instantiation, whose residues mod q
// follow the rules of p: canonical, forward transform outputs in [0, 2q).
//
// Layout: 8 slots per thread. Radix-8 passes over bits 0-2 (the thread's own slots
// 8t..8t+7), 3-5, 6-8 ..., the top pass taking the remaining 1-3 bits; each round trip
// fuses the top inverse pass, the centred lift and the top forward pass. In LDS, slot e
// sits at word e + 8 (e >> 6), with its two 4-word halves swapped where bit 6 of e is
// set: every pass then reads a thread's group at one base plus immediate offsets, the
// word accesses of the passes over bits 3-5, 6-8 and 9-10 hit 64 distinct banks, and the
// 16-byte own-slot accesses of 16 consecutive lanes cover all 64 banks once.
#include <hip/hip_runtime.h>

#include "chacha.hpp"
#include "sm.hpp"

namespace inspire {

namespace {

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
// [0, 2m) -> [0, m)
__device__ __forceinline__ uint32_t red(uint32_t x, uint32_t m) { return umin(x, x - m); }
__device__ __forceinline__ uint32_t shoup(uint32_t a, Pair32 w, uint32_t m) { return a * w.w - __umulhi(a, w.wp) * m; }
// a * b * 2^-32 mod m in [0, 2m) for any a and b < m; minv = -m^-1 mod 2^32. The low
// words of a*b and k*m cancel, carrying exactly when a*b's low word is nonzero.
__device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b, uint32_t m, uint32_t minv) {
    const uint32_t lo = a * b, hi = __umulhi(a, b);
    return hi + __umulhi(lo * minv, m) + (lo != 0u);
}

// mod q, lazy: GS on [0, 2q) -> [0, 2q); CT (Harvey) on [0, 4q) -> [0, 4q)
__device__ __forceinline__ void gs_q(uint32_t &x, uint32_t &y, Pair32 w, uint32_t q) {
    const uint32_t s = x + y, d = x - y + 2 * q;
    x = red(s, 2 * q);
    y = shoup(d, w, q);
}
__device__ __forceinline__ void ct_q(uint32_t &x, uint32_t &y, Pair32 w, uint32_t q) {
    const uint32_t u = red(x, 2 * q), t = shoup(y, w, q);
    x = u + t;
    y = u - t + 2 * q;
}
// mod p (4p overflows a word): GS canonical -> canonical; CT (Harvey) on [0, 2p) -> [0, 2p),
// since every consumer of a forward transform mod p (a Montgomery or Shoup product) takes
// any word
__device__ __forceinline__ void gs_p(uint32_t &x, uint32_t &y, Pair32 w, uint32_t p) {
    const uint32_t s = x + y, d = x - y + p;
    x = red(s, p);
    y = red(shoup(d, w, p), p);
}
__device__ __forceinline__ void ct_p(uint32_t &x, uint32_t &y, Pair32 w, uint32_t p) {
    const uint32_t u = red(x, p), t = red(shoup(y, w, p), p);
    x = u + t;
    y = u - t + p;
}

template <int LOGN>
struct Plan {
    static constexpr int N = 1 << LOGN, T = N / 8;
    static constexpr int BT = 3 * ((LOGN - 1) / 3);  // base bit of the top pass
    static constexpr int RT = LOGN - BT;             // its radix bits (1-3)
};

__device__ __forceinline__ int swz(int e) { return (e ^ (((e >> 6) & 1) << 2)) + ((e >> 6) << 3); }
// words of one transform buffer
template <int LOGN>
struct Lds {
    static constexpr int words = (1 << LOGN) + (1 << LOGN) / 8;
};

// A value the optimiser cannot see through, once per loop trip: loads and unpacking
// that depend on it (the twiddles, the rotation offsets) are issued where they are used
// instead of being hoisted out of the loop into hundreds of live registers.
__device__ __forceinline__ int opaque_s(int x) {
    asm volatile("" : "+s"(x));
    return x;
}
__device__ __forceinline__ uint32_t opaque_v(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}
// hi of a pass over bits >= 6 is the same on every lane of a wave (64 lanes, or a
// whole workgroup of n/8 <= 64 threads aligned to its size): keep it scalar
template <int B>
__device__ __forceinline__ int pass_hi(int gid) {
    const int hi = gid >> B;
    if constexpr (B >= 6) return __builtin_amdgcn_readfirstlane(hi);
    return hi;
}

// Slot of group element k of the thread's group gi in a pass over bits [B, B+R).
template <int LOGN, int B, int R>
__device__ __forceinline__ int gpos(int t, int gi, int k) {
    const int gid = t + gi * Plan<LOGN>::T;
    const int lo = gid & ((1 << B) - 1), hi = gid >> B;
    return (hi << (B + R)) | lo | (k << B);
}

template <int LOGN, int B, int R>
__device__ __forceinline__ void lds_load(const uint32_t *X, uint32_t v[8], int t) {
#pragma unroll
    for (int gi = 0; gi < (8 >> R); gi++)
#pragma unroll
        for (int k = 0; k < (1 << R); k++) v[gi * (1 << R) + k] = X[swz(gpos<LOGN, B, R>(t, gi, k))];
}
template <int LOGN, int B, int R>
__device__ __forceinline__ void lds_store(uint32_t *X, const uint32_t v[8], int t) {
#pragma unroll
    for (int gi = 0; gi < (8 >> R); gi++)
#pragma unroll
        for (int k = 0; k < (1 << R); k++) X[swz(gpos<LOGN, B, R>(t, gi, k))] = v[gi * (1 << R) + k];
}
// own slots 8t..8t+7: two 16-byte accesses
__device__ __forceinline__ void own_load(const uint32_t *X, uint32_t v[8], int t) {
    const uint4 a = *reinterpret_cast<const uint4 *>(X + swz(8 * t));
    const uint4 b = *reinterpret_cast<const uint4 *>(X + swz(8 * t + 4));
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void own_store(uint32_t *X, const uint32_t v[8], int t) {
    *reinterpret_cast<uint4 *>(X + swz(8 * t)) = make_uint4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<uint4 *>(X + swz(8 * t + 4)) = make_uint4(v[4], v[5], v[6], v[7]);
}
// Device-memory accesses through the global address space: pointers read from a job entry are
// generic to the compiler, and a generic (flat) load also counts against the LDS counter, so
// every LDS wait of the transform behind it would wait for HBM as well.
typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) U32x4 *GVec;
typedef __attribute__((address_space(1))) U32x4 *GVecOut;
__device__ __forceinline__ void gload8(const uint32_t *g, uint32_t v[8]) {
    const GVec p = (GVec)g;
    const U32x4 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void gstore8(uint32_t *g, const uint32_t v[8]) {
    const GVecOut p = (GVecOut)g;
    p[0] = U32x4{v[0], v[1], v[2], v[3]};
    p[1] = U32x4{v[4], v[5], v[6], v[7]};
}

// Inverse (Gentleman-Sande) stages B .. B+R-1 on the thread's groups. The twiddle of
// stage s for the block holding slot e is tw[(n >> (s+1)) + (e >> (s+1))], the index
// the reference's inverse loop walks (arith.rs:247-272).
template <int LOGN, int B, int R, bool P>
__device__ __forceinline__ void inv_pass(uint32_t v[8], int t, const Pair32 *__restrict__ tw, uint32_t m) {
    constexpr int N = 1 << LOGN;
#pragma unroll
    for (int gi = 0; gi < (8 >> R); gi++) {
        const int hi = pass_hi<B>(t + gi * Plan<LOGN>::T);
        uint32_t *x = v + gi * (1 << R);
#pragma unroll
        for (int s = 0; s < R; s++)
#pragma unroll
            for (int k = 0; k < (1 << R); k++) {
                if (k & (1 << s)) continue;
                const Pair32 w = tw[(N >> (B + s + 1)) + (hi << (R - s - 1)) + (k >> (s + 1))];
                if (P) gs_p(x[k], x[k + (1 << s)], w, m);
                else gs_q(x[k], x[k + (1 << s)], w, m);
            }
    }
}
// Forward (Cooley-Tukey) stages B+R-1 .. B, same twiddle index (arith.rs:188-205).
template <int LOGN, int B, int R, bool P>
__device__ __forceinline__ void fwd_pass(uint32_t v[8], int t, const Pair32 *__restrict__ tw, uint32_t m) {
    constexpr int N = 1 << LOGN;
#pragma unroll
    for (int gi = 0; gi < (8 >> R); gi++) {
        const int hi = pass_hi<B>(t + gi * Plan<LOGN>::T);
        uint32_t *x = v + gi * (1 << R);
#pragma unroll
        for (int s = R - 1; s >= 0; s--)
#pragma unroll
            for (int k = 0; k < (1 << R); k++) {
                if (k & (1 << s)) continue;
                const Pair32 w = tw[(N >> (B + s + 1)) + (hi << (R - s - 1)) + (k >> (s + 1))];
                if (P) ct_p(x[k], x[k + (1 << s)], w, m);
                else ct_q(x[k], x[k + (1 << s)], w, m);
            }
    }
}

// Ordering between two passes whose groups span slot bits below HI. Each 64-lane wave
// owns the 512 slots 512w .. 512w+511 in every pass that stays below bit 9 (8 slots per
// lane: the own pass, and the passes over bits 3-5 and 6-8 whose groups are 8 and 64
// consecutive lanes), so those exchanges only need the wave's own LDS order; a pass
// reaching bit 9 or above mixes the waves and takes the workgroup barrier.
template <int HI>
__device__ __forceinline__ void pass_sync() {
    if constexpr (HI <= 9) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}
// top bit of the pass that follows the pass over bits [B, B+3) going up
template <int LOGN, int B>
struct NextUp {
    static constexpr int value = (B + 3 < Plan<LOGN>::BT) ? B + 6 : LOGN;
};

// radix-8 passes from bit B up to (not including) the top pass, through LDS
template <int LOGN, int B, bool P>
__device__ __forceinline__ void inv_mid(uint32_t *X, const Pair32 *tw, uint32_t m, int t) {
    if constexpr (B < Plan<LOGN>::BT) {
        uint32_t v[8];
        lds_load<LOGN, B, 3>(X, v, t);
        inv_pass<LOGN, B, 3, P>(v, t, tw, m);
        lds_store<LOGN, B, 3>(X, v, t);
        pass_sync<NextUp<LOGN, B>::value>();
        inv_mid<LOGN, B + 3, P>(X, tw, m, t);
    }
}
// radix-8 passes from bit B down to bit 3, through LDS
template <int LOGN, int B, bool P>
__device__ __forceinline__ void fwd_mid(uint32_t *X, const Pair32 *tw, uint32_t m, int t) {
    if constexpr (B >= 3) {
        uint32_t v[8];
        lds_load<LOGN, B, 3>(X, v, t);
        fwd_pass<LOGN, B, 3, P>(v, t, tw, m);
        lds_store<LOGN, B, 3>(X, v, t);
        pass_sync<B + 3>();
        fwd_mid<LOGN, B - 3, P>(X, tw, m, t);
    }
}

// centred lifts (mod.rs:333-356)
__device__ __forceinline__ uint32_t lift_qp(uint32_t x, const SmCtxDev &c) {
    const uint32_t v = red(x, c.q);  // [0, 2q) -> canonical
    return v > c.half_q ? v + c.p_minus_q : v;
}
__device__ __forceinline__ uint32_t lift_pq(uint32_t r, const SmCtxDev &c) {
    const uint32_t x = r - __umulhi(r, c.q_barrett) * c.q;  // r mod q, in [0, 2q)
    return x + (r > c.half_p ? c.neg_off : 0u);             // [0, 3q)
}

// The same lift for a ciphertext prime above 2^32 / 6, canonical: r mod q, less p mod q when
// r > p/2 (neg_off = (q - p mod q) mod q).
__device__ __forceinline__ uint32_t lift_pq_wide(uint32_t r, const SmCtxDev &c) {
    const uint32_t x = red(r - __umulhi(r, c.q_barrett) * c.q, c.q);  // r mod q
    return r > c.half_p ? red(x + c.neg_off, c.q) : x;
}

// iNTT of the own-slot registers v, centred lift, NTT: QP maps q -> p, else p -> q.
// Leaves v in the other prime's transform domain (mod p in [0, 2p), mod q in [0, 4q), or
// [0, 2q) under WQ). Inputs mod q: [0, 2q), canonical under WQ.
template <int LOGN, bool QP, bool WQ, class Mid>
__device__ __forceinline__ void round_trip(uint32_t v[8], uint32_t *X, const SmCtxDev &c, const Pair32 *twi,
                                           const Pair32 *twf, int t, Mid &&mid) {
    constexpr int BT = Plan<LOGN>::BT, RT = Plan<LOGN>::RT;
    // canonical-residue transforms: always mod p, mod q under WQ
    constexpr bool PI = QP ? WQ : true, PF = QP ? true : WQ;
    const uint32_t mi = QP ? c.q : c.p, mf = QP ? c.p : c.q;
    inv_pass<LOGN, 0, 3, PI>(v, t, twi, mi);
    own_store(X, v, t);
    pass_sync<NextUp<LOGN, 0>::value>();
    inv_mid<LOGN, 3, PI>(X, twi, mi, t);
    {
        uint32_t u[8];
        lds_load<LOGN, BT, RT>(X, u, t);
        inv_pass<LOGN, BT, RT, PI>(u, t, twi, mi);
#pragma unroll
        for (int k = 0; k < 8; k++) u[k] = QP ? lift_qp(u[k], c) : WQ ? lift_pq_wide(u[k], c) : lift_pq(u[k], c);
        fwd_pass<LOGN, BT, RT, PF>(u, t, twf, mf);
        lds_store<LOGN, BT, RT>(X, u, t);
    }
    pass_sync<LOGN>();
    mid();  // issued here: loads the caller needs after the transform
    fwd_mid<LOGN, BT - 3, PF>(X, twf, mf, t);
    own_load(X, v, t);
    fwd_pass<LOGN, 0, 3, PF>(v, t, twf, mf);
}

// Forward transform of the coefficients in X; returns the own slots in v.
template <int LOGN, bool P>
__device__ __forceinline__ void forward_from_lds(uint32_t *X, uint32_t v[8], const Pair32 *tw, uint32_t m, int t) {
    constexpr int BT = Plan<LOGN>::BT, RT = Plan<LOGN>::RT;
    {
        uint32_t u[8];
        lds_load<LOGN, BT, RT>(X, u, t);
        fwd_pass<LOGN, BT, RT, P>(u, t, tw, m);
        lds_store<LOGN, BT, RT>(X, u, t);
    }
    pass_sync<LOGN>();
    fwd_mid<LOGN, BT - 3, P>(X, tw, m, t);
    own_load(X, v, t);
    fwd_pass<LOGN, 0, 3, P>(v, t, tw, m);
}

// Unscaled inverse transform of the own-slot registers v into X (coefficient order,
// swizzled), barrier included. Mod q outputs are in [0, 2q), mod p canonical.
template <int LOGN, bool P>
__device__ __forceinline__ void inverse_to_lds(uint32_t v[8], uint32_t *X, const Pair32 *tw, uint32_t m, int t) {
    constexpr int BT = Plan<LOGN>::BT, RT = Plan<LOGN>::RT;
    inv_pass<LOGN, 0, 3, P>(v, t, tw, m);
    own_store(X, v, t);
    pass_sync<NextUp<LOGN, 0>::value>();
    inv_mid<LOGN, 3, P>(X, tw, m, t);
    {
        uint32_t u[8];
        lds_load<LOGN, BT, RT>(X, u, t);
        inv_pass<LOGN, BT, RT, P>(u, t, tw, m);
        lds_store<LOGN, BT, RT>(X, u, t);
    }
    __syncthreads();
}

__device__ __forceinline__ int bitrev_n(int x, int logn) { return (int)(__brev((uint32_t)x) >> (32 - logn)); }

// Slot of f that slot j of tau_t(f) reads: slot j evaluates at psi^(2 brv(j) + 1).
__device__ __forceinline__ int rot_slot(uint32_t t, int j, int logn) {
    const uint32_t two_n_mask = (2u << logn) - 1;
    const uint32_t e = (t * (2u * (uint32_t)bitrev_n(j, logn) + 1u)) & two_n_mask;
    return bitrev_n((int)(e >> 1), logn);
}

// ---------------------------------------------------------------- respond

// Four workgroups per CU (128 registers per lane, no spills). With a launch's jobs in seat order
// the seat rows the chain re-reads each step come from L2, and four unspilled workgroups per CU
// beat five with 17 registers spilled (23.6-24.5 against 25.9-26.1 ms per launch of 2560).
template <int LOGN>
struct SmOccupancy {
    // workgroups per CU the register budget is set for: 4 waves per SIMD at every ring up to
    // 2048 (n/8 <= 256 threads), and at 4096 and 8192 (512 and 1024 threads per workgroup)
    static constexpr int blocks = LOGN <= 11 ? 4 : LOGN == 12 ? 2 : 1;
};
template <int LOGN, bool WQ>
__global__ __launch_bounds__((1 << LOGN) / 8, SmOccupancy<LOGN>::blocks) void k_sm_respond(SmCtxDev c, SmBatchDev bt) {
    constexpr int N = 1 << LOGN, T = N / 8, NW = Lds<LOGN>::words;
    // three buffers (27 KiB): the transform buffer X also serves the rotation of V (PV), the
    // accumulators keep PQ and PP
    __shared__ __align__(16) uint32_t L[3 * NW];
    __shared__ int rejected;
    uint32_t *X = L, *PV = L, *PQ = L + NW, *PP = L + 2 * NW;
    const int t = threadIdx.x;
    const SmJob job = bt.jobs[blockIdx.x];
    const int qi = job.qi;
    const uint32_t q = c.q, p = c.p;

    // --- a = expand_uniform(ChaCha20(seed), q) (server.rs:408), one block of 8 draws per thread
    if (t == 0) rejected = 0;
    __syncthreads();
    {
        uint32_t key[8];
        seed_key(bt.seeds + (size_t)qi * 32, key);
        uint32_t w[16];
        chacha20_block(key, (uint64_t)t, 0, w);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint64_t u = (uint64_t)w[2 * i] | (uint64_t)w[2 * i + 1] << 32;
            if (u >= c.zone_q) rejected = 1;
            X[swz(8 * t + i)] = reduce64(u, q, c.q_recip);
        }
        __syncthreads();
        if (rejected) {
            // a draw in the short last interval (probability < 2^-34 each): replay the
            // stream serially with rejection, as the reference's loop does
            if (t == 0)
                replay_expand_uniform(key, 0, 1, N, &q, &c.zone_q, &c.q_recip,
                                      [&](int, int k, uint32_t v) { X[swz(k)] = v; });
            __syncthreads();
        }
    }
    // NTT(a), lazy in [0, 4q) ([0, 2q) under WQ), parked in the query's slice of the mask
    // output: like the seat's rows, it is read once per step, just ahead of its use, instead of
    // pinning registers for all steps. V: [0, 2q), canonical under WQ.
    uint32_t *apark = bt.mask_out + (size_t)qi * N + 8 * t;
    uint32_t V[8], BQ[8], BP[8];
    {
        uint32_t A[8], S[8];
        forward_from_lds<LOGN, WQ>(X, A, c.twf_q, q, t);
        gstore8(apark, A);
        gload8(job.step1 + (size_t)(c.gamma - 1) * N + 8 * t, S);
#pragma unroll
        for (int k = 0; k < 8; k++) V[k] = red(mont(A[k], S[k], q, c.qninv), WQ ? q : 2 * q);
    }
    const uint32_t *seat = job.seat + 8 * t;

    // --- the rotation by tau_g as LDS offsets (two 16-bit offsets per word)
    uint32_t rot[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        rot[k] = (uint32_t)swz(rot_slot(c.g, 8 * t + 2 * k, LOGN)) |
                 (uint32_t)swz(rot_slot(c.g, 8 * t + 2 * k + 1, LOGN)) << 16;

    // The accumulators wait in their permutation buffers (own slots) between uses rather than
    // in registers across the transforms: B_p is gathered after the first transform and stored
    // back during the second, past its first barrier; B_q is gathered after the second and stored
    // after the next step's first barrier, so no store meets another thread's gather.
    uint32_t bq[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        bq[k] = 0;
        BP[k] = 0;
    }
    own_store(PP, BP, t);

    for (int i = c.gamma - 2; i >= 0; i--) {
        const int z = opaque_s(0);
        const Pair32 *twi_q = c.twi_q + z, *twf_q = c.twf_q + z, *twi_p = c.twi_p + z, *twf_p = c.twf_p + z;
        // Y = tau_g(V); B <- tau_g(B) as each accumulator is next used
        own_store(PV, V, t);
        __syncthreads();
        own_store(PQ, bq, t);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int r = (int)((opaque_v(rot[k >> 1]) >> (16 * (k & 1))) & 0xffffu);
            V[k] = PV[r];
        }
        __syncthreads();  // every gather from PV is done before the transform reuses it
        // tp = NTT_p(lift(iNTT_q(Y)))
        uint32_t v[8], S[8], A[8], WQR[8], KQ[8], WP[8], KP[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = V[k];
        // the seat's rows are loaded after the transform, the table row (streamed from HBM)
        // during it: rows loaded early stay live across the transform, and at the 128-register
        // budget of four workgroups per CU each one adds spills
        round_trip<LOGN, true, WQ>(v, X, c, twi_q, twf_p, t, [&] {});
        {
            const uint32_t *sp = seat + opaque_s(0);
            gload8(sp + 2 * N, WP);
            gload8(sp + 3 * N, KP);
        }
        uint32_t bp[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int r = (int)((opaque_v(rot[k >> 1]) >> (16 * (k & 1))) & 0xffffu);
            bp[k] = red(PP[r] + red(mont(v[k], KP[k], p, c.pninv), p), p);
            v[k] = red(mont(v[k], WP[k], p, c.pninv), p);
        }
        // r = NTT_q(lift(iNTT_p(tp * w_p)))
        round_trip<LOGN, false, WQ>(v, X, c, twi_p, twf_q, t, [&] {
            own_store(PP, bp, t);  // every thread's gather from PP is behind the transform's barriers
            const int z2 = opaque_s(0);
            gload8(job.step1 + (size_t)i * N + 8 * t + z2, S);
        });
        {
            const int z2 = opaque_s(0);
            gload8(apark + z2, A);
            gload8(seat + z2, WQR);
            gload8(seat + N + z2, KQ);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t t1 = mont(A[k], S[k], q, c.qninv);
            const uint32_t t2 = mont(V[k], WQR[k], q, c.qninv);
            const int r = (int)((opaque_v(rot[k >> 1]) >> (16 * (k & 1))) & 0xffffu);
            if constexpr (WQ) {
                // every sum of two residues below q: canonical operands, sums below 2q < 2^32
                const uint32_t t3 = shoup(red(t2, q) + q - red(v[k], q), c.pinv_n, q);  // (0, 2q)
                bq[k] = red(PQ[r], q) + red(mont(V[k], KQ[k], q, c.qninv), q);       // [0, 2q)
                V[k] = red(red(t1, q) + red(t3, q), q);
            } else {
                const uint32_t t3 = shoup(t2 + 4 * q - v[k], c.pinv_n, q);  // (0, 6q): 6q < 2^32
                bq[k] = red(PQ[r] + mont(V[k], KQ[k], q, c.qninv), 2 * q);
                V[k] = red(t1 + t3, 2 * q);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) BQ[k] = WQ ? red(bq[k], q) : bq[k];  // the WQ inverse takes canonical words
    own_load(PP, BP, t);  // stored by this thread in the last step
    __syncthreads();

    // --- body fold (server.rs:398-410): fold[k] = sum_m c_m * nibble(m, k), c_0 = centred(b_0),
    // c_m = -centred(b_{n-m}); i64 sums, reduced mod q
    constexpr int FK = (N / 2 + T - 1) / T;  // body coefficients per thread, nb <= gamma <= n/2
    uint32_t F[FK];
    {
        const uint32_t *body = bt.bodies + (size_t)qi * N;
        int32_t *cm = reinterpret_cast<int32_t *>(X);
        for (int m = t; m < N; m += T) {
            const uint32_t v = body[m == 0 ? 0 : N - m];
            const int32_t cen = v > c.half_q ? (int32_t)v - (int32_t)q : (int32_t)v;
            cm[m] = m == 0 ? cen : -cen;
        }
        __syncthreads();
        const int W = c.cell_bytes / 4, R = T / W;  // words per cell, rows in flight
        const int w = t % W, r0 = t / W;
        int64_t s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(job.cells);
        for (int m = r0; m < N; m += R) {
            const uint32_t word = cw[(size_t)m * W + w];
            const int32_t cv = cm[m];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int32_t nib = (int32_t)(word << (28 - 4 * b)) >> 28;
                s[b] += (int64_t)cv * nib;  // |cv| <= q/2 < 2^30: the product needs 64 bits
            }
        }
        int64_t *part = reinterpret_cast<int64_t *>(PQ);  // R x 8W = 8T i64 over PQ and PP
#pragma unroll
        for (int b = 0; b < 8; b++) part[r0 * 8 * W + 8 * w + b] = s[b];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < FK; j++) {
            const int k = t + j * T;
            F[j] = 0;
            if (k < c.nb) {
                int64_t f = 0;
                for (int r = 0; r < R; r++) f += part[r * 8 * W + k];
                const int64_t m = f % (int64_t)q;
                F[j] = (uint32_t)(m < 0 ? m + q : m);
            }
        }
        __syncthreads();
    }

    // --- mask = iNTT_q(V_0); body = mod_down(iNTT_q(B_q), iNTT_p(B_p)) + fold
    inverse_to_lds<LOGN, WQ>(V, X, c.twi_q, q, t);
    inverse_to_lds<LOGN, WQ>(BQ, PQ, c.twi_q, q, t);
    inverse_to_lds<LOGN, true>(BP, PP, c.twi_p, p, t);
    if (!bt.wire_out) {
        uint32_t *mask = bt.mask_out + (size_t)qi * N;
        for (int j = t; j < N; j += T) mask[j] = red(X[swz(j)], q);
        uint32_t *bout = bt.body_out + (size_t)qi * c.nb;
#pragma unroll
        for (int j = 0; j < FK; j++) {
            const int k = t + j * T;
            if (k < c.nb) {
                const uint32_t xq = red(PQ[swz(k)], q);
                const uint32_t xp = PP[swz(k)];
                // (0, 4q); (0, 2q) under WQ
                const uint32_t d = WQ ? xq + q - lift_pq_wide(xp, c) : xq + 3 * q - lift_pq(xp, c);
                bout[k] = red(red(shoup(d, c.pinv, q), q) + F[j], q);
            }
        }
        return;
    }
    // the response bytes: the final mask in place in X, the body in PQ (each thread its own
    // coefficients), then one 4-byte word per lane and pass
    for (int j = t; j < N; j += T) X[swz(j)] = red(X[swz(j)], q);
#pragma unroll
    for (int j = 0; j < FK; j++) {
        const int k = t + j * T;
        if (k < c.nb) {
            const uint32_t xq = red(PQ[swz(k)], q);
            const uint32_t xp = PP[swz(k)];
            const uint32_t d = WQ ? xq + q - lift_pq_wide(xp, c) : xq + 3 * q - lift_pq(xp, c);
            PQ[swz(k)] = red(red(shoup(d, c.pinv, q), q) + F[j], q);
        }
    }
    __syncthreads();
    const SmWire WL = sm_wire(N, c.nb, q);
    // byte p of `count` coefficients packed LSB-first at WL.w bits (w >= 8: a byte spans at
    // most two coefficients); zero past the last one
    auto tight = [&](const uint32_t *V, uint32_t count, uint32_t p) -> uint32_t {
        const uint32_t bit = 8 * p, i = bit / WL.w, sh = bit - i * WL.w;
        if (i >= count) return 0u;
        uint64_t v = V[swz((int)i)] >> sh;
        if (WL.w - sh < 8 && i + 1 < count) v |= (uint64_t)V[swz((int)i + 1)] << (WL.w - sh);
        return (uint32_t)(v & 0xff);
    };
    uint32_t *out = reinterpret_cast<uint32_t *>(bt.wire_out + (size_t)qi * bt.wire_stride);
    const uint32_t words = (WL.total + 3) / 4;
    for (uint32_t wi = t; wi < words; wi += T) {
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t o = 4 * wi + k;
            const uint32_t byte = o < WL.mask_bytes ? tight(X, N, o)
                                  : o < WL.total     ? tight(PQ, (uint32_t)c.nb, o - WL.mask_bytes)
                                                    : 0u;
            word |= byte << (8 * k);
        }
        out[wi] = word;
    }
}

// ---------------------------------------------------------------- table

// cols[j * n + m] = signed nibble j of cell m (mod.rs:366-371)
__global__ void k_sm_columns(const uint8_t *cells, int n, int cell_bytes, int8_t *cols) {
    const int m = blockIdx.x;
    for (int b = threadIdx.x; b < cell_bytes; b += blockDim.x) {
        const uint8_t byte = cells[(size_t)m * cell_bytes + b];
        cols[(size_t)(2 * b) * n + m] = (int8_t)((int8_t)(uint8_t)(byte << 4) >> 4);
        cols[(size_t)(2 * b + 1) * n + m] = (int8_t)((int8_t)byte >> 4);
    }
}

// S'_i = (n gamma)^-1 NTT_q(P_i), P_i = sum_j X^(j * g^(n-i)) h_j (step1_row_impl,
// server.rs:330-383, whose tau_{g^i} gather this frame drops). One row i per workgroup.
template <int LOGN, bool WQ>
__global__ __launch_bounds__((1 << LOGN) / 8) void k_sm_table(SmCtxDev c, const int8_t *cols, uint32_t *step1) {
    constexpr int N = 1 << LOGN, T = N / 8;
    __shared__ __align__(16) uint32_t X[Lds<LOGN>::words];
    const int i = blockIdx.x, t = threadIdx.x;
    const uint32_t two_n = 2u * N;
    uint32_t e = 1, base = c.g % two_n;
    for (uint32_t k = (uint32_t)((N - i) % N); k; k >>= 1) {
        if (k & 1) e = e * base % two_n;
        base = base * base % two_n;
    }
    int32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < c.gamma; j++) {
        const uint32_t kk = (uint32_t)j * e % two_n;
        const bool flip = kk >= (uint32_t)N;
        const int sh = (int)(flip ? kk - N : kk);
        const int8_t *col = cols + (size_t)j * N;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int cpos = t + k * T;
            int m = cpos - sh;
            bool neg = flip;
            if (m < 0) { m += N; neg = !neg; }
            const int32_t v = col[m];
            acc[k] += neg ? -v : v;
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) X[swz(t + k * T)] = acc[k] < 0 ? (uint32_t)(acc[k] + (int32_t)c.q) : (uint32_t)acc[k];
    __syncthreads();
    uint32_t v[8];
    forward_from_lds<LOGN, WQ>(X, v, c.twf_q, c.q, t);
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = red(shoup(v[k], c.table_scale, c.q), c.q);
    gstore8(step1 + (size_t)i * N + 8 * t, v);
}

// ---------------------------------------------------------------- seat

// Seat of one key (server.rs:177-203): the key mask from ChaCha20(seed) on stream
// "key mask", q limb then p limb, each draw rejected in the short last interval
// (mod.rs:216-237, 313-330); both key limbs; all in transform form as Montgomery operands.
template <int LOGN, bool WQ>
__global__ __launch_bounds__((1 << LOGN) / 8) void k_sm_seat(SmCtxDev c, const uint8_t *mask_seed,
                                                            const uint32_t *key_q, const uint32_t *key_p,
                                                            uint32_t *seat) {
    constexpr int N = 1 << LOGN, T = N / 8;
    constexpr uint64_t kStream = 0x6b6579206d61736bULL;  // big-endian "key mask"
    constexpr int NW = Lds<LOGN>::words;
    __shared__ __align__(16) uint32_t L[4 * NW];
    __shared__ int rejected;
    uint32_t *XQ = L, *XP = L + NW, *KQ = L + 2 * NW, *KP = L + 3 * NW;
    const int t = threadIdx.x;
    const uint32_t q = c.q, p = c.p;
    uint32_t key[8], w[16];
    seed_key(mask_seed, key);
    if (t == 0) rejected = 0;
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int blk = t + h * T;  // draws 8 blk .. 8 blk + 7
        chacha20_block(key, (uint64_t)blk, kStream, w);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint64_t u = (uint64_t)w[2 * i] | (uint64_t)w[2 * i + 1] << 32;
            const int dd = 8 * blk + i;
            if (dd < N) {
                if (u >= c.zone_q) rejected = 1;
                XQ[swz(dd)] = reduce64(u, q, c.q_recip);
            } else {
                if (u >= c.zone_p) rejected = 1;
                XP[swz(dd - N)] = reduce64(u, p, c.p_recip);
            }
        }
    }
    for (int j = t; j < N; j += T) {
        KQ[swz(j)] = key_q[j];
        KP[swz(j)] = key_p[j];
    }
    __syncthreads();
    if (rejected) {
        // a draw in the short last interval of either limb: replay both limbs serially
        if (t == 0) {
            const uint32_t m[2] = {q, p};
            const uint64_t zone[2] = {c.zone_q, c.zone_p}, recip[2] = {c.q_recip, c.p_recip};
            replay_expand_uniform(key, kStream, 2, N, m, zone, recip,
                                  [&](int l, int k, uint32_t v) { (l ? XP : XQ)[swz(k)] = v; });
        }
        __syncthreads();
    }
    uint32_t v[8];
    forward_from_lds<LOGN, WQ>(XQ, v, c.twf_q, q, t);  // w_q
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = red(shoup(v[k], c.seat_wq, q), q);
    gstore8(seat + 8 * t, v);
    forward_from_lds<LOGN, WQ>(KQ, v, c.twf_q, q, t);  // key_q
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = red(shoup(v[k], c.seat_kq, q), q);
    gstore8(seat + N + 8 * t, v);
    forward_from_lds<LOGN, true>(XP, v, c.twf_p, p, t);  // w_p
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = red(shoup(v[k], c.seat_p, p), p);
    gstore8(seat + 2 * N + 8 * t, v);
    forward_from_lds<LOGN, true>(KP, v, c.twf_p, p, t);  // key_p
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = red(shoup(v[k], c.seat_p, p), p);
    gstore8(seat + 3 * N + 8 * t, v);
}

template <int LOGN>
hipError_t table_t(const SmCtxDev &c, const uint8_t *cells, int8_t *cols, uint32_t *step1, hipStream_t s) {
    hipLaunchKernelGGL(k_sm_columns, dim3(c.n), dim3(256), 0, s, cells, c.n, c.cell_bytes, cols);
    if (c.wide_q)
        hipLaunchKernelGGL((k_sm_table<LOGN, true>), dim3(c.gamma), dim3((1 << LOGN) / 8), 0, s, c, cols, step1);
    else
        hipLaunchKernelGGL((k_sm_table<LOGN, false>), dim3(c.gamma), dim3((1 << LOGN) / 8), 0, s, c, cols, step1);
    return hipGetLastError();
}
template <int LOGN>
hipError_t seat_t(const SmCtxDev &c, const uint8_t *seed, const uint32_t *kq, const uint32_t *kp, uint32_t *seat,
                  hipStream_t s) {
    if (c.wide_q)
        hipLaunchKernelGGL((k_sm_seat<LOGN, true>), dim3(1), dim3((1 << LOGN) / 8), 0, s, c, seed, kq, kp, seat);
    else
        hipLaunchKernelGGL((k_sm_seat<LOGN, false>), dim3(1), dim3((1 << LOGN) / 8), 0, s, c, seed, kq, kp, seat);
    return hipGetLastError();
}
template <int LOGN>
hipError_t respond_t(const SmCtxDev &c, const SmBatchDev &b, hipStream_t s) {
    if (c.wide_q)
        hipLaunchKernelGGL((k_sm_respond<LOGN, true>), dim3(b.count), dim3((1 << LOGN) / 8), 0, s, c, b);
    else
        hipLaunchKernelGGL((k_sm_respond<LOGN, false>), dim3(b.count), dim3((1 << LOGN) / 8), 0, s, c, b);
    return hipGetLastError();
}

}  // namespace

#define SM_DISPATCH(fn, ...)                       \
    switch (c.logn) {                              \
        case 4: return fn<4>(__VA_ARGS__);         \
        case 5: return fn<5>(__VA_ARGS__);         \
        case 6: return fn<6>(__VA_ARGS__);         \
        case 7: return fn<7>(__VA_ARGS__);         \
        case 8: return fn<8>(__VA_ARGS__);         \
        case 9: return fn<9>(__VA_ARGS__);         \
        case 10: return fn<10>(__VA_ARGS__);       \
        case 11: return fn<11>(__VA_ARGS__);       \
        case 12: return fn<12>(__VA_ARGS__);       \
        case 13: return fn<13>(__VA_ARGS__);       \
        default: return hipErrorInvalidValue;      \
    }

hipError_t sm_launch_table(const SmCtxDev &c, const uint8_t *cells, int8_t *cols, uint32_t *step1, hipStream_t s) {
    SM_DISPATCH(table_t, c, cells, cols, step1, s)
}
hipError_t sm_launch_seat(const SmCtxDev &c, const uint8_t *mask_seed, const uint32_t *key_q, const uint32_t *key_p,
                          uint32_t *seat, hipStream_t s) {
    SM_DISPATCH(seat_t, c, mask_seed, key_q, key_p, seat, s)
}
hipError_t sm_launch_respond(const SmCtxDev &c, const SmBatchDev &b, hipStream_t s) {
    if (b.count == 0) return hipSuccess;
    SM_DISPATCH(respond_t, c, b, s)
}

}  // namespace inspire
