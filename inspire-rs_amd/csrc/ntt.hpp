// Negacyclic NTT over Z_q[X]/(X^n + 1) for one ring element held in LDS,
// computed cooperatively by a 256-thread workgroup.
//
// Same transform as the reference (src/math/ntt.rs:578-640): Cooley-Tukey forward
// with the bit-reversed psi ladder psi_pow[m] = psi^brv(m), Gentleman-Sande inverse
// with the psi^-1 ladder; slot j holds f(psi^(2*brv(j)+1)). Values are standard-form
// residues (the reference carries Montgomery form; outputs agree once lifted).
//
// Three radix-2 stages are fused per LDS round trip (radix-8 groups held in
// registers), so a 2048-point transform is 4 rounds: stages on bits 10-8, 7-5,
// 4-2, 1-0 forward and the mirror order inverse. The inverse leaves out the n^-1
// factor: callers fold it into the operands that feed the transform.
#pragma once
#include "modq.hpp"

namespace inspire {

constexpr int kThreads = 256;

// LDS index of logical element e: one u64 of padding per 32 elements keeps the
// strided radix-8 group loads off a single bank.
__device__ __forceinline__ int lds_ix(int e) { return e + (e >> 5); }
template <int LOGN>
struct LdsSize { static constexpr int value = (1 << LOGN) + ((1 << LOGN) >> 5); };

// One forward pass over bits BT, BT-1, ..., BT-NB+1.
template <int LOGN, int BT, int NB>
__device__ __forceinline__ void fwd_pass(uint64_t *x, const uint64_t *__restrict__ psi) {
    constexpr int N = 1 << LOGN;
    constexpr int L = BT - NB + 1;
    constexpr int G = N >> NB;
    constexpr int M0 = N >> (BT + 1);
    for (int g = threadIdx.x; g < G; g += kThreads) {
        int lo = g & ((1 << L) - 1), hi = g >> L;
        int base = (hi << (BT + 1)) | lo;
        uint64_t v[1 << NB];
#pragma unroll
        for (int k = 0; k < (1 << NB); k++) v[k] = x[lds_ix(base + (k << L))];
#pragma unroll
        for (int s = 0; s < NB; s++) {
            const int half = 1 << (NB - 1 - s);
#pragma unroll
            for (int k = 0; k < (1 << NB); k++) {
                if (k & half) continue;
                uint64_t w = psi[(M0 << s) + (hi << s) + (k >> (NB - s))];
                uint64_t u = v[k], t = mulmod(v[k + half], w);
                v[k] = addmod(u, t);
                v[k + half] = submod(u, t);
            }
        }
#pragma unroll
        for (int k = 0; k < (1 << NB); k++) x[lds_ix(base + (k << L))] = v[k];
    }
}

// One inverse pass over bits B, B+1, ..., B+NB-1.
template <int LOGN, int B, int NB>
__device__ __forceinline__ void inv_pass(uint64_t *x, const uint64_t *__restrict__ psi_inv) {
    constexpr int N = 1 << LOGN;
    constexpr int G = N >> NB;
    for (int g = threadIdx.x; g < G; g += kThreads) {
        int lo = g & ((1 << B) - 1), hi = g >> B;
        int base = (hi << (B + NB)) | lo;
        uint64_t v[1 << NB];
#pragma unroll
        for (int k = 0; k < (1 << NB); k++) v[k] = x[lds_ix(base + (k << B))];
#pragma unroll
        for (int s = 0; s < NB; s++) {
            const int d = 1 << s;
            const int m = N >> (B + s + 1);
#pragma unroll
            for (int k = 0; k < (1 << NB); k++) {
                if (k & d) continue;
                uint64_t w = psi_inv[m + (hi << (NB - 1 - s)) + (k >> (s + 1))];
                uint64_t u = v[k], t = v[k + d];
                v[k] = addmod(u, t);
                v[k + d] = mulmod(submod(u, t), w);
            }
        }
#pragma unroll
        for (int k = 0; k < (1 << NB); k++) x[lds_ix(base + (k << B))] = v[k];
    }
}

template <int LOGN, int BT>
__device__ __forceinline__ void fwd_from(uint64_t *x, const uint64_t *__restrict__ psi) {
    if constexpr (BT >= 0) {
        constexpr int NB = BT + 1 >= 3 ? 3 : BT + 1;
        fwd_pass<LOGN, BT, NB>(x, psi);
        __syncthreads();
        fwd_from<LOGN, BT - NB>(x, psi);
    }
}

template <int LOGN, int B>
__device__ __forceinline__ void inv_from(uint64_t *x, const uint64_t *__restrict__ psi_inv) {
    if constexpr (B < LOGN) {
        constexpr int NB = LOGN - B >= 3 ? 3 : LOGN - B;
        inv_pass<LOGN, B, NB>(x, psi_inv);
        __syncthreads();
        inv_from<LOGN, B + NB>(x, psi_inv);
    }
}

// Forward NTT of the LDS array x (standard-form in, standard-form out).
// Caller must __syncthreads() before (x fully written); ends with a barrier.
template <int LOGN>
__device__ __forceinline__ void ntt_forward(uint64_t *x, const uint64_t *__restrict__ psi) {
    fwd_from<LOGN, LOGN - 1>(x, psi);
}

// Inverse NTT without the n^-1 factor. Same barrier contract as ntt_forward.
template <int LOGN>
__device__ __forceinline__ void ntt_inverse_noscale(uint64_t *x, const uint64_t *__restrict__ psi_inv) {
    inv_from<LOGN, 0>(x, psi_inv);
}

}  // namespace inspire
